"""Oracle renders of seeded workload documents in worker processes (CPU only).

The full-size configurations take the scalar oracle seconds to minutes per
document, so the tests that check them render the oracle side in a spawn pool.
Workers rebuild each document from its seed (nothing large is pickled on the way
in) and import only numpy, the ABI mirror and the oracle bindings.
"""
from __future__ import annotations

import multiprocessing as mp
import os


class KeepFilterGroup:
    """full_filter_doc's `keep`: filter group `fg` and every `stride`-th plain path,
    starting at plain item `phase`."""

    def __init__(self, fg: int, stride: int, phase: int):
        self.fg, self.stride, self.phase = fg, stride, phase
        self.nf = self.np = 0

    def __call__(self, item, is_filter):
        if is_filter:
            self.nf += 1
            return self.nf - 1 == self.fg
        self.np += 1
        return (self.np - 1) % self.stride == self.phase


def full_filter_part(args):
    """(size, n_paths, seed, fg, stride, phase) -> the document and its (w, h)."""
    size, n_paths, seed, fg, stride, phase = args
    from resvg_amd.workloads import full_filter_doc
    return full_filter_doc(seed=seed, size=size, n_paths=n_paths, keep=KeepFilterGroup(fg, stride, phase))


def _render_full_filter_part(args):
    import numpy as np
    from oracle import bindings as orc
    from resvg_amd import abi
    sc, (w, h) = full_filter_part(args)
    px = np.zeros((h, w, 4), np.uint8)
    orc.render(sc.to_abi().abi, abi.Transform.identity(), px)
    return px


def batch_docs(n_docs, seed, **kw):
    from resvg_amd.workloads import document_batch
    return document_batch(n_docs, seed=seed, **kw)


def _render_batch_doc(args):
    import numpy as np
    from oracle import bindings as orc
    from resvg_amd import abi
    n_docs, seed, kw, i = args
    sc, (w, h) = batch_docs(n_docs, seed, **kw)[i]
    px = np.zeros((h, w, 4), np.uint8)
    orc.render(sc.to_abi().abi, abi.Transform.identity(), px)
    return px


def _pool(n):
    workers = max(1, min(n, 16, len(os.sched_getaffinity(0))))
    return mp.get_context("spawn").Pool(workers)


def oracle_full_filter_parts(parts):
    """Oracle renders of full_filter_part(args) for every args in `parts`, in parallel."""
    with _pool(len(parts)) as p:
        return p.map(_render_full_filter_part, parts)


def oracle_batch_docs(n_docs, seed, indices, **kw):
    with _pool(len(indices)) as p:
        return p.map(_render_batch_doc, [(n_docs, seed, kw, i) for i in indices])
