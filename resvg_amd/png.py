"""PNG files around the device's scanlines (tiny-skia's Pixmap::encode_png, which
resvg's CLI writes its output with, crates/resvg/src/main.rs:104-117).

The device demultiplies and filters the rows (``rvk_png_filter_rows``,
``Context.encode_png``); the host deflates them (zlib, in parallel pieces) and adds
the chunks: an 8-bit RGBA ``IHDR``, one ``IDAT`` and ``IEND``, each with its CRC-32.
"""
from __future__ import annotations

import os
import struct
import zlib
from concurrent.futures import ThreadPoolExecutor

SIGNATURE = b"\x89PNG\r\n\x1a\n"
_CHUNK = 1 << 20  # bytes deflated per task
_WINDOW = 1 << 15  # deflate's window: each task starts from the preceding 32 KiB


def deflate(data: bytes, level: int = 6, workers: int | None = None) -> bytes:
    """A zlib stream of `data` deflated in 1 MiB pieces on host threads (zlib releases
    the GIL): each piece is a raw deflate stream primed with the preceding 32 KiB as
    its dictionary and ended by a sync flush (an empty stored block, byte-aligned), the
    last by the final block; the pieces follow a zlib header and end with the Adler-32
    of all of `data` -- one valid stream, decodable by any inflater."""
    if len(data) <= _CHUNK:
        return zlib.compress(data, level)
    view = memoryview(data)
    starts = list(range(0, len(data), _CHUNK))

    def piece(i: int) -> bytes:
        s = starts[i]
        kw = {"zdict": bytes(view[max(0, s - _WINDOW):s])} if s > 0 else {}
        c = zlib.compressobj(level, zlib.DEFLATED, -15, **kw)
        out = c.compress(view[s:s + _CHUNK])
        return out + c.flush(zlib.Z_FINISH if i == len(starts) - 1 else zlib.Z_SYNC_FLUSH)

    n = workers or max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(n) as ex:
        parts = list(ex.map(piece, range(len(starts))))
    return b"\x78\x9c" + b"".join(parts) + struct.pack(">I", zlib.adler32(data) & 0xFFFFFFFF)


def chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def container(width: int, height: int, idat: bytes) -> bytes:
    """A PNG file of 8-bit RGBA (colour type 6, no interlace) with a deflated IDAT stream."""
    return (SIGNATURE + chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", idat) + chunk(b"IEND", b""))


def from_rows(width: int, height: int, rows: bytes, level: int = 6) -> bytes:
    """A PNG file from the IDAT stream's uncompressed rows (a filter-type byte and
    4 * width bytes per row), deflated at zlib `level`."""
    if len(rows) != height * (4 * width + 1):
        raise ValueError("rows hold height * (4 * width + 1) bytes")
    return container(width, height, deflate(rows, level))
