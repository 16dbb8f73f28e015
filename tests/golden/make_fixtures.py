"""Copies the reference's own known-answer PNGs used by the fixture tests, and the
raster images the image fixtures draw.

The PNGs are resvg's test references (crates/resvg/tests/tests/{filters, masking,
paint-servers, painting, shapes, structure}/..., MIT or Apache-2.0, Copyright the
Resvg Authors).  Only PNG data files are copied — no
reference source.  Run in a checkout that has the reference tree mounted; the
copies are committed so the tests never read the reference at run time.
"""
import base64
import re
import shutil
import sys
from pathlib import Path

SRC = Path("/root/reference/crates/resvg/tests/tests/filters")
DST = Path(__file__).resolve().parent / "resvg_png"

NAMES = [
    "feGaussianBlur/simple-case", "feGaussianBlur/small-stdDeviation", "feGaussianBlur/stdDeviation=5-0",
    "feOffset/simple-case", "feOffset/negative-offset",
    "feTile/simple-case", "feTile/empty-region", "feTile/with-region", "feTile/with-subregion-1",
    "feTile/with-subregion-2", "feTile/with-subregion-3", "feTile/complex-transform",
    "feComposite/operator=arithmetic", "feComposite/operator=over", "feComposite/operator=in",
    "feComposite/operator=xor",
    "feBlend/mode=multiply", "feBlend/mode=screen", "feBlend/mode=hue", "feBlend/mode=color-burn",
    "feBlend/mode=normal", "feBlend/mode=darken", "feBlend/mode=lighten",
    "feFlood/seagreen",
    "feComponentTransfer/type=gamma-on-blue", "feComponentTransfer/type=table-on-blue",
    "feComponentTransfer/type=linear-on-blue", "feComponentTransfer/type=discrete-on-blue",
    "feTurbulence/type=fractalNoise", "feTurbulence/no-attributes", "feTurbulence/seed=20",
    "feTurbulence/stitchTiles=stitch", "feTurbulence/numOctaves=5", "feTurbulence/baseFrequency=0.05-0.01",
    "filter/subregion-and-primitiveUnits=objectBoundingBox-1", "filter/subregion-and-primitiveUnits=objectBoundingBox-2",
    "feImage/link-to-an-element", "feImage/link-to-g", "feImage/link-to-use", "feImage/link-to-an-element-with-opacity",
    "feImage/link-to-an-element-with-transform", "feImage/link-on-an-element-with-transform",
    "feImage/link-on-an-element-with-complex-transform", "feImage/link-to-an-element-outside-defs-1",
    "feImage/link-to-an-element-outside-defs-2", "feImage/with-x-y", "feImage/with-x-y-and-protruding-subregion-1",
    "feImage/with-x-y-and-protruding-subregion-2", "feImage/chained-feImage",
]
# fixtures outside filters/ (paths relative to crates/resvg/tests/tests)
OTHER = [
    "painting/fill-rule/evenodd", "painting/fill-rule/nonzero",
    "paint-servers/linearGradient/default-attributes", "paint-servers/linearGradient/gradientUnits=userSpaceOnUse",
    "paint-servers/linearGradient/spreadMethod=reflect", "paint-servers/linearGradient/spreadMethod=repeat",
    "paint-servers/linearGradient/many-stops", "paint-servers/linearGradient/gradientTransform",
    "paint-servers/radialGradient/default-attributes", "paint-servers/radialGradient/gradientUnits=userSpaceOnUse",
    "paint-servers/radialGradient/spreadMethod=reflect", "paint-servers/radialGradient/fr=0.5",
    "paint-servers/radialGradient/fr=0.2", "paint-servers/radialGradient/fr=0.7",
    "paint-servers/radialGradient/fx-resolving-2", "paint-servers/radialGradient/fy-resolving-2",
    "paint-servers/radialGradient/attributes-via-xlink-href",
    "paint-servers/radialGradient/attributes-via-xlink-href-only-required",
    "paint-servers/radialGradient/attributes-via-xlink-href-complex-order",
    "paint-servers/radialGradient/gradientUnits=userSpaceOnUse-with-percent",
    "paint-servers/radialGradient/focal-point-correction",
    "painting/opacity/50percent",
    "painting/mix-blend-mode/multiply", "painting/mix-blend-mode/screen", "painting/mix-blend-mode/hue",
    "painting/mix-blend-mode/darken", "painting/mix-blend-mode/color-burn", "painting/mix-blend-mode/normal",
    "shapes/circle/simple-case", "shapes/ellipse/simple-case", "shapes/rect/rounded-rect",
    "painting/stroke-linejoin/miter", "painting/stroke-linejoin/round", "painting/stroke-linejoin/bevel",
    "painting/stroke-linejoin/arcs", "painting/stroke-linejoin/miter-clip",
    "painting/stroke-linecap/butt", "painting/stroke-linecap/round", "painting/stroke-linecap/square",
    "painting/stroke-linecap/open-path-with-butt", "painting/stroke-linecap/open-path-with-round",
    "painting/stroke-linecap/open-path-with-square", "painting/stroke-linecap/zero-length-path-with-butt",
    "painting/stroke-linecap/zero-length-path-with-round", "painting/stroke-linecap/zero-length-path-with-square",
    "shapes/path/M-C", "shapes/path/M-Q", "shapes/path/M-C-S", "shapes/path/M-Q-T", "shapes/path/M-L-L-Z",
    "painting/stroke-miterlimit/default", "painting/stroke-miterlimit/valid-value",
    "painting/stroke/control-points-clamping-1",
    "masking/clipPath/simple-case", "masking/clipPath/clip-rule=evenodd", "masking/clipPath/multiple-children",
    "masking/clipPath/overlapped-shapes-with-evenodd", "masking/clipPath/clip-path-on-child",
    "masking/clipPath/clip-path-on-children", "masking/clipPath/nested-clip-path",
    "masking/clipPath/clipPathUnits=objectBoundingBox", "masking/clipPath/transform-on-clipPath",
    "masking/clipPath/clip-path-with-transform",
    "masking/mask/simple-case", "masking/mask/color-interpolation=linearRGB", "masking/mask/mask-type=alpha",
    "masking/mask/maskUnits=userSpaceOnUse-with-rect", "masking/mask/maskUnits=userSpaceOnUse-with-percent",
    "masking/mask/nested-objectBoundingBox", "masking/mask/mask-on-self", "masking/mask/mask-on-child",
    "masking/mask/on-group-with-transform",
    "painting/stroke-dasharray/0-n-with-butt-caps", "painting/stroke-dasharray/0-n-with-round-caps",
    "painting/stroke-dasharray/0-n-with-square-caps", "painting/stroke-dasharray/n-0",
    "painting/stroke-dasharray/negative-sum", "painting/stroke-dasharray/negative-values",
    "painting/stroke-dasharray/zero-sum", "painting/stroke-dasharray/none",
    "painting/stroke-dasharray/comma-ws-separator", "painting/stroke-dasharray/even-count",
    "painting/stroke-dasharray/odd-count", "painting/stroke-dasharray/ws-separator",
    "painting/stroke-dasharray/on-a-circle", "painting/stroke-dasharray/mm-units",
    "painting/stroke-dasharray/em-units", "painting/stroke-dasharray/percent-units",
    "painting/stroke-dasharray/multiple-subpaths",
    "painting/stroke-dashoffset/default", "painting/stroke-dashoffset/em-units",
    "painting/stroke-dashoffset/mm-units", "painting/stroke-dashoffset/negative-value",
    "painting/stroke-dashoffset/percent-units", "painting/stroke-dashoffset/px-units",
    "paint-servers/pattern/simple-case", "paint-servers/pattern/overflow=visible",
    "paint-servers/pattern/with-patternTransform", "paint-servers/pattern/with-x-and-y",
    "paint-servers/pattern/patternUnits=userSpaceOnUse-with-percent",
    "paint-servers/pattern/transform-and-patternTransform", "paint-servers/pattern/with-viewBox",
    "paint-servers/pattern/preserveAspectRatio", "paint-servers/pattern/tiny-pattern-upscaled",
    "paint-servers/pattern/patternUnits=objectBoundingBox", "paint-servers/pattern/patternContentUnits=objectBoundingBox",
    "paint-servers/pattern/patternContentUnits-with-viewBox", "paint-servers/pattern/nested-objectBoundingBox",
    "paint-servers/pattern/pattern-on-child",
    "structure/image/external-png", "structure/image/embedded-png", "structure/image/with-transform",
    "structure/image/raster-image-and-size-with-odd-numbers", "structure/image/image-with-float-size-scaling",
    "structure/image/preserveAspectRatio=xMidYMid-meet", "structure/image/preserveAspectRatio=xMinYMin-meet",
    "structure/image/preserveAspectRatio=xMaxYMax-meet", "structure/image/preserveAspectRatio=xMidYMid-slice",
    "structure/image/preserveAspectRatio=xMinYMin-slice", "structure/image/preserveAspectRatio=xMaxYMax-slice",
    "structure/image/preserveAspectRatio=none",
]
# raster images the image fixtures draw: the reference's resource PNGs, and the PNGs
# embedded as data URIs in fixture SVGs (the base64 payload decoded; data only)
IMAGES_DST = Path(__file__).resolve().parent / "images"
RESOURCE_PNGS = ["image.png", "image-63x61.png"]
EMBEDDED_PNGS = ["structure/image/embedded-png", "structure/image/with-transform",
                 "structure/image/image-with-float-size-scaling", "structure/image/preserveAspectRatio=none"]


def main() -> int:
    DST.mkdir(exist_ok=True)
    for n in NAMES:
        src = SRC / f"{n}.png"
        if not src.exists():
            print("missing", src, file=sys.stderr)
            return 1
        shutil.copyfile(src, DST / (n.replace("/", "__") + ".png"))
    for n in OTHER:
        src = SRC.parent / f"{n}.png"
        if not src.exists():
            print("missing", src, file=sys.stderr)
            return 1
        shutil.copyfile(src, DST / (n.replace("/", "__") + ".png"))
    IMAGES_DST.mkdir(exist_ok=True)
    for n in RESOURCE_PNGS:
        shutil.copyfile(SRC.parents[1] / "resources" / n, IMAGES_DST / n)
    for n in EMBEDDED_PNGS:
        text = (SRC.parent / f"{n}.svg").read_text()
        payloads = re.findall(r"data:image/png;base64,([A-Za-z0-9+/=\s]+)", text)
        if not payloads:
            print("no embedded PNG in", n, file=sys.stderr)
            return 1
        data = base64.b64decode("".join(payloads[0].split()))
        (IMAGES_DST / (n.split("/")[-1] + ".png")).write_bytes(data)
    note = ("PNG test references copied from linebender/resvg 0.48.1, from\n"
            "crates/resvg/tests/tests/{filters,masking,paint-servers,painting,shapes,structure}\n"
            "(file names: the fixture path with \"/\" replaced by \"__\"),\n"
            "Copyright 2018-2025 the Resvg Authors, dual-licensed MIT OR Apache-2.0.\n")
    (DST / "LICENSE-NOTE.txt").write_text(note)
    (IMAGES_DST / "LICENSE-NOTE.txt").write_text(
        "Raster images from linebender/resvg 0.48.1: crates/resvg/tests/resources/{image.png,image-63x61.png},\n"
        "and the PNGs embedded as data URIs in crates/resvg/tests/tests/structure/image/*.svg (named after\n"
        "their fixture), Copyright 2018-2025 the Resvg Authors, dual-licensed MIT OR Apache-2.0.\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
