"""Write fact_tables.inc: the FACT camera's pixel positions in chid order and
the FACT mirror reflectivity, as numbers, from the FACT propagation app of
the reference (merlict_fact/apps/fact_propagation.cpp: fact_mirror_reflectivity
and the pixel_x / pixel_y arrays of main).  Run where the reference sources
are (python gen_fact_tables.py [path to fact_propagation.cpp]); the output is
committed, so the library builds without them.

The pixel coordinates are float32 millimetres in the app (std::array<float>),
widened to double and divided by 1e3 when the pixels are placed; they are
kept as float literals here so the library repeats that conversion."""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

DEFAULT = "/root/reference/merlict_fact/apps/fact_propagation.cpp"


def parse(path: str):
    src = Path(path).read_text()
    px = re.search(r"pixel_x\s*=\s*\{([^}]*)\}", src).group(1)
    py = re.search(r"pixel_y\s*=\s*\{([^}]*)\}", src).group(1)
    x = np.array([float(v) for v in px.split(",")], dtype=np.float32)
    y = np.array([float(v) for v in py.split(",")], dtype=np.float32)
    body = re.search(r"fact_mirror_reflectivity\(\)\s*\{(.*?)return refl", src, re.S).group(1)
    refl = [(float(a), float(b)) for a, b in re.findall(r"\{\s*([0-9.eE+-]+)\s*,\s*([0-9.eE+-]+)\s*\}", body)]
    return x, y, refl


def float_literal(v: np.float32) -> str:
    s = np.format_float_positional(v, unique=True, trim="0") if np.isfinite(v) else "0"
    if "." not in s:
        s += ".0"
    assert np.float32(float(s)) == v
    return s + "f"


def main() -> int:
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    x, y, refl = parse(path)
    assert len(x) == len(y) == 1440, (len(x), len(y))
    out = ["// FACT camera and mirror tables (generated by gen_fact_tables.py; do not edit).",
           "// Pixel centres of chid 0..1439 in millimetres (float, as the FACT",
           "// propagation app stores them), and the facet reflectivity (wavelength m,",
           "// reflectivity) of the app's fact_mirror_reflectivity().",
           "#pragma once", "", "namespace mlgpu {", "",
           f"constexpr int kFactNumPixels = {len(x)};"]
    for name, arr in (("kFactPixelXmm", x), ("kFactPixelYmm", y)):
        out.append(f"constexpr float {name}[{len(arr)}] = {{")
        vals = [float_literal(v) for v in arr]
        for i in range(0, len(vals), 8):
            out.append("    " + ", ".join(vals[i:i + 8]) + ",")
        out.append("};")
    out.append(f"constexpr int kFactMirrorReflectivityPoints = {len(refl)};")
    out.append(f"constexpr double kFactMirrorReflectivity[{len(refl)}][2] = {{")
    for a, b in refl:
        out.append(f"    {{{a!r}, {b!r}}},")
    out.append("};")
    out += ["", "}  // namespace mlgpu", ""]
    dst = Path(__file__).resolve().parent / "fact_tables.inc"
    dst.write_text("\n".join(out))
    print(f"wrote {dst}: {len(x)} pixels, {len(refl)} reflectivity points")
    return 0


if __name__ == "__main__":
    sys.exit(main())
