#!/usr/bin/env python3
"""Convert the reference's Geant4-DNA cross-section tables into a binary table file.

Source: public/cross_sections.wgsl (generated upstream by tools/convert_g4data.py from
G4EMLOW 8.8; Geant4 Software License, cite the Geant4-DNA collaboration).  The WGSL
file is the data contract of the transport path: every ``const NAME=array<f32,N>(...)``
and every scalar ``const NAME=value`` is carried over unchanged.

Literal semantics follow WGSL: a decimal literal is an abstract float (f64) that is
converted to f32 by round-to-nearest, i.e. decimal -> nearest f64 -> nearest f32.

Output (written next to this script):
  cross_sections.bin   little-endian blocks, fixed order (TABLE_ORDER below), no header
  cross_sections.json  manifest: name -> {offset (bytes), count}, scalar constants,
                       sha256 of the .bin, and the source file's sha256

This script runs only where the reference checkout exists; the committed .bin/.json are
what the library and tests load.
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

SRC_DEFAULT = "/root/reference/public/cross_sections.wgsl"

# Block order of the binary file.  Sizes are checked against the WGSL declarations.
TABLE_ORDER = [
    ("XE", 100), ("XI", 100), ("XC", 100), ("XL", 100),
    ("XSF0", 100), ("XSF1", 100), ("XSF2", 100), ("XSF3", 100), ("XSF4", 100),
    ("XEF0", 100), ("XEF1", 100), ("XEF2", 100), ("XEF3", 100), ("XEF4", 100),
    ("XAE", 25), ("XAC", 625),
    ("XWE", 58),
    ("XWC0", 14500), ("XWC1", 14500), ("XWC2", 14500), ("XWC3", 14500), ("XWC4", 14500),
    ("XWT0", 14500), ("XWT1", 14500), ("XWT2", 14500), ("XWT3", 14500), ("XWT4", 14500),
    ("XVE", 38), ("XVS", 38), ("XVMF", 342), ("VIB_LEV", 9),
]
FLOAT_SCALARS = ["LOG_XE0", "INV_LOG_XE_STEP", "LOG_XAE0", "INV_LOG_XAE_STEP",
                 "LOG_XWE0", "INV_LOG_XWE_STEP"]
UINT_SCALARS = ["XN", "N_AE", "N_AC", "N_WE", "N_WR", "N_VIB"]


def wgsl_f32(tok: str) -> np.float32:
    return np.float32(np.float64(tok.strip()))


def parse(text: str):
    arrays = {}
    for m in re.finditer(r"const ([A-Z_0-9]+)=array<f32,(\d+)>\(([^)]*)\)", text):
        name, n, body = m.group(1), int(m.group(2)), m.group(3)
        vals = np.array([wgsl_f32(t) for t in body.split(",")], dtype=np.float32)
        if vals.size != n:
            raise ValueError(f"{name}: declared {n}, found {vals.size}")
        arrays[name] = vals
    fscal, uscal = {}, {}
    for name in FLOAT_SCALARS:
        m = re.search(rf"const {name}=([-+0-9.eE]+);", text)
        fscal[name] = float(wgsl_f32(m.group(1)))
    for name in UINT_SCALARS:
        m = re.search(rf"const {name}=(\d+)u;", text)
        uscal[name] = int(m.group(1))
    return arrays, fscal, uscal


def main(src: str = SRC_DEFAULT, outdir: str = None) -> None:
    outdir = outdir or os.path.dirname(os.path.abspath(__file__))
    raw = open(src, "rb").read()
    arrays, fscal, uscal = parse(raw.decode("utf-8"))
    blob = bytearray()
    manifest = {"tables": {}, "f32_scalars": fscal, "u32_scalars": uscal,
                "source": "public/cross_sections.wgsl",
                "source_sha256": hashlib.sha256(raw).hexdigest(),
                "license": "G4EMLOW-derived data, Geant4 Software License"}
    for name, n in TABLE_ORDER:
        a = arrays[name]
        if a.size != n:
            raise ValueError(f"{name}: expected {n}, got {a.size}")
        manifest["tables"][name] = {"offset": len(blob), "count": n}
        blob += a.astype("<f4").tobytes()
    # scalar block appended after the tables, in FLOAT_SCALARS order
    manifest["tables"]["SCALARS"] = {"offset": len(blob), "count": len(FLOAT_SCALARS)}
    blob += np.array([fscal[k] for k in FLOAT_SCALARS], dtype="<f4").tobytes()
    manifest["bin_sha256"] = hashlib.sha256(bytes(blob)).hexdigest()
    manifest["bin_bytes"] = len(blob)
    with open(os.path.join(outdir, "cross_sections.bin"), "wb") as f:
        f.write(bytes(blob))
    with open(os.path.join(outdir, "cross_sections.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(f"wrote {len(blob)} bytes, sha256 {manifest['bin_sha256'][:16]}")


if __name__ == "__main__":
    main(*sys.argv[1:])
