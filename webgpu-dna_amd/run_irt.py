#!/usr/bin/env python3
"""IRT chemistry on a saved rad_buf dump, on the GPU — the counterpart of the reference's
tools/run_irt.cjs (usage and output contract of run_irt.cjs:1-79).

  python run_irt.py <dump.bin> <n_therm> <E_eV> [seed]

dump.bin: raw little-endian float32 (x, y, z, pid*8+code) per radical, as POSTed to the
reference's dump server (tools/dump_server.cjs:38-54).  Prints one JSON line
{"type": "result", "timeline": [...], "n_reacted", "chem_n", "t_wall", "n_repaired",
 "rxn_info": [{label, count, sigma, rc}]} on stdout (sigma and rc as the worker's toFixed(4) and
toFixed(3) strings, irt-worker.js:690-697) and a human-readable summary on stderr.
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def write_dump(path, rad_buf):
    """Save a (n, 4) float32 radical buffer in the .bin dump format."""
    import numpy as np
    np.ascontiguousarray(rad_buf, "<f4").tofile(path)


def main(argv):
    if len(argv) < 3:
        print("usage: run_irt.py <dump_file> <n_therm> <E_eV> [seed]", file=sys.stderr)
        return 1
    import numpy as np
    from dna_amd.host import ensure_pipelines, run_chemistry_worker
    dump, n_therm = argv[1], int(argv[2])
    E = int(argv[3]) if len(argv) > 3 else 10000
    seed = int(argv[4]) if len(argv) > 4 else 1
    if not n_therm:
        print("usage: run_irt.py <dump_file> <n_therm> <E_eV> [seed]", file=sys.stderr)
        return 1
    rad = np.fromfile(dump, dtype="<f4").reshape(-1, 4)
    print(f"[run_irt] {dump}: {rad.shape[0]} radicals ({rad.nbytes / 1e6:.2f} MB)  n_therm={n_therm}  E={E} eV",
          file=sys.stderr)
    if ensure_pipelines(0) is None:
        print("[run_irt] no GPU device", file=sys.stderr)
        return 2
    res = run_chemistry_worker(rad, rad.shape[0], n_therm, E, seed=seed)
    out = {"type": "result", "timeline": res["timeline"], "n_reacted": res["n_reacted"],
           "chem_n": res["chem_n"], "t_wall": res["t_wall"], "n_repaired": 0, "rxn_info": res["rxn_info"]}
    sys.stdout.write(json.dumps(out) + "\n")
    print(f"[worker] {res['n_reacted']} reactions in {res['t_wall'] / 1000:.3f}s", file=sys.stderr)
    for cp in res["timeline"]:
        print(f"  {cp['label']:8s} G(OH)={cp['G_OH']:7.3f}  G(eaq)={cp['G_eaq']:7.3f}  G(H)={cp['G_H']:7.3f}  "
              f"G(H2O2)={cp['G_H2O2']:7.3f}  G(H2)={cp['G_H2']:7.3f}", file=sys.stderr)
    print("  reactions fired:", file=sys.stderr)
    for rx in res["rxn_info"]:
        print(f"    {rx['label']:28s} count={rx['count']:7d}  σ={rx['sigma']}nm  rc={rx['rc']}nm", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
