#!/usr/bin/env python3
"""Extract the reference's own fixtures and published validation numbers into
tests/golden/reference_fixtures.json.

Reads (in the reference checkout, as data, nothing is executed):
  validation/webgpu-results.json            reference harness numbers at 10 keV
  tests/fixtures/geant4-reference.json      Geant4 11.3 dnaphysics/chem6 reference values
  validation/g4_per_event.csv               Geant4 per-event ions/exc/path/edep, 4096 x 10 keV
  validation/g4_mfp.csv                     Geant4 per-process mean free paths
  experiments/results/2026-05-07/level-1/E1..E4 json   table peak values
  experiments/results/2026-05-07/level-2/E5 json       CSDA statistics
  experiments/results/2026-05-07/level-4/E10 json      IRT G(1 us) per energy
  tests/unit/cross-sections.test.ts:27-33               Geant4 sigma reference bands (numbers)
  experiments/results/2026-05-08/level-0/B1-harness-liveness.json  the reference harness's
      100 eV x 1024 results-table row (src/ui/table.ts:49-71 cell order) and its log line
The output is committed; tests never read the reference checkout.
"""
import csv
import json
import math
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")


def j(path):
    with open(os.path.join(REF, path)) as f:
        return json.load(f)


def main():
    fx = {}
    fx["webgpu_results"] = j("validation/webgpu-results.json")
    fx["geant4_reference"] = j("tests/fixtures/geant4-reference.json")
    rows = list(csv.DictReader(open(os.path.join(REF, "validation/g4_per_event.csv"))))
    ev = {}
    for col in ("ions", "exc", "path_nm", "edep_eV"):
        v = [float(r[col]) for r in rows]
        m = sum(v) / len(v)
        sd = math.sqrt(sum((x - m) ** 2 for x in v) / (len(v) - 1))
        ev[col] = {"mean": m, "std": sd, "sem": sd / math.sqrt(len(v)), "n": len(v)}
    fx["g4_per_event_summary"] = ev
    fx["g4_mfp"] = [dict(r) for r in csv.DictReader(open(os.path.join(REF, "validation/g4_mfp.csv")))]
    peaks = {}
    for name in ("E1-ion-xs-match", "E2-exc-xs-match", "E3-elastic-xs-match", "E4-vib-xs-match"):
        s = j(f"experiments/results/2026-05-07/level-1/{name}.json")["summary"]
        peaks[name] = s["peakSigmaWgslNm2"]
    fx["table_peaks"] = peaks
    fx["E5_summary"] = j("experiments/results/2026-05-07/level-2/E5-csda-vs-g4-ntuple.json")["summary"]
    e6 = j("experiments/results/2026-05-07/level-2/E6-mfp-vs-g4-ntuple.json")["rows"]
    fx["E6_rows"] = [{k: r[k] for k in ("eLoEv", "eHiEv", "eMidEv", "sigmaXi", "sigmaXc", "sigmaXl",
                                         "mfpG4MeanNm")} for r in e6]
    fx["E10_per_energy"] = j("experiments/results/2026-05-07/level-4/E10-irt-vs-karamitros.json")[
        "summary"]["perEnergyHeadline"]
    ts = open(os.path.join(REF, "tests/unit/cross-sections.test.ts")).read()
    g4ref = {}
    for m in re.finditer(r"(\d+):\s*\{\s*ion:\s*([0-9.e-]+),\s*el:\s*([0-9.e-]+),\s*total:\s*([0-9.e-]+)", ts):
        g4ref[m.group(1)] = {"ion": float(m.group(2)), "el": float(m.group(3)), "total": float(m.group(4))}
    fx["xs_test_g4_ref"] = g4ref
    b1 = j("experiments/results/2026-05-08/level-0/B1-harness-liveness.json")
    row = b1["rows"][0]
    cells = row["firstRowCells"]
    names = ["E", "n_therm", "n_esc", "mean_total", "estar_csda", "csda_ratio", "mean_prod", "mean_sp",
             "estar_sp", "sp_ratio", "mean_ions", "sec_per_pri", "cons_pct", "G_OH", "G_eaq", "G_H",
             "ssb_dir", "ssb_ind", "dsb", "dsb_per_gy_gbp", "dt_ms"]
    assert len(cells) == len(names)
    log = row["logTail"] if isinstance(row["logTail"], str) else " ".join(row["logTail"])
    m = re.search(r"E=100 eV: .*?rad_n=(\d+) E_cons=([0-9.]+)%", log)
    fx["B1_100eV_row"] = {"np": 1024, "cells": dict(zip(names, cells)), "rad_n": int(m.group(1)),
                          "log_E_cons_pct": m.group(2),
                          "source": "experiments/results/2026-05-08/level-0/B1-harness-liveness.json"}
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
