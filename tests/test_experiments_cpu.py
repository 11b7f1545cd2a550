"""The validation-ledger emitters (dna_amd.experiments) on the CPU: E6 and E6b driven by a
stand-in context whose table lookups are the CPU restatement's (the GPU ones are
bit-identical, tests/test_xs_probe_gpu.py), checked against the reference's own E6/E6b
artifacts: the same bins, MFP ratios within 1% of the reference run's (its sigmas were
interpolated log-log in JavaScript, these with the transport's log-linear lookup) and every
row passing its bar, in the reference's artifact layout."""
import json
import os
import sys

import numpy as np

import oracle as O
from dna_amd import experiments as X

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_fixtures.json")))


class OracleLookups:
    def xs_probe(self, which, E, r=None):
        return O.xs(which, np.asarray(E, np.float32))


def test_artifact_layout_and_e6_against_reference_run():
    a = X.e6(OracleLookups(), REF)
    assert list(a) == ["meta", "env", "status", "diagnosis", "summary", "rows"]
    assert a["meta"]["protocol"] == "E6-mfp-vs-g4-ntuple" and a["status"] == "pass"
    ref_rows = {(r["eLoEv"], r["eHiEv"]): r for r in REF["E6_rows"]}
    assert len(a["rows"]) == len(ref_rows) == 6
    for r in a["rows"]:
        rr = ref_rows[(r["eLoEv"], r["eHiEv"])]
        assert r["mfpG4MeanNm"] == rr["mfpG4MeanNm"]
        mfp_ref = 1.0 / (X.N_WATER_PER_NM3 * (rr["sigmaXi"] + rr["sigmaXc"] + rr["sigmaXl"]))
        assert abs(r["mfpWgslNm"] / mfp_ref - 1) < 0.01
    json.dumps(a)


def test_e6b_bands():
    a = X.e6b(OracleLookups(), REF)
    assert a["status"] == "pass", [r for r in a["rows"] if r["status"] != "pass"]
    exc = [r["ratio"] for r in a["rows"] if r["process"] == "excitation"]
    assert exc and all(1.8 <= x <= 3.2 for x in exc)
    json.dumps(a)
