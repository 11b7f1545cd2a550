"""The host mirror (runValidation / runAtEnergy / runChemistryWorker / scoring) running
end to end on the GPU, checked against the CPU restatement on the same inputs."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params
from dna_amd import host

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_run_validation_two_energies():
    logs = []
    out = host.run_validation({"np": 128, "boxNm": 15000.0, "ceEV": 7.4, "energies": [1000, 10000],
                               "log": lambda m, c="": logs.append((m, c))})
    rows, res = out["rows"], out["results"]
    assert [r["E"] for r in rows] == [1000, 10000]
    dna = DNATarget()
    p = default_params(10000.0, 128, dna=dna)
    o = O.transport(p)
    r = rows[1]
    m = o["results"]["esc"] == 0
    assert r["n_therm"] == int(m.sum())
    assert r["mean_total"] == pytest.approx(float(o["results"]["path"][m].astype(np.float64).mean()), rel=1e-12)
    assert r["total_deposited_eV"] == o["dose_sum"] / 100.0
    assert r["rad_n_stored"] == o["rad_n"] and r["kernel_dna_hits"] == int(o["counters"][4])
    # chemistry + scoring at 10 keV, equal to the CPU restatement (seed 1, LCG from 0x12345678)
    co = O.irt(o["rad"], int(m.sum()), 10000.0, [0.001, 0.01, 0.1, 1, 10, 100, 1000], seed=1)
    assert [tl["G_OH"] for tl in r["chem_result"]["timeline"]] == [co.G[c][0] for c in range(7)]
    so = O.score(dna, o["rad"], lcg_state=0x12345678)
    assert r["damage"]["ssb_dir"] == so["ssb_count"] and r["damage"]["dsb"] == so["dsb"]
    assert set(res) >= {"tracking", "preChemistry", "chemistry1Us", "dnaDamage"}
    assert any("DAMAGE" in m for m, _ in logs)


def test_run_irt_cli(tmp_path):
    sys.path.insert(0, os.path.join(ROOT, "webgpu-dna_amd"))
    from dna_amd import Context
    ctx = Context(0)
    t = ctx.transport(default_params(1000.0, 64))
    dump = tmp_path / "rad.bin"
    t["rad"].astype("<f4").tofile(dump)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "webgpu-dna_amd", "run_irt.py"), str(dump), "64", "1000"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["type"] == "result" and len(line["timeline"]) == 7
    o = O.irt(t["rad"], 64, 1000.0, [0.001, 0.01, 0.1, 1, 10, 100, 1000], seed=1)
    assert line["timeline"][-1]["G_OH"] == o.G[6][0]
    assert [rx["count"] for rx in line["rxn_info"]] == list(o.rxn_counts)
    assert line["rxn_info"][3]["sigma"] == "0.1715" and line["rxn_info"][5]["rc"] == "-0.711"
    # the Node CLI (the reference's own runtime for run_irt.cjs) prints the same line
    import shutil
    node = shutil.which("node")
    if node:
        r2 = subprocess.run([node, os.path.join(ROOT, "webgpu-dna_amd", "js", "run_irt.js"), str(dump), "64", "1000"],
                            capture_output=True, text=True, timeout=120)
        assert r2.returncode == 0, r2.stderr
        js = json.loads(r2.stdout.strip().splitlines()[-1])
        for d in (line, js):
            d.pop("t_wall")
        assert js == line
