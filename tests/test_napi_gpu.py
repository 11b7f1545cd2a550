"""The reference's host API in JavaScript (webgpu-dna_amd/js/dna_host.js) on the GPU
through the Node-API addon: runValidation at 1 and 10 keV from Node gives the same
transport, chemistry and damage results as the CPU restatement on the same inputs."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
CK = [0.001, 0.01, 0.1, 1, 10, 100, 1000]


@pytest.mark.skipif(NODE is None, reason="node is missing")
def test_js_run_validation_matches_oracle():
    print("node run_validation.js --np 96 --energies 1000,10000", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "webgpu-dna_amd", "js", "run_validation.js"), "--np", "96",
                        "--energies", "1000,10000"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    rows = out["rows"]
    assert [row["E"] for row in rows] == [1000, 10000]
    dna = DNATarget()
    for row in rows:
        E = float(row["E"])
        o = O.transport(default_params(E, 96, dna=dna))
        m = o["results"]["esc"] == 0
        assert row["n_therm"] == int(m.sum())
        assert row["mean_total"] == pytest.approx(float(o["results"]["path"][m].astype(np.float64).mean()), rel=1e-12)
        assert row["mean_ions"] == pytest.approx(float(o["results"]["ni"][m].astype(np.float64).mean()), rel=1e-12)
        assert row["total_deposited_eV"] == o["dose_sum"] / 100.0
        assert row["rad_n_stored"] == o["rad_n"]
        assert row["rad_OH"] == int(o["counters"][0]) and row["kernel_dna_hits"] == int(o["counters"][4])
        if E == 10000.0:
            co = O.irt(o["rad"], int(m.sum()), 10000.0, CK, seed=1)
            assert [tl["G_OH"] for tl in row["chem_result"]["timeline"]] == [co.G[c][0] for c in range(7)]
            assert [tl["G_H2"] for tl in row["chem_result"]["timeline"]] == [co.G[c][4] for c in range(7)]
            so = O.score(dna, o["rad"], lcg_state=0x12345678)
            assert row["damage"]["ssb_dir"] == so["ssb_count"] and row["damage"]["dsb"] == so["dsb"]
    res = out["results"]
    assert res["nPrimaries"] == 96 and res["primaryEnergyEv"] == 10000
    assert set(res) >= {"tracking", "preChemistry", "chemistry1Us", "dnaDamage", "karamitros2011"}


@pytest.mark.skipif(NODE is None, reason="node is missing")
def test_js_snapshot_export_and_indirect_ssb_match_python(tmp_path):
    """generateSnapshotAt10keV from Node writes the same WGDNA4D bytes as the Python host,
    and runValidation with indirectSSB gives the same damage (direct, indirect, DSB)."""
    from dna_amd import host, splat
    out = str(tmp_path / "js.bin")
    print("node generateSnapshotAt10keV np=96 + runValidation indirectSSB np=256", flush=True)
    script = (f"const s=require({json.dumps(os.path.join(ROOT, 'webgpu-dna_amd', 'js', 'splat.js'))}); "
              f"const h=require({json.dumps(os.path.join(ROOT, 'webgpu-dna_amd', 'js', 'dna_host.js'))}); "
              f"(async()=>{{ const p=await s.exportSnapshotAt10keV({{np:96, boxNm:15000, ceEV:7.4, snapshot_n:6000}}, "
              f"{json.dumps(out)}); "
              f"const v=await h.runValidation({{np:256, boxNm:15000, ceEV:7.4, energies:[10000], indirectSSB:true}}); "
              f"console.log(JSON.stringify({{p, d:v.rows[0].damage, r:v.results.dnaDamage}})); }})()"
              f".catch(e=>{{console.error(e); process.exit(1);}});")
    r = subprocess.run([NODE, "-e", script], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    js = json.loads(r.stdout.strip().splitlines()[-1])
    py = splat.generate_snapshot_at_10kev({"np": 96, "boxNm": 15000.0, "ceEV": 7.4, "snapshot_n": 6000})
    assert open(out, "rb").read() == py["blob"]
    v = host.run_validation({"np": 256, "boxNm": 15000.0, "ceEV": 7.4, "energies": [10000], "indirectSSB": True})
    d = v["rows"][0]["damage"]
    assert (js["d"]["ssb_dir"], js["d"]["ssb_ind"], js["d"]["dsb"]) == (d["ssb_dir"], d["ssb_ind"], d["dsb"])
    assert js["r"]["SSB_ind"] == d["ssb_ind"] == v["results"]["dnaDamage"]["SSB_ind"]
