"""The Node-API addon and the JavaScript host mirror, without a GPU: the addon loads and
exports its functions, a missing device gives null (src/gpu/device.ts:12-21), argument
errors throw, and buildDNATarget in V8 (the reference's JS engine) gives the same f32
fiber arrays as the Python host bit for bit."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "webgpu-dna_amd", "lib", "dna_amd.node")
HOST = os.path.join(ROOT, "webgpu-dna_amd", "js", "dna_host.js")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON),
                                reason="node or the built addon is missing")


def _node(script, timeout=60):
    r = subprocess.run([NODE, "-e", script], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _no_gpu():
    try:
        import torch
        return torch.cuda.device_count() == 0
    except Exception:
        return True


def test_addon_exports():
    out = json.loads(_node(f"const a=require({json.dumps(ADDON)}); "
                           "console.log(JSON.stringify({k:Object.keys(a), v:a.version()}))"))
    assert set(out["k"]) == {"version", "lastError", "ctxCreate", "ctxDestroy", "runTransport", "runIrt",
                             "score", "hashCells", "setOption", "getOption", "irtOhFinal", "irtParticles",
                             "scoreOh", "irtReactions"}
    assert "gfx950" in out["v"]


@pytest.mark.skipif(not _no_gpu(), reason="a GPU is present")
def test_no_device_is_null_not_faked():
    out = json.loads(_node(f"const a=require({json.dumps(ADDON)}); const c=a.ctxCreate(0); "
                           "console.log(JSON.stringify({c:c, e:a.lastError()}))"))
    assert out["c"] is None and "device" in out["e"]
    r = subprocess.run([NODE, os.path.join(ROOT, "webgpu-dna_amd", "js", "run_validation.js"), "--np", "4",
                        "--energies", "1000"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "no GPU device" in r.stderr


def test_bad_context_throws():
    out = _node(f"const a=require({json.dumps(ADDON)}); "
                "try { a.runTransport({}, {n:1}); console.log('no throw'); } "
                "catch (e) { console.log(e.constructor.name + ':' + e.message); }")
    assert out.startswith("TypeError:") and "context" in out


def test_dna_target_v8_matches_python_bits(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "webgpu-dna_amd"))
    from dna_amd import DNATarget
    base = str(tmp_path)
    _node(f"const h=require({json.dumps(HOST)}); const fs=require('fs'); const t=h.buildDNATarget(); "
          f"for (const k of ['fy','fz','rbb0','rbb1']) fs.writeFileSync({json.dumps(base)}+'/'+k, Buffer.from(t[k].buffer)); "
          f"fs.writeFileSync({json.dumps(base)}+'/meta.json', JSON.stringify({{n_bp:t.n_bp, x0:t.x0, n_bp_per:t.n_bp_per}}));")
    t = DNATarget()
    meta = json.load(open(os.path.join(base, "meta.json")))
    assert meta == {"n_bp": t.n_bp, "x0": t.x0, "n_bp_per": t.n_bp_per}
    for k in ("fy", "fz", "rbb0", "rbb1"):
        js = np.fromfile(os.path.join(base, k), np.float32)
        assert np.array_equal(js.view(np.uint32), getattr(t, k).view(np.uint32)), k


def test_snapshot_blob_js_matches_python_bytes(tmp_path):
    """splat.js (V8) and dna_amd.splat read the same particle table into the same snapshots
    and pack the same WGDNA4D bytes; the JS parser reads them back."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "webgpu-dna_amd"))
    from dna_amd import DNATarget
    from dna_amd.splat import pack_snapshot_blob, snapshots_from_particles
    rng = np.random.default_rng(4)
    n = 3000
    parts = np.zeros((n, 6))
    parts[:, :3] = rng.normal(0, 300, (n, 3)).astype(np.float32)
    parts[:, 3] = rng.integers(0, 500, n) * 8 + rng.choice([0, 1, 2, 3, 4, 6], n)
    born = rng.random(n) < 0.3
    parts[:, 4] = np.where(born, 10 ** rng.uniform(-3, 3, n), 0.0)
    die = rng.random(n) < 0.6
    parts[:, 5] = np.where(die, parts[:, 4] + 10 ** rng.uniform(-3, 3, n), np.inf)
    parts[:7, 4] = [0.0, 0.001, 0.01, 1.0, 0.0, 0.1, 1000.0]           # births and deaths on checkpoints
    parts[:7, 5] = [0.001, 0.01, np.inf, 10.0, 1000.0, 0.1, np.inf]
    rad0 = rng.normal(0, 50, (500, 4)).astype(np.float32)
    base = str(tmp_path)
    parts.tofile(base + "/parts.f64")
    rad0.tofile(base + "/rad0.f32")
    script = (f"const s=require({json.dumps(os.path.join(ROOT, 'webgpu-dna_amd', 'js', 'splat.js'))}); "
              f"const h=require({json.dumps(HOST)}); const fs=require('fs'); "
              f"const rd=(f,T)=>{{const b=fs.readFileSync({json.dumps(base)}+'/'+f); "
              f"return new T(b.buffer.slice(b.byteOffset,b.byteOffset+b.length));}}; "
              f"const snaps=s.snapshotsFromParticles(rd('parts.f64',Float64Array), 2000); "
              f"const blob=s.packSnapshotBlob(snaps, h.buildDNATarget(), rd('rad0.f32',Float32Array), 400, 10000); "
              f"fs.writeFileSync({json.dumps(base)}+'/blob.bin', blob); "
              f"const b=s.parseSnapshotBlob(blob); "
              f"console.log(JSON.stringify({{n:b.num_snaps, sn:b.snap_n, labels:b.snapshots.map(x=>x.label), "
              f"alive:b.snapshots.map(x=>x.alive.reduce((a,c)=>a+c,0)), r0:b.rad0.rad_n_initial}}));")
    out = json.loads(_node(script))
    ss = snapshots_from_particles(parts, 2000)
    ref = pack_snapshot_blob(ss, DNATarget(), rad0, 400, 10000)
    got = open(base + "/blob.bin", "rb").read()
    assert got == ref
    assert out == {"n": 8, "sn": 2000, "labels": [s["label"] for s in ss],
                   "alive": [int(s["alive"].sum()) for s in ss], "r0": 400}


def test_reaction_table_and_to_fixed_match_v8():
    """The reaction table from the addon equals the CPU restatement's, and the Python host's
    toFixed gives V8's strings (rxn_info sigma / rc, irt-worker.js:690-694) on the table and
    on values where rounding is delicate (binary midpoints, negative zero, tiny values)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "webgpu-dna_amd"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from dna_amd.host import to_fixed
    vals = [0.5, 1.5, 2.5, -0.0, 0.0, 1.0005, 1.0015, 0.125, 0.0625, -0.7109, 2.5e-5, 1e-9, 123.45675,
            0.1715, 0.71, 1.0117, -0.7112]
    out = json.loads(_node(f"const a=require({json.dumps(ADDON)}); const t=Array.from(a.irtReactions()); "
                           f"const v={json.dumps(vals)}.concat([-0]); "
                           "console.log(JSON.stringify({t, f3:v.map(x=>x.toFixed(3)), f4:v.map(x=>x.toFixed(4))}))"))
    tab = np.array(out["t"]).reshape(9, 6)
    assert np.array_equal(tab, O.irt_params())
    vals = vals + [-0.0]
    assert [to_fixed(x, 3) for x in vals] == out["f3"]
    assert [to_fixed(x, 4) for x in vals] == out["f4"]
