"""The IRT's particle table on the GPU (dna_irt_particles after an IRT run with
DNA_OPT_IRT_FINAL) against the CPU restatement, bit for bit: position, rad_buf code, birth
and death time of every particle slot, from a 10 keV transport and from synthetic tracks
given as host radicals.  Then the WGDNA4D blob of the host mirror (the snapshot capture of
run_chemistry_worker and generate_snapshot_at_10kev): snapshots whose per-species alive
counts equal the IRT's tallies, and a blob that parses back."""
import numpy as np
import pytest

import oracle as O
from dna_amd import default_params
from dna_amd.abi import DNA_COPY_RESULTS, DNA_OPT_IRT_FINAL
from dna_amd.splat import SNAPSHOT_LABELS, SNAPSHOT_TIMES_NS, parse_snapshot_blob, snapshots_from_particles

pytestmark = pytest.mark.gpu
CK = list(SNAPSHOT_TIMES_NS[1:])


def irt_final(ctx, rad, rad_n, n_therm, E, seed):
    ctx.set_option(DNA_OPT_IRT_FINAL, 1)
    try:
        return ctx.irt(rad, rad_n, n_therm, E, CK, seed=seed)
    finally:
        ctx.set_option(DNA_OPT_IRT_FINAL, 0)


def test_particles_exact_10kev(ctx):
    p = default_params(10000.0, 64)
    print("IRT particle table, 10 keV x 64", flush=True)
    t = ctx.transport(p, copy=DNA_COPY_RESULTS | 4)
    n_therm = int((t["results"]["esc"] == 0).sum())
    c = irt_final(ctx, None, t["rad_n"], n_therm, 10000.0, 7)
    g = ctx.irt_particles()
    o, ref = O.irt_particles(t["rad"], n_therm, 10000.0, CK, seed=7, n_threads=8)
    assert g.shape == ref.shape
    assert np.array_equal(g.view(np.uint64), ref.view(np.uint64))
    assert int(np.isfinite(g[:, 5]).sum()) == 2 * c.n_reacted
    ss = snapshots_from_particles(g, snapshot_n=g.shape[0])
    code = ss[0]["pos"][:, 3]
    for k, s in enumerate(ss[1:]):
        assert [int(s["alive"][code == q].sum()) for q in (0, 1, 2, 4)] == [c.tally[k][j] for j in range(4)]


def test_particles_exact_host_input(ctx):
    rng = np.random.default_rng(5)
    n = 2000
    rad = np.zeros((n, 4), np.float32)
    rad[:, :3] = rng.normal(0, 3, (n, 3))
    pid = rng.integers(0, 40, n)
    rad[:, 3] = pid * 8 + rng.choice([0, 1, 2, 3, 6], n)
    print("IRT particle table, 2000 host radicals in 40 tracks", flush=True)
    irt_final(ctx, rad, n, 40, 1000.0, 3)
    g = ctx.irt_particles()
    _, ref = O.irt_particles(rad, 40, 1000.0, CK, seed=3)
    assert np.array_equal(g.view(np.uint64), ref.view(np.uint64))


def test_particles_need_the_option(ctx):
    p = default_params(1000.0, 8)
    t = ctx.transport(p, copy=DNA_COPY_RESULTS)
    ctx.irt(None, t["rad_n"], 8, 1000.0, CK, seed=1)
    with pytest.raises(Exception, match="DNA_OPT_IRT_FINAL"):
        ctx.irt_particles()


def test_snapshot_blob_at_10kev():
    from dna_amd import host, splat
    print("WGDNA4D blob, 10 keV x 128", flush=True)
    r = splat.generate_snapshot_at_10kev({"np": 128, "boxNm": 15000.0, "ceEV": 7.4, "snapshot_n": 20000})
    b = parse_snapshot_blob(r["blob"])
    assert b["num_snaps"] == 8 == r["numCheckpoints"]
    assert [s["label"] for s in b["snapshots"]] == list(SNAPSHOT_LABELS)
    assert b["snap_n"] == 20000 and b["rad0"]["rad_n_initial"] == 20000 and b["energy_eV"] == 10000.0
    alive = [int(s["alive"].sum()) for s in b["snapshots"]]
    assert alive[0] > 0 and alive[1] >= alive[-1] > 0
    assert set(np.unique(b["snapshots"][0]["pos"][:, 3])) <= {0.0, 1.0, 2.0, 3.0, 4.0, 6.0}
    assert host.ensure_pipelines(128) is not None
