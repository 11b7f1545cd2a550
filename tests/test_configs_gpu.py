"""The batched workloads at their full sizes on the GPU.

* 65,536 primaries with log-uniform energies in [100 eV, 30 keV], isotropic, started
  uniformly in a 2 um cube, in one persistent launch: sampled primaries are bit-exact
  against the CPU restatement (each primary's RNG stream depends only on its global index,
  buffers.ts:94), energy conservation holds in fixed point over the whole batch, and the
  per-energy-bin CSDA and ion counts are written to the test log.
* 512 independent IRT instances, each the species list of one 1-20 keV primary
  (log-uniform), at the 16 log-spaced times from 1 ps to 1 us plus the reference's 7
  checkpoints: every instance's tallies are exact against the CPU restatement.
* G(1 us) at 3, 5 and 20 keV (4096 primaries each) against the reference's E10 rows
  (experiments/results/2026-05-07/level-4/E10-irt-vs-karamitros.json): within 2%, or
  2.5 sigma of the batch-mean Monte Carlo error.  Together with the 1 and 10 keV checks of
  the CPU restatement this pins the chemistry at every published energy.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from dna_amd import default_params
from dna_amd import surrogate as S
from dna_amd.abi import DNA_COPY_ALL, DNA_COPY_RESULTS

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FX = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_fixtures.json")))
E5_CSDA_SEED = 0x54524B01          # experiments/lib/seeds.mjs:17
E10_IRT_SEED = 0x43484D01          # experiments/lib/seeds.mjs:26
CK7 = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0]


def test_batched_uneven_transport_65536():
    from dna_amd import Context
    ctx = Context(0)
    n = 65536
    E = S.log_uniform_energies(n, E5_CSDA_SEED, 100.0, 30000.0)
    p = default_params(1000.0, n, start_half=1000.0, seed=E5_CSDA_SEED)
    print("transport 65536 primaries, 0.1-30 keV, 2 um start cube", flush=True)
    g = ctx.transport(p, energies=E, copy=DNA_COPY_ALL)
    res = g["results"]
    # energy conservation: every deposit is truncated to 0.01 eV, nothing escapes the box
    assert int((res["esc"] != 0).sum()) == 0 and g["rad_dropped"] == 0 and g["sec_dropped"] == 0
    dep = g["dose_sum"] / 100.0
    assert 0.999 * E.astype(np.float64).sum() < dep <= E.astype(np.float64).sum()
    # sampled primaries bit-exact against the CPU restatement run on just those primaries
    pid = (np.floor(g["rad"][:, 3].astype(np.float64) + 0.5) // 8).astype(np.int64)
    for first in (0, 31000, n - 200):
        o = O.transport(p, energies=E, first=first, count=200, n_threads=8)
        gr = res[first:first + 200]
        for f in ("path", "prod", "fE", "mx"):
            assert np.array_equal(gr[f].view(np.uint32), o["results"][f].view(np.uint32)), (first, f)
        for f in ("ni", "nx", "esc"):
            assert np.array_equal(gr[f], o["results"][f]), (first, f)
        a, b = np.searchsorted(pid, first), np.searchsorted(pid, first + 200)
        assert np.array_equal(g["rad"][a:b].view(np.uint32), o["rad"].view(np.uint32)), first
    # per-energy-bin CSDA and primary ionisations (20 log bins), for the log
    edges = 100.0 * 300.0 ** (np.arange(21) / 20.0)
    k = np.clip(np.searchsorted(edges, E, side="right") - 1, 0, 19)
    rows = []
    for b in range(20):
        m = k == b
        rows.append((round(float(np.sqrt(edges[b] * edges[b + 1])), 1), int(m.sum()),
                     round(float(res["path"][m].astype(np.float64).mean()), 2),
                     round(float(res["ni"][m].astype(np.float64).mean()), 2)))
    print("per-bin (E_mid eV, n, CSDA nm, ions/primary):", rows, flush=True)
    csda = np.array([r[2] for r in rows])
    assert np.all(np.diff(csda) > 0)
    ctx.close()


def test_batched_irt_512_instances(ctx):
    n = 512
    E = S.log_uniform_energies(n, E10_IRT_SEED, 1000.0, 20000.0)
    p = S.transport_params(n, seed=E10_IRT_SEED)
    print("transport 512 primaries 1-20 keV for IRT instances", flush=True)
    t = ctx.transport(p, energies=E)
    rad = t["rad"]
    pid = (np.floor(rad[:, 3].astype(np.float64) + 0.5) // 8).astype(np.int64)
    off = np.searchsorted(pid, np.arange(n + 1)).astype(np.uint64)
    tc = sorted(set(np.round(S.checkpoints_16(), 12).tolist()) | set(CK7))
    assert len(tc) == 19
    nth = np.ones(n, np.uint32)
    print(f"irt batch: 512 instances, {len(tc)} checkpoints", flush=True)
    g = ctx.irt_batch(rad, off, nth, E, tc, seed=21)
    o = O.irt_batch(rad, off, nth, E, tc, seed=21, n_threads=8)
    sizes = np.diff(off)
    assert sizes.min() > 50 and sizes.max() > 10 * sizes.min()     # uneven instances
    for i in range(n):
        gt = np.array([[g[i].tally[c][s] for s in range(5)] for c in range(len(tc))])
        ot = np.array([[o[i].tally[c][s] for s in range(5)] for c in range(len(tc))])
        assert np.array_equal(gt, ot), i
        assert list(g[i].rxn_counts) == list(o[i].rxn_counts), i


@pytest.mark.parametrize("E", [3000.0, 5000.0, 20000.0])
def test_g_1us_against_e10(ctx, E):
    row = [r for r in FX["E10_per_energy"] if r["eEv"] == int(E)][0]["g"]
    p = default_params(E, 4096)
    print(f"G(1 us) at {E} eV, 4096 primaries", flush=True)
    t = ctx.transport(p, copy=DNA_COPY_RESULTS | 4)
    nth = int((t["results"]["esc"] == 0).sum())
    rad = t["rad"]
    pid = (np.floor(rad[:, 3].astype(np.float64) + 0.5) // 8).astype(np.int64)
    Gs = []
    for b in range(8):               # 8 independent sub-batches for the MC error
        sel = (pid % 8) == b
        r = ctx.irt(rad[sel], None, nth // 8, E, [1000.0], seed=100 + b)
        Gs.append([r.G[0][s] for s in range(5)])
    Gs = np.array(Gs)
    G, sem = Gs.mean(0), Gs.std(0, ddof=1) / np.sqrt(8)
    print(f"  G = {G.round(4).tolist()} sem = {sem.round(4).tolist()} E10 = {row}", flush=True)
    for s, key in enumerate(("OH", "eaq", "H", "H2O2", "H2")):
        want = row[key]
        assert abs(G[s] - want) < max(0.02 * want, 2.5 * sem[s]), (key, G[s], want, sem[s])


def test_captured_irt_batch_replays_match_eager(ctx):
    """The 512-instance batch captured as one hipGraph (dna_irt_plan): every replay, with
    the program's own eager work and synchronisations in between, equals the eager
    dna_run_irt_batch result for the same seed, and the first replay equals the oracle's
    batch (oracle_irt_batch) instance by instance."""
    n = 512
    E = S.log_uniform_energies(n, E10_IRT_SEED, 1000.0, 20000.0)
    p = S.transport_params(n, seed=E10_IRT_SEED)
    t = ctx.transport(p, energies=E)
    rad = t["rad"]
    pid = (np.floor(rad[:, 3].astype(np.float64) + 0.5) // 8).astype(np.int64)
    off = np.searchsorted(pid, np.arange(n + 1)).astype(np.uint64)
    tc = S.checkpoints_16()
    nth = np.ones(n, np.uint32)
    print("irt plan from the resident radicals", flush=True)
    plan = ctx.irt_plan(None, off, nth, E, tc)
    try:
        print("oracle batch, seed 21", flush=True)
        ob = O.irt_batch(rad, off, nth, E, tc, seed=21, n_threads=8)
        first = plan.run(21)
        for i in range(n):
            a = [[first[i].tally[c][s] for s in range(5)] for c in range(16)]
            b = [[ob[i].tally[c][s] for s in range(5)] for c in range(16)]
            assert a == b, i
            assert list(first[i].rxn_counts) == list(ob[i].rxn_counts), i
        for k, seed in enumerate((21, 22, 21, 23)):
            print(f"replay {k} seed {seed}", flush=True)
            gp = plan.run(seed)
            ge = ctx.irt_batch(rad, off, nth, E, tc, seed=seed)
            ctx.transport(default_params(1000.0, 64), copy=DNA_COPY_RESULTS)     # eager work in between
            for i in range(n):
                a = [[gp[i].tally[c][s] for s in range(5)] for c in range(16)]
                b = [[ge[i].tally[c][s] for s in range(5)] for c in range(16)]
                assert a == b, (seed, i)
                assert list(gp[i].rxn_counts) == list(ge[i].rxn_counts), (seed, i)
                assert gp[i].n_tracks == ge[i].n_tracks
    finally:
        plan.close()
