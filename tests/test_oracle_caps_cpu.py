"""Transport caps in the CPU restatement: the reference's MAX_SEC / MAX_RAD with
*_dropped counts (src/gpu/dispatch.ts:124-126,190-192; primary.wgsl:253,364).

Here a cap keeps the first max_sec secondaries / max_rad radical entries in canonical
(primary, emission) order, so the capped run is fixed by the uncapped one:
  * the primaries' own histories do not change;
  * the kept radical buffer of a max_rad cap is the uncapped buffer's prefix;
  * under a max_sec cap each primary's radical block is a prefix of its uncapped block
    (its own emissions come first, then its secondaries' in emission order), and the
    whole block for every primary whose secondaries all fall below the cap.
The per-primary secondary counts come from single-primary runs."""
import numpy as np

import oracle as O
from dna_amd import default_params


def _blocks(rad):
    pid = np.floor(rad[:, 3].astype(np.float64) / 8.0).astype(np.int64)
    return {int(k): rad[pid == k] for k in np.unique(pid)}


def test_rad_cap_keeps_the_canonical_prefix():
    p = default_params(1000.0, 48)
    full = O.transport(p, n_threads=4)
    q = default_params(1000.0, 48)
    q.max_rad = full["rad_n"] // 3
    cap = O.transport(q, n_threads=4)
    assert cap["rad_n"] == q.max_rad and cap["rad_n_raw"] == full["rad_n"]
    assert np.array_equal(cap["rad"].view(np.uint32), full["rad"][: q.max_rad].view(np.uint32))
    assert np.array_equal(cap["results"].view(np.uint32), full["results"].view(np.uint32))
    assert np.array_equal(cap["counters"], full["counters"])      # emitted, as counters[7]
    assert cap["dose_sum"] == full["dose_sum"]


def test_sec_cap_transports_the_first_secondaries():
    n = 40
    p = default_params(1000.0, n)
    full = O.transport(p, n_threads=4)
    nsec = np.array([O.transport(p, first=i, count=1)["sec_n"] for i in range(n)])
    assert nsec.sum() == full["sec_n"]
    K = int(full["sec_n"] // 2)
    q = default_params(1000.0, n)
    q.max_sec = K
    cap = O.transport(q, n_threads=4)
    assert cap["sec_n"] == K and cap["sec_n_raw"] == full["sec_n"]
    assert np.array_equal(cap["results"].view(np.uint32), full["results"].view(np.uint32))
    assert cap["counters"][6] == full["counters"][6]               # secondaries emitted
    assert cap["dose_sum"] < full["dose_sum"]                      # dropped secondaries deposit nothing
    fb, cb = _blocks(full["rad"]), _blocks(cap["rad"])
    before = np.concatenate([[0], np.cumsum(nsec)])
    for pid, blk in fb.items():
        got = cb.get(pid, np.empty((0, 4), np.float32))
        assert got.shape[0] <= blk.shape[0]
        assert np.array_equal(got.view(np.uint32), blk[: got.shape[0]].view(np.uint32)), pid
        if before[pid + 1] <= K:
            assert got.shape[0] == blk.shape[0], pid
        elif before[pid] >= K:
            assert got.shape[0] < blk.shape[0], pid                   # its secondaries dropped
