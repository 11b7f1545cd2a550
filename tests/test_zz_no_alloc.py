"""dna_score_noalloc (scoring with its scratch in a fixed per-GPU array) against dna_score,
through the C ABI: counts, dsb_expected and the LCG state after the call must be equal, on
two calls in a row, the second on a smaller input.  Inputs are the radicals resident from
a transport (rad_dev = NULL) and a caller-owned device buffer, allocated here with the HIP
runtime the library is bound to in this process (opened by its soname once the library is
loaded, so a runtime loaded earlier by another package is the one used)."""
import ctypes as C

import numpy as np
import pytest

from dna_amd import DNATarget, default_params
from dna_amd.abi import DnaScoreOut, DnaTarget
from dna_amd.lib import lib

pytestmark = pytest.mark.gpu


def _noalloc(ctx, dna, rad_dev, n, lcg):
    L = lib()
    L.dna_score_noalloc.argtypes = [C.c_void_p, C.POINTER(DnaTarget), C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint32), C.POINTER(DnaScoreOut)]
    t = dna.to_c()
    st = C.c_uint32(lcg)
    o = DnaScoreOut()
    rc = L.dna_score_noalloc(ctx.h, C.byref(t), rad_dev, n, C.byref(st), C.byref(o))
    assert rc == 0, L.dna_last_error()
    assert not o.hits
    return dict(ssb_count=o.ssb_count, candidates=o.candidates, in_reach=o.in_reach, dsb=o.dsb,
                ssb0=o.ssb0, ssb1=o.ssb1, dsb_expected=o.dsb_expected, lcg_state=st.value)


def _same(a, b):
    for k in ("ssb_count", "candidates", "in_reach", "dsb", "ssb0", "ssb1", "lcg_state"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert a["dsb_expected"] == b["dsb_expected"]


def test_noalloc_resident_two_calls(ctx):
    """10 keV x 1024 (about 1.3 M radicals: two chunks of 2^20), then a smaller prefix of the
    same radicals."""
    dna = DNATarget()
    p = default_params(10000.0, 1024, dna=dna)
    print("noalloc scoring: 10 keV x 1024 resident, then a prefix", flush=True)
    g = ctx.transport(p, copy=0)
    n = g["rad_n"]
    for m, seed in ((n, 0x12345678), (n // 3, 0x0BADF00D)):
        ref = ctx.score(dna, None, m, lcg_state=seed)
        got = _noalloc(ctx, dna, None, m, seed)
        _same(got, ref)
        assert got["in_reach"] > 0


def test_noalloc_device_buffer_spanning_chunks(ctx):
    """A caller-owned device buffer of synthetic radicals on and near the fibers, larger than
    one chunk (duplicate runs across the chunk boundary included), then a smaller one."""
    dna = DNATarget()
    lib()                                        # the library first: its HIP runtime is loaded
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    rng = np.random.default_rng(11)
    n = (1 << 20) + 300_000
    fy, fz = dna.fy.astype(np.float64), dna.fz.astype(np.float64)
    f = rng.integers(0, fy.size, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(0.6, 1.4, n)
    rad = np.empty((n, 4), np.float32)
    rad[:, 0] = rng.uniform(-dna.L_nm / 2, dna.L_nm / 2, n)
    rad[:, 1] = fy[f] + r * np.cos(ang)
    rad[:, 2] = fz[f] + r * np.sin(ang)
    rad[:, 3] = (np.arange(n) // 50 * 8).astype(np.float32)
    b = 1 << 20                                  # a run of equal positions across the chunk edge
    rad[b - 3:b + 4, :3] = rad[b - 3, :3]
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), rad.nbytes) == 0
    try:
        assert hip.hipMemcpy(buf, rad.ctypes.data, rad.nbytes, 1) == 0        # host to device
        print(f"noalloc scoring: device buffer of {n} radicals, then {n // 5}", flush=True)
        for m, seed in ((n, 0x12345678), (n // 5, 77)):
            ref = ctx.score(dna, rad[:m], lcg_state=seed)
            got = _noalloc(ctx, dna, buf, m, seed)
            _same(got, ref)
            assert got["ssb_count"] > 0
    finally:
        hip.hipFree(buf)


def test_noalloc_rejects_oversized_target(ctx):
    dna = DNATarget(grid_N=33)                   # 1089 fibers > 1024
    p = default_params(1000.0, 8)
    ctx.transport(p, copy=0)
    L = lib()
    L.dna_score_noalloc.argtypes = [C.c_void_p, C.POINTER(DnaTarget), C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint32), C.POINTER(DnaScoreOut)]
    t = dna.to_c()
    st = C.c_uint32(1)
    o = DnaScoreOut()
    assert L.dna_score_noalloc(ctx.h, C.byref(t), None, 0, C.byref(st), C.byref(o)) == 5   # DNA_ERR_CAPACITY
    assert b"fixed scratch" in L.dna_last_error()
