"""ctypes wrapper of oracle/liboracle.so — the CPU restatement of the reference chain.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product (webgpu-dna_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "webgpu-dna_amd"))
from dna_amd.abi import DnaChemOut, DnaIndirectOut, DnaParams, DnaResult, DnaScoreOut, DnaTarget  # noqa: E402

TABLES = os.path.join(ROOT, "webgpu-dna_amd", "data", "cross_sections.bin")
LIB = os.path.join(HERE, "liboracle.so")


class OracleTransportOut(C.Structure):
    _fields_ = [("results", C.c_void_p), ("dose", C.c_void_p), ("rad_buf", C.c_void_p),
                ("rad_n", C.c_uint64), ("sec_n", C.c_uint64),
                ("counters", C.c_uint32 * 8), ("sec_stats", C.c_uint32 * 8),
                ("dbg", C.c_uint32 * 8), ("dose_sum", C.c_uint64),
                ("rad_n_raw", C.c_uint64), ("sec_n_raw", C.c_uint64)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_load_tables.argtypes = [C.c_char_p]
        L.oracle_fmath32.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_fmath64.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_xs.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_xs_el_cos.argtypes = [C.c_float, C.c_float]
        L.oracle_xs_el_cos.restype = C.c_float
        L.oracle_sample_W.argtypes = [C.c_float, C.c_uint32, C.c_float]
        L.oracle_sample_W.restype = C.c_float
        L.oracle_sample_vib_mode.argtypes = [C.c_float, C.c_float]
        L.oracle_sample_vib_mode.restype = C.c_uint32
        L.oracle_meesungnoen_sigma.argtypes = [C.c_float]
        L.oracle_meesungnoen_sigma.restype = C.c_float
        L.oracle_seed_primary.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
        L.oracle_rng_draws.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_transport_range.argtypes = [C.POINTER(DnaParams), C.c_void_p, C.c_uint32,
                                             C.c_uint32, C.c_int, C.POINTER(OracleTransportOut)]
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_irt.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_float, C.c_void_p,
                                 C.c_int, C.c_uint64, C.c_int, C.POINTER(DnaChemOut)]
        L.oracle_irt_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_int,
                                       C.c_void_p]
        L.oracle_irt_params.argtypes = [C.c_void_p]
        L.oracle_irt_sample.argtypes = [C.c_double, C.c_int, C.c_uint64, C.c_uint32,
                                        C.c_uint32, C.c_uint32]
        L.oracle_irt_sample.restype = C.c_double
        L.oracle_score.argtypes = [C.POINTER(DnaTarget), C.c_void_p, C.c_uint64,
                                   C.POINTER(C.c_uint32), C.c_void_p, C.POINTER(DnaScoreOut)]
        L.oracle_cluster.argtypes = [C.POINTER(DnaTarget), C.c_void_p, C.POINTER(DnaScoreOut)]
        L.oracle_irt_oh.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_float, C.c_void_p, C.c_int,
                                    C.c_uint64, C.c_int, C.POINTER(DnaChemOut), C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_uint64)]
        L.oracle_score_oh.argtypes = [C.POINTER(DnaTarget), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint32), C.c_void_p, C.POINTER(DnaScoreOut),
                                      C.POINTER(DnaIndirectOut)]
        L.oracle_hash_cells.argtypes = [C.c_void_p, C.c_uint64, C.c_float, C.c_uint32,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        rc = L.oracle_load_tables(TABLES.encode())
        if rc != 0:
            raise RuntimeError(f"oracle: cannot load {TABLES} (rc={rc})")
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def fmath32(fn, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().oracle_fmath32(fn, _ptr(x), _ptr(y), x.size)
    return y


def fmath64(fn, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().oracle_fmath64(fn, _ptr(x), _ptr(y), x.size)
    return y


def xs(which, E):
    E = np.ascontiguousarray(E, np.float32)
    width = {0: 3, 1: 5, 2: 5, 3: 1, 4: 1}[which]
    out = np.empty((E.size, width), np.float32)
    lib().oracle_xs(which, _ptr(E), _ptr(out), E.size)
    return out


def seed_primary(np_, seed):
    st = np.empty(4 * np_, np.uint32)
    lib().oracle_seed_primary(np_, seed, _ptr(st))
    return st.reshape(np_, 4)


def rng_draws(state4, n):
    state4 = np.ascontiguousarray(state4, np.uint32)
    raw = np.empty(n, np.uint32)
    rf = np.empty(n, np.float32)
    lib().oracle_rng_draws(_ptr(state4), n, _ptr(raw), _ptr(rf))
    return raw, rf


def transport(params, energies=None, first=0, count=None, n_threads=0):
    """Phase A+B on the CPU.  Returns a dict of numpy arrays (canonical rad order)."""
    if count is None:
        count = params.n - first
    e = None if energies is None else np.ascontiguousarray(energies, np.float32)
    o = OracleTransportOut()
    L = lib()
    rc = L.oracle_transport_range(C.byref(params), _ptr(e), first, count, n_threads, C.byref(o))
    if rc:
        raise RuntimeError(f"oracle_transport rc={rc}")
    nv = params.vc ** 3
    res = np.ctypeslib.as_array(C.cast(o.results, C.POINTER(DnaResult)), (count,)).copy()
    dose = np.ctypeslib.as_array(C.cast(o.dose, C.POINTER(C.c_uint32)), (nv,)).copy()
    rad = np.ctypeslib.as_array(C.cast(o.rad_buf, C.POINTER(C.c_float)),
                                (max(o.rad_n, 1) * 4,))[: o.rad_n * 4].copy().reshape(-1, 4)
    out = dict(results=res, dose=dose, rad=rad, rad_n=int(o.rad_n), sec_n=int(o.sec_n),
               rad_n_raw=int(o.rad_n_raw), sec_n_raw=int(o.sec_n_raw),
               counters=np.array(o.counters[:], np.uint32),
               sec_stats=np.array(o.sec_stats[:], np.uint32),
               dbg=np.array(o.dbg[:], np.uint32), dose_sum=int(o.dose_sum))
    for p in (o.results, o.dose, o.rad_buf):
        L.oracle_free(p)
    return out


def irt(rad, n_therm, E_eV, t_ckpt, seed=1, n_threads=0):
    rad = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
    tc = np.ascontiguousarray(t_ckpt, np.float64)
    o = DnaChemOut()
    rc = lib().oracle_irt(_ptr(rad), rad.shape[0], n_therm, E_eV, _ptr(tc), tc.size, seed,
                          n_threads, C.byref(o))
    if rc:
        raise RuntimeError(f"oracle_irt rc={rc}")
    return o


def irt_oh(rad, n_therm, E_eV, t_ckpt, seed=1, n_threads=0):
    """oracle_irt plus the OH alive at 1 us ([n, 4] float32, track order)."""
    rad = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
    tc = np.ascontiguousarray(t_ckpt, np.float64)
    o = DnaChemOut()
    p, n = C.c_void_p(), C.c_uint64()
    rc = lib().oracle_irt_oh(_ptr(rad), rad.shape[0], n_therm, E_eV, _ptr(tc), tc.size, seed,
                             n_threads, C.byref(o), C.byref(p), C.byref(n))
    if rc:
        raise RuntimeError(f"oracle_irt_oh rc={rc}")
    oh = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(max(n.value, 1) * 4,))[:n.value * 4]
    oh = oh.reshape(-1, 4).copy()
    lib().oracle_free(p)
    return o, oh


def irt_particles(rad, n_therm, E_eV, t_ckpt, seed=1, n_threads=0):
    """oracle_irt plus every particle slot at the end: [n, 6] float64 (x, y, z,
    pid*8+code, t_birth, t_death)."""
    rad = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
    tc = np.ascontiguousarray(t_ckpt, np.float64)
    o = DnaChemOut()
    p, n = C.c_void_p(), C.c_uint64()
    L = lib()
    L.oracle_irt_particles.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_float, C.c_void_p, C.c_int,
                                       C.c_uint64, C.c_int, C.POINTER(DnaChemOut), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_uint64)]
    rc = L.oracle_irt_particles(_ptr(rad), rad.shape[0], n_therm, E_eV, _ptr(tc), tc.size, seed,
                                n_threads, C.byref(o), C.byref(p), C.byref(n))
    if rc:
        raise RuntimeError(f"oracle_irt_particles rc={rc}")
    a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(max(n.value, 1) * 6,))[:n.value * 6]
    a = a.reshape(-1, 6).copy()
    L.oracle_free(p)
    return o, a


def irt_batch(rad, inst_off, n_therm, E, t_ckpt, seed=1, n_threads=0):
    rad = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
    off = np.ascontiguousarray(inst_off, np.uint64)
    nt = np.ascontiguousarray(n_therm, np.uint32)
    Ee = np.ascontiguousarray(E, np.float32)
    tc = np.ascontiguousarray(t_ckpt, np.float64)
    n_inst = off.size - 1
    outs = (DnaChemOut * n_inst)()
    rc = lib().oracle_irt_batch(_ptr(rad), _ptr(off), n_inst, _ptr(nt), _ptr(Ee), _ptr(tc),
                                tc.size, seed, n_threads, outs)
    if rc:
        raise RuntimeError(f"oracle_irt_batch rc={rc}")
    return list(outs)


def irt_params():
    o = np.empty(54, np.float64)
    lib().oracle_irt_params(_ptr(o))
    return o.reshape(9, 6)


def meesungnoen_sigma(k):
    """meesungnoen_sigma (primary.wgsl:90-98) of the transport restatement, FP32."""
    return lib().oracle_meesungnoen_sigma(float(k))


def irt_sample(r0, ri, seed, track, a, b):
    return lib().oracle_irt_sample(r0, ri, seed, track, a, b)


def score(target, rad, lcg_state=0x12345678, want_hits=False):
    rad = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
    t = target.to_c()
    st = C.c_uint32(lcg_state)
    hits = np.zeros(2 * target.n_bp, np.uint8) if want_hits else None
    o = DnaScoreOut()
    lib().oracle_score(C.byref(t), _ptr(rad), rad.shape[0], C.byref(st), _ptr(hits), C.byref(o))
    d = dict(ssb_count=o.ssb_count, candidates=o.candidates, in_reach=o.in_reach, dsb=o.dsb,
             ssb0=o.ssb0, ssb1=o.ssb1, dsb_expected=o.dsb_expected, lcg_state=st.value)
    if want_hits:
        d["hits"] = hits
    return d


def score_oh(target, rad, oh, lcg_state=0x12345678, want_hits=False):
    """Indirect SSB from the OH list, then direct, combined hits and clustering."""
    rad = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
    oh = np.ascontiguousarray(oh, np.float32).reshape(-1, 4)
    t = target.to_c()
    st = C.c_uint32(lcg_state)
    hits = np.zeros(2 * target.n_bp, np.uint8) if want_hits else None
    o, ind = DnaScoreOut(), DnaIndirectOut()
    lib().oracle_score_oh(C.byref(t), _ptr(rad), rad.shape[0], _ptr(oh), oh.shape[0], C.byref(st),
                          _ptr(hits), C.byref(o), C.byref(ind))
    d = dict(ssb_count=o.ssb_count, candidates=o.candidates, in_reach=o.in_reach, dsb=o.dsb,
             ssb0=o.ssb0, ssb1=o.ssb1, dsb_expected=o.dsb_expected, lcg_state=st.value,
             indirect=dict(ssb0=ind.ssb0, ssb1=ind.ssb1, candidates=ind.candidates, in_reach=ind.in_reach))
    if want_hits:
        d["hits"] = hits
    return d


def hash_cells(pos4, cell_nm=1.5, bits=23):
    pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    n = pos4.shape[0]
    key = np.empty(n, np.uint32)
    start = np.empty((1 << bits) + 1, np.uint32)
    order = np.empty(n, np.uint32)
    lib().oracle_hash_cells(_ptr(pos4), n, cell_nm, bits, _ptr(key), _ptr(start), _ptr(order))
    return key, start, order


def cluster(target, hits):
    hits = np.ascontiguousarray(hits, np.uint8)
    t = target.to_c()
    o = DnaScoreOut()
    lib().oracle_cluster(C.byref(t), _ptr(hits), C.byref(o))
    return dict(dsb=o.dsb, ssb0=o.ssb0, ssb1=o.ssb1, dsb_expected=o.dsb_expected)
