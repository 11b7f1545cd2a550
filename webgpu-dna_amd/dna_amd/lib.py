"""Loader and thin object wrapper of libdna_amd.so (the HIP C ABI of include/dna_amd.h).

There is no fallback: if the library or a GPU is missing, calls raise.

Handles still open when the interpreter exits are closed by an atexit hook, surrogates and
captured plans before their contexts, while this module and the HIP runtime are intact
(at interpreter shutdown, module globals are cleared in an order that `__del__` cannot
rely on).
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from .abi import (DNA_COPY_ALL, DnaChemOut, DnaIndirectOut, DnaParams, DnaResult, DnaScoreOut, DnaTarget,
                  DnaTransportOut)

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libdna_amd.so")
TABLE_PATH = os.path.join(PKG, "data", "cross_sections.bin")

_lib = None
_open = weakref.WeakSet()         # Context / Surrogate / IrtPlan objects with a live handle


@atexit.register
def _close_open_handles():
    objs = list(_open)
    for o in objs:                 # dependants first: they use their context's stream
        if not isinstance(o, Context):
            try:
                o.close()
            except Exception:
                pass
    for o in objs:
        if isinstance(o, Context):
            try:
                o.close()
            except Exception:
                pass


def irt_reactions():
    """The IRT's reaction table, [9, 6] float64: sigma, sigma_eff, rc (nm), k_dif, k_act
    (nm^3/ns), contact probability (dna_irt_reactions; needs no device)."""
    out = np.empty((9, 6), np.float64)
    _check("dna_irt_reactions", lib().dna_irt_reactions(_ptr(out)))
    return out


class DnaError(RuntimeError):
    def __init__(self, fn, rc, msg):
        super().__init__(f"{fn} failed (status {rc}): {msg}")
        self.rc = rc


def lib():
    """Load libdna_amd.so (built by webgpu-dna_amd/build.py); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: run webgpu-dna_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
    L.dna_last_error.restype = C.c_char_p
    L.dna_version.restype = C.c_char_p
    L.dna_ctx_create.argtypes = [C.c_int, C.c_char_p, C.POINTER(vp)]
    L.dna_ctx_destroy.argtypes = [vp]
    L.dna_run_transport.argtypes = [vp, C.POINTER(DnaParams), vp, u32, C.POINTER(DnaTransportOut)]
    L.dna_run_irt.argtypes = [vp, vp, u64, u32, C.c_float, vp, C.c_int, u64, C.POINTER(DnaChemOut)]
    L.dna_run_irt_batch.argtypes = [vp, vp, vp, u32, vp, vp, vp, C.c_int, u64, vp]
    L.dna_irt_stage.argtypes = [vp, vp, u64, u32, C.c_float, vp, C.c_int]
    L.dna_irt_finish.argtypes = [vp, u64, C.POINTER(DnaChemOut)]
    L.dna_irt_launch.argtypes = [vp, u64]
    L.dna_irt_wait.argtypes = [vp, C.c_int, vp]
    L.dna_score.argtypes = [vp, C.POINTER(DnaTarget), vp, u64, C.POINTER(u32), C.c_int,
                            C.POINTER(DnaScoreOut)]
    L.dna_hash_cells.argtypes = [vp, vp, u64, C.c_float, u32, vp, vp, vp]
    L.dna_irt_plan_create.argtypes = [vp, vp, vp, u32, vp, vp, vp, C.c_int, C.POINTER(vp)]
    L.dna_irt_plan_run.argtypes = [vp, u64, vp]
    L.dna_irt_plan_destroy.argtypes = [vp]
    L.dna_irt_plan_destroy.restype = None
    L.dna_set_option.argtypes = [vp, C.c_int, C.c_int]
    L.dna_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.dna_xs_probe.argtypes = [vp, C.c_int, vp, vp, u64, vp]
    L.dna_irt_oh_final.argtypes = [vp, vp, u64, C.POINTER(u64)]
    L.dna_irt_particles.argtypes = [vp, vp, u64, C.POINTER(u64)]
    L.dna_irt_reactions.argtypes = [vp]
    L.dna_score_oh.argtypes = [vp, C.POINTER(DnaTarget), vp, u64, vp, u64, C.POINTER(u32), C.c_int,
                               C.POINTER(DnaScoreOut), C.POINTER(DnaIndirectOut)]
    L.dna_score_noalloc.argtypes = [vp, C.POINTER(DnaTarget), vp, u64, C.POINTER(u32), C.POINTER(DnaScoreOut)]
    L.dna_run_irt_tracks.argtypes = [vp, vp, u64, u32, C.c_float, vp, u32, vp, C.c_int, u64, vp,
                                     C.POINTER(DnaChemOut)]
    ip = C.POINTER(C.c_int)
    L.dna_surrogate_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, u64, C.POINTER(vp)]
    L.dna_surrogate_destroy.argtypes = [vp]
    L.dna_surrogate_destroy.restype = None
    L.dna_surrogate_shape.argtypes = [vp, ip, ip, ip, ip, C.POINTER(u64)]
    L.dna_surrogate_get_params.argtypes = [vp, vp]
    L.dna_surrogate_set_params.argtypes = [vp, vp]
    L.dna_surrogate_set_norm.argtypes = [vp, vp, vp, vp, vp]
    L.dna_surrogate_forward.argtypes = [vp, vp, C.c_int, vp]
    L.dna_surrogate_grad.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_double), vp]
    L.dna_surrogate_train.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                      u64, C.POINTER(C.c_double)]
    L.dna_track_features.argtypes = [vp, vp, vp, u32, vp, vp]
    L.dna_surrogate_predict.argtypes = [vp, vp, u32, C.c_float, vp, vp, vp, C.POINTER(u32), vp]
    _lib = L
    return L


EXPORTED_SYMBOLS = ["dna_ctx_create", "dna_ctx_destroy", "dna_last_error", "dna_version",
                    "dna_run_transport", "dna_run_irt", "dna_run_irt_batch", "dna_score",
                    "dna_hash_cells", "dna_irt_plan_create", "dna_irt_plan_run",
                    "dna_irt_plan_destroy", "dna_set_option", "dna_get_option", "dna_run_irt_tracks", "dna_surrogate_create",
                    "dna_surrogate_destroy", "dna_surrogate_shape", "dna_surrogate_get_params",
                    "dna_surrogate_set_params", "dna_surrogate_set_norm", "dna_surrogate_forward",
                    "dna_surrogate_grad", "dna_surrogate_train", "dna_track_features",
                    "dna_surrogate_predict", "dna_xs_probe", "dna_irt_oh_final", "dna_score_oh", "dna_irt_particles",
                    "dna_irt_reactions", "dna_score_noalloc"]


def _check(fn, rc):
    if rc != 0:
        raise DnaError(fn, rc, lib().dna_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


class Context:
    """One dna_ctx (one GPU).  Output arrays returned by methods are copies."""

    def __init__(self, device=0, table_path=TABLE_PATH):
        L = lib()
        h = C.c_void_p()
        _check("dna_ctx_create", L.dna_ctx_create(device, table_path.encode(), C.byref(h)))
        self.h = h
        _open.add(self)

    def close(self):
        if self.h:
            lib().dna_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- launch options (DNA_OPT_*)
    def set_option(self, option, value):
        _check("dna_set_option", lib().dna_set_option(self.h, option, value))

    def get_option(self, option):
        v = C.c_int()
        _check("dna_get_option", lib().dna_get_option(self.h, option, C.byref(v)))
        return v.value

    # ---- table lookups of the transport step (dna_xs_probe)
    XS_WIDTH = {0: 3, 1: 5, 2: 5, 3: 1, 4: 1, 5: 1}

    def xs_probe(self, which, E, r=None):
        """Evaluate a transport lookup on the GPU: 0 xs_all -> [n,3] (ion, exc, el nm^2),
        1 shell fractions [n,5], 2 excitation fractions [n,5], 3 vibrational total [n,1],
        4 DEA [n,1], 5 elastic cos(theta) at (E, r) [n,1]."""
        E = np.ascontiguousarray(E, np.float32).ravel()
        rr = None if r is None else np.ascontiguousarray(r, np.float32).ravel()
        out = np.empty((E.size, self.XS_WIDTH[which]), np.float32)
        _check("dna_xs_probe", lib().dna_xs_probe(self.h, which, _ptr(E), None if rr is None else _ptr(rr),
                                                  E.size, _ptr(out)))
        return out

    # ---- Phase A + B
    def transport(self, params, energies=None, copy=DNA_COPY_ALL):
        e = None if energies is None else np.ascontiguousarray(energies, np.float32)
        o = DnaTransportOut()
        _check("dna_run_transport", lib().dna_run_transport(self.h, C.byref(params), _ptr(e), copy,
                                                            C.byref(o)))
        out = dict(counters=np.array(o.counters[:], np.uint32),
                   sec_stats=np.array(o.sec_stats[:], np.uint32),
                   dbg=np.array(o.dbg[:], np.uint32), dose_sum=int(o.dose_sum),
                   rad_n=int(o.rad_n_stored), rad_n_raw=int(o.rad_n_raw),
                   rad_dropped=int(o.rad_dropped), sec_n=int(o.sec_n),
                   sec_n_raw=int(o.sec_n_raw), sec_dropped=int(o.sec_dropped),
                   t_prim_ms=o.t_prim_ms, t_sec_ms=o.t_sec_ms, t_total_ms=o.t_total_ms)
        n = params.n
        if o.results:
            out["results"] = np.ctypeslib.as_array(C.cast(o.results, C.POINTER(DnaResult)), (n,)).copy()
        if o.dose:
            nv = params.vc ** 3
            out["dose"] = np.ctypeslib.as_array(C.cast(o.dose, C.POINTER(C.c_uint32)), (nv,)).copy()
        if o.rad_buf:
            m = out["rad_n"]
            out["rad"] = np.ctypeslib.as_array(C.cast(o.rad_buf, C.POINTER(C.c_float)),
                                               (m * 4,)).copy().reshape(m, 4)
        return out

    # ---- Phase C
    def irt(self, rad, rad_n, n_therm, E_eV, t_ckpt, seed=1):
        """rad: (n,4) float32 array, or None for the device-resident radicals."""
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        tc = np.ascontiguousarray(t_ckpt, np.float64)
        o = DnaChemOut()
        n = rad_n if r is None else r.shape[0]
        _check("dna_run_irt", lib().dna_run_irt(self.h, _ptr(r), n, n_therm, E_eV, _ptr(tc),
                                                tc.size, seed, C.byref(o)))
        return o

    def irt_stage(self, rad, rad_n, n_therm, E_eV, t_ckpt):
        """irt() in two calls (dna_irt_stage / dna_irt_finish): the grouping and preparation
        now, the IRT kernels at irt_finish(seed).  rad None = the device-resident radicals,
        which must not change before irt_finish."""
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        tc = np.ascontiguousarray(t_ckpt, np.float64)
        n = rad_n if r is None else r.shape[0]
        _check("dna_irt_stage", lib().dna_irt_stage(self.h, _ptr(r), n, n_therm, E_eV, _ptr(tc), tc.size))

    def irt_finish(self, seed=1):
        o = DnaChemOut()
        _check("dna_irt_finish", lib().dna_irt_finish(self.h, seed, C.byref(o)))
        return o

    def irt_launch(self, seed=1):
        """irt_finish in two calls: the staged IRT's kernels enqueued now."""
        _check("dna_irt_launch", lib().dna_irt_launch(self.h, seed))

    def irt_wait(self, loop_only=False):
        """Wait for the launched IRT: until its event loop starts (loop_only, returns None),
        or until it is done (returns its DnaChemOut)."""
        if loop_only:
            _check("dna_irt_wait", lib().dna_irt_wait(self.h, 0, None))
            return None
        o = DnaChemOut()
        _check("dna_irt_wait", lib().dna_irt_wait(self.h, 1, C.byref(o)))
        return o

    def irt_batch(self, rad, inst_off, n_therm, E, t_ckpt, seed=1):
        r = np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        off = np.ascontiguousarray(inst_off, np.uint64)
        nt = np.ascontiguousarray(n_therm, np.uint32)
        Ee = np.ascontiguousarray(E, np.float32)
        tc = np.ascontiguousarray(t_ckpt, np.float64)
        n_inst = off.size - 1
        outs = (DnaChemOut * n_inst)()
        _check("dna_run_irt_batch", lib().dna_run_irt_batch(self.h, _ptr(r), _ptr(off), n_inst,
                                                            _ptr(nt), _ptr(Ee), _ptr(tc), tc.size,
                                                            seed, outs))
        return list(outs)

    def irt_tracks(self, rad, rad_n, n_prim, n_therm, E_eV, t_ckpt, seed=1, select=None,
                   per_track=True):
        """IRT on the tracks of primaries pid < n_prim (those with select[pid] != 0 when given).
        rad: (n,4) float32 array, or None for the rad_n device-resident radicals.
        Returns (aggregate DnaChemOut, per-primary species counts [n_prim, n_ckpt, 5] or None)."""
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        sel = None if select is None else np.ascontiguousarray(select, np.uint8)
        if sel is not None and sel.size != n_prim:
            raise ValueError("select must have n_prim entries")
        tc = np.ascontiguousarray(t_ckpt, np.float64)
        tt = np.zeros((n_prim, tc.size, 5), np.uint32) if per_track else None
        o = DnaChemOut()
        n = rad_n if r is None else r.shape[0]
        _check("dna_run_irt_tracks", lib().dna_run_irt_tracks(
            self.h, _ptr(r), n, n_therm, E_eV, _ptr(sel), n_prim, _ptr(tc), tc.size, seed,
            _ptr(tt), C.byref(o)))
        return o, tt

    def irt_plan(self, rad, inst_off, n_therm, E, t_ckpt):
        """A captured batch (IrtPlan); rad None = the resident radicals."""
        return IrtPlan(self, rad, inst_off, n_therm, E, t_ckpt)

    # ---- scoring
    def score(self, target, rad=None, rad_n=0, lcg_state=0x12345678, want_hits=False):
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        t = target.to_c()
        st = C.c_uint32(lcg_state)
        o = DnaScoreOut()
        n = rad_n if r is None else r.shape[0]
        _check("dna_score", lib().dna_score(self.h, C.byref(t), _ptr(r), n, C.byref(st),
                                            1 if want_hits else 0, C.byref(o)))
        d = dict(ssb_count=int(o.ssb_count), candidates=int(o.candidates),
                 in_reach=int(o.in_reach), dsb=int(o.dsb), ssb0=int(o.ssb0), ssb1=int(o.ssb1),
                 dsb_expected=o.dsb_expected, lcg_state=st.value, t_ms=o.t_ms)
        if want_hits and o.hits:
            d["hits"] = np.ctypeslib.as_array(C.cast(o.hits, C.POINTER(C.c_uint8)),
                                              (2 * target.n_bp,)).copy()
        return d

    def irt_oh_final(self, copy=True):
        """The OH alive at 1 us of the last irt() made with DNA_OPT_IRT_FINAL = 1:
        [n, 4] float32 in track order (copy=True), or just n (copy=False; the list stays
        on the device for score_oh(oh=None))."""
        n = C.c_uint64()
        _check("dna_irt_oh_final", lib().dna_irt_oh_final(self.h, None, 0, C.byref(n)))
        if not copy:
            return n.value
        out = np.empty((max(n.value, 1), 4), np.float32)
        _check("dna_irt_oh_final", lib().dna_irt_oh_final(self.h, _ptr(out), n.value, C.byref(n)))
        return out[:n.value]

    def irt_particles(self):
        """Every particle slot of the last irt() made with DNA_OPT_IRT_FINAL = 1: [n, 6]
        float64 (x, y, z, pid*8+code, t_birth, t_death), track order."""
        n = C.c_uint64()
        _check("dna_irt_particles", lib().dna_irt_particles(self.h, None, 0, C.byref(n)))
        out = np.empty((max(n.value, 1), 6), np.float64)
        _check("dna_irt_particles", lib().dna_irt_particles(self.h, _ptr(out), n.value, C.byref(n)))
        return out[:n.value]

    def score_oh(self, target, rad=None, rad_n=0, oh=None, n_oh=0, lcg_state=0x12345678, want_hits=False):
        """score() with the indirect-SSB pass on the alive-OH list first (oh=None: the
        list resident from irt_oh_final)."""
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        q = None if oh is None else np.ascontiguousarray(oh, np.float32).reshape(-1, 4)
        t = target.to_c()
        st = C.c_uint32(lcg_state)
        o, ind = DnaScoreOut(), DnaIndirectOut()
        n = rad_n if r is None else r.shape[0]
        m = n_oh if q is None else q.shape[0]
        _check("dna_score_oh", lib().dna_score_oh(self.h, C.byref(t), _ptr(r), n, _ptr(q), m, C.byref(st),
                                                  1 if want_hits else 0, C.byref(o), C.byref(ind)))
        d = dict(ssb_count=int(o.ssb_count), candidates=int(o.candidates),
                 in_reach=int(o.in_reach), dsb=int(o.dsb), ssb0=int(o.ssb0), ssb1=int(o.ssb1),
                 dsb_expected=o.dsb_expected, lcg_state=st.value, t_ms=o.t_ms,
                 indirect=dict(ssb0=int(ind.ssb0), ssb1=int(ind.ssb1), candidates=int(ind.candidates),
                               in_reach=int(ind.in_reach)))
        if want_hits and o.hits:
            d["hits"] = np.ctypeslib.as_array(C.cast(o.hits, C.POINTER(C.c_uint8)),
                                              (2 * target.n_bp,)).copy()
        return d

    # ---- spatial hash
    def hash_cells(self, pos4, cell_nm=1.5, bits=23):
        p = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        n = p.shape[0]
        key = np.empty(n, np.uint32)
        start = np.empty((1 << bits) + 1, np.uint32)
        order = np.empty(n, np.uint32)
        _check("dna_hash_cells", lib().dna_hash_cells(self.h, _ptr(p), n, cell_nm, bits, _ptr(key),
                                                      _ptr(start), _ptr(order)))
        return key, start, order


class IrtPlan:
    """dna_irt_plan: IRT instances planned once and captured as a hipGraph; run(seed)
    replays it."""

    def __init__(self, ctx, rad, inst_off, n_therm, E, t_ckpt):
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        off = np.ascontiguousarray(inst_off, np.uint64)
        self._keep = (np.ascontiguousarray(n_therm, np.uint32), np.ascontiguousarray(E, np.float32),
                      np.ascontiguousarray(t_ckpt, np.float64))
        nt, Ee, tc = self._keep
        self.n_inst = off.size - 1
        h = C.c_void_p()
        _check("dna_irt_plan_create", lib().dna_irt_plan_create(ctx.h, _ptr(r), _ptr(off), self.n_inst, _ptr(nt),
                                                                _ptr(Ee), _ptr(tc), tc.size, C.byref(h)))
        self.h, self.ctx = h, ctx
        _open.add(self)

    def run(self, seed):
        outs = (DnaChemOut * self.n_inst)()
        _check("dna_irt_plan_run", lib().dna_irt_plan_run(self.h, seed, outs))
        return list(outs)

    def close(self):
        if self.h:
            lib().dna_irt_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """dna_surrogate: the BF16 MLP chemistry surrogate on the context's GPU."""

    def __init__(self, ctx, n_in=70, n_hidden=256, n_layers=3, n_out=80, seed=1):
        h = C.c_void_p()
        _check("dna_surrogate_create", lib().dna_surrogate_create(ctx.h, n_in, n_hidden, n_layers,
                                                                  n_out, seed, C.byref(h)))
        self.h, self.ctx = h, ctx
        _open.add(self)
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        npar = C.c_uint64()
        _check("dna_surrogate_shape", lib().dna_surrogate_shape(h, C.byref(a), C.byref(b), C.byref(c),
                                                                C.byref(d), C.byref(npar)))
        self.n_in, self.n_hidden, self.n_layers, self.n_out = a.value, b.value, c.value, d.value
        self.n_params = npar.value
        self.n_in_pad = (self.n_in + 31) // 32 * 32

    def close(self):
        if self.h:
            lib().dna_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layer_dims(self):
        """[(din, dout)] per layer, as laid out in the flat parameter vector."""
        dims = []
        for l in range(self.n_layers + 1):
            di = self.n_in_pad if l == 0 else self.n_hidden
            do = 2 * self.n_out if l == self.n_layers else self.n_hidden
            dims.append((di, do))
        return dims

    def unflatten(self, flat):
        """Split a flat parameter vector into ([W_l (din, dout)], [b_l (dout,)])."""
        Ws, bs, off = [], [], 0
        for di, do in self.layer_dims():
            Ws.append(flat[off:off + di * do].reshape(di, do))
            off += di * do
        for di, do in self.layer_dims():
            bs.append(flat[off:off + do])
            off += do
        return Ws, bs

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        _check("dna_surrogate_get_params", lib().dna_surrogate_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter vector has the wrong size")
        _check("dna_surrogate_set_params", lib().dna_surrogate_set_params(self.h, _ptr(p)))

    def set_norm(self, x_mean, x_std, y_mean, y_std):
        a = [np.ascontiguousarray(v, np.float32) for v in (x_mean, x_std, y_mean, y_std)]
        _check("dna_surrogate_set_norm", lib().dna_surrogate_set_norm(self.h, *[_ptr(v) for v in a]))
        self.norm = a

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.empty((X.shape[0], 2 * self.n_out), np.float32)
        _check("dna_surrogate_forward", lib().dna_surrogate_forward(self.h, _ptr(X), X.shape[0], _ptr(Y)))
        return Y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        _check("dna_surrogate_grad", lib().dna_surrogate_grad(self.h, _ptr(X), _ptr(T), X.shape[0],
                                                              C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, X, T, batch=512, epochs=10, lr=1e-3, weight_decay=0.0, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = C.c_double()
        _check("dna_surrogate_train", lib().dna_surrogate_train(self.h, _ptr(X), _ptr(T), X.shape[0],
                                                                batch, epochs, lr, weight_decay, seed,
                                                                C.byref(loss)))
        return loss.value

    def features(self, E, rad=None, inst_off=None):
        """Per-track features [n, 70]; rad None = the primaries resident from the last transport."""
        Ee = np.ascontiguousarray(E, np.float32)
        n = Ee.size
        F = np.empty((n, 70), np.float32)
        r = None if rad is None else np.ascontiguousarray(rad, np.float32).reshape(-1, 4)
        off = None if inst_off is None else np.ascontiguousarray(inst_off, np.uint64)
        _check("dna_track_features", lib().dna_track_features(self.h, _ptr(r), _ptr(off), n, _ptr(Ee),
                                                              _ptr(F)))
        return F

    def predict(self, E, sigma_max, want_values=True):
        """Surrogate chemistry for the resident primaries: (mu, sigma, fallback, n_fallback, t_ms)."""
        Ee = np.ascontiguousarray(E, np.float32)
        n = Ee.size
        mu = np.empty((n, self.n_out), np.float32) if want_values else None
        sg = np.empty((n, self.n_out), np.float32) if want_values else None
        fb = np.empty(n, np.uint8)
        nf = C.c_uint32()
        t = np.zeros(3, np.float32)
        _check("dna_surrogate_predict", lib().dna_surrogate_predict(self.h, _ptr(Ee), n, sigma_max,
                                                                    _ptr(mu), _ptr(sg), _ptr(fb),
                                                                    C.byref(nf), _ptr(t)))
        return mu, sg, fb, nf.value, t
