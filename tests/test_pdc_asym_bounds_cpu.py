"""The cancellation-free bounds on SamplePDC's f = 1 - a sqrt(pi X) erfcx(z) for z >= 6
(irt.hip: f_asym, used by pdc_fast and pdc_refine) against the FP64 value the exact path
computes (irt-worker.js:101-103 with the worker's erfcx, irt-worker.js:53-88, as restated by
the oracle: include/dna_irt_math.h dna_erfcx).

For large z the FP64 expression cancels: f falls to 1e-5 .. 1e-14 while its two terms are
near 1.  The device bounds it as (b + a X (1 - S)) / (b + a X), S = sqrt(pi) z erfcx(z),
with the asymptotic series of 1 - S (four terms, remainder at most the fifth) in FP32,
then widens by a relative margin (1e-5 on the first decision's FP32 inputs, 4e-6 on the
second decision's FP64-rounded ones) and 1e-13 absolute (the FP64 expression's own
rounding).  Restated here in numpy float32, operation by operation, and checked on
random (a, b, X) over and beyond the ranges the IRT meets (a 0.5-20, b 0-500,
z 6-1e8): the FP64 f lies inside every interval, the intervals are as tight as the margins
and the remainder term allow, and a series with its r^4 term dropped or with a sign
flipped fails containment (the check has teeth)."""
import ctypes as C

import numpy as np

import oracle as O

F = np.float32


def _oracle_erfcx(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    dp = C.POINTER(C.c_double)
    O.lib().oracle_fmath64(15, x.ctypes.data_as(dp), y.ctypes.data_as(dp), C.c_int64(x.size))
    return y


def _f_ref(a, b, X):
    """the exact path's FP64 f, in the worker's operation order"""
    sqX = np.sqrt(X)
    return 1.0 - a * np.sqrt(np.pi * X) * _oracle_erfcx(b / sqX + a * sqX)


def _f_asym(aX, b, z, c4=F(6.5625), s3=1.0):
    aX, b, z = aX.astype(F), b.astype(F), z.astype(F)
    r = (F(1) / (z * z).astype(F)).astype(F)
    inner = (c4 * r).astype(F)
    inner = (F(1.875) - inner).astype(F) if s3 > 0 else (F(1.875) + inner).astype(F)
    inner = (F(0.75) - (r * inner).astype(F)).astype(F)
    inner = (F(0.5) - (r * inner).astype(F)).astype(F)
    t4 = (r * inner).astype(F)
    rm = (F(29.53125) * r * r * r * r * r).astype(F)
    den = (F(1) / (aX + b).astype(F)).astype(F)
    lo = ((b + (aX * t4).astype(F)).astype(F) * den).astype(F)
    hi = ((b + (aX * (t4 + rm).astype(F)).astype(F)).astype(F) * den).astype(F)
    return lo, hi


def _fast_inputs(a, b, X):
    """pdc_fast's FP32 a X, b and z (z from FP32 sqrt and reciprocal)"""
    Xf, af, bf = X.astype(F), a.astype(F), b.astype(F)
    sq = np.sqrt(Xf).astype(F)
    z = ((bf * (F(1) / sq).astype(F)).astype(F) + (af * sq).astype(F)).astype(F)
    return (af * Xf).astype(F), bf, z


def _refine_inputs(a, b, X):
    """pdc_refine's FP64 a X, b and z rounded to FP32"""
    sqX = np.sqrt(X)
    return (a * X).astype(F), b.astype(F), (b / sqX + a * sqX).astype(F)


def _intervals(a, b, X, **kw):
    out = []
    for inputs, m in ((_fast_inputs, 1e-5), (_refine_inputs, 4e-6)):
        aX, bf, z = inputs(a, b, X)
        lo, hi = _f_asym(aX, bf, z, **kw)
        lo = (lo * F(1 - m)).astype(F) - F(1e-13)
        hi = (hi * F(1 + m)).astype(F) + F(1e-13)
        out.append((lo.astype(np.float64), hi.astype(np.float64), z))
    return out


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    a = np.exp(rng.uniform(np.log(0.5), np.log(20.0), n))
    b = np.exp(rng.uniform(np.log(1e-4), np.log(500.0), n))
    b[: n // 20] = 0.0
    sqX = np.exp(rng.uniform(np.log(1e-2), np.log(1e8), n))
    X = sqX * sqX
    z = b / np.sqrt(X) + a * np.sqrt(X)
    keep = (z >= 6.05) & (z <= 1e8)
    return a[keep], b[keep], X[keep]


def test_fp64_f_inside_both_intervals():
    a, b, X = _samples(400_000, 3)
    assert a.size > 100_000
    f = _f_ref(a, b, X)
    for lo, hi, z in _intervals(a, b, X):
        ok = z >= 6.0
        assert np.all(lo[ok] <= f[ok]) and np.all(f[ok] <= hi[ok])


def test_intervals_are_tight():
    """width / f is at most: twice the relative margin, plus the remainder term's share
    945 r^5 / (32 t4) <= 63 r^4 (3.7e-5 at z = 6), plus the absolute margins' share
    2e-13 / f, plus 1e-6 of FP32 rounding"""
    a, b, X = _samples(200_000, 5)
    f = _f_ref(a, b, X)
    assert (f > 1e-8).sum() > 10_000
    ok = f > 1e-12                      # below, the FP64 value itself is mostly rounding
    a, b, X, f = a[ok], b[ok], X[ok], f[ok]
    for (lo, hi, z), m in zip(_intervals(a, b, X), (1e-5, 4e-6)):
        r = 1.0 / (z.astype(np.float64) ** 2)
        w = (hi - lo) / f
        assert np.all(w <= 2 * m + 63.0 * r ** 4 + 2e-13 / f + 1e-6)
        assert np.max(w[(f > 1e-7) & (z > 10)]) < 3e-5


def test_planted_mistakes_fail_containment():
    a, b, X = _samples(100_000, 7)
    f = _f_ref(a, b, X)
    for kw in ({"c4": F(0.0)}, {"s3": -1.0}):
        bad = 0
        for lo, hi, z in _intervals(a, b, X, **kw):
            ok = z >= 6.0
            bad += int(np.sum((lo[ok] > f[ok]) | (f[ok] > hi[ok])))
        assert bad > 0, kw
