"""The FP32 erfcx bounds of the IRT's SamplePDC second decision (irt.hip: erfcx_bounds)
against the FP64 erfcx the exact path uses (include/dna_irt_math.h: dna_erfcx, the oracle's
restatement of the worker's erfcx, irt-worker.js:58-88).

The device bounds are e (1 -/+ 1e-5) with e the FP64 routine's own Chebyshev series
evaluated in FP32 below z = 6, and the alternating asymptotic series
sqrt(pi) z erfcx(z) = 1 - 1/(2z^2) + 3/(4z^4) - 15/(8z^6) + 105/(16z^8) with the first
omitted term 945/(32 z^10) as its remainder bound from z = 6 on.  Both are restated here in
numpy (float32 arithmetic operation by operation, as the device's with -ffp-contract=off)
and checked on a dense grid: the FP32 series stays within 1e-6 of the FP64 routine (a tenth
of the margin; the device's expf adds at most a few ulp), and the asymptotic interval holds
the FP64 routine's value.  A dropped or mistyped coefficient or a wrong sign in the series
moves the result far outside these bounds."""
import ctypes as C

import numpy as np

import oracle as O

CHEB = [-1.3026537197817094, 6.4196979235649026e-1, 1.9476473204185836e-2,
        -9.561514786808631e-3, -9.46595344482036e-4, 3.66839497852761e-4,
        4.2523324806907e-5, -2.0278578112534e-5, -1.624290004647e-6,
        1.303655835580e-6, 1.5626441722e-8, -8.5238095915e-8, 6.529054439e-9,
        5.059343495e-9, -9.91364156e-10, -2.27365122e-10, 9.6467911e-11,
        2.394038e-12, -6.886027e-12, 8.94487e-13, 3.13092e-13, -1.12708e-13,
        3.81e-16, 7.106e-15, -1.523e-15, -9.4e-17, 1.21e-16, -2.8e-17]
F = np.float32


def _oracle_erfcx(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    dp = C.POINTER(C.c_double)
    O.lib().oracle_fmath64(15, x.ctypes.data_as(dp), y.ctypes.data_as(dp), C.c_int64(x.size))
    return y


def _cheb32(z):
    c = [F(v) for v in CHEB]
    z = z.astype(F)
    t = (F(2) / (F(2) + z)).astype(F)
    ty = (F(4) * t - F(2)).astype(F)
    d = np.zeros_like(z)
    dd = np.zeros_like(z)
    for j in range(27, 0, -1):
        tmp = d
        d = ((ty * d).astype(F) - dd).astype(F)
        d = (d + c[j]).astype(F)
        dd = tmp
    arg = ((F(0.5) * (c[0] + (ty * d).astype(F)).astype(F)).astype(F) - dd).astype(F)
    return (t * np.exp(arg)).astype(F)


def _asym_bounds(z):
    z = z.astype(F)
    r = (F(1) / (z * z).astype(F)).astype(F)
    inner = (F(1.875) - (r * F(6.5625)).astype(F)).astype(F)
    inner = (F(0.75) - (r * inner).astype(F)).astype(F)
    inner = (F(0.5) - (r * inner).astype(F)).astype(F)
    sm = (F(1) - (r * inner).astype(F)).astype(F)
    rem = (F(29.53125) * r * r * r * r * r).astype(F)
    q = (F(0.56418958) / z).astype(F)
    return ((sm - rem) * q * F(0.99999)).astype(np.float64), ((sm + rem) * q * F(1.00001)).astype(np.float64)


def test_fp32_chebyshev_within_a_tenth_of_the_margin():
    z = np.linspace(0.0, 6.0, 400_001, dtype=np.float32)[:-1]
    e = _cheb32(z).astype(np.float64)
    ref = _oracle_erfcx(z.astype(np.float64))
    rel = np.abs(e - ref) / ref
    assert rel.max() < 1e-6, (rel.max(), z[rel.argmax()])
    # the margin is real: the bounds hold the FP64 value
    assert np.all(e * 0.99999 <= ref) and np.all(ref <= e * 1.00001)


def test_asymptotic_interval_holds_the_fp64_value():
    z = np.geomspace(6.0, 2.0e4, 200_000).astype(np.float32)
    lo, hi = _asym_bounds(z)
    ref = _oracle_erfcx(z.astype(np.float64))
    assert np.all(lo <= ref) and np.all(ref <= hi)
    # and it is tight: under 1e-4 wide relative to the value everywhere past z = 6
    assert np.max((hi - lo) / ref) < 1e-4


def test_bounds_are_not_vacuous_where_the_first_bracket_is_loose():
    """At z ~ 0.5-3 the Abramowitz & Stegun 7.1.13 bracket of the first decision is several
    percent wide; the second decision's interval is below 3e-5 there."""
    z = np.linspace(0.5, 3.0, 1001, dtype=np.float32)
    e = _cheb32(z).astype(np.float64)
    ref = _oracle_erfcx(z.astype(np.float64))
    lo_as = 2 / np.sqrt(np.pi) / (z + np.sqrt(z.astype(np.float64) ** 2 + 2))
    hi_as = 2 / np.sqrt(np.pi) / (z + np.sqrt(z.astype(np.float64) ** 2 + 4 / np.pi))
    assert np.all(lo_as < ref) and np.all(ref <= hi_as)
    assert np.min((hi_as - lo_as) / ref) > 1e-2
    assert np.max(2.1e-5 * e / ref) < 3e-5
