"""Population storage formats through the device encoders
(lbm_storage_roundtrip), value by value.

F16C (sign, 4 exponent bits with bias 17, 11 fraction bits, subnormals) is
checked against the format's definition: the value of every one of its
32768 non-negative codes is written down from that definition, and each
input must come back as the code value nearest to it, ties to the even
code, saturating at the largest finite value, sign kept.  Inputs are FP32
values (the kernels encode FP32 results), which the kernels hold in units
of 2^-110: below 2^-16 (the format's subnormals) such a value is an FP32
denormal, already rounded to a multiple of 2^-149, so a value within 2^-39
of a tie there takes the tie's rounding (numpy's FP32 product |x| 2^-110
states the same first rounding here).  F16S is
checked against numpy's IEEE half conversion of x * 2^12, F32 against
numpy's float32.  All bars are exact equality.
"""
import numpy as np
import pytest

import lbm_amd

pytestmark = pytest.mark.gpu

BIAS = 17


def f16c_values():
    """Value of code c = e << 11 | m (c < 2^15) by the format's definition."""
    c = np.arange(1 << 15)
    e, m = c >> 11, (c & 0x7FF).astype(np.float64)
    return np.where(e == 0, m / 2048.0 * 2.0 ** (1 - BIAS), (1.0 + m / 2048.0) * 2.0 ** (e - BIAS))


def f16c_nearest(x):
    """Nearest code value to each float32 x, ties to the even code, from
    the FP32 product |x| 2^-110."""
    vals = f16c_values()
    with np.errstate(under="ignore"):
        a = (np.abs(x) * np.float32(2.0 ** -110)).astype(np.float64) * 2.0 ** 110
    hi = np.clip(np.searchsorted(vals, a), 1, vals.size - 1)
    lo = hi - 1
    dlo, dhi = a - vals[lo], vals[hi] - a
    pick = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo % 2 == 0, lo, hi)))
    pick = np.where(a >= vals[-1], vals.size - 1, pick)
    return np.copysign(vals[pick], x.astype(np.float64))


def _inputs(rng):
    vals = f16c_values()
    mids = 0.5 * (vals[1:] + vals[:-1])  # ties: 13 significant bits, exact in FP32
    m32 = mids.astype(np.float32)
    assert np.array_equal(m32.astype(np.float64), mids)
    up = np.nextafter(m32, np.float32(np.inf)).astype(np.float64)
    dn = np.nextafter(m32, np.float32(-np.inf)).astype(np.float64)
    rand = np.exp2(rng.uniform(-32.0, 0.0, 40000)) * rng.choice([-1.0, 1.0], 40000)
    big = np.array([0.4999, 0.5, 0.75, 1.0, 2.0, 1e3, 1e30, -0.5, -1e30])
    small = np.array([0.0, -0.0, 2.0 ** -30, 2.0 ** -28, 3 * 2.0 ** -29, 2.0 ** -27, 2.0 ** -16])
    x = np.concatenate([vals, -vals, mids, -mids, up, dn, rand, big, small])
    return x.astype(np.float32).astype(np.float64)


def test_f16c_matches_format_definition():
    rng = np.random.default_rng(11)
    x = _inputs(rng)
    got = lbm_amd.storage_roundtrip(x, "f16c")
    want = f16c_nearest(x.astype(np.float32))
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches, first x={x[bad[0]]!r} got {got[bad[0]]!r} want {want[bad[0]]!r}"
    # sign of zero kept
    z = lbm_amd.storage_roundtrip(np.array([-0.0]), "f16c")
    assert np.signbit(z[0])


def test_f16c_relative_error_bound():
    """One significant bit more than IEEE half: relative error <= 2^-12 over
    the normal range 2^-16 <= |x| < 0.5, against <= 2^-11 for F16S."""
    rng = np.random.default_rng(12)
    x = np.exp2(rng.uniform(-16.0, -1.0, 100000)) * rng.choice([-1.0, 1.0], 100000)
    x = x.astype(np.float32).astype(np.float64)
    e_c = np.abs(lbm_amd.storage_roundtrip(x, "f16c") - x) / np.abs(x)
    e_s = np.abs(lbm_amd.storage_roundtrip(x, "f16s") - x) / np.abs(x)
    assert e_c.max() <= 2.0 ** -12
    assert e_s.max() <= 2.0 ** -11
    assert e_c.mean() < 0.6 * e_s.mean()


def test_f16s_matches_ieee_half():
    """Inputs are FP32 values (the kernels convert FP32 results), so the
    scaling by 2^12 is exact and the half conversion is the one rounding."""
    rng = np.random.default_rng(13)
    x = np.concatenate([np.exp2(rng.uniform(-40.0, 4.0, 40000)) * rng.choice([-1.0, 1.0], 40000),
                        [0.0, 15.99, 16.0, 20.0, -20.0]])
    x = x.astype(np.float32).astype(np.float64)
    got = lbm_amd.storage_roundtrip(x, "f16s")
    K = 2.0 ** 12
    with np.errstate(over="ignore"):
        want = np.float16(np.float32(x * K)).astype(np.float64) / K
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches: " + ", ".join(
        f"x*K={x[i] * K!r} got*K={got[i] * K!r} want*K={want[i] * K!r}" for i in bad[:6])


def test_f32_and_f64_roundtrip():
    rng = np.random.default_rng(14)
    x = rng.standard_normal(10000) * 10.0 ** rng.uniform(-12, 2, 10000)
    assert np.array_equal(lbm_amd.storage_roundtrip(x, "f32"), np.float32(x).astype(np.float64))
    assert np.array_equal(lbm_amd.storage_roundtrip(x, "f64"), x)
