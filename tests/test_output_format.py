"""The number format of the output files (no GPU needed).

The library's ES14.6 fast path and its general ESw.d path are compared with
an independent restatement of the Fortran edit descriptor on Python's
correctly rounded conversion (tests/fortran_io.py), over values chosen to
break a careless formatter: exact rounding ties (round half to even), values
next to every decade boundary, 3-digit exponents (where Fortran drops the
'E'), subnormals, signed zeros, infinities and NaN, and random bit patterns
over the whole double range.
"""
import math

import numpy as np
import pytest

import libcfd2lcs_amd as L
from fortran_io import es


def _hard_values():
    rng = np.random.default_rng(7)
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 1e-5, 9.9999995e5, 9.9999994999e5, 9.99999950001e5, 12345675.0,
            12345665.0, 1234567.5, 1234568.5, 0.125, 2.5e-7, 5e-324, -5e-324, 2.2250738585072014e-308,
            1.7976931348623157e308, -1.7976931348623157e308, 1e100, 9.9999995e99, 1e-100, 1e99, 1e-99,
            math.pi, -math.e, float("inf"), float("-inf"), float("nan")]
    for e in range(-320, 309):
        p = 10.0 ** e if e > -308 else float(f"1e{e}")
        if p == 0 or math.isinf(p):
            continue
        for q in (p, p * 9.9999995, p * 9.999999499999, p * 9.9999995000001, p * 1.0000005, p * 0.99999995):
            if math.isfinite(q) and q != 0:
                vals += [q, np.nextafter(q, 0.0), np.nextafter(q, np.inf), -q]
    # exact ties at the 7th digit: m.5 * 10^k with the product exact in binary
    for m in rng.integers(1000000, 9999999, 2000):
        vals.append((int(m) * 10 + 5) * 1.0)          # 8-digit integers ending in 5
        vals.append((int(m) * 10 + 5) / 1024.0 * 1024.0)
        vals.append(float(int(m) * 10 + 5) / 16.0)     # exact dyadic fractions
    # the doubles nearest to 7th-digit ties and to decade bounds, at random
    # exponents, and their neighbours
    for m, e in zip(rng.integers(1000000, 9999999, 20000), rng.integers(-300, 300, 20000)):
        for q in (float(f"{int(m)}.5e{int(e)}"), float(f"9999999.5e{int(e)}"), float(f"999999.95e{int(e)}")):
            if math.isfinite(q) and q != 0:
                vals += [q, float(np.nextafter(q, 0.0)), float(np.nextafter(q, np.inf))]
    # random bit patterns (finite), and float32 values as the sp library writes them
    bits = rng.integers(0, 2**63 - 1, 200000, dtype=np.int64)
    d = bits.view(np.float64)
    vals += [float(x) for x in d[np.isfinite(d)]]
    f32 = rng.standard_normal(50000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 50000).astype(
        np.float32)
    vals += [float(x) for x in f32 if np.isfinite(x)]
    vals += list(rng.standard_normal(100000) * 10.0 ** rng.uniform(-12, 12, 100000))
    return np.array(vals, dtype=np.float64)


@pytest.mark.parametrize("w,d", [(14, 6), (18, 4)])
def test_es_matches_fortran_rules(w, d):
    try:
        L.load("dp")
    except RuntimeError as e:
        pytest.skip(str(e))
    v = _hard_values()
    got = L.format_es(v, w, d)
    bad = [(float(x), g, es(x, w, d)) for x, g in zip(v, got) if g != es(x, w, d)]
    assert not bad, f"{len(bad)} mismatches, e.g. {bad[:5]}"


def test_es_fixed_examples():
    try:
        L.load("dp")
    except RuntimeError as e:
        pytest.skip(str(e))
    cases = [
        (1.0, "  1.000000E+00"),
        (-1.0, " -1.000000E+00"),
        (0.0, "  0.000000E+00"),
        (-0.0, " -0.000000E+00"),
        (12345675.0, "  1.234568E+07"),   # tie, rounds to even
        (12345665.0, "  1.234566E+07"),   # tie, rounds to even
        (1.5e100, "  1.500000+100"),      # 3-digit exponent drops the E
        (-2.5e-200, " -2.500000-200"),
        (float("inf"), "      Infinity"),
        (float("-inf"), "     -Infinity"),
        (float("nan"), "           NaN"),
    ]
    got = L.format_es([c[0] for c in cases], 14, 6)
    assert got == [c[1] for c in cases]
    assert L.format_es([1.5], 18, 4) == ["        1.5000E+00"]
