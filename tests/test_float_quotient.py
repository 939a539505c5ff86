"""The FP32 form divides the three velocity components of a substep by the parcel's
shared denominator 1 + f through one double reciprocal:

    a / den  ==  (float)((double)a * RN64(1 / (double)den))

for float a, den whenever the quotient is zero or outside the FP32 subnormal range
(the kernels fall back to the float division there).  These checks exercise the
identity with IEEE arithmetic on the CPU: random pairs over the full exponent range,
denominators of the drag form 1 + f, and quotients placed next to float rounding
midpoints, where a wrong rounding would show first.  The GPU kernels' results are
compared bit for bit with the oracle's float division in the parity tests.
"""
import numpy as np

F32_MIN_NORMAL = np.float32(2.0 ** -126)


def via_reciprocal(a, den):
    r = 1.0 / den.astype(np.float64)
    return (a.astype(np.float64) * r).astype(np.float32)


def check(a, den):
    a = a.astype(np.float32)
    den = den.astype(np.float32)
    with np.errstate(all="ignore"):
        want = a / den
        got = via_reciprocal(a, den)
    normal = np.isfinite(want) & ((np.abs(want) >= F32_MIN_NORMAL) | (want == 0))
    bad = normal & (got.view(np.uint32) != want.view(np.uint32))
    assert not bad.any(), (a[bad][:5], den[bad][:5], got[bad][:5], want[bad][:5])
    return int(normal.sum())


def test_random_pairs_full_exponent_range():
    rng = np.random.default_rng(20)
    n = 0
    for _ in range(8):
        bits = rng.integers(0, 2 ** 32, size=2_000_000, dtype=np.uint64).astype(np.uint32)
        a = bits.view(np.float32)
        den = rng.integers(0, 2 ** 32, size=bits.size, dtype=np.uint64).astype(np.uint32).view(np.float32)
        n += check(a, den)
    assert n > 5_000_000


def test_drag_denominators():
    """den = 1 + f with f over the drag range of the workloads (1e-4 .. 1e6) and
    numerators v + f u of the size of velocities."""
    rng = np.random.default_rng(21)
    f = (10.0 ** rng.uniform(-4, 6, size=4_000_000)).astype(np.float32)
    den = np.float32(1) + f
    a = (rng.standard_normal(f.size) * 10.0 ** rng.uniform(-8, 2, size=f.size)).astype(np.float32)
    assert check(a, den) > 3_900_000


def test_quotients_next_to_midpoints():
    """a = RN32(t * den) for t within a few double ulps of a float midpoint: the exact
    a / den then lies as close to the midpoint as two floats allow."""
    rng = np.random.default_rng(22)
    n = 2_000_000
    den = (rng.uniform(1.0, 2.0, size=n) * 2.0 ** rng.integers(-40, 40, size=n)).astype(np.float32)
    q = (rng.uniform(1.0, 2.0, size=n) * 2.0 ** rng.integers(-40, 40, size=n)).astype(np.float32)
    mid = q.astype(np.float64) + np.spacing(q).astype(np.float64) / 2
    t = mid * (1.0 + rng.integers(-4, 5, size=n) * 2.0 ** -53)
    a = (t * den.astype(np.float64)).astype(np.float32)
    assert check(a, den) > 1_900_000


def test_signed_zeros_and_exact_quotients():
    den = np.array([1.0, 3.0, 1.5, 1e6, 2.0 ** 20, 7.0], dtype=np.float32)
    for a0 in (0.0, -0.0):
        a = np.full(den.size, a0, dtype=np.float32)
        got = via_reciprocal(a, den)
        assert (np.signbit(got) == np.signbit(a / den)).all() and (got == 0).all()
    a = np.array([3.0, 9.0, -4.5, 5e6, 1.0, 49.0], dtype=np.float32)
    np.testing.assert_array_equal(via_reciprocal(a, den), a / den)


def three(a3, den):
    """The kernels' rule for the three components of a substep: the rounded products,
    and the float division for a component whose product is below 2^-126 and whose
    numerator is nonzero, taken only when the smallest product is below 2^-126."""
    r = 1.0 / den.astype(np.float64)
    out = [(a.astype(np.float64) * r).astype(np.float32) for a in a3]
    small = np.minimum(np.minimum(np.abs(out[0]), np.abs(out[1])), np.abs(out[2])) < F32_MIN_NORMAL
    with np.errstate(all="ignore"):
        for k in range(3):
            fix = small & (np.abs(out[k]) < F32_MIN_NORMAL) & (a3[k] != 0)
            out[k] = np.where(fix, a3[k] / den, out[k])
    return out


def check_three(a3, den):
    a3 = [a.astype(np.float32) for a in a3]
    den = den.astype(np.float32)
    got = three(a3, den)
    with np.errstate(all="ignore"):
        for k in range(3):
            want = a3[k] / den
            ok = np.isnan(want) | (got[k].view(np.uint32) == want.view(np.uint32))
            assert ok.all(), (k, a3[k][~ok][:5], den[~ok][:5], got[k][~ok][:5], want[~ok][:5])


def test_three_component_rule_against_ieee_division():
    """Every result equals the IEEE float quotient, subnormal and zero results included,
    with components mixed so the shared branch is taken for some and not others."""
    rng = np.random.default_rng(23)
    n = 2_000_000
    den = (np.float32(1) + (10.0 ** rng.uniform(-4, 6, size=n)).astype(np.float32)).astype(np.float32)
    a3 = []
    for k in range(3):
        kind = rng.integers(0, 5, size=n)
        mag = np.where(kind == 0, 10.0 ** rng.uniform(-45, -36, size=n),       # subnormal-range quotients
              np.where(kind == 1, 10.0 ** rng.uniform(-38, -30, size=n),       # near the normal threshold
              np.where(kind == 2, 0.0, 10.0 ** rng.uniform(-8, 2, size=n))))   # zeros, velocities
        a3.append((mag * rng.choice([-1.0, 1.0], size=n)).astype(np.float32))
    check_three(a3, den)


def test_three_component_rule_at_the_normal_threshold():
    """Numerators placed so the products land within a few ulps of 2^-126, on both sides
    (a product that rounds up to 2^-126 keeps the product path)."""
    rng = np.random.default_rng(24)
    n = 2_000_000
    den = (rng.uniform(1.0, 4.0, size=n)).astype(np.float32)
    t = 2.0 ** -126 * (1.0 + rng.integers(-40, 41, size=n) * 2.0 ** -24)
    a = (t * den.astype(np.float64)).astype(np.float32)
    zero = np.zeros(n, dtype=np.float32)
    check_three([a, zero, -a], den)
    check_three([a, np.full(n, 1.0, np.float32), a[::-1].copy()], den)


def test_three_component_rule_on_exact_subnormal_midpoints():
    """a / den exactly halfway between two subnormals with den = 2 D (D odd > 1): the
    double product misses the midpoint by the rounding of 1 / den and, for these D (found
    by a search over odd D < 400), rounds to the wrong neighbour for many j -- the
    quotients the fallback exists for."""
    j = np.arange(0, 4096, dtype=np.float64)
    wrong = 0
    for D in (49, 75, 77, 91, 93, 99, 117, 123):
        den = np.full(j.size, 2.0 * D, dtype=np.float32)
        a = ((2 * j + 1) * D * 2.0 ** -149).astype(np.float32)
        assert np.all(a.astype(np.float64) == (2 * j + 1) * D * 2.0 ** -149)  # representable
        r = 1.0 / den.astype(np.float64)
        wrong += int(np.sum((a.astype(np.float64) * r).astype(np.float32) != a / den))
        check_three([a, a[::-1].copy(), np.ones(j.size, np.float32)], den)
    assert wrong > 1000  # the product path alone would be wrong on many of them


# ---- FP64: one reciprocal per parcel and an FMA correction -------------------------
# The FP64 kernels divide by 1 + f as q0 = RN(a r), r = RN(1 / den), then
# RN(q0 + RN(a - q0 den) r) with both steps FMAs, falling back to the IEEE division for
# all three components when a product is below 2^-969, zero, above 2^1000 or NaN.
# numpy has no FMA: the FMA is emulated exactly with rationals (one rounding of the
# exact x y + z, which float(Fraction) performs correctly).

from fractions import Fraction
import math


def fma_exact(x, y, z):
    """RN(x y + z) with one rounding; an exact zero takes IEEE's sign (the common sign
    when x y and z are zeros of the same sign, +0 otherwise)."""
    v = Fraction(x) * Fraction(y) + Fraction(z)
    if v != 0:
        return float(v)
    p_neg = (math.copysign(1.0, x) < 0) != (math.copysign(1.0, y) < 0)
    if x * y == 0.0 and z == 0.0 and p_neg == (math.copysign(1.0, z) < 0):
        return math.copysign(0.0, z)
    return 0.0


def three64(a3, den):
    """The kernels' rule: q0 = RN(a r), t = RN(q0 den - a), RN(q0 - t r), unless a
    quotient is nonzero and below 2^-969, or above 2^1000."""
    r = 1.0 / den
    q = [a * r for a in a3]
    tiny = any(qk != 0.0 and abs(qk) < 2.0 ** -969 for qk in q)
    if tiny or max(abs(v) for v in q) > 2.0 ** 1000:
        return [ak / den for ak in a3]
    return [fma_exact(-fma_exact(qk, den, -ak), r, qk) for ak, qk in zip(a3, q)]


def draw64(rng, kind):
    if kind == 0:
        return float(rng.uniform(-1, 1) * 10.0 ** rng.uniform(-320, -285))  # tiny and subnormal
    if kind == 1:
        return 0.0 if rng.random() < 0.5 else -0.0
    if kind == 2:
        return float(rng.uniform(-1, 1) * 10.0 ** rng.uniform(-20, 3))
    return float(rng.uniform(-1, 1) * 10.0 ** rng.uniform(-280, 280))


def near_midpoint64(rng, den):
    """A numerator whose exact quotient lies within an ulp fraction of a rounding midpoint."""
    q = float(rng.uniform(1, 2) * 2.0 ** int(rng.integers(-60, 10)))
    return float(Fraction(q + math.ulp(q) / 2) * Fraction(den))


def test_fp64_rule_against_ieee_division():
    rng = np.random.default_rng(31)
    bad = 0
    for _ in range(20000):
        den = 1.0 + float(10.0 ** rng.uniform(-8, 7))
        a3 = [draw64(rng, int(rng.integers(0, 4))) if rng.random() < 0.6 else near_midpoint64(rng, den)
              for _ in range(3)]
        got = three64(a3, den)
        for ak, gk in zip(a3, got):
            want = ak / den
            bad += np.float64(gk).view(np.uint64) != np.float64(want).view(np.uint64)
    assert bad == 0


def test_fp64_rule_keeps_signed_zeros():
    """Zero numerators stay on the FMA path (no fallback) and keep their sign."""
    for den in (1.0, 1.5, 98.0, 1.0 + 2.0 ** -40, 3.7e5):
        for a in (0.0, -0.0):
            got = three64([a, 1.0, -2.5], den)
            assert math.copysign(1.0, got[0]) == math.copysign(1.0, a) and got[0] == 0.0
            assert got[1] == 1.0 / den and got[2] == -2.5 / den


def test_fp64_correction_is_needed():
    """The uncorrected product RN(a r) misses the IEEE quotient on a large share of
    near-midpoint numerators, so the agreement above pins the FMA step."""
    rng = np.random.default_rng(32)
    wrong = 0
    for _ in range(4000):
        den = 1.0 + float(10.0 ** rng.uniform(-8, 7))
        a = near_midpoint64(rng, den)
        wrong += (a * (1.0 / den)) != a / den
    assert wrong > 400
