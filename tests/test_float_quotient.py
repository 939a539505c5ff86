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
