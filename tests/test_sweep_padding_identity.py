"""The grouped-row sweep (csrc/precond.hip, sweep_stage_row) stores -0.0 in
a row's empty product slots and adds every slot, and +0.0 where the form
subtracts.  That keeps each row's sum bitwise equal to the sum over its
entries only, because s + (-0.0) and s - (+0.0) are s in IEEE-754 binary64
round-to-nearest for every s, signed zeros, subnormals, infinities and the
largest finite values included."""
import numpy as np


def _specials():
    f = np.finfo(np.float64)
    v = [0.0, -0.0, 1.0, -1.0, f.tiny, -f.tiny, f.smallest_subnormal, -f.smallest_subnormal,
         f.max, -f.max, np.inf, -np.inf, f.eps, 1.0 + f.eps, -3.5e-300]
    rng = np.random.default_rng(11)
    r = rng.standard_normal(2000) * 10.0 ** rng.integers(-300, 300, 2000)
    return np.concatenate([np.array(v), r])


def test_negative_zero_is_the_additive_identity():
    s = _specials()
    out = s + np.float64(-0.0)
    assert out.tobytes() == s.tobytes()


def test_positive_zero_is_the_subtractive_identity():
    s = _specials()
    out = s - np.float64(0.0)
    assert out.tobytes() == s.tobytes()


def test_the_other_zeros_are_not():
    # the choice matters: +0.0 added, or -0.0 subtracted, turns -0.0 into +0.0
    assert np.signbit(np.float64(-0.0) + np.float64(0.0)) == False  # noqa: E712
    assert np.signbit(np.float64(-0.0) - np.float64(-0.0)) == False  # noqa: E712
