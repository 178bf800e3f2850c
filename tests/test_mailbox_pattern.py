"""The persistent substep posts residual partials by value and treats the
all-ones 64-bit pattern as "not posted yet".  The partials are doubles formed
from f32 values (f32 sums widened to double, then added in double), so the
pattern must be one that widening cannot produce and that double addition of
widened values cannot produce: a widened f32 keeps the low 29 significand bits
zero, NaN included, and IEEE addition returns a widened operand's NaN (quieted)
or the default NaN, never one with those bits set."""
import numpy as np

EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)
LOW29 = np.uint64((1 << 29) - 1)


def test_widened_f32_never_sets_the_low_29_bits():
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2**32, size=1_000_000, dtype=np.uint64).astype(np.uint32)
    special = np.array([0xFFFFFFFF, 0x7FFFFFFF, 0xFFC00000, 0x7FC00000, 0x7F800000, 0xFF800000,
                        0x00000001, 0x80000001, 0x7F7FFFFF, 0xFF7FFFFF, 0, 0x80000000],
                       dtype=np.uint32)
    f = np.concatenate([bits, special]).view(np.float32)
    with np.errstate(invalid="ignore"):
        d = f.astype(np.float64).view(np.uint64)
    assert np.all((d & LOW29) == 0)
    assert not np.any(d == EMPTY)


def test_sums_of_widened_values_never_produce_the_empty_pattern():
    # finite sums stay finite or overflow to inf; NaN operands propagate a
    # widened payload (low bits zero) or the default NaN
    rng = np.random.default_rng(1)
    a = rng.standard_normal(100_000).astype(np.float32).astype(np.float64)
    nans = np.array([0xFFFFFFFF, 0x7FFFFFFF, 0x7FC00001], dtype=np.uint32).view(np.float32)
    a[:3] = nans.astype(np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        s = np.cumsum(a)
        t = a + a[::-1]
    for v in (s, t, np.array([np.inf - np.inf, 0.0 * np.inf])):
        u = v.view(np.uint64)
        assert not np.any(u == EMPTY)
        assert np.all((u[np.isnan(v)] & LOW29) == 0)
