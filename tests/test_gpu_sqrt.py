"""The function evaluation's square root (eval_dev.h sqrt_rn) against the
host's correctly rounded one, bit for bit: random inputs over many binades,
exact squares, and inputs within a few ulps of the rounding midpoints, where
the fast acceptance test must hand over to the exact midpoint tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _near_midpoints(rng, n):
    """Doubles m whose root lies near a midpoint y + 2^-53 between two
    doubles y in [1, 2), and their neighbours, scaled by even powers of two."""
    from fractions import Fraction

    out = []
    ys = rng.uniform(1.0, 2.0, n)
    for i, y in enumerate(ys):
        s = 1 if i % 2 else -1
        m = float((Fraction(float(y)) + Fraction(s, 2 ** 53)) ** 2)
        if not (1.0 <= m < 4.0):
            continue
        for x in (m, np.nextafter(m, 0.0), np.nextafter(m, 8.0)):
            out.append(float(x) * 4.0 ** int(rng.integers(-40, 40)))
    return np.array(out, dtype=np.float64)


def test_sqrt_matches_host_bit_for_bit(ctx):
    rng = np.random.default_rng(11)
    parts = [
        np.exp(rng.uniform(-60.0, 60.0, 1_000_000)),
        rng.uniform(0.0, 4.0, 500_000),
        np.square(rng.integers(1, 2 ** 26, 100_000).astype(np.float64)),
        _near_midpoints(rng, 60_000),
        np.array([0.0, 1.0, 4.0, 2.0, np.nextafter(1.0, 2.0), np.nextafter(4.0, 0.0), np.inf, 5e-324, 2.2250738585072014e-308]),
    ]
    x = np.concatenate(parts)
    got = ctx.debug_sqrt(x)
    want = np.sqrt(x)
    bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
    assert bad.size == 0, (x[bad[:5]], got[bad[:5]], want[bad[:5]])
