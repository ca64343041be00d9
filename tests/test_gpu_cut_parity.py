"""Device per-tet cut (rin_debug_cuts) against the oracle's cut, word for word
on the serialized records: random planes/materials and degenerate integer
configurations (zeros at corners, coincident and duplicated planes)."""
import numpy as np
import pytest

import support as S

pytestmark = pytest.mark.gpu


def _cases(rng, k, n, degenerate):
    out = []
    for _ in range(n):
        if degenerate:
            rows = rng.integers(-2, 3, size=(k, 4)).astype(np.float64)
            if k >= 2 and rng.random() < 0.3:
                rows[-1] = rows[0] * rng.choice([1.0, 2.0, -1.0])
        else:
            rows = rng.standard_normal((k, 4))
        out.append(rows)
    return out


# mode 0: IA thread-per-tet cut, 1: MI thread-per-tet cut, 2: IA wave-cooperative cut
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("degenerate", [False, True])
def test_cut_records_match_oracle(ctx, mode, degenerate):
    mi = 1 if mode == 1 else 0
    rng = np.random.default_rng(100 + 10 * mi + degenerate)
    cases = []
    for k in range(1, 9 if mode == 2 else 7):
        cases += _cases(rng, k, 120 if k < 5 else 40, degenerate)
    dev = ctx.debug_cuts(mode, cases)
    checked = 0
    for i, rows in enumerate(cases):
        try:
            ref = S.oracle_cut(mi, rows)
        except RuntimeError:
            continue
        assert not isinstance(dev[i], int), ("device cut failed", i, rows.tolist(), dev[i])
        assert np.array_equal(dev[i], ref), ("record mismatch", i, rows.tolist())
        checked += 1
    assert checked > 0.9 * len(cases)
