"""The device scans beyond their single-pass sizes (rin_debug_scan): the
one-level scan (tile sums added up by each applying workgroup) ends at 4,096
tiles, one top-level workgroup scans up to 32,768 tile sums, and the
single-field scan recurses beyond that; the three-field scan (512-entry
tiles) has the same first two levels and runs three single-field scans
beyond them.  Each size ends in a partial tile, and a device-side count below
the capacity (as the pipeline passes) scans only the counted prefix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _excl(a):
    out = np.zeros(len(a), dtype=np.uint64)
    np.cumsum(a[:-1], dtype=np.uint64, out=out[1:])
    return out


@pytest.mark.parametrize("n", [2048 * 4096 - 5, 2048 * 4096 + 2047, 2048 * 32768 + 3001])
def test_scan_u32_levels(ctx, n):
    rng = np.random.default_rng(n % 1000)
    a = rng.integers(0, 4, n, dtype=np.uint32)  # totals stay below 2^32
    out, tot = ctx.debug_scan(1, 1, a, n)
    want = _excl(a)
    np.testing.assert_array_equal(out[0], want.astype(np.uint32))
    assert int(tot[0]) == int(want[-1]) + int(a[-1])


@pytest.mark.parametrize("n,stride", [(512 * 4096 + 77, 3), (512 * 20000 + 5, 3), (512 * 32768 + 9, 1)])
def test_scan3_levels(ctx, n, stride):
    rng = np.random.default_rng(n % 997)
    f3 = rng.integers(0, 9, (3, n), dtype=np.uint32)
    data = f3.reshape(-1) if stride == 1 else np.ascontiguousarray(f3.T).reshape(-1)
    out, tot = ctx.debug_scan(3, stride, data, n)
    for f in range(3):
        want = _excl(f3[f])
        np.testing.assert_array_equal(out[f], want.astype(np.uint32))
        assert int(tot[f]) == int(want[-1]) + int(f3[f][-1])


@pytest.mark.parametrize("fields", [1, 3])
def test_scan_device_count(ctx, fields):
    """A device-side count below the capacity: the prefix is scanned and the
    totals cover only it."""
    n, count = 512 * 5000 + 3, 512 * 4500 + 1
    rng = np.random.default_rng(7)
    f = rng.integers(0, 5, (fields, n), dtype=np.uint32)
    stride = 1 if fields == 1 else 3
    data = f.reshape(-1) if fields == 1 else np.ascontiguousarray(f.T).reshape(-1)
    out, tot = ctx.debug_scan(fields, stride, data, n, count)
    for q in range(fields):
        want = _excl(f[q][:count])
        np.testing.assert_array_equal(out[q][:count], want.astype(np.uint32))
        assert int(tot[q]) == int(f[q][:count].sum())
