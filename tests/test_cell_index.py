"""The reflection pass splits a 0-based cell id into (x, y) with a float32 reciprocal of the
grid width and one fix-up step (cell_xy in k_hourly.hip).  numpy float32 arithmetic is the
same IEEE single rounding (conversion, correctly rounded reciprocal, product, truncation), so
this checks the claim the library's scene bound rests on: exact on every grid of at most
2^22 rows, and not beyond it."""
import numpy as np
import pytest

F32 = np.float32
MAX_ROWS = 1 << 22


def cell_xy(ids, w):
    inv_w = F32(1.0) / F32(w)
    q = (ids.astype(F32) * inv_w).astype(np.int64)      # truncation towards zero, as (int)
    r = ids - q * w
    q = np.where(r < 0, q - 1, np.where(r >= w, q + 1, q))
    r = np.where(r < 0, r + w, np.where(r >= w, r - w, r))
    return r, q


def _edge_ids(w, rows):
    """The first and last cell of every row and the cell before each row start: where a
    truncated estimate is most often one row off."""
    y = np.arange(rows, dtype=np.int64)
    ids = np.concatenate([y * w, y * w + (w - 1), np.maximum(y * w - 1, 0)])
    return ids[ids < w * rows]


@pytest.mark.parametrize("w", [1, 2, 3, 5, 7, 8, 100, 255])
def test_exact_on_every_row_at_the_row_bound(w):
    rows = MAX_ROWS
    assert w * (rows + 2) <= 1 << 30          # a supported grid (32-bit DDA byte offsets)
    ids = _edge_ids(w, rows)
    x, y = cell_xy(ids, w)
    np.testing.assert_array_equal(x, ids % w)
    np.testing.assert_array_equal(y, ids // w)


@pytest.mark.parametrize("w,h", [(8192, 8192), (32766, 32766), (16384, 4096), (4096, 65536), (1000, 1000)])
def test_exact_on_wide_grids(w, h):
    rng = np.random.default_rng(w + h)
    ids = np.concatenate([_edge_ids(w, h), rng.integers(0, w * h, 2_000_000)])
    x, y = cell_xy(ids, w)
    np.testing.assert_array_equal(x, ids % w)
    np.testing.assert_array_equal(y, ids // w)


def test_one_fix_up_is_not_enough_far_beyond_the_bound():
    """A 3-wide grid of 3.5e8 rows (allowed by the byte-offset bound alone) gets rows wrong,
    which is why scenes of more than 2^22 rows are refused."""
    w, rows = 3, 350_000_000
    ids = np.random.default_rng(1).integers(0, w * rows, 200_000)
    x, y = cell_xy(ids, w)
    assert (y != ids // w).any()
