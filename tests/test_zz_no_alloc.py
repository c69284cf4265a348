"""ih_dot_noalloc (no device allocation, a fixed scratch array per GPU) gives
ih_dot's result bit for bit, on two calls in a row, the second on a smaller
box, so the second call reads only partials it wrote itself."""
import numpy as np
import pytest

import impact_hip as ih

pytestmark = pytest.mark.gpu


def _case(N, seed):
    c = ih.Case(N, (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    rng = np.random.default_rng(seed)
    c.upload(ih.TMP1, np.ascontiguousarray(rng.standard_normal(c.shape)))
    c.upload(ih.TMP2, np.ascontiguousarray(rng.standard_normal(c.shape)))
    return c


def test_dot_noalloc_equals_dot_twice_shrinking():
    big = _case((257, 129, 129), 11)
    small = _case((33, 17, 17), 12)
    for c, lo, hi in ((big, big.Sp, big.Np), (small, small.Sp, small.Np), (small, (2, 3, 4), (20, 9, 11))):
        print(f"no-alloc dot: grid {c.N}, box {tuple(lo)}..{tuple(hi)}", flush=True)
        want = c.dot(lo, hi, ih.TMP1, ih.TMP2)
        got = c.dot_noalloc(lo, hi, ih.TMP1, ih.TMP2)
        assert got == want, (got, want)
        assert got != 0.0
    big.close()
    small.close()
