"""The captured training set does not depend on the order in which the
device appended its rows: content_order gives the same sequence for any
permutation of the same rows, and a different row set gives a different
one."""
import numpy as np

from code_mpm_amd.surrogate import content_order


def test_content_order_ignores_arrival_order():
    rng = np.random.default_rng(3)
    rows = rng.normal(size=(5000, 14))
    rows[100] = rows[7]                     # a duplicate row
    rows[200, 3] = -0.0                     # signed zeros keep their bits
    want = rows[content_order(rows)]
    for seed in range(3):
        p = np.random.default_rng(seed).permutation(rows.shape[0])
        got = rows[p][content_order(rows[p])]
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_content_order_depends_on_contents():
    rng = np.random.default_rng(4)
    rows = rng.normal(size=(2000, 14))
    other = rows.copy()
    other[5, 13] = np.nextafter(other[5, 13], np.inf)
    a, b = content_order(rows), content_order(other)
    assert sorted(a.tolist()) == list(range(2000)) and sorted(b.tolist()) == list(range(2000))
    assert not np.array_equal(a, b)
