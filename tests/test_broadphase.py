"""Broad phase: the listed pairs are exactly the brute-force set within the list margin, on each backend.

The set the reference's grid must find (GridTest :25-44 checks it through positions; here it is checked directly):
every pair o > a with |xa - xo|^2 < ((ra + ro) * 1.02)^2 in f32, and nothing else.
"""

import numpy as np
import pytest

import scene as sc

pkg = sc.pkg
BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def brute_force(x, y, z, r):
    m = np.float32(1) + np.float32(0.02)
    want = set()
    for i in range(len(x)):
        d0, d1, d2_ = x[i] - x, y[i] - y, z[i] - z
        d2 = d0 * d0 + (d1 * d1 + d2_ * d2_)
        near = (r[i] + r) * m
        for j in np.nonzero(d2 < near * near)[0]:
            if j > i:
                want.add((i, int(j)))
    return want


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("seed", [5, 23])
def test_listed_pairs_are_the_brute_force_set(backend, seed):
    s = sc.crowd(seed)
    s.gy = 0
    s.substeps = 1
    s.solve = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
    x, y, z, r = s.x.copy(), s.y.copy(), s.z.copy(), s.r.copy()
    s.start(backend).advance(1)
    cnt = s.ints("contactCount")
    listed = int(cnt[0])
    a = s.ints("contactA")[:listed]
    b = s.ints("contactB")[:listed]
    got = list(zip(a.tolist(), b.tolist()))
    s.close()
    want = brute_force(x, y, z, r)
    assert len(got) == len(set(got)), "a pair listed twice"
    assert set(got) == want
    # Rows by lower sphere, each in the 27-cell walk's order: a is non-decreasing.
    assert np.all(np.diff(a) >= 0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_empty_and_single_sphere_lists(backend):
    s = sc.Scene(1).uniform(0.1, 1)
    s.solve = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
    s.x[0] = s.y[0] = s.z[0] = 0.5
    s.start(backend).advance(2)
    assert int(s.ints("contactCount")[0]) == 0
    s.close()
