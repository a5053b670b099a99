"""Pool/window streaming of the oracle vs plain pull streaming.

ASL streams by sliding one window per direction over a pool and copying the
window back when it would leave the pool (src/num/aslLBGK.cxx:129-185,237-250).
With the relaxation rate 0 (nu = inf gives w = 1/(3 nu + 1/2) = 0, so
fNew = fOld exactly) one step must be the pull f_i(j) <- f_i(j - c_i.T) on the
linear index, bitwise, across the window copies.  Stale pool memory is NaN in
the oracle; a cell that reads it turns NaN in every direction (rho is NaN),
so such cells are left out and the count of compared values is checked.
"""
import numpy as np
import pytest

from oracle import OracleLattice, template


@pytest.mark.parametrize("q,size", [(19, (8, 10, 12)), (15, (5, 7, 6)), (27, (6, 5, 7))])
def test_one_step_is_linear_pull_across_copies(q, size):
    L = OracleLattice(size, q, fp64=True, nu=float("inf"))
    S = tuple(s + 2 for s in size)
    n = int(np.prod(S))
    rng = np.random.default_rng(1)
    f0 = rng.uniform(0.5, 1.5, size=(q,) + S)
    L.write_f(f0)
    c, _, _ = template(q)
    T = np.array([S[1] * S[2], S[2], 1])
    off = c @ T
    inc = -off
    max_inc = inc.max()
    pool = 2 * n + 4 * max_inc
    steps = int(3 * pool / max_inc)  # several copies of the fastest windows
    # walls on every face keep stale (NaN) ghost data away from the interior;
    # snapshots are taken after the BC fill, which is what the next pull reads
    b = L.add_slice_bc(L.NOSLIP, 63)
    L.execute_bc(b)
    prev = L.f().reshape(q, n)
    compared = 0
    interior = np.zeros(S, dtype=bool)
    interior[1:-1, 1:-1, 1:-1] = True
    interior = interior.reshape(n)
    for _ in range(steps):
        L.execute()
        cur = L.f().reshape(q, n)
        assert not np.isnan(cur[:, interior]).any()
        L.execute_bc(b)
        after_bc = L.f().reshape(q, n)
        ok_cell = ~np.isnan(cur).any(axis=0)
        for i in range(q):
            j = np.arange(n)
            src = j - off[i]
            m = (src >= 0) & (src < n) & ok_cell
            m &= ~np.isnan(prev[i, np.clip(src, 0, n - 1)])
            assert np.array_equal(cur[i, m], prev[i, src[m]]), "direction %d" % i
            compared += int(m.sum())
        prev = after_bc
    assert compared > steps * q * interior.sum()  # every interior value, every step


def test_copies_happen():
    # the test above only means something if the fastest window is copied
    size = (8, 10, 12)
    S = tuple(s + 2 for s in size)
    n = int(np.prod(S))
    c, _, _ = template(19)
    inc = -(c @ np.array([S[1] * S[2], S[2], 1]))
    pool = 2 * n + 4 * inc.max()
    first_copy = (pool - n) // inc.max()
    assert first_copy < int(3 * pool / inc.max())


def test_rest_at_rest_is_exact():
    # uniform rest state: every interior cell stays exactly at f_i = 1
    L = OracleLattice((6, 6, 6), 19, fp64=True, nu=0.1)
    L.init_f(1.0, (0.0, 0.0, 0.0))
    b = L.add_slice_bc(L.NOSLIP, 63)
    L.execute_bc(b)
    L.run(50, [b])
    f = L.f()[:, 1:-1, 1:-1, 1:-1]
    r = L.rho()[1:-1, 1:-1, 1:-1]
    u = L.velocity()[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(f - 1.0).max() < 1e-15
    assert np.abs(r - 1.0).max() < 1e-15
    assert np.abs(u).max() < 1e-16
