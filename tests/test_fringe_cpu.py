"""The hex donor search of the oracle (the extension of interpolation_tet to
trilinear hexes, which the reference skips) pinned without the oracle's own
formulas: closed-form local coordinates on axis-aligned grids with a
brute-force first-index donor over every element, affine (rotated) grids
whose local coordinates are known by construction, distorted grids through
points mapped from known local coordinates, and exact interpolation of
linear fields."""
import numpy as np

from oversetdevfort_amd import mesh as M


def _linear_field(x, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=3)
    B = rng.normal(size=(3, 3))
    return a + x @ B, (a, B)


def _corners(m, e):
    return m.xg[m.ien[e]]


def _map(X, xi):
    """x(xi) = sum_a N_a(xi) X_a with the linhexshl corner signs."""
    s = 2 * M.HEX_CORNERS - 1
    N = np.prod(1 + s * xi[None, :], axis=1) / 8
    return N @ X


def test_axis_aligned_donors_match_brute_force(oracle_mod):
    O = oracle_mod
    m = M.structured_hex(4, 3, 2, lo=(0.0, 0.0, 0.0), hi=(1.0, 0.75, 0.5))
    h = 0.25
    rng = np.random.default_rng(11)
    inner = rng.uniform([0, 0, 0], [1.0, 0.75, 0.5], size=(60, 3))
    faces = (np.floor(inner / h) * h)            # points on cell faces and edges
    faces[:, 1] = inner[:, 1]
    outside = np.array([[1.1, 0.2, 0.2], [-0.01, 0.1, 0.1], [0.5, 0.8, 0.1]])
    xr = np.concatenate([inner, m.xg, faces, outside])
    ug, (a, B) = _linear_field(m.xg)
    zero = np.zeros_like(m.xg)
    donor, val = O.interp_hex(m, 1, zero, ug, xr, np.zeros_like(xr))
    for j, p in enumerate(xr):
        want = -1
        for e in range(m.nelem):
            xi = 2.0 * (p - _corners(m, e).min(axis=0)) / h - 1.0
            if np.all(np.abs(xi) <= 1.0 + 1e-10):
                want = e
                break
        assert donor[j] == want, (j, p, donor[j], want)
    inside = donor >= 0
    assert inside.sum() == len(xr) - 3
    np.testing.assert_allclose(val[inside], a + xr[inside] @ B, rtol=0, atol=1e-12)
    assert np.all(val[~inside] == 0.0)


def test_rotated_affine_grid_local_coordinates(oracle_mod):
    O = oracle_mod
    m = M.structured_hex(3, 3, 3, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0))
    th = np.deg2rad(23.0)
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m.xg = m.xg @ R.T + np.array([0.3, -0.2, 0.5])
    rng = np.random.default_rng(4)
    es = rng.integers(0, m.nelem, size=50)
    xis = rng.uniform(-0.95, 0.95, size=(50, 3))
    xr = np.array([_map(_corners(m, e), xi) for e, xi in zip(es, xis)])
    for e, xi, p in zip(es, xis, xr):
        ok, got = O.hex_local(_corners(m, e), p)
        assert ok
        np.testing.assert_allclose(got, xi, rtol=0, atol=1e-13)
    ug, (a, B) = _linear_field(m.xg)
    zero = np.zeros_like(m.xg)
    donor, val = O.interp_hex(m, 1, zero, ug, xr, np.zeros_like(xr))
    assert np.array_equal(donor, es)
    np.testing.assert_allclose(val, a + xr @ B, rtol=0, atol=1e-12)


def test_distorted_grid_and_displacement(oracle_mod):
    """Jittered (non-affine) hexes in a displaced configuration x + d: points
    mapped from known local coordinates are found in their element, and the
    receiver displacement dr moves the point as xr + dr."""
    O = oracle_mod
    m = M.structured_hex(3, 3, 3, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter=0.08,
                         seed=9)
    rng = np.random.default_rng(12)
    dg = 0.01 * rng.normal(size=m.xg.shape)
    es = rng.integers(0, m.nelem, size=40)
    xis = rng.uniform(-0.9, 0.9, size=(40, 3))
    X = m.xg + dg
    p = np.array([_map(X[m.ien[e]], xi) for e, xi in zip(es, xis)])
    dr = 0.002 * rng.normal(size=p.shape)
    xr = p - dr
    ug = rng.normal(size=m.xg.shape)
    donor, val = O.interp_hex(m, 1, dg, ug, xr, dr)
    assert np.array_equal(donor, es)
    s = 2 * M.HEX_CORNERS - 1
    for j, (e, xi) in enumerate(zip(es, xis)):
        N = np.prod(1 + s * xi[None, :], axis=1) / 8
        np.testing.assert_allclose(val[j], N @ ug[m.ien[e]], rtol=0, atol=1e-11)


def test_source_id_selects_elements(oracle_mod):
    O = oracle_mod
    m = M.structured_hex(2, 1, 1, lo=(0.0, 0.0, 0.0), hi=(1.0, 0.5, 0.5))
    m.elm_id = np.array([2, 1], np.int32)      # only the second cell is a source
    xr = np.array([[0.25, 0.25, 0.25], [0.75, 0.25, 0.25], [0.5, 0.25, 0.25]])
    ug = np.ones_like(m.xg)
    donor, _ = O.interp_hex(m, 1, np.zeros_like(m.xg), ug, xr, np.zeros_like(xr))
    assert list(donor) == [-1, 1, 1]
