"""ofd_fringe_interp_noalloc -- the donor search and interpolation of
ofd_fringe_interp (interpolation_tet, interpolation.f90:5-76, and its hex
extension) with its scratch in a static device array instead of memory the
library allocates -- called through the C ABI and compared bit for bit with
ofd_fringe_interp on the same context: donors and values.

Two calls in a row, the second with a smaller input: the full overset pair
(96x48x48 background + rotated 32^3 component, 253,952 hexes, near the
array's limit) with its 6,146 component receivers, then the 1,543 background
fringe receivers, then a small tet mesh in a second context.  A search that
needs more scratch than the array holds is refused with OFD_ERR_ARG."""
import ctypes

import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M

pytestmark = pytest.mark.gpu


def _soa(a):
    return np.ascontiguousarray(np.asarray(a, np.float64).T)


def _noalloc(ctx, src_id, xr, dr):
    lib = L.load()
    xr_, dr_ = _soa(xr), _soa(dr)
    n = xr_.shape[1]
    donor = np.full(n, -7, np.int32)
    val = np.full((3, n), np.nan)
    rc = lib.ofd_fringe_interp_noalloc(ctx.h, int(src_id), n, xr_.ctypes.data,
                                       dr_.ctypes.data, donor.ctypes.data, val.ctypes.data)
    return rc, donor, val.T.copy()


def _same(ctx, src_id, xr, dr, label):
    rc, d1, v1 = _noalloc(ctx, src_id, xr, dr)
    L.check(rc)
    d0, v0 = ctx.interpolation(src_id, xr, dr)
    found = int((d0 > 0).sum())
    print(f"no-alloc fringe {label}: {len(xr)} receivers, {found} donors", flush=True)
    assert found > 0
    assert np.array_equal(d1, d0)
    assert np.array_equal(v1, v0)


def test_noalloc_fringe_matches_allocating_entry():
    from oversetdevfort_amd import solver as S
    prob = C.overset()
    m = prob.mesh
    rng = np.random.default_rng(3)
    ug = rng.normal(size=(m.nnode, 3))
    print(f"no-alloc fringe: overset pair, {m.nelem} hexes", flush=True)
    ctx = S.Context(0)
    ctx.upload_mesh(m)
    ctx.set_field("UG", ug)
    comp, bsrc = prob.fringe["component"]
    back, csrc = prob.fringe["background"]
    xr = m.xg[comp]
    _same(ctx, bsrc, xr, np.zeros_like(xr), "component (call 1)")
    xb = m.xg[back]
    _same(ctx, csrc, xb, 1e-4 * rng.normal(size=xb.shape), "background (call 2, smaller)")
    # more receivers than the array holds beside this mesh: refused
    big = np.zeros((200_000, 3))
    rc, _, _ = _noalloc(ctx, bsrc, big, big)
    assert rc == L.OFD_ERR_ARG
    assert b"scratch" in L.load().ofd_last_error()
    ctx.close()
    # a smaller mesh in a second context on the same GPU (tets)
    hm = M.structured_hex(4, 4, 3, hi=(1.0, 1.0, 0.75), jitter=0.1, seed=2)
    tm = M.hex_to_tets(hm)
    tm.elm_id = rng.integers(1, 3, size=tm.nelem).astype(np.int32)
    ctx = S.Context(0)
    ctx.upload_mesh(tm)
    ctx.set_field("UG", rng.normal(size=(tm.nnode, 3)))
    xt = rng.uniform([-0.05, -0.05, -0.05], [1.05, 1.05, 0.8], size=(120, 3))
    _same(ctx, 1, xt, np.zeros_like(xt), "tets (call 3, smaller)")
    ctx.close()
