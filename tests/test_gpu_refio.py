"""Restart files through a GPU context (writeSol / readSol,
iosol.f90:122-194,377-454): a weak-wall cavity written to the reference's
part file, read back into a problem, stepped once on the GPU; the restart
written from that context read into a fresh one gives the same old fields bit
for bit, with writeSol's choice of the old temperature.
"""
import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M
from oversetdevfort_amd import refio as R

pytestmark = pytest.mark.gpu


def test_restart_round_trip_through_contexts(tmp_path):
    from oversetdevfort_amd import solver as S
    base = C.weak_cavity(4, seed=3)
    m = base.mesh
    bound = R.bound_from_faces(m, 5, *base.weak_faces)
    R.write_part(tmp_path / R.part_name(1), m, None, [bound])
    prob = C.from_part_file(str(tmp_path / R.part_name(1)), base.params, weak_face_ids=(5,))
    assert np.array_equal(prob.mesh.ien, m.ien) and np.array_equal(prob.mesh.xg, m.xg)
    assert np.array_equal(prob.weak_faces[0], base.weak_faces[0])
    prob.ibc, prob.bcval, prob.fields = base.ibc, base.bcval, base.fields
    ctx = S.setup(prob)
    S.begin_step(ctx)
    S.newton_step(ctx, rtol=1e-10, precision=L.OFD_FP64, max_cycles=200)
    S.write_restart(ctx, tmp_path / R.restart_name(1), 1, 0.01)
    fresh = S.setup(prob)
    step, time = S.read_restart(fresh, tmp_path / R.restart_name(1))
    assert (step, time) == (1, 0.01)
    for k, f in (("UG", "UGOLD"), ("ACG", "ACGOLD"), ("PG", "PGOLD"), ("DG", "DGOLD"),
                 ("TGOLD", "TGOLD")):
        assert np.array_equal(fresh.get_field(f), ctx.get_field(k)), k
    assert np.abs(ctx.get_field("UG")).max() > 0.1
    ctx.close()
    fresh.close()
