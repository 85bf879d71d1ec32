"""GPU: several molecules evaluated from one recorded HIP graph (qmckl_amd_batch_*).

Every replay is checked against an eager evaluation of the same coordinates by separate
contexts (bit for bit: the same kernels on the same inputs), over several replays with new
coordinates and eager work of the program between them.
"""
import numpy as np
import pytest

import qmckl_amd as q
from qmckl_amd import system as S
from qmckl_amd.batch import Batch

pytestmark = pytest.mark.gpu


def _contexts(mols):
    out = []
    for s, x, W in mols:
        ctx = q.context_create()
        S.load(ctx, s)
        S.set_coords(ctx, s, x, W)
        q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, q.MO_BF16X3))
        out.append(ctx)
    return out


def test_batch_replays_match_eager():
    mols = [S.molecule_workload(ne, walk_num=8, seed=100 + ne) for ne in (10, 37, 61)]
    ctxs = _contexts(mols)
    ref_ctxs = _contexts(mols)
    coord = [q.DeviceBuffer.from_array(x) for _, x, _ in mols]
    mo = [q.DeviceBuffer(8 * W * s.elec_num * 5 * s.mo_num_kept) for s, _, W in mols]
    val = [q.DeviceBuffer(8 * W) for _, _, W in mols]
    gl = [q.DeviceBuffer(8 * W * 4 * s.elec_num) for s, _, W in mols]
    b = Batch(ctxs, coord, mo, val, gl)
    rng = np.random.default_rng(5)
    for rep in range(3):
        xs = [np.ascontiguousarray(x + rng.normal(0.0, 0.05, x.shape)) for _, x, _ in mols]
        for buf, x in zip(coord, xs):
            buf.upload(x)
        q.device_synchronize()
        b.replay()
        b.synchronize()
        for (s, _, W), rc, x, m, v, g in zip(mols, ref_ctxs, xs, mo, val, gl):
            # eager evaluation of the same coordinates (the program's own work between replays)
            S.set_coords(rc, s, x, W)
            Ne, nmo = s.elec_num, s.mo_num_kept
            ref_mo = q.get_array(rc, "qmckl_get_mo_basis_mo_vgl", W * Ne * 5 * nmo)
            ref_v = q.get_array(rc, "qmckl_get_jastrow_champ_value", W)
            ref_gl = q.get_array(rc, "qmckl_get_jastrow_champ_gl", W * 4 * Ne)
            np.testing.assert_array_equal(m.to_array(ref_mo.shape), ref_mo, err_msg=f"replay {rep} {s.name} mo")
            np.testing.assert_array_equal(v.to_array(ref_v.shape), ref_v, err_msg=f"replay {rep} {s.name} value")
            np.testing.assert_array_equal(g.to_array(ref_gl.shape), ref_gl, err_msg=f"replay {rep} {s.name} gl")
    b.destroy()
    # the contexts are usable on their own again
    s, x, W = mols[0]
    S.set_coords(ctxs[0], s, x, W)
    assert np.isfinite(q.get_array(ctxs[0], "qmckl_get_jastrow_champ_value", W)).all()
    for c in ctxs + ref_ctxs:
        q.context_destroy(c)
