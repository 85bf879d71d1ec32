"""GPU: qmckl_amd_get_jastrow_champ_device_noalloc, the Jastrow evaluation that allocates
nothing (its intermediates in the library's static device scratch), against the device
getters it mirrors (qmckl_amd_get_jastrow_champ_value_device / _gl_device): bit for bit, on
two calls in a row, the second with fewer walkers, for the LDS-resident e-e-n path and the
pass-based one; sizes beyond the scratch and missing device tables are refused before any
launch."""
import numpy as np
import pytest

import qmckl_amd as q
from qmckl_amd import system as S

pytestmark = pytest.mark.gpu


def _reference(s, x, W):
    """value and GL through the device getters of a fresh context."""
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    L = q.lib()
    Ne = s.elec_num
    xd = q.DeviceBuffer.from_array(x)
    val = q.DeviceBuffer(8 * W)
    gl = q.DeviceBuffer(8 * W * 4 * Ne)
    q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
    q.check(ctx, L.qmckl_amd_get_jastrow_champ_value_device(ctx, val.ptr, W))
    q.check(ctx, L.qmckl_amd_get_jastrow_champ_gl_device(ctx, gl.ptr, W * 4 * Ne))
    q.synchronize(ctx)
    out = val.to_array(W), gl.to_array((W, 4, Ne))
    q.context_destroy(ctx)
    return out


@pytest.mark.parametrize("case", ["medium", "pass_based"])
def test_noalloc_matches_device_getters(case):
    if case == "medium":
        s, x, W = S.workload("medium", walk_num=256)
        W2 = 96
    else:
        s, x, W = S.jastrow_workload(20, 200, 200, 5, 5, 6, walk_num=4)
        W2 = 3
    Ne = s.elec_num
    L = q.lib()
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    xd = q.DeviceBuffer.from_array(x)
    val = q.DeviceBuffer(8 * W)
    gl = q.DeviceBuffer(8 * W * 4 * Ne)
    # the device tables are not uploaded yet: refused, nothing launched
    assert L.qmckl_amd_get_jastrow_champ_device_noalloc(ctx, W, xd.ptr, x.size, val.ptr, gl.ptr, W * 4 * Ne) == \
        q.QMCKL_NOT_PROVIDED
    q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)  # uploads the tables
    for w in (W, W2):
        xw = np.ascontiguousarray(x.reshape(3, W, Ne)[:, :w].reshape(3, w * Ne))
        xd.upload(xw)
        q.device_synchronize()
        print(f"{case}: noalloc call with {w} walkers", flush=True)
        q.check(ctx, L.qmckl_amd_get_jastrow_champ_device_noalloc(ctx, w, xd.ptr, xw.size, val.ptr, gl.ptr,
                                                                   w * 4 * Ne))
        q.synchronize(ctx)
        ref_v, ref_gl = _reference(s, xw, w)
        np.testing.assert_array_equal(val.to_array(w), ref_v)
        np.testing.assert_array_equal(gl.to_array((w, 4, Ne)), ref_gl)
    # the context's own results are untouched (its coordinates are still the W walkers')
    np.testing.assert_array_equal(q.get_array(ctx, "qmckl_get_jastrow_champ_value", W), _reference(s, x, W)[0])
    q.context_destroy(ctx)


def test_noalloc_refuses_sizes_beyond_the_scratch():
    s, x, W = S.workload("medium", walk_num=8)
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
    L = q.lib()
    big = 8192  # 8192 walkers of 150 electrons need more than 64 MiB of intermediates
    n = 3 * big * s.elec_num
    xd = q.DeviceBuffer(8)
    # sizes are checked before any pointer is used: the small buffers are never touched
    assert L.qmckl_amd_get_jastrow_champ_device_noalloc(ctx, big, xd.ptr, n, xd.ptr, xd.ptr,
                                                       4 * big * s.elec_num) == 2
    assert "64 MiB" in q.last_error(ctx)
    q.context_destroy(ctx)
