"""GPU: contexts filled by qmckl_trexio_read from the reference's TREXIO test files (be2 and
chbrclf, TEXT back end) evaluate the AO and MO VGL of the oracle built from the same data
(read back through the getters), at 1e-12 of scale, in FP64; and the 3xBF16 product of chbrclf
stays within the split-precision bounds of the hand-built chbrclf.
"""
import os
import tarfile

import numpy as np
import pytest

import oracle
import qmckl_amd as q
from qmckl_amd import system as S
from test_gpu_parity import vgl_close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def paths(tmp_path_factory):
    d = tmp_path_factory.mktemp("trexio")
    with tarfile.open(os.path.join(GOLDEN, "trexio", "chbrclf.tar.xz")) as t:
        t.extractall(d, filter="data")
    return {"be2": os.path.join(GOLDEN, "trexio", "be2"), "chbrclf": str(d / "chbrclf")}


@pytest.mark.parametrize("which,W", [("be2", 64), ("chbrclf", 8)])
def test_trexio_orbitals_match_the_oracle(paths, which, W):
    print(f"trexio {which} W={W}", flush=True)
    ctx = q.context_create()
    q.trexio_read(ctx, paths[which])
    s = S.from_context(ctx)
    x = S.electron_coords(np.random.default_rng(5), s, W)
    S.set_coords(ctx, s, x, W)
    P = W * s.elec_num
    ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num).reshape(P, 5, -1)
    ao_ref = oracle.ao_vgl(s, x)
    vgl_close(ao, ao_ref, 1e-12, f"ao_vgl {which}")
    mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * s.mo_num).reshape(P, 5, -1)
    ref = oracle.mo_vgl(s.mo_coef, ao_ref)
    vgl_close(mo, ref, 1e-12, f"mo_vgl {which}")
    q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, q.MO_BF16X3))
    mo3 = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * s.mo_num).reshape(P, 5, -1)
    for c, bound in ((0, 4e-7), (1, 4e-7), (2, 4e-7), (3, 4e-7), (4, 1e-6)):
        err = np.abs(mo3[:, c] - ref[:, c]).max() / np.abs(ref[:, c]).max()
        assert err <= bound, (which, c, err)
    q.context_destroy(ctx)
