"""Shapes of the fused AO->MO product (column tiles per pass, waves per workgroup): every
built shape gives the same MO VGL bit for bit (each output element is accumulated over the
same 32-AO chunks by the same matrix instructions, whatever the shape), in FP64 and in
3xBF16, and the FP64 result matches the oracle to 1e-12 of scale per component.

Systems: the medium workload's 75 MOs (5 column tiles: one pass by default) and a molecule of
the batched workload with 300 electrons and 150 MOs (10 column tiles: one 8-wave pass by
default, two 4-wave passes with 8 tiles, three with 4).
"""
import ctypes

import numpy as np
import pytest

import oracle
import qmckl_amd as q
from qmckl_amd import system as S
from test_gpu_parity import vgl_close

pytestmark = pytest.mark.gpu


def _shape(ctx):
    g, w, p = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    q.check(ctx, q.lib().qmckl_amd_get_mo_tile(ctx, ctypes.byref(g), ctypes.byref(w), ctypes.byref(p)))
    return g.value, w.value, p.value


def _mo(ctx, s, W):
    P, nmo = W * s.elec_num, s.mo_num_kept
    return q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo).reshape(P, 5, nmo)


CASES = {
    # workload, walkers, default shape, other shapes (col tiles, waves)
    "medium": (lambda: S.workload("medium", walk_num=2), (5, 4, 1), [(5, 8), (3, 4), (2, 4), (8, 8)]),
    "molecule-300e": (lambda: S.molecule_workload(300, walk_num=4), (10, 8, 1),
                      [(8, 4), (4, 4), (5, 8), (16, 8), (12, 8), (7, 4)]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_product_shapes_agree_bitwise_and_with_oracle(name):
    make, default, others = CASES[name]
    s, x, W = make()
    print(f"{name}: Ne={s.elec_num} Nao={s.ao_num} Nmo={s.mo_num_kept} W={W}", flush=True)
    ctx = q.context_create()
    try:
        S.load(ctx, s)
        S.set_coords(ctx, s, x, W)
        ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
        for mode in (q.MO_FP64, q.MO_BF16X3):
            q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, mode))
            q.check(ctx, q.lib().qmckl_amd_set_mo_tile(ctx, 0, 0))
            assert _shape(ctx) == default
            base = _mo(ctx, s, W)
            if mode == q.MO_FP64:
                vgl_close(base, ref, 1e-12, f"{name} default shape")
            for g, w in others:
                q.check(ctx, q.lib().qmckl_amd_set_mo_tile(ctx, g, w))
                shape = _shape(ctx)
                nct = -(-s.mo_num_kept // 16)
                assert shape[0] == min(g, nct) and shape[2] == -(-nct // shape[0])
                got = _mo(ctx, s, W)
                print(f"  mode {mode} shape {shape}: max |diff| {np.abs(got - base).max():.1e}", flush=True)
                assert np.array_equal(got, base), f"{name} mode {mode} shape {shape} differs from the default"
        q.check(ctx, q.lib().qmckl_amd_set_mo_tile(ctx, 0, 0))
    finally:
        q.context_destroy(ctx)


def test_autotune_picks_built_shapes_and_keeps_results():
    """qmckl_amd_autotune: a built shape per context, the same MO VGL bit for bit as before
    tuning, and the same choice when the buckets are tuned again (kept per process)."""
    mols = [S.molecule_workload(ne, walk_num=16, seed=300 + ne) for ne in (60, 150, 300)]
    ctxs = []
    try:
        before = []
        for s, x, W in mols:
            ctx = q.context_create()
            ctxs.append(ctx)
            S.load(ctx, s)
            S.set_coords(ctx, s, x, W)
            q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, q.MO_BF16X3))
            before.append(_mo(ctx, s, W))
        arr = np.asarray(ctxs, np.int64)
        shapes = np.zeros((len(ctxs), 3), np.int32)
        q.check(ctxs[0], q.lib().qmckl_amd_autotune(arr.ctypes.data, len(ctxs), shapes.ctypes.data))
        for (s, x, W), ctx, sh, ref in zip(mols, ctxs, shapes, before):
            nct = -(-s.mo_num_kept // 16)
            G, nw, passes = (int(v) for v in sh)
            print(f"  {s.name}: {nct} column tiles -> {G} per pass on {nw} waves, {passes} passes", flush=True)
            assert (nw == 4 and 1 <= G <= 8) or (nw == 8 and 5 <= G <= 16)
            assert G <= nct and passes == -(-nct // G) and _shape(ctx) == (G, nw, passes)
            assert np.array_equal(_mo(ctx, s, W), ref), f"{s.name}: tuned shape changes the result"
        again = np.zeros_like(shapes)
        q.check(ctxs[0], q.lib().qmckl_amd_autotune(arr.ctypes.data, len(ctxs), again.ctypes.data))
        assert np.array_equal(again, shapes)
    finally:
        for ctx in ctxs:
            q.context_destroy(ctx)


VALUE_CASES = {
    # workload, shapes (col tiles, waves; 0 = default) to run the values-only product with
    "medium": (lambda: S.workload("medium", walk_num=4), [(0, 0), (8, 8), (2, 4)]),
    "molecule-300e": (lambda: S.molecule_workload(300, walk_num=4), [(0, 0), (4, 4)]),
    "slater": (lambda: S.slater_workload(walk_num=8), [(0, 0), (1, 4)]),
}


@pytest.mark.parametrize("name", list(VALUE_CASES))
def test_values_only_product(name):
    """qmckl_get_mo_basis_mo_value with no current MO VGL runs the values-only product (AO
    values without derivatives, one row tile per column tile): in every precision and shape
    its values equal the value rows of the full MO VGL bit for bit, the device getter gives
    the same, and the FP64 values match the oracle to 1e-12 of scale."""
    make, shapes = VALUE_CASES[name]
    s, x, W = make()
    P, nmo = W * s.elec_num, s.mo_num_kept
    print(f"{name}: Ne={s.elec_num} Nao={s.ao_num} Nmo={nmo} W={W}", flush=True)
    ctx = q.context_create()
    dev = q.DeviceBuffer(8 * P * nmo)
    try:
        S.load(ctx, s)
        ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))[:, 0, :]
        for mode in (q.MO_FP64, q.MO_FP32, q.MO_BF16X3):
            q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, mode))
            for g, w in shapes:
                q.check(ctx, q.lib().qmckl_amd_set_mo_tile(ctx, g, w))
                S.set_coords(ctx, s, x, W)  # no current MO VGL: the values-only product runs
                val = q.get_array(ctx, "qmckl_get_mo_basis_mo_value", P * nmo).reshape(P, nmo)
                S.set_coords(ctx, s, x, W)
                q.check(ctx, q.lib().qmckl_amd_get_mo_basis_mo_value_device(ctx, dev.ptr, P * nmo))
                q.check(ctx, q.lib().qmckl_amd_synchronize(ctx))
                dval = dev.to_array((P, nmo))
                vgl = _mo(ctx, s, W)
                print(f"  mode {mode} shape {_shape(ctx)}: max |value - vgl row| "
                      f"{np.abs(val - vgl[:, 0, :]).max():.1e}", flush=True)
                assert np.array_equal(val, vgl[:, 0, :]), f"{name} mode {mode} shape {(g, w)}"
                assert np.array_equal(dval, val)
                if mode == q.MO_FP64:
                    assert np.abs(val - ref).max() <= 1e-12 * np.abs(ref).max()
        q.check(ctx, q.lib().qmckl_amd_set_mo_tile(ctx, 0, 0))
    finally:
        dev.free()
        q.context_destroy(ctx)
