"""GPU: several molecules evaluated from one recorded HIP graph (qmckl_amd_batch_*).

* Every replay is checked against an eager evaluation of the same coordinates by separate
  contexts (bit for bit: the same kernels on the same inputs), over several replays with new
  coordinates and eager work of the program between them.
* Molecules of 10 to 400 electrons (both e-e-n kernels: the LDS-resident one and, above
  about 240 electrons, the pass-based one) replayed in 3xBF16 mode against the CPU oracle:
  MO VGL within 1.5e-7 of max |component| (1e-6 on the Laplacian), Jastrow factors 1e-12,
  GL 1e-11 of scale.
* The full batched workload (64 molecules of 10..400 electrons, 256 walkers each) as a
  property check: finite MO VGL, no NaN, finite GL of e^J where e^J is below 1e250, two
  replays of the same coordinates bit-identical.
* A replay is refused once a context changed after the recording.
"""
import numpy as np
import pytest

import oracle
import qmckl_amd as q
from qmckl_amd import system as S
from qmckl_amd.batch import Batch
from test_gpu_configs import SPLIT_BOUND, assert_split_contract

# the molecules reach 1.0e-7 on a gradient component: the FP32 sums of the leading bf16
# product inside the matrix instructions (about 2^-24 of the 32-term chunk sums) set that
# floor; 1.5e-7 leaves room for the chunk composition
SPLIT_BOUND_MOLECULES = (1.5e-7,) * 4 + SPLIT_BOUND[4:]
from test_gpu_parity import _compare_jastrow

pytestmark = pytest.mark.gpu


def _contexts(mols, mode=q.MO_BF16X3):
    out = []
    for s, x, W in mols:
        ctx = q.context_create()
        S.load(ctx, s)
        S.set_coords(ctx, s, x, W)
        q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, mode))
        out.append(ctx)
    return out


def _buffers(mols):
    coord = [q.DeviceBuffer.from_array(x) for _, x, _ in mols]
    mo = [q.DeviceBuffer(8 * W * s.elec_num * 5 * s.mo_num_kept) for s, _, W in mols]
    val = [q.DeviceBuffer(8 * W) for _, _, W in mols]
    gl = [q.DeviceBuffer(8 * W * 4 * s.elec_num) for s, _, W in mols]
    return coord, mo, val, gl


def test_batch_of_values_alone():
    """A batch recorded for the Jastrow values alone (no GL buffers, no MO) on fresh contexts,
    whose value getters would run the values-only evaluation: the recording evaluates the
    whole Jastrow as its graph does, and the replayed values equal the eager values bit for
    bit on two replays."""
    mols = [S.molecule_workload(ne, walk_num=8, seed=200 + ne) for ne in (12, 45)]
    ctxs = _contexts(mols)
    ref_ctxs = _contexts(mols)
    coord, _, val, _ = _buffers(mols)
    b = Batch(ctxs, coord, None, val, None)
    try:
        for rep in range(2):
            xs = [np.ascontiguousarray(x + 0.03 * rep) for _, x, _ in mols]
            for buf, x in zip(coord, xs):
                buf.upload(x)
            q.device_synchronize()
            b.replay()
            b.synchronize()
            for (s, _, W), rc, x, v in zip(mols, ref_ctxs, xs, val):
                S.set_coords(rc, s, x, W)
                ref_v = q.get_array(rc, "qmckl_get_jastrow_champ_value", W)
                assert np.array_equal(v.to_array((W,)), ref_v), f"replay {rep} {s.name}"
    finally:
        b.destroy()
        for ctx in ctxs + ref_ctxs:
            q.context_destroy(ctx)
        for buf in coord + val:
            buf.free()


def test_batch_replays_match_eager():
    mols = [S.molecule_workload(ne, walk_num=8, seed=100 + ne) for ne in (10, 37, 61)]
    ctxs = _contexts(mols)
    ref_ctxs = _contexts(mols)
    coord, mo, val, gl = _buffers(mols)
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


def test_batch_molecules_vs_oracle():
    """Molecules of 10..400 electrons replayed from one graph (3xBF16 MO, FP64 Jastrow)
    against the oracle on the same coordinates."""
    mols = [S.molecule_workload(ne, walk_num=2, seed=300 + ne) for ne in (10, 61, 150, 247, 313, 400)]
    ctxs = _contexts(mols)
    coord, mo, val, gl = _buffers(mols)
    b = Batch(ctxs, coord, mo, val, gl)
    b.replay()
    b.synchronize()
    outs = [(m.to_array((W * s.elec_num, 5, s.mo_num_kept)), v.to_array(W), g.to_array((W, 4, s.elec_num)))
            for (s, _, W), m, v, g in zip(mols, mo, val, gl)]
    b.destroy()
    for (s, x, W), ctx, (m, v, g) in zip(mols, ctxs, outs):
        print(f"{s.name}: Nn={s.nucl_num} Nao={s.ao_num} Nmo={s.mo_num_kept}", flush=True)
        ref_mo = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
        assert_split_contract(m, ref_mo, f"{s.name} batch bf16x3", SPLIT_BOUND_MOLECULES)
        ref = oracle.jastrow(s, x, W)
        # factors and factor GLs from the same context (its device results are the replay's)
        got = {
            "value": v, "gl": g,
            **{f"factor_{k}": q.get_array(ctx, f"qmckl_get_jastrow_champ_factor_{k}", W) for k in ("ee", "en", "een")},
            **{f"factor_{k}_gl": q.get_array(ctx, f"qmckl_get_jastrow_champ_factor_{k}_gl", W * 4 * s.elec_num)
               .reshape(W, 4, s.elec_num) for k in ("ee", "en", "een")},
        }
        _compare_jastrow(got, ref, f"{s.name} batch")
    for c in ctxs:
        q.context_destroy(c)


@pytest.mark.timeout(900)
def test_full_batched_workload_properties():
    """64 molecules of 10..400 electrons, 256 walkers each (the batched benchmark): finite MO
    VGL, e^J never NaN and its GL finite wherever e^J < 1e250, and two replays of the same
    coordinates give bit-identical outputs (compared through 64-bit checksums)."""
    mols = S.batched_workload(64, 256)
    ctxs = _contexts(mols)
    coord, mo, val, gl = _buffers(mols)
    b = Batch(ctxs, coord, mo, val, gl)

    def sums():
        out = []
        for (s, _, W), m, v, g in zip(mols, mo, val, gl):
            a = m.to_array(W * s.elec_num * 5 * s.mo_num_kept)
            assert np.isfinite(a).all(), s.name
            vv = v.to_array(W)
            gg = g.to_array((W, 4, s.elec_num))
            # e^J overflows for the larger molecules (J of several hundred with the recipe's
            # parameters); below 1e250 its GL is finite
            assert not np.isnan(vv).any() and np.isfinite(gg[vv < 1e250]).all(), s.name
            out.append(tuple(int(z.view(np.uint64).sum(dtype=np.uint64)) for z in (a, vv, gg.ravel())))
        return out

    b.replay()
    b.synchronize()
    first = sums()
    b.replay()
    b.synchronize()
    assert sums() == first
    b.destroy()
    print(f"{len(mols)} molecules, electrons {min(s.elec_num for s, _, _ in mols)}.."
          f"{max(s.elec_num for s, _, _ in mols)}", flush=True)
    for c in ctxs:
        q.context_destroy(c)


def test_replay_refused_after_a_context_changed():
    """Changing a context's walker count or MO precision after recording makes the replay
    return QMCKL_FAILURE without launching the graph."""
    mols = [S.molecule_workload(ne, walk_num=4, seed=200 + ne) for ne in (12, 20)]
    ctxs = _contexts(mols)
    coord, mo, val, gl = _buffers(mols)
    b = Batch(ctxs, coord, mo, val, gl)
    b.replay()
    b.synchronize()
    L = q.lib()
    q.check(ctxs[1], L.qmckl_amd_set_mo_precision(ctxs[1], q.MO_FP64))
    with pytest.raises(q.QmcklError) as e:
        b.replay()
    assert e.value.code == 101
    q.check(ctxs[1], L.qmckl_amd_set_mo_precision(ctxs[1], q.MO_BF16X3))
    b.replay()  # back to the recorded state
    b.synchronize()
    s, x, W = mols[0]
    x2 = np.ascontiguousarray(x.reshape(3, W, s.elec_num)[:, :2].reshape(3, 2 * s.elec_num))
    S.set_coords(ctxs[0], s, x2, 2)
    with pytest.raises(q.QmcklError) as e:
        b.replay()
    assert e.value.code == 101
    b.destroy()
    for c in ctxs:
        q.context_destroy(c)
