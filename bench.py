"""Walker-evaluation throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload medium] [--walkers N]
                    [--mode bf16x3|fp32|fp64] [--no-cpu-baseline]

One step is one walker evaluation of every walker of the batch through the C-ABI, with
the inputs already in device memory:
    qmckl_amd_set_electron_coord_device    (qmckl_set_electron_coord)
    qmckl_amd_get_mo_basis_mo_vgl_device   (AO VGL -> MO VGL)
    qmckl_amd_get_jastrow_champ_value_device
    qmckl_amd_get_jastrow_champ_gl_device
The step is timed for the FP64 product (the reference's arithmetic) and for the
split-precision AO->MO product (--mode, default bf16x3); `value` is the split-precision
throughput.  The MO error of the split product against the FP64 one is reported per VGL
component, together with per-kernel times measured with HIP events on the library's
stream, the roofline position of the dominant kernel and the CPU oracle's throughput on a
bounded sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SLOTS = ["ao_vgl", "mo_gemm", "jastrow_pair", "jastrow_een", "jastrow_total"]
PEAK = {"hbm_GBs": 8000.0, "fp64_TFs": 78.6, "fp32_TFs": 157.3, "bf16_TFs": 2500.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def per_walker_costs(s):
    """Algorithmic work per walker: AO bytes = 8 (5 Ne Nao + 3 Ne); MO flops = 2*5*Ne*Nao*Nmo;
    e-e-n flops = 2 Ne^2 Nn * 5 T with T = c(c-1)/2 (value + 4 derivative components)."""
    Ne, Nao, Nmo, Nn, c = s.elec_num, s.ao_num, s.mo_num_kept, s.nucl_num, s.cord
    T = c * (c - 1) // 2
    return {
        "ao_bytes": 8.0 * (5 * Ne * Nao + 3 * Ne),
        "mo_flops": 2.0 * 5 * Ne * Nao * Nmo,
        "een_flops": 2.0 * Ne * Ne * Nn * 5 * T,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="medium")
    ap.add_argument("--walkers", type=int, default=None)
    ap.add_argument("--mode", default="bf16x3", choices=["bf16x3", "fp32", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--derivatives", action="store_true",
                    help="instead of the walker evaluation, time the Jastrow parameter derivatives and "
                         "nuclear forces (all of their getters) on the 14-electron system, 4096 walkers")
    ap.add_argument("--batched", action="store_true",
                    help="instead of one molecule, evaluate 64 molecules of 10..400 electrons (256 walkers "
                         "each) from one recorded HIP graph, against the same steps launched eagerly")
    ap.add_argument("--molecules", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched workload: keep the default product shapes")
    ap.add_argument("--copy-mb", type=float, default=0.0,
                    help="copy this many MiB device-to-device once before timing (a copy kernel of "
                         "known size, the scale for HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: walkers are independent and this build measures one GPU's "
              f"share (--gpus {args.gpus} is not supported).", file=sys.stderr)
        sys.exit(2)

    import qmckl_amd as q
    from qmckl_amd import system as S

    L = q.lib()
    if args.derivatives:
        bench_derivatives(args, q, S, L)
        return
    if args.batched:
        bench_batched(args, q, S, L)
        return
    log(f"building workload {args.workload}")
    s, x, W = S.workload(args.workload, walk_num=args.walkers)
    Ne, Nmo = s.elec_num, s.mo_num_kept
    P = W * Ne
    log(f"Nn={s.nucl_num} Ne={Ne} Nao={s.ao_num} Nmo={Nmo} W={W} precision={s.precision}")
    ctx = q.context_create()
    S.load(ctx, s)
    S.set_coords(ctx, s, x, W)
    n_mo, n_gl = P * 5 * Nmo, W * 4 * Ne
    xd = q.DeviceBuffer.from_array(x)
    mo = q.DeviceBuffer(8 * n_mo)
    val = q.DeviceBuffer(8 * W)
    gl = q.DeviceBuffer(8 * n_gl)
    if args.copy_mb > 0:
        nb = int(args.copy_mb * 2**20)
        src, dst = q.DeviceBuffer(nb), q.DeviceBuffer(nb)
        dst.copy_from(src)
        q.device_synchronize()
        log(f"copied {nb} bytes device-to-device")
        src.free()
        dst.free()
    q.device_synchronize()

    def step():
        q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
        q.check(ctx, L.qmckl_amd_get_mo_basis_mo_vgl_device(ctx, mo.ptr, n_mo))
        q.check(ctx, L.qmckl_amd_get_jastrow_champ_value_device(ctx, val.ptr, W))
        q.check(ctx, L.qmckl_amd_get_jastrow_champ_gl_device(ctx, gl.ptr, n_gl))

    def timed(mode):
        q.check(ctx, L.qmckl_amd_set_mo_precision(ctx, mode))
        for _ in range(args.warmup):
            step()
        q.synchronize(ctx)
        q.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        q.synchronize(ctx)
        q.device_synchronize()
        return (time.perf_counter() - t0) / args.steps

    def jastrow_sum():
        parts = [q.get_array(ctx, f"qmckl_get_jastrow_champ_factor_{k}", W) for k in ("ee", "en", "een")]
        return parts[0] + parts[1] + parts[2]

    def kinetic_surrogate():
        """-1/2 sum_i (lap_i J + |grad_i J|^2) per walker: the part of the Jastrow-only local
        energy surrogate E_L = -1/2 sum_i (lap J + |grad J|^2) + V_ee + V_en + V_nn that the
        evaluation computes (the potentials depend on the coordinates only)."""
        g = sum(q.get_array(ctx, f"qmckl_get_jastrow_champ_factor_{k}_gl", n_gl).reshape(W, 4, Ne)
                for k in ("ee", "en", "een"))
        return -0.5 * (g[:, 3] + g[:, 0] ** 2 + g[:, 1] ** 2 + g[:, 2] ** 2).sum(axis=1)

    def kernel_times(mode, n=3, fn=step):
        q.check(ctx, L.qmckl_amd_set_mo_precision(ctx, mode))
        fn()
        q.synchronize(ctx)
        ms = np.zeros(5)
        cnt = np.zeros(5, np.int64)
        q.check(ctx, L.qmckl_amd_set_profiling(ctx, 1))
        q.check(ctx, L.qmckl_amd_get_kernel_times(ctx, q.ptr(ms), q.ptr(cnt), 5))
        for _ in range(n):
            fn()
        q.check(ctx, L.qmckl_amd_get_kernel_times(ctx, q.ptr(ms), q.ptr(cnt), 5))
        q.check(ctx, L.qmckl_amd_set_profiling(ctx, 0))
        return {SLOTS[i]: {"ms_per_step": ms[i] / n, "launches_per_step": int(cnt[i]) // n} for i in range(5)}

    modes = {"fp64": q.MO_FP64, "fp32": q.MO_FP32, "bf16x3": q.MO_BF16X3}
    log("timing FP64 path")
    t64 = timed(q.MO_FP64)
    mo64 = mo.to_array((P, 5, Nmo))
    J64 = jastrow_sum()
    K64 = kinetic_surrogate()
    log(f"FP64: {t64 * 1e3:.3f} ms/step")
    k64 = kernel_times(q.MO_FP64)
    log(f"timing {args.mode} path")
    tsp = timed(modes[args.mode])
    log(f"{args.mode}: {tsp * 1e3:.3f} ms/step")
    ksp = kernel_times(modes[args.mode])
    q.synchronize(ctx)
    # qmckl_get_mo_basis_mo_value's own path: the values-only product (no AO derivatives)
    mv = q.DeviceBuffer(8 * P * Nmo)

    def value_step():
        q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
        q.check(ctx, L.qmckl_amd_get_mo_basis_mo_value_device(ctx, mv.ptr, P * Nmo))

    kval = kernel_times(modes[args.mode], fn=value_step)["mo_gemm"]
    mv.free()
    log(f"values-only MO product ({args.mode}): {kval['ms_per_step']:.3f} ms")

    # qmckl_get_jastrow_champ_value's own path: the values-only e-e-n evaluation (after the
    # first call, which follows GL requests and so evaluates everything)
    def jvalue_step():
        q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
        q.check(ctx, L.qmckl_amd_get_jastrow_champ_value_device(ctx, val.ptr, W))

    kjv = kernel_times(modes[args.mode], fn=jvalue_step)["jastrow_een"]
    log(f"values-only Jastrow: {kjv['ms_per_step']:.3f} ms")
    # MO error of the split product per VGL component: max |split - FP64| / max |FP64| over the
    # component (the contract: <= 1e-7 value and gradient, <= 1e-6 Laplacian), and the
    # largest per-entry relative error |split - FP64| / |FP64| over the entries with
    # |FP64| >= 1e-3 max |FP64| of the component
    a = mo.to_array((P, 5, Nmo))
    mo_err, mo_rel = [], []
    for c in range(5):
        ref_c = mo64[:, c]
        scale = np.abs(ref_c).max()
        d = np.abs(a[:, c] - ref_c)
        mo_err.append(float(d.max() / scale))
        big = np.abs(ref_c) >= 1e-3 * scale
        mo_rel.append(float((d[big] / np.abs(ref_c[big])).max()))
        log(f"MO error component {c}: {mo_err[-1]:.3e} of max, {mo_rel[-1]:.3e} per entry")
    del a, mo64
    # e^J and its GL as the step left them: at the medium workload's recipe parameters J is
    # several hundred, so e^J reaches ~1e300 and some GL entries overflow (the reference's
    # arithmetic gives the same); parity is asserted on J and the factor GLs
    ev, eg = val.to_array((W,)), gl.to_array((W, 4, Ne))
    J_now = jastrow_sum()
    expj = {"J_min": float(J_now.min()), "J_max": float(J_now.max()),
            "value_nonfinite": int((~np.isfinite(ev)).sum()), "gl_nonfinite": int((~np.isfinite(eg)).sum()),
            "gl_entries": int(eg.size),
            "note": "e^J = exp(J) overflows FP64 range near J = 709; the workload recipe's random "
                    "parameters put J at several hundred, so some GL entries of e^J are inf"}
    del ev, eg
    dJ = float(np.abs(J_now - J64).max())
    dEL = float(np.abs(kinetic_surrogate() - K64).max())
    ao_density = ao_nonzero_density(q, S, s, x, W)

    costs = per_walker_costs(s)
    # dominant kernel of the split path and its roofline position
    dom = max(ksp, key=lambda k: ksp[k]["ms_per_step"])
    t_dom = ksp[dom]["ms_per_step"] / max(1, ksp[dom]["launches_per_step"]) * 1e-3
    if dom == "jastrow_een":
        work = costs["een_flops"] * W
        roof = {"bound": "mfma", "achieved": work / t_dom / 1e12, "peak": PEAK["fp64_TFs"], "unit": "TFLOP/s"}
        roof["note"] = ("FP64 FMA work of the e-e-n contraction priced against the 78.6 TFLOP/s FP64 peak; "
                        "where the LDS-resident kernel applies the launch also computes the e-e and e-n "
                        "terms (their few flops not counted)")
    elif dom == "mo_gemm":
        work = costs["mo_flops"] * W / max(1, ksp[dom]["launches_per_step"])
        peak = {"bf16x3": PEAK["bf16_TFs"] / 6, "fp32": PEAK["fp32_TFs"], "fp64": PEAK["fp64_TFs"]}[args.mode]
        roof = {"bound": "mfma", "achieved": work / t_dom / 1e12, "peak": peak, "unit": "TFLOP/s"}
    else:
        work = costs["ao_bytes"] * W / max(1, ksp[dom]["launches_per_step"])
        roof = {"bound": "hbm", "achieved": work / t_dom / 1e9, "peak": PEAK["hbm_GBs"], "unit": "GB/s"}
    roof["kernel"] = dom
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = measured_traffic(args.workload, W, args.mode, dom)
    if roof["traffic"] is not None:
        roof["traffic_source"] = ("profiles/hbm_traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                  "passes of this kernel at this workload, each scaled by its reading for "
                                  "a copy of known size (not measured in this run)")

    # The drop-in boundary with host buffers: the same step through the reference's
    # getters (coordinates from and results to host memory, over PCIe).  Reported beside
    # `value`, never as it.
    q.check(ctx, L.qmckl_amd_set_mo_precision(ctx, modes[args.mode]))
    host_mo = np.empty(n_mo)
    host_val = np.empty(W)
    host_gl = np.empty(n_gl)

    def host_step():
        q.check(ctx, L.qmckl_set_electron_coord(ctx, b"T", W, q.ptr(x), x.size))
        q.check(ctx, L.qmckl_get_mo_basis_mo_vgl(ctx, q.ptr(host_mo), n_mo))
        q.check(ctx, L.qmckl_get_jastrow_champ_value(ctx, q.ptr(host_val), W))
        q.check(ctx, L.qmckl_get_jastrow_champ_gl(ctx, q.ptr(host_gl), n_gl))

    host_step()
    t0 = time.perf_counter()
    nh = max(1, min(args.steps, 3))
    for _ in range(nh):
        host_step()
    th = (time.perf_counter() - t0) / nh
    log(f"host-buffer boundary ({args.mode}): {th * 1e3:.3f} ms/step")
    del host_mo

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(s, x, W, args.cpu_seconds)

    out = {
        "metric": "walker-evals/sec (AO+MO+Jastrow VGL), split-prec vs FP64, at stated MO error",
        "value": W / tsp,
        "unit": "walker-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tsp * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 (AO, Jastrow) + {args.mode} (MO product)",
        "data": "synthetic (seeded)",
        "config": {
            "workload": f"{args.workload}: {s.nucl_num} nuclei, {Ne} electrons, {s.ao_num} AOs, "
                        f"{Nmo} occupied MOs, {W} walkers, Jastrow 5/5/5, precision {s.precision}",
            "walkers": W, "elec_num": Ne, "ao_num": s.ao_num, "mo_num": Nmo, "nucl_num": s.nucl_num,
            "mo_precision": args.mode,
        },
        "fp64_path": {"value": W / t64, "ms_per_step": t64 * 1e3, "kernels": k64,
                      "mo_product": mo_rate(costs, W, k64, "fp64")},
        "split_path": {"kernels": ksp, "speedup_vs_fp64": t64 / tsp,
                       "mo_product": mo_rate(costs, W, ksp, args.mode)},
        "mo_values_only": {"ms_per_launch": kval["ms_per_step"] / max(1, kval["launches_per_step"]),
                           "note": "qmckl_get_mo_basis_mo_value without a current MO VGL: AO values "
                                   "without derivatives, a fifth of the product's matrix work"},
        "jastrow_values_only": {"een_ms_per_launch": kjv["ms_per_step"] / max(1, kjv["launches_per_step"]),
                                "note": "qmckl_get_jastrow_champ_value alone: the value component of the "
                                        "e-e-n products only, no GL (LDS-resident and balanced pass-based "
                                        "kernels; else the full evaluation)"},
        "mo_err_vs_fp64_max_normalised": dict(zip(["value", "dx", "dy", "dz", "lap"], mo_err)),
        "mo_rel_err_vs_fp64_above_1e-3_of_max": dict(zip(["value", "dx", "dy", "dz", "lap"], mo_rel)),
        "jastrow_max_abs_dJ_vs_fp64": dJ,
        "local_energy_surrogate_max_abs_dE_vs_fp64": dEL,
        "ao_nonzero_density": ao_density,
        "exp_J": expj,
        "host_buffer_path": {"value": W / th, "ms_per_step": th * 1e3,
                             "note": "same step through the host-pointer getters (PCIe-inclusive)"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    q.context_destroy(ctx)
    print(json.dumps(out))


DERIV_GETTERS = [
    ("factor_ee_pderiv", "qmckl_get_jastrow_champ_factor_ee_pderiv"),
    ("factor_ee_gl_pderiv", "qmckl_get_jastrow_champ_factor_ee_gl_pderiv"),
    ("factor_en_pderiv", "qmckl_get_jastrow_champ_factor_en_pderiv"),
    ("factor_en_gl_pderiv", "qmckl_get_jastrow_champ_factor_en_gl_pderiv"),
    ("factor_een_pderiv", "qmckl_get_jastrow_champ_factor_een_pderiv"),
    ("factor_een_gl_pderiv", "qmckl_get_jastrow_champ_factor_een_gl_pderiv"),
    ("forces_jastrow_en", "qmckl_get_forces_jastrow_en"),
    ("forces_jastrow_en_g", "qmckl_get_forces_jastrow_en_g"),
    ("forces_jastrow_en_l", "qmckl_get_forces_jastrow_en_l"),
    ("forces_jastrow_een", "qmckl_get_forces_jastrow_een"),
    ("forces_jastrow_een_g", "qmckl_get_forces_jastrow_een_g"),
    ("forces_jastrow_een_l", "qmckl_get_forces_jastrow_een_l"),
]


def bench_derivatives(args, q, S, L):
    """One step: new coordinates, then every parameter-derivative and force getter (results
    to host buffers) for all walkers; the 14-electron system, 4096 walkers by default."""
    W = args.walkers or 4096
    s, x, W = S.workload("small", walk_num=W)
    Ne, Nn = s.elec_num, s.nucl_num
    ctx = q.context_create()
    S.load(ctx, s, mo=False)
    S.set_coords(ctx, s, x, W)
    d = S.dim_c_vector(s.cord)
    T = s.type_nucl_num
    sizes = {
        "factor_ee_pderiv": s.bord + 1, "factor_ee_gl_pderiv": (s.bord + 1) * Ne * 4,
        "factor_en_pderiv": T * (s.aord + 1), "factor_en_gl_pderiv": T * (s.aord + 1) * Ne * 4,
        "factor_een_pderiv": T * d, "factor_een_gl_pderiv": T * d * Ne * 4,
        "forces_jastrow_en": W * Nn * 3, "forces_jastrow_en_g": W * Nn * 3 * Ne * 3,
        "forces_jastrow_en_l": W * Nn * 3, "forces_jastrow_een": W * Nn * 3,
        "forces_jastrow_een_g": W * 3 * Nn * 3 * Ne, "forces_jastrow_een_l": W * 3 * Nn,
    }
    bufs = {k: np.empty(n) for k, n in sizes.items()}
    xs = [x, np.ascontiguousarray(x + 1e-3)]

    def step(it):
        xx = xs[it % 2]
        q.check(ctx, L.qmckl_set_electron_coord(ctx, b"T", W, q.ptr(xx), xx.size))
        per = {}
        for key, fn in DERIV_GETTERS:
            t0 = time.perf_counter()
            q.check(ctx, getattr(L, fn)(ctx, q.ptr(bufs[key]), bufs[key].size))
            per[key] = time.perf_counter() - t0
        return per

    for it in range(args.warmup):
        step(it)
    totals = {k: 0.0 for k, _ in DERIV_GETTERS}
    t0 = time.perf_counter()
    for it in range(args.steps):
        for k, v in step(it).items():
            totals[k] += v
    t = (time.perf_counter() - t0) / args.steps

    # the adjoint: L = sum_w [ln Psi_J + E_L surrogate] and its gradient in a, b, c and the
    # nuclear coordinates, one call per step after new coordinates (forward + adjoint kernels)
    from qmckl_amd.autograd import jastrow_adjoint

    def adj_step(it):
        xx = xs[it % 2]
        q.check(ctx, L.qmckl_set_electron_coord(ctx, b"T", W, q.ptr(xx), xx.size))
        return jastrow_adjoint(ctx, 1.0, 1.0)

    for it in range(args.warmup):
        adj_step(it)
    t0 = time.perf_counter()
    for it in range(args.steps):
        adj_step(it)
    ta = (time.perf_counter() - t0) / args.steps
    log(f"adjoint: {ta * 1e3:.3f} ms/step")
    q.context_destroy(ctx)
    out = {
        "metric": "walker-derivative-evals/sec (Jastrow a/b/c parameter derivatives with GL, and forces with GL)",
        "value": W / t, "unit": "walker-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": f"small: 2 nuclei, {Ne} electrons, Jastrow 5/5/5, {W} walkers, all 12 derivative "
                               f"and force getters per step (host buffers)", "walkers": W},
        "getter_ms_per_step": {k: v / args.steps * 1e3 for k, v in totals.items()},
        "adjoint": {"value": W / ta, "ms_per_step": ta * 1e3,
                    "what": "L = sum_w [ln Psi_J + E_L surrogate] and dL/d(a, b, c, R) by the adjoint kernel "
                            "(qmckl_amd_get_jastrow_champ_adjoint: forward pass, adjoint, walker sum, host copy)"},
    }
    print(json.dumps(out))


def bench_batched(args, q, S, L):
    """One step: the device step (coordinates in, MO VGL, Jastrow value and GL out) of every
    molecule of the batch, as one graph replay.  Timed for the split-precision product
    (--mode) and for the FP64 one (a second recording), with the same steps launched one call
    at a time beside them; the split product's MO error against the FP64 one per VGL
    component (max over the molecules of max |split - FP64| / max |FP64|)."""
    from qmckl_amd.batch import Batch

    W = args.walkers or 256
    log(f"building {args.molecules} molecules")
    mols = S.batched_workload(args.molecules, W)
    mode = {"fp64": q.MO_FP64, "fp32": q.MO_FP32, "bf16x3": q.MO_BF16X3}[args.mode]
    ctxs, coord, mo, val, gl = [], [], [], [], []
    for s, x, _ in mols:
        ctx = q.context_create()
        S.load(ctx, s)
        S.set_coords(ctx, s, x, W)
        ctxs.append(ctx)
        coord.append(q.DeviceBuffer.from_array(x))
        mo.append(q.DeviceBuffer(8 * W * s.elec_num * 5 * s.mo_num_kept))
        val.append(q.DeviceBuffer(8 * W))
        gl.append(q.DeviceBuffer(8 * W * 4 * s.elec_num))
    q.device_synchronize()
    n_walk = W * len(mols)
    ne = [s.elec_num for s, _, _ in mols]
    nao = [s.ao_num for s, _, _ in mols]

    ctx_arr = np.asarray(ctxs, np.int64)
    tuned = {}

    def set_mode(m):
        for ctx in ctxs:
            q.check(ctx, L.qmckl_amd_set_mo_precision(ctx, m))
        if not args.no_autotune:  # product shapes per size bucket, timed once per process
            shapes = np.zeros((len(ctxs), 3), np.int32)
            t0 = time.perf_counter()
            q.check(ctxs[0], L.qmckl_amd_autotune(ctx_arr.ctypes.data, len(ctxs), shapes.ctypes.data))
            q.device_synchronize()
            tuned[m] = {"seconds": time.perf_counter() - t0,
                        "shapes": sorted({(int(-(-s.mo_num_kept // 16)), *map(int, sh))
                                          for (s, _, _), sh in zip(mols, shapes)})}

    def eager_step():
        for (s, _, _), ctx, cb, m, v, g in zip(mols, ctxs, coord, mo, val, gl):
            Ne = s.elec_num
            q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, cb.ptr, 3 * W * Ne))
            q.check(ctx, L.qmckl_amd_get_mo_basis_mo_vgl_device(ctx, m.ptr, W * Ne * 5 * s.mo_num_kept))
            q.check(ctx, L.qmckl_amd_get_jastrow_champ_value_device(ctx, v.ptr, W))
            q.check(ctx, L.qmckl_amd_get_jastrow_champ_gl_device(ctx, g.ptr, W * 4 * Ne))

    def timed(fn, sync):
        for _ in range(args.warmup):
            fn()
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        sync()
        return (time.perf_counter() - t0) / args.steps

    def graph_time(m):
        set_mode(m)
        b = Batch(ctxs, coord, mo, val, gl)
        t = timed(b.replay, b.synchronize)
        b.destroy()
        return t

    log("timing FP64 graph")
    t64 = graph_time(q.MO_FP64)
    log(f"FP64 graph: {t64 * 1e3:.3f} ms/step")
    mo64 = [m.to_array(W * s.elec_num * 5 * s.mo_num_kept).reshape(-1, 5, s.mo_num_kept)
            for (s, _, _), m in zip(mols, mo)]
    set_mode(mode)
    log("timing eager launches")
    t_eager = timed(eager_step, q.device_synchronize)
    log(f"eager: {t_eager * 1e3:.3f} ms/step")
    log(f"timing {args.mode} graph")
    t_graph = graph_time(mode)
    log(f"{args.mode} graph: {t_graph * 1e3:.3f} ms/step")
    err = np.zeros(5)
    for (s, _, _), m, ref in zip(mols, mo, mo64):
        a = m.to_array(W * s.elec_num * 5 * s.mo_num_kept).reshape(-1, 5, s.mo_num_kept)
        for c in range(5):
            err[c] = max(err[c], float(np.abs(a[:, c] - ref[:, c]).max() / np.abs(ref[:, c]).max()))
    del mo64
    for c in ctxs:
        q.context_destroy(c)
    out = {
        "metric": "walker-evals/sec (AO+MO+Jastrow VGL), batched uneven molecules in one HIP graph",
        "value": n_walk / t_graph, "unit": "walker-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_graph * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"f64 (AO, Jastrow) + {args.mode} (MO product)", "data": "synthetic (seeded)",
        "config": {"workload": f"{len(mols)} molecules, electrons {min(ne)}..{max(ne)} (mean {np.mean(ne):.0f}), "
                               f"AOs {min(nao)}..{max(nao)}, {W} walkers each, Jastrow 5/5/5",
                   "walkers": n_walk, "molecules": len(mols), "mo_precision": args.mode},
        "fp64_graph": {"value": n_walk / t64, "ms_per_step": t64 * 1e3},
        "speedup_vs_fp64": t64 / t_graph,
        "mo_err_vs_fp64_max_normalised": dict(zip(["value", "dx", "dy", "dz", "lap"], err.tolist())),
        "eager_path": {"value": n_walk / t_eager, "ms_per_step": t_eager * 1e3},
        "graph_speedup_vs_eager": t_eager / t_graph,
        "autotune": None if args.no_autotune else {
            {q.MO_FP64: "fp64", q.MO_FP32: "fp32", q.MO_BF16X3: "bf16x3"}[m]: {
                "seconds": round(v["seconds"], 3),
                "shapes (column tiles: per pass, waves, passes)": [list(x) for x in v["shapes"]]}
            for m, v in tuned.items()},
    }
    print(json.dumps(out))


def measured_traffic(workload, W, mode, kernel):
    """HBM bytes per launch of `kernel`, from the committed FETCH_SIZE/WRITE_SIZE counter
    summary (profiles/hbm_traffic.json, each counter scaled by its reading for a copy of
    known size), when it covers this workload; else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("walkers") != W:
        return None
    k = t.get("kernels", {}).get(f"{kernel}:{mode}") or t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


def mo_rate(costs, W, kernels, mode):
    """Rate of the fused AO->MO kernel in dense-equivalent flops, F_MO = 2*5*Ne*Nao*Nmo per
    walker, and for 3xBF16 in bf16 matrix flops (6 cross products per dense flop), each against
    the matching MI355X peak.  The kernel skips whole AO chunks that are screened out at all the
    points of a tile, so on large molecules it executes fewer flops than the dense count and the
    dense-equivalent fraction can pass 1: a speed figure, not a utilisation."""
    k = kernels["mo_gemm"]
    t = k["ms_per_step"] * 1e-3
    if t <= 0:
        return None
    tf = costs["mo_flops"] * W / t / 1e12
    out = {"dense_equiv_TFs": tf}
    if mode == "bf16x3":
        out.update({"bf16_products_TFs": 6 * tf, "dense_equiv_frac_of_bf16_peak": 6 * tf / PEAK["bf16_TFs"]})
    else:
        peak = PEAK["fp64_TFs"] if mode == "fp64" else PEAK["fp32_TFs"]
        out["dense_equiv_frac_of_peak"] = tf / peak
    return out


def ao_nonzero_density(q, S, s, x, W):
    """Fraction of the AO values that are not exactly zero after screening, from the
    library's AO getter on the first walkers of the workload (at most 2400 points): the
    fused kernels multiply the dense AO tile, the reference skips the zero AOs."""
    Ne = s.elec_num
    nw = max(1, min(W, 2400 // Ne))
    xs = np.ascontiguousarray(x.reshape(3, W, Ne)[:, :nw, :].reshape(3, nw * Ne))
    c = q.context_create()
    try:
        S.load(c, s, jastrow=False, mo=False)
        S.set_coords(c, s, xs, nw)
        ao = q.get_array(c, "qmckl_get_ao_basis_ao_vgl", nw * Ne * 5 * s.ao_num).reshape(nw * Ne, 5, s.ao_num)
        return {"value": float(np.count_nonzero(ao[:, 0, :])) / (nw * Ne * s.ao_num), "points": nw * Ne}
    finally:
        q.context_destroy(c)


def cpu_baseline(s, x, W, seconds):
    """The oracle (FP64 C restatement, OpenMP over walkers/points) on a bounded sample of
    walkers of the same workload: AO VGL + MO VGL + Jastrow value and GL."""
    import oracle

    Ne = s.elec_num
    coef = oracle.kept_mo_coef(s)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def run(nw):
        xs = np.ascontiguousarray(x.reshape(3, W, Ne)[:, :nw, :].reshape(3, nw * Ne))
        t0 = time.perf_counter()
        ao = oracle.ao_vgl(s, xs)
        oracle.mo_vgl(coef, ao)
        oracle.jastrow(s, xs, nw, want=("value", "gl"))
        return time.perf_counter() - t0

    nw = max(1, threads)
    t = run(nw)
    log(f"cpu oracle: {nw} walkers in {t:.2f} s")
    while t < seconds / 3 and nw * 2 <= W:
        nw *= 2
        t = run(nw)
        log(f"cpu oracle: {nw} walkers in {t:.2f} s")
    out = {"value": nw / t, "unit": "walker-evals/s", "cores": threads, "kind": "port",
           "sample": f"{nw} walkers of the same workload (AO VGL + MO VGL + Jastrow value/GL), "
                     f"{t:.1f} s on {threads} OpenMP threads",
           "what": "the oracle: a C restatement of the reference's doc code path (naive loops, no BLAS, "
                   "no zero-AO skipping, no sparse P tensors), not the reference's HPC build"}
    # one thread, on a smaller sample (the OpenMP runtime of the oracle library)
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        n1 = 1
        t1 = run(n1)
        while t1 < seconds / 6 and n1 * 2 <= W:
            n1 *= 2
            t1 = run(n1)
        gomp.omp_set_num_threads(threads)
        log(f"cpu oracle, 1 thread: {n1} walkers in {t1:.2f} s")
        out["single_thread"] = {"value": n1 / t1, "cores": 1, "sample": f"{n1} walkers, {t1:.1f} s"}
    except OSError:
        out["single_thread"] = None
    return out


if __name__ == "__main__":
    main()
