"""Differentiation of the CHAMP Jastrow through the library's adjoint kernel.

``jastrow_adjoint(ctx, weight_j, weight_el)`` calls ``qmckl_amd_get_jastrow_champ_adjoint``
(include/qmckl.h): the walker sum L = sum_w [weight_j ln Psi_J + weight_el E_L] over the
context's walkers, with the Jastrow-only local-energy surrogate
E_L = -1/2 sum_i (lap_i J + |grad_i J|^2) + V_ee + V_en + V_nn, and its gradient in the Jastrow
parameters a, b, c and the nuclear coordinates.

``JastrowSurrogate`` wraps it as a ``torch.autograd.Function`` of CPU float64 tensors
(a [type][aord+1], b [bord+1], c [type][dim_c], nuclear coordinates [3][nucl]): forward
evaluates L on the GPU for a system and walker coordinates; backward returns the adjoint's
gradient.  Tensors cross the C-ABI as host buffers (torch's own HIP runtime and the
library's are not mixed).
"""

import copy

import numpy as np

from . import check, lib, ptr
from . import system as S


def jastrow_adjoint(ctx, weight_j=1.0, weight_el=1.0):
    """(L, dL/da [type][aord+1], dL/db [bord+1], dL/dc [type][dim_c], dL/dR [nucl][3])."""
    import ctypes

    L = lib()
    T = ctypes.c_int64()
    check(ctx, L.qmckl_get_jastrow_champ_type_nucl_num(ctx, ctypes.byref(T)))
    ao, bo, co, dc, nn = (ctypes.c_int64() for _ in range(5))
    check(ctx, L.qmckl_get_jastrow_champ_aord_num(ctx, ctypes.byref(ao)))
    check(ctx, L.qmckl_get_jastrow_champ_bord_num(ctx, ctypes.byref(bo)))
    check(ctx, L.qmckl_get_jastrow_champ_cord_num(ctx, ctypes.byref(co)))
    check(ctx, L.qmckl_get_jastrow_champ_dim_c_vector(ctx, ctypes.byref(dc)))
    check(ctx, L.qmckl_get_nucleus_num(ctx, ctypes.byref(nn)))
    T, ao, bo, co, dc, nn = T.value, ao.value, bo.value, co.value, dc.value, nn.value
    value = np.zeros(1)
    ga = np.zeros((T, ao + 1))
    gb = np.zeros(bo + 1)
    gc = np.zeros((T, dc if co >= 2 else 0))
    gr = np.zeros((nn, 3))
    gc_buf = gc if gc.size else np.zeros(1)
    check(ctx, L.qmckl_amd_get_jastrow_champ_adjoint(ctx, float(weight_j), float(weight_el), ptr(value),
                                                     ptr(ga), ga.size, ptr(gb), gb.size, ptr(gc_buf), gc.size,
                                                     ptr(gr), gr.size), "qmckl_amd_get_jastrow_champ_adjoint")
    return value[0], ga, gb, gc, gr


def surrogate(s, coord_T, walk_num, weight_j=1.0, weight_el=1.0):
    """jastrow_adjoint for System s (its parameters and nuclei) at walker coordinates
    coord_T [3][walk_num*elec_num], on a context of its own."""
    from . import context_create, context_destroy

    ctx = context_create()
    try:
        S.load(ctx, s, mo=False)
        S.set_coords(ctx, s, coord_T, walk_num)
        return jastrow_adjoint(ctx, weight_j, weight_el)
    finally:
        context_destroy(ctx)


def _function():
    import torch

    class JastrowSurrogate(torch.autograd.Function):
        """L(a, b, c, R) = sum_w [weight_j ln Psi_J + weight_el E_L] for a System's walkers;
        gradients by the library's adjoint kernel."""

        @staticmethod
        def forward(ctx, a, b, c, nucl_coord, system, coord_T, walk_num, weight_j=1.0, weight_el=1.0):
            s = copy.deepcopy(system)
            s.a_vector = a.detach().cpu().double().numpy().copy()
            s.b_vector = b.detach().cpu().double().numpy().copy()
            if s.cord > 1:
                s.c_vector = c.detach().cpu().double().numpy().copy()
            s.nucl_coord = nucl_coord.detach().cpu().double().numpy().copy()
            val, ga, gb, gc, gr = surrogate(s, coord_T, walk_num, weight_j, weight_el)
            ctx.save_for_backward(torch.from_numpy(ga), torch.from_numpy(gb),
                                  torch.from_numpy(gc.reshape(c.shape) if s.cord > 1 else np.zeros(c.shape)),
                                  torch.from_numpy(np.ascontiguousarray(gr.T)))
            return torch.tensor(val, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            ga, gb, gc, gr = ctx.saved_tensors
            return gout * ga, gout * gb, gout * gc, gout * gr, None, None, None, None, None

    return JastrowSurrogate


_FN = None


def JastrowSurrogate():
    """The torch.autograd.Function class (torch is imported on first use)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN
