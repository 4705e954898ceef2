"""ctypes binding of liblpm.so (the C ABI declared in include/lpm.h) and the torch-facing operators built on it.

Every operator here runs the HIP kernels of liblpm.so on the tensors' GPU; there is no CPU path.  A missing
library or a CPU tensor raises at once.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpm.so")

LPM_OK, LPM_E_BADARG, LPM_E_HIP, LPM_E_NOT_SPD, LPM_E_NONFINITE, LPM_E_UNSUPPORTED, LPM_E_FULL = range(7)
F32, F64 = 0, 1
DIRECTED, PACKED = 0, 1


class Bcast(ctypes.Structure):
    _fields_ = [("sb", ctypes.c_int64), ("sl", ctypes.c_int64)]


class Grid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("h", ctypes.c_double), ("bonds", ctypes.c_void_p), ("bb", Bcast)]


class CgOpts(ctypes.Structure):
    _fields_ = [("precond", ctypes.c_int32), ("maxit", ctypes.c_int32), ("check", ctypes.c_int32),
                ("graph", ctypes.c_int32), ("tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("max_outer", ctypes.c_int32), ("pad_", ctypes.c_int32)]


_LIB = None
P = ctypes.c_void_p
I32 = ctypes.c_int32
GP = ctypes.POINTER(Grid)
SIGS = {
    "lpm_last_error": (ctypes.c_char_p, []),
    "lpm_version": (I32, []),
    "lpm_profile_ku": (None, [I32]),
    "lpm_set_ku_tiling": (None, [I32, I32]),
    "lpm_set_ku_kernel": (None, [I32]),
    "lpm_packed_len": (ctypes.c_int64, [GP]),
    "lpm_profile_ku_read": (None, [P, P]),
    "lpm_bond_words": (I32, [GP, I32, P, P, P]),
    "lpm_apply": (I32, [GP, I32, I32, I32, I32, P, Bcast, P, P, Bcast, P, P]),
    "lpm_pack_coef": (I32, [GP, I32, I32, I32, P, P, P, P, ctypes.c_double, P, P]),
    "lpm_block_jacobi": (I32, [GP, I32, I32, I32, I32, I32, P, P, Bcast, P, Bcast, P, I32, P]),
    "lpm_post": (I32, [GP, I32, I32, I32, I32, P, Bcast, P, P, Bcast, P, P, P, P, P, P, P]),
    "lpm_bond_grad": (I32, [GP, I32, I32, I32, I32, P, P, P, P, ctypes.c_double, P, P, P, P]),
    "lpm_post_vjp": (I32, [GP, I32, I32, I32, I32, P, P, P, P, P, P, P, P, P, P, P]),
    "lpm_pcg_workspace": (ctypes.c_size_t, [GP, I32, I32]),
    "lpm_pcg_solve": (I32, [GP, I32, I32, I32, I32, P, P, Bcast, P, Bcast, P, Bcast, P, P, I32,
                            ctypes.POINTER(CgOpts), P, P, P, P, P]),
    "lpm_pcg_snapshots": (I32, [GP, I32, I32, I32, I32, P, P, Bcast, P, Bcast, P, Bcast, P, P, I32,
                                ctypes.POINTER(CgOpts), I32, P, P, P, P, P, P, P, P]),
    "lpm_cg2_persistent_workspace": (ctypes.c_size_t, [I32, I32]),
    "lpm_cg2_persistent": (I32, [I32, I32, P, P, P, P, P, ctypes.c_double, I32, I32, P, P, P, P, P]),
    "lpm_cg2_pipelined": (I32, [I32, I32, P, P, P, P, P, ctypes.c_double, I32, I32, P, P, P, P, P]),
    "lpm_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "lpm_pcg_solve_noalloc": (I32, [GP, I32, I32, I32, I32, P, P, Bcast, P, Bcast, P, Bcast, P, P, I32,
                                    ctypes.POINTER(CgOpts), P, P, P, P]),
    "lpm_serve_footprint": (ctypes.c_size_t, [I32, I32]),
    "lpm_serve_create": (I32, [I32, ctypes.c_size_t, I32, ctypes.POINTER(ctypes.c_void_p)]),
    "lpm_serve_destroy": (None, [P]),
    "lpm_serve_admit": (I32, [P, I32, I32, ctypes.c_double, P, P, P, P, ctypes.c_double, ctypes.c_double, I32, P,
                              ctypes.POINTER(I32)]),
    "lpm_serve_step": (I32, [P, ctypes.c_int64, I32, P]),
    "lpm_serve_set_fused": (None, [I32]),
    "lpm_set_ku_box": (None, [I32]),
    "lpm_serve_poll": (I32, [P, P, P, P, P, P]),
    "lpm_serve_result": (I32, [P, I32, P, P]),
    "lpm_serve_release": (I32, [P, I32]),
    "lpm_serve_arena_free": (ctypes.c_size_t, [P]),
    "lpm_refined_workspace": (ctypes.c_size_t, [GP, I32]),
    "lpm_pcg_refined": (I32, [GP, I32, I32, I32, P, P, Bcast, I32, P, P, Bcast, P, Bcast, P, Bcast, P, P, I32,
                              ctypes.POINTER(CgOpts), P, P, P, P, P, P]),
}


def lib():
    """The loaded liblpm.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblpm.so is not built (run build.py in %s); the LPM solve path has no CPU "
                               "fallback" % HERE)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc == LPM_OK:
        return
    msg = lib().lpm_last_error().decode(errors="replace")
    if rc == LPM_E_NOT_SPD:
        raise FloatingPointError(msg)
    if rc == LPM_E_BADARG:
        raise ValueError(msg)
    if rc == LPM_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError("liblpm error %d: %s" % (rc, msg))


def stream_of(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def dtype_code(dt):
    if dt == torch.float32:
        return F32
    if dt == torch.float64:
        return F64
    raise TypeError("LPM kernels compute in float32 or float64, got %s" % dt)


def require_gpu(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise RuntimeError("the LPM solve path runs on the GPU only (got a tensor on %s)" % t.device)


def grid_of(dims, h, bonds=None, bb=None):
    """lpm_grid of a box; `bonds` (a device tensor of bond words, kept alive by the caller) with strides bb."""
    dims = tuple(int(d) for d in dims)
    g = Grid()
    g.dim = len(dims)
    g.nx, g.ny = dims[0], dims[1]
    g.nz = dims[2] if len(dims) == 3 else 1
    g.h = float(h)
    g.bonds = bonds.data_ptr() if bonds is not None else None
    g.bb = bb if bb is not None else Bcast(0, 0)
    return g


def nn_of(dim):
    return 18 if dim == 3 else 8


# ------------------------------------------------------------------------------------------------ batching
def batch2(shape):
    """Leading batch shape -> (nb, nl) view dims; at most two batch dimensions."""
    if len(shape) == 0:
        return 1, 1
    if len(shape) == 1:
        return 1, int(shape[0])
    if len(shape) == 2:
        return int(shape[0]), int(shape[1])
    raise ValueError("at most two batch dimensions are supported (got %s)" % (tuple(shape),))


def bview(t, batch, inner):
    """Expand t (batch', *inner) to (batch, *inner) and return (tensor, Bcast) with the element strides of the
    (b, l) indices; the inner block is made contiguous."""
    nin = len(inner)
    if tuple(t.shape[t.dim() - nin:]) != tuple(inner):
        raise ValueError("operand shape %s does not end in %s" % (tuple(t.shape), tuple(inner)))
    inner_c = True
    exp_st = 1
    for k in range(nin - 1, -1, -1):
        if inner[k] != 1 and t.stride(t.dim() - nin + k) != exp_st:
            inner_c = False
        exp_st *= inner[k]
    if not inner_c:
        t = t.contiguous()
    lead = t.shape[:t.dim() - nin]
    while len(lead) < len(batch):
        t = t.unsqueeze(0)
        lead = t.shape[:t.dim() - nin]
    e = t.expand(tuple(batch) + tuple(inner))
    nbd = len(batch)
    if nbd == 0:
        return e, Bcast(0, 0)
    if nbd == 1:
        return e, Bcast(0, e.stride(0) if batch[0] > 1 else 0)
    return e, Bcast(e.stride(0) if batch[0] > 1 else 0, e.stride(1) if batch[1] > 1 else 0)


def dof_mask(free, batch, N, d, device):
    """Reference masks -> per-DOF uint8 (batch, N, d) view: free (..., N) per particle (3D torch convention) or
    (..., N, d) per DOF; None = all free."""
    if free is None:
        return None, Bcast(0, 0)
    f = torch.as_tensor(free, device=device)
    if not (f.dim() >= 2 and tuple(f.shape[-2:]) == (N, d)):
        f = f.unsqueeze(-1).expand(*f.shape, d)
    m = (f != 0).to(torch.uint8)
    return bview(m, batch, (N, d))
