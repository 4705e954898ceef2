"""Torch-facing operators of the LPM solve path.  Each one hands device pointers to liblpm.so (HIP, gfx950).

Shapes follow the reference: vectors (..., N, d), directed coefficients (..., N, NN) with NN = 18 (3D) or 8 (2D),
at most two broadcast batch dimensions (B, L) (coarse3d/torch_lpm3d.py:212-227).  `dims` is (nx, ny[, nz]).
"""
import ctypes

import numpy as np
import torch

import _lpmlib as L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _dims(dims):
    dims = tuple(int(d) for d in dims)
    N = int(np.prod(dims))
    return dims, N, len(dims), L.nn_of(len(dims))


def packed_guard(dims):
    """Zero guard values before each packed plane: one z-plane plus one row plus one site (2D: one row plus one
    site), rounded up to 64 values (256-byte aligned planes in FP32)."""
    dims = tuple(int(d) for d in dims)
    g = dims[0] * dims[1] + dims[0] + 1 if len(dims) == 3 else dims[0] + 1
    return (g + 63) // 64 * 64


def packed_len(dims):
    """Plane stride P: the leading guard, the N particles and at least as many trailing guards, rounded up to 64."""
    n = int(np.prod([int(d) for d in dims])) + 2 * packed_guard(dims)
    return (n + 63) // 64 * 64


def _bshape(bonds):
    return [] if bonds is None else [bonds.shape[:-1]]


def _bonds_view(bonds, batch, N):
    """Bond words (..., N) int32 -> (tensor, Bcast) for the (b, l) systems of `batch`; (None, zero) if absent."""
    if bonds is None:
        return None, L.Bcast(0, 0)
    L.require_gpu(bonds)
    if bonds.dtype != torch.int32:
        raise TypeError("bond words are int32 (uint32 bit patterns), got %s" % bonds.dtype)
    return L.bview(bonds, batch, (N,))


def _bonds_sys(bonds, batch, N, spc):
    """Bond words for the gradient kernels, whose system s = cb * spc + j of the full batch reads word row s."""
    if bonds is None:
        return None, L.Bcast(0, 0)
    L.require_gpu(bonds)
    if bonds.dtype != torch.int32:
        raise TypeError("bond words are int32 (uint32 bit patterns), got %s" % bonds.dtype)
    nb, nl = L.batch2(batch)
    w = bonds.expand(tuple(batch) + (N,)).reshape(nb * nl, N).contiguous()
    return w, L.Bcast(spc * N, N)


def bond_words(dims, solid):
    """Bond words of removed particles: solid (..., N) bool/uint8 -> int32 (..., N), bit k of particle p set iff
    slot k stays in the box and p and its slot-k neighbour are both solid (cement3d/torch_v1.py:36-53)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(solid)
    lead = solid.shape[:-1]
    nbat = int(np.prod(lead)) if len(lead) else 1
    s = (solid.reshape(nbat, N) != 0).to(torch.uint8).contiguous()
    w = torch.empty((nbat, N), device=solid.device, dtype=torch.int32)
    g = L.grid_of(dims, 1.0)
    L.check(L.lib().lpm_bond_words(ctypes.byref(g), nbat, _ptr(s), _ptr(w), L.stream_of(solid)))
    return w.reshape(tuple(lead) + (N,))


def _coef_inner(dims, layout):
    """Per-batch shape of a coefficient array: directed (N, NN) or packed (NN/2, P)."""
    NN = L.nn_of(len(dims))
    if layout == L.DIRECTED:
        return (int(np.prod(dims)), NN)
    return (NN // 2, packed_len(dims))


def apply(dims, h, u, Kn, Tv, layout=L.DIRECTED, bonds=None):
    """K u for broadcast batches: u (..., N, d), Kn/Tv (..., N, NN) directed or (..., NN/2, N) packed; optional
    bond words (..., N) of removed particles (bond_words)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(u, Kn, Tv)
    dt = u.dtype
    if Kn.dtype != dt or Tv.dtype != dt:
        raise TypeError("u, Kn and Tv must share one dtype")
    ci = _coef_inner(dims, layout)
    batch = torch.broadcast_shapes(u.shape[:-2], Kn.shape[:-2], Tv.shape[:-2], *_bshape(bonds))
    nb, nl = L.batch2(batch)
    ue, bu = L.bview(u, batch, (N, d))
    ke, bc = L.bview(Kn, batch, ci)
    te, bt = L.bview(Tv, batch, ci)
    if (bc.sb, bc.sl) != (bt.sb, bt.sl):
        raise ValueError("Kn and Tv must broadcast the same way")
    out = torch.empty(tuple(batch) + (N, d), device=u.device, dtype=dt)
    we, bb = _bonds_view(bonds, batch, N)
    g = L.grid_of(dims, h, we, bb)
    L.check(L.lib().lpm_apply(ctypes.byref(g), nb, nl, L.dtype_code(dt), layout, _ptr(ue), bu, _ptr(ke), _ptr(te),
                              bc, _ptr(out), L.stream_of(u)))
    return out


def post(dims, h, u, Kn, Tv, ku=False, stress=False, energy=False, dl=False, force=False, layout=L.DIRECTED,
         bonds=None):
    """Per-particle outputs; returns a dict with the requested entries (batch = broadcast of u and Kn)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(u, Kn, Tv)
    dt = u.dtype
    ci = _coef_inner(dims, layout)
    batch = torch.broadcast_shapes(u.shape[:-2], Kn.shape[:-2], Tv.shape[:-2], *_bshape(bonds))
    nb, nl = L.batch2(batch)
    ue, bu = L.bview(u, batch, (N, d))
    ke, bc = L.bview(Kn, batch, ci)
    te, _ = L.bview(Tv, batch, ci)
    dev = u.device
    B = tuple(batch)
    work = torch.empty(B + (N, 4), device=dev, dtype=dt)
    ns = 6 if d == 3 else 3
    out = dict(ku=torch.empty(B + (N, d), device=dev, dtype=dt) if ku else None,
               stress=torch.empty(B + (N, ns), device=dev, dtype=dt) if stress else None,
               energy=torch.empty(B + (N,), device=dev, dtype=dt) if energy else None,
               dl=torch.empty(B + (N, NN), device=dev, dtype=dt) if dl else None,
               force=torch.empty(B + (N, NN), device=dev, dtype=dt) if force else None)
    we, bb = _bonds_view(bonds, batch, N)
    g = L.grid_of(dims, h, we, bb)
    L.check(L.lib().lpm_post(ctypes.byref(g), nb, nl, L.dtype_code(dt), layout, _ptr(ue), bu, _ptr(ke), _ptr(te), bc,
                             _ptr(work), _ptr(out["ku"]), _ptr(out["stress"]), _ptr(out["energy"]), _ptr(out["dl"]),
                             _ptr(out["force"]), L.stream_of(u)))
    out["shell"] = work
    return out


def block_jacobi(dims, h, Kn, Tv, free=None, out_dtype=None, raw=False, batch=None, layout=L.DIRECTED, bonds=None):
    """Block-Jacobi inverse blocks (batch, N, 6|3) symmetric, or the diagonal blocks (batch, N, d, d) if raw."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(Kn, Tv)
    ci = _coef_inner(dims, layout)
    shapes = [Kn.shape[:-2]]
    if free is not None:
        f = torch.as_tensor(free)
        shapes.append(f.shape[:-2] if (f.dim() >= 2 and tuple(f.shape[-2:]) == (N, d)) else f.shape[:-1])
    if batch is not None:
        shapes.append(tuple(batch))
    shapes += _bshape(bonds)
    bshape = torch.broadcast_shapes(*shapes)
    nb, nl = L.batch2(bshape)
    ke, bc = L.bview(Kn, bshape, ci)
    te, _ = L.bview(Tv, bshape, ci)
    me, bm = L.dof_mask(free, bshape, N, d, Kn.device)
    odt = out_dtype or Kn.dtype
    per = (d, d) if raw else ((6,) if d == 3 else (3,))
    out = torch.empty(tuple(bshape) + (N,) + per, device=Kn.device, dtype=odt)
    we, bb = _bonds_view(bonds, bshape, N)
    g = L.grid_of(dims, h, we, bb)
    L.check(L.lib().lpm_block_jacobi(ctypes.byref(g), nb, nl, L.dtype_code(Kn.dtype), L.dtype_code(odt), layout,
                                     _ptr(ke), _ptr(te), bc, _ptr(me), bm, _ptr(out), 1 if raw else 0,
                                     L.stream_of(Kn)))
    return out


def sym_from_full(M):
    """(..., d, d) -> symmetric storage (..., 6 | 3) in the kernels' order (xx yy zz yz xz xy | xx yy xy)."""
    if M.shape[-1] == 3:
        return torch.stack([M[..., 0, 0], M[..., 1, 1], M[..., 2, 2], 0.5 * (M[..., 1, 2] + M[..., 2, 1]),
                            0.5 * (M[..., 0, 2] + M[..., 2, 0]), 0.5 * (M[..., 0, 1] + M[..., 1, 0])], -1)
    return torch.stack([M[..., 0, 0], M[..., 1, 1], 0.5 * (M[..., 0, 1] + M[..., 1, 0])], -1)


def full_from_sym(S):
    if S.shape[-1] == 6:
        r0 = torch.stack([S[..., 0], S[..., 5], S[..., 4]], -1)
        r1 = torch.stack([S[..., 5], S[..., 1], S[..., 3]], -1)
        r2 = torch.stack([S[..., 4], S[..., 3], S[..., 2]], -1)
        return torch.stack([r0, r1, r2], -2)
    r0 = torch.stack([S[..., 0], S[..., 2]], -1)
    r1 = torch.stack([S[..., 2], S[..., 1]], -1)
    return torch.stack([r0, r1], -2)


def pack(dims, Kn, Tv, out_dtype=None, rtol=1e-12):
    """Directed (..., N, NN) -> packed forward planes (..., NN/2, P) with zero guards (P = packed_len(dims));
    returns (kp, tp, n_asymmetric_bonds per batch)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(Kn, Tv)
    lead = Kn.shape[:-2]
    nbat = int(np.prod(lead)) if len(lead) else 1
    k = Kn.reshape(nbat, N, NN).contiguous()
    t = Tv.reshape(nbat, N, NN).contiguous()
    odt = out_dtype or Kn.dtype
    P = packed_len(dims)
    kp = torch.empty((nbat, NN // 2, P), device=Kn.device, dtype=odt)
    tp = torch.empty_like(kp)
    asym = torch.zeros(nbat, device=Kn.device, dtype=torch.int32)
    g = L.grid_of(dims, 1.0)
    L.check(L.lib().lpm_pack_coef(ctypes.byref(g), nbat, L.dtype_code(Kn.dtype), L.dtype_code(odt), _ptr(k), _ptr(t),
                                  _ptr(kp), _ptr(tp), float(rtol), _ptr(asym), L.stream_of(Kn)))
    return kp.reshape(tuple(lead) + (NN // 2, P)), tp.reshape(tuple(lead) + (NN // 2, P)), asym


def _coef_batches(batch, Kn, Tv, ci):
    """Coefficient batching of a gradient kernel: systems s = cb * spc + j share coefficient batch cb.  Returns
    (ncb, spc, kk, tt, general) with kk, tt (ncb, *ci) contiguous; general means one set per system, whose
    gradients still have to be summed to the broadcast shape of Kn."""
    nb, nl = L.batch2(batch)
    kb2 = (1,) * (len(batch) - (Kn.dim() - 2)) + tuple(Kn.shape[:-2])
    if len(batch) == 2 and kb2 == (nb, 1):
        return nb, nl, Kn.reshape(nb, *ci).contiguous(), Tv.reshape(nb, *ci).contiguous(), False
    if all(s == 1 for s in kb2):
        return 1, nb * nl, Kn.reshape(1, *ci).contiguous(), Tv.reshape(1, *ci).contiguous(), False
    kk = Kn.expand(tuple(batch) + ci).reshape(nb * nl, *ci).contiguous()
    tt = Tv.expand(tuple(batch) + ci).reshape(nb * nl, *ci).contiguous()
    return nb * nl, 1, kk, tt, True


def _sum_coef_grad(g, batch, Kn, N, NN, general):
    kb = tuple(Kn.shape[:-2])
    if general:
        kb2 = (1,) * (len(batch) - len(kb)) + kb
        g = g.reshape(tuple(batch) + (N, NN)).sum_to_size(kb2 + (N, NN))
    return g.reshape(kb + (N, NN))


def bond_grad(dims, h, Kn, Tv, u, lam, scale=1.0, layout=L.DIRECTED, bonds=None):
    """scale * d/d(Kn, Tv) sum_s lam_s . K u_s per directed slot, reduced to the broadcast shape of Kn.
    u, lam (..., N, d) with the full batch; Kn (..., N, NN)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(Kn, Tv, u, lam)
    dt = u.dtype
    batch = torch.broadcast_shapes(u.shape[:-2], lam.shape[:-2], Kn.shape[:-2], *_bshape(bonds))
    ci = _coef_inner(dims, layout)
    nb, nl = L.batch2(batch)
    ue = u.expand(tuple(batch) + (N, d)).contiguous()
    le = lam.expand(tuple(batch) + (N, d)).contiguous()
    ncb, spc, kk, tt, general = _coef_batches(batch, Kn, Tv, ci)
    work = torch.empty((2, nb * nl, N, 4), device=u.device, dtype=dt)
    gk = torch.empty((ncb, N, NN), device=u.device, dtype=dt)
    gt = torch.empty_like(gk)
    we, bb = _bonds_sys(bonds, batch, N, spc)
    g = L.grid_of(dims, h, we, bb)
    L.check(L.lib().lpm_bond_grad(ctypes.byref(g), ncb, spc, L.dtype_code(dt), layout, _ptr(kk), _ptr(tt), _ptr(ue),
                                  _ptr(le), float(scale), _ptr(work), _ptr(gk), _ptr(gt), L.stream_of(u)))
    return _sum_coef_grad(gk, batch, Kn, N, NN, general), _sum_coef_grad(gt, batch, Kn, N, NN, general)


def post_vjp(dims, h, u, Kn, Tv, g_ku=None, g_stress=None, g_energy=None, layout=L.DIRECTED, bonds=None):
    """Vector-Jacobian product of post()'s (ku, stress, energy) for the given cotangents (None = zero).
    Returns (gu with u's batch broadcast to the full batch, gKn, gTv directed and reduced to Kn's shape)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(u, Kn, Tv)
    dt = u.dtype
    batch = torch.broadcast_shapes(u.shape[:-2], Kn.shape[:-2], Tv.shape[:-2], *_bshape(bonds))
    ci = _coef_inner(dims, layout)
    nb, nl = L.batch2(batch)
    B = tuple(batch)
    ns = 6 if d == 3 else 3
    ue = u.expand(B + (N, d)).contiguous()
    cot = []
    for gg, tail in ((g_ku, (N, d)), (g_stress, (N, ns)), (g_energy, (N,))):
        cot.append(None if gg is None else gg.to(dt).expand(B + tail).contiguous())
    ncb, spc, kk, tt, general = _coef_batches(batch, Kn, Tv, ci)
    work = torch.empty((nb * nl, N, 4 + 2 * NN), device=u.device, dtype=dt)
    gu = torch.empty(B + (N, d), device=u.device, dtype=dt)
    gk = torch.empty((ncb, N, NN), device=u.device, dtype=dt)
    gt = torch.empty_like(gk)
    we, bb = _bonds_sys(bonds, batch, N, spc)
    g = L.grid_of(dims, h, we, bb)
    L.check(L.lib().lpm_post_vjp(ctypes.byref(g), ncb, spc, L.dtype_code(dt), layout, _ptr(kk), _ptr(tt), _ptr(ue),
                                 _ptr(cot[0]), _ptr(cot[1]), _ptr(cot[2]), _ptr(work), _ptr(gu), _ptr(gk), _ptr(gt),
                                 L.stream_of(u)))
    return gu, _sum_coef_grad(gk, batch, Kn, N, NN, general), _sum_coef_grad(gt, batch, Kn, N, NN, general)


class SolveInfo:
    def __init__(self, iters, relres, conv, outer=None, snaps=None, snap_iters=None):
        self.iters, self.relres, self.conv, self.outer = iters, relres, conv, outer
        self.snaps, self.snap_iters = snaps, snap_iters

    @property
    def it(self):
        return int(self.iters.max()) if self.iters.size else 0

    @property
    def relmax(self):
        return float(self.relres.max()) if self.relres.size else 0.0


def _opts(precond, tol, maxit, check, graph, inner_tol=1e-5, max_outer=20):
    o = L.CgOpts()
    o.precond = 1 if precond else 0
    o.maxit = int(maxit)
    o.check = int(check)
    o.graph = 1 if graph else 0
    o.tol = float(tol)
    o.inner_tol = float(inner_tol)
    o.max_outer = int(max_outer)
    return o


def pcg(dims, h, b, Kn, Tv, minv=None, free=None, tol=1e-10, maxit=20000, x0=None, check=10, graph=False,
        layout=L.DIRECTED, bonds=None, tols=None):
    """Batched (P)CG on K_ff x = b_f.  b (B, L, N, d) or broadcastable; minv symmetric blocks (…, N, 6|3) or None
    (plain CG).  Returns x (batch, N, d) and SolveInfo (per-system iterations, relative residual, convergence).
    tols: a descending tolerance list (multi-tolerance solve, timing/t2d/solver2d.py:349-421): the solve runs until
    every level is met (tol is not used) and info.snaps (batch, ntol, N, d) / info.snap_iters (nb, nl, ntol) hold the
    iterate of the first iteration meeting each level (-1: never met, final iterate)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(b, Kn, Tv)
    dt = b.dtype
    ci = _coef_inner(dims, layout)
    shapes = [b.shape[:-2], Kn.shape[:-2]]
    if minv is not None:
        shapes.append(minv.shape[:-2])
    if free is not None:
        f = torch.as_tensor(free)
        shapes.append(f.shape[:-2] if (f.dim() >= 2 and tuple(f.shape[-2:]) == (N, d)) else f.shape[:-1])
    shapes += _bshape(bonds)
    batch = torch.broadcast_shapes(*shapes)
    nb, nl = L.batch2(batch)
    be = b.expand(tuple(batch) + (N, d)).contiguous()
    ke, bc = L.bview(Kn, batch, ci)
    te, _ = L.bview(Tv, batch, ci)
    me, bm = L.dof_mask(free, batch, N, d, b.device)
    if minv is not None:
        mi, bmi = L.bview(minv.to(dt), batch, (N, 6 if d == 3 else 3))
    else:
        mi, bmi = None, L.Bcast(0, 0)
    if x0 is not None:
        x = x0.to(dt).expand(tuple(batch) + (N, d)).contiguous().clone()
    else:
        x = torch.empty(tuple(batch) + (N, d), device=b.device, dtype=dt)
    we, bb = _bonds_view(bonds, batch, N)
    g = L.grid_of(dims, h, we, bb)
    nsys = nb * nl
    wbytes = L.lib().lpm_pcg_workspace(ctypes.byref(g), nsys, L.dtype_code(dt))
    work = torch.empty(int(wbytes), device=b.device, dtype=torch.uint8)
    iters = np.zeros(nsys, np.int32)
    rel = np.zeros(nsys, np.float64)
    conv = np.zeros(nsys, np.int32)
    o = _opts(minv is not None, tol, maxit, check, graph)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    if tols is None:
        L.check(L.lib().lpm_pcg_solve(ctypes.byref(g), nb, nl, L.dtype_code(dt), layout, _ptr(ke), _ptr(te), bc,
                                      _ptr(me), bm, _ptr(mi), bmi, _ptr(be), _ptr(x), 1 if x0 is not None else 0,
                                      ctypes.byref(o), _ptr(work), vp(iters), vp(rel), vp(conv), L.stream_of(b)))
        return x, SolveInfo(iters.reshape(nb, nl), rel.reshape(nb, nl), conv.reshape(nb, nl))
    tl = np.ascontiguousarray(np.asarray(tols, np.float64).ravel())
    snaps = torch.empty(tuple(batch) + (len(tl), N, d), device=b.device, dtype=dt)
    sit = np.zeros(nsys * len(tl), np.int32)
    L.check(L.lib().lpm_pcg_snapshots(ctypes.byref(g), nb, nl, L.dtype_code(dt), layout, _ptr(ke), _ptr(te), bc,
                                      _ptr(me), bm, _ptr(mi), bmi, _ptr(be), _ptr(x), 1 if x0 is not None else 0,
                                      ctypes.byref(o), len(tl), vp(tl), _ptr(snaps), vp(sit), _ptr(work), vp(iters),
                                      vp(rel), vp(conv), L.stream_of(b)))
    return x, SolveInfo(iters.reshape(nb, nl), rel.reshape(nb, nl), conv.reshape(nb, nl), snaps=snaps,
                        snap_iters=sit.reshape(nb, nl, len(tl)))


class Cg2Persistent:
    """Plain FP64 CG of one 2D square-lattice system in one cooperative launch (cg2plain.cg_mf,
    timing/t2d/cg2plain.py:11-49).  Kn, Tv (N, 8) directed FP64, free (N, 2) per-DOF mask (nonzero = free);
    solve(b) returns x (N, 2) from a zero start and SolveInfo.  The workspace is allocated once per instance;
    workgroups = 0 picks the launch shape automatically.  pipelined=True runs the pipelined form of the same CG
    (one grid-wide reduction per iteration, overlapped with the product; lpm_cg2_pipelined): the same iterates in
    exact arithmetic, rounding differs."""

    def __init__(self, dims, Kn, Tv, free, workgroups=0, pipelined=False):
        nx, ny = int(dims[0]), int(dims[1])
        N = nx * ny
        self.dims, self.workgroups, self.pipelined = (nx, ny), int(workgroups), bool(pipelined)
        dev = Kn.device
        self.kn = Kn.to(torch.float64).reshape(N, 8).contiguous()
        self.tv = Tv.to(torch.float64).reshape(N, 8).contiguous()
        self.mask = (free.reshape(N, 2) != 0).to(torch.uint8).contiguous()
        self.work = torch.empty(int(L.lib().lpm_cg2_persistent_workspace(nx, ny)), device=dev, dtype=torch.uint8)

    def solve(self, b, tol=1e-8, maxit=200000, out=None):
        nx, ny = self.dims
        N = nx * ny
        be = b.to(torch.float64).reshape(N, 2).contiguous()
        x = out if out is not None else torch.empty((N, 2), device=be.device, dtype=torch.float64)
        it, conv = ctypes.c_int32(0), ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        fn = L.lib().lpm_cg2_pipelined if self.pipelined else L.lib().lpm_cg2_persistent
        L.check(fn(nx, ny, _ptr(self.kn), _ptr(self.tv), _ptr(self.mask), _ptr(be), _ptr(x), float(tol), int(maxit),
                   self.workgroups, _ptr(self.work), ctypes.byref(it), ctypes.byref(rel), ctypes.byref(conv),
                   L.stream_of(be)))
        return x, SolveInfo(np.array([[it.value]]), np.array([[rel.value]]), np.array([[conv.value]]))


def pcg_refined(dims, h, b, Kn64, Tv64, Kn32, Tv32, minv32=None, free=None, tol=1e-8, maxit=20000, inner_tol=1e-5,
                max_outer=30, x0=None, check=10, graph=False, layout64=L.DIRECTED, layout32=L.DIRECTED, bonds=None):
    """FP32 (P)CG inner solves inside an FP64 iterative refinement; b FP64 (…, N, d).  Returns x FP64 and SolveInfo
    (relres = true FP64 relative residual, iters = total inner iterations, outer = refinement steps)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(b, Kn64, Tv64, Kn32, Tv32)
    if b.dtype != torch.float64 or Kn64.dtype != torch.float64 or Kn32.dtype != torch.float32:
        raise TypeError("refined solve: b and the FP64 operator in float64, the inner operator in float32")
    ci64 = _coef_inner(dims, layout64)
    ci32 = _coef_inner(dims, layout32)
    shapes = [b.shape[:-2], Kn64.shape[:-2], Kn32.shape[:-2]]
    if minv32 is not None:
        shapes.append(minv32.shape[:-2])
    if free is not None:
        f = torch.as_tensor(free)
        shapes.append(f.shape[:-2] if (f.dim() >= 2 and tuple(f.shape[-2:]) == (N, d)) else f.shape[:-1])
    shapes += _bshape(bonds)
    batch = torch.broadcast_shapes(*shapes)
    nb, nl = L.batch2(batch)
    be = b.expand(tuple(batch) + (N, d)).contiguous()
    k64, bc64 = L.bview(Kn64, batch, ci64)
    t64, _ = L.bview(Tv64, batch, ci64)
    k32, bc32 = L.bview(Kn32, batch, ci32)
    t32, _ = L.bview(Tv32, batch, ci32)
    me, bm = L.dof_mask(free, batch, N, d, b.device)
    if minv32 is not None:
        mi, bmi = L.bview(minv32.to(torch.float32), batch, (N, 6 if d == 3 else 3))
    else:
        mi, bmi = None, L.Bcast(0, 0)
    if x0 is not None:
        x = x0.to(torch.float64).expand(tuple(batch) + (N, d)).contiguous().clone()
    else:
        x = torch.empty(tuple(batch) + (N, d), device=b.device, dtype=torch.float64)
    we, bb = _bonds_view(bonds, batch, N)
    g = L.grid_of(dims, h, we, bb)
    nsys = nb * nl
    work = torch.empty(int(L.lib().lpm_refined_workspace(ctypes.byref(g), nsys)), device=b.device, dtype=torch.uint8)
    iters = np.zeros(nsys, np.int32)
    rel = np.zeros(nsys, np.float64)
    conv = np.zeros(nsys, np.int32)
    outer = np.zeros(1, np.int32)
    o = _opts(minv32 is not None, tol, maxit, check, graph, inner_tol, max_outer)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L.check(L.lib().lpm_pcg_refined(ctypes.byref(g), nb, nl, layout64, _ptr(k64), _ptr(t64), bc64, layout32, _ptr(k32),
                                    _ptr(t32), bc32, _ptr(me), bm, _ptr(mi), bmi, _ptr(be), _ptr(x),
                                    1 if x0 is not None else 0, ctypes.byref(o), _ptr(work), vp(iters), vp(rel),
                                    vp(outer), vp(conv), L.stream_of(b)))
    return x, SolveInfo(iters.reshape(nb, nl), rel.reshape(nb, nl), conv.reshape(nb, nl), int(outer[0]))


# ------------------------------------------------------------------------------------------------ CSR assembly
_OFFS = {3: np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
                      (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1),
                      (0, -1, -1), (0, 1, -1), (0, -1, 1)]),
         2: np.array([(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)])}
_PATTERNS = {}


def csr_pattern(dims):
    """Block pattern of K on a full box, the structure of the reference's assembly (coarse3d/lin3d.py:306-376):
    particle q is a column block of row p if q = p, q is a bond neighbour of p, or q is a neighbour of a bond
    neighbour i of p in the shell of the bond (p, i).  Returns (offsets (O, d) sorted by linear offset,
    present (N, O) bool, rowptr (N+1,) int64 of column blocks).  Host-side index arithmetic, cached per box."""
    dims = tuple(int(v) for v in dims)
    if dims in _PATTERNS:
        return _PATTERNS[dims]
    d = len(dims)
    offs = _OFFS[d]
    NN = len(offs)
    nsh = 6 if d == 3 else 4
    shell = np.arange(NN) >= nsh
    ijk = np.indices(dims[::-1]).reshape(d, -1)[::-1].T                     # (N, d), x fastest
    ext = np.array(dims)
    inbox = lambda c: np.all((c >= 0) & (c < ext), axis=-1)
    cand = {(0,) * d: np.ones(len(ijk), bool)}
    for a in range(NN):
        ia = ijk + offs[a]
        va = inbox(ia)
        key = tuple(offs[a])
        cand[key] = cand.get(key, False) | va
        for b in range(NN):
            if shell[b] != shell[a]:
                continue
            o = offs[a] + offs[b]
            key = tuple(o)
            cand[key] = cand.get(key, False) | (va & inbox(ia + offs[b]))
    strides = np.cumprod([1] + list(dims[:-1]))
    keys = sorted(cand, key=lambda o: int(np.dot(o, strides)))
    offsets = np.array(keys, np.int64)
    present = np.stack([cand[k] for k in keys], 1)
    rowptr = np.zeros(len(ijk) + 1, np.int64)
    np.cumsum(present.sum(1), out=rowptr[1:])
    _PATTERNS[dims] = (offsets, present, rowptr)
    return _PATTERNS[dims]


def stiffness_csr(dims, h, Kn, Tv, colours_per_pass=None, bonds=None):
    """Assembled K (d N x d N) in CSR form for one symmetric coefficient set Kn, Tv (N, NN) on the GPU.

    K is probed with the batched matrix-free operator: two particles closer than 5 sites along every axis never
    share a row's pattern, so with colours c(q) = (q_x mod 5, q_y mod 5[, q_z mod 5]) the product of K with the
    indicator of (colour c, component b) holds in row (p, a) exactly the entry K[(p, a), (q, b)] of the single
    column particle q of colour c in p's pattern.  5^d x d probe vectors are applied in batches; entries are
    gathered into the pattern of csr_pattern (columns ascending).  Returns (indptr int64, indices int32, data)
    as host NumPy arrays; data in Kn's dtype.  With bond words (N,) of removed particles the values are those of
    the reduced lattice (entries through removed bonds are zero in the full-box pattern)."""
    dims, N, d, NN = _dims(dims)
    L.require_gpu(Kn, Tv)
    if Kn.shape != (N, NN) or Tv.shape != (N, NN):
        raise ValueError("stiffness_csr: one coefficient set (N, NN) expected")
    dev, dt = Kn.device, Kn.dtype
    offsets, present, rowptr = csr_pattern(dims)
    O = len(offsets)
    ijk = np.indices(dims[::-1]).reshape(d, -1)[::-1].T
    colour_of = np.zeros(N, np.int64)
    for ax in range(d):
        colour_of += (ijk[:, ax] % 5) * 5 ** ax
    ncol = 5 ** d
    strides = np.cumprod([1] + list(dims[:-1]))
    lin = offsets @ strides                                                 # (O,)
    p_idx, o_idx = np.nonzero(present)                                      # row-major: p, then ascending offset
    q_idx = p_idx + lin[o_idx]
    c_idx = colour_of[q_idx]
    vals = torch.empty((len(p_idx), d, d), device=dev, dtype=dt)            # [block, a, b]
    cpp = colours_per_pass or max(1, min(ncol, int(2e9 // max(1, N * d * d * 8 * 2))))
    colour_t = torch.as_tensor(colour_of, device=dev)
    for c0 in range(0, ncol, cpp):
        cs = torch.arange(c0, min(ncol, c0 + cpp), device=dev)
        probe = torch.zeros((len(cs), d, N, d), device=dev, dtype=dt)       # [colour, b, particle, component]
        hit = (colour_t[None, :] == cs[:, None]).to(dt)                     # (nc, N)
        for b in range(d):
            probe[:, b, :, b] = hit
        Y = apply(dims, h, probe, Kn, Tv, bonds=bonds)                     # [colour, b, p, a]
        sel = np.nonzero((c_idx >= c0) & (c_idx < c0 + len(cs)))[0]
        if len(sel) == 0:
            continue
        st = torch.as_tensor(sel, device=dev)
        cc = torch.as_tensor(c_idx[sel] - c0, device=dev)
        pp = torch.as_tensor(p_idx[sel], device=dev)
        vals[st] = Y[cc, :, pp, :].transpose(1, 2)                          # (blocks, a, b)
    # CSR row (p, a) holds the blocks of p in offset order, each with d columns b
    nblk = np.diff(rowptr)
    blk_start = np.repeat(rowptr[:-1], nblk)                                # first block of the row's particle
    local = np.arange(len(p_idx)) - blk_start                               # block number within the row
    rowlen = np.repeat(d * nblk, d)
    indptr = np.zeros(d * N + 1, np.int64)
    np.cumsum(rowlen, out=indptr[1:])
    data = torch.empty(int(indptr[-1]), device=dev, dtype=dt)
    indices = np.empty(int(indptr[-1]), np.int32)
    for a in range(d):
        base = indptr[d * p_idx + a] + d * local                            # (blocks,)
        for b in range(d):
            pos = base + b
            indices[pos] = d * q_idx + b
            data[torch.as_tensor(pos, device=dev)] = vals[:, a, b]
    return indptr, indices, data.cpu().numpy()
