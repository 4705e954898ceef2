"""Batched, differentiable 3D simple-cubic LPM on MI355X: the interface of coarse3d/torch_lpm3d.py backed by the HIP
kernels of liblpm.so.

Operator (coarse3d/torch_lpm3d.py:63-83): dL_ik = n_k.(u_i - u_q), F_ik = 2 Kn dL + 1/2 (TS_i^c + TS_q^c)
+ 1/2 Tv (S_i^c + S_q^c), K u = sum_k F_ik n_k; one fused kernel per K u.  Solves are batched block-Jacobi PCG over
all (specimen x load) systems with per-system convergence (coarse3d/torch_lpm3d.py:109-148).  LinSolve keeps the
reference's adjoint contract (coarse3d/torch_lpm3d.py:151-175): backward = one adjoint PCG plus the closed-form
bond-parameter gradient kernel (no autograd graph through the operator).

With mixed=True an SCT in float64 solves with FP32 inner PCG iterations inside an FP64 iterative refinement
(true residual and solution update in FP64).  Symmetric coefficient sets are packed once per solve into
forward-slot planes, which halves the coefficient traffic of every K u in the iteration.

SCTV / CoarseTV / FineRefTV (cement3d/torch_v1.py:18-235) add per-specimen removed particles (voids): a bond exists
only if both ends are solid.  The validity travels to the kernels as one 32-bit bond word per particle (bit k = slot
k is a bond) instead of a (B, 1, N, 18) float tensor; every kernel (K u, shell sums, stress opp-flag rule, block
diagonal, gradients) tests the bit.
"""
import numpy as np
import torch

import _lpmlib as L
import lpm_ops as ops
from geom3d import LatticeGeom, SC_SHELL, LOADS, PLATES, GDIR, prescribed, free_surface_cells  # noqa: F401

def inv3(D):
    """Inverse of (..., 3, 3) blocks (cofactor form, as coarse3d/torch_lpm3d.py:22-33)."""
    return torch.linalg.inv(D)


class SCT:
    """Matrix-free LPM operator on a full-box lattice: 3D simple cubic (geom3d.LatticeGeom) or, with the same
    methods, the 2D square lattice (torch_lpm.SquareLattice).  symmetric_only: refuse coefficient sets with
    Kn[i,k] != Kn[q,k^1] in solves (the 2D dense reference assembles the energy-form K, which equals this
    operator only for symmetric sets)."""

    def __init__(self, lat, device="cuda", dtype=torch.float64, mixed=False, pack=True, graph=False,
                 symmetric_only=False):
        dims = tuple(int(v) for v in lat.dims)
        if getattr(lat, "N", None) != int(np.prod(dims)):
            raise NotImplementedError("the MI355X operator supports full-box lattices (no inactive sites)")
        self.lat, self.dev, self.dt = lat, torch.device(device), dtype
        L.require_gpu(torch.empty(0, device=self.dev))
        L.lib()
        self.N = lat.N
        self.dims = dims
        self.d = len(dims)
        self.h = float(lat.h)
        self.mixed, self.pack_coef, self.graph = bool(mixed), bool(pack), bool(graph)
        self.symmetric_only = bool(symmetric_only)
        NN = lat.NN
        nbr = lat.nbr.copy()
        nbr[nbr < 0] = self.N
        self.nbr = torch.as_tensor(nbr, device=self.dev)
        self.valid = torch.as_tensor(lat.valid, device=self.dev, dtype=dtype)
        self.ndir = torch.as_tensor(lat.ndir, device=self.dev, dtype=dtype)
        self.slot = torch.arange(NN, device=self.dev)[None, :]
        self.opp = torch.as_tensor(np.arange(NN) ^ 1, device=self.dev)[None, :]
        self.shell = torch.as_tensor(np.asarray(lat.shell), device=self.dev)
        self.last_info = None
        self.bonds = None      # bond words (B|1, 1, N) int32 of removed particles, or None (full box)

    # ------------------------------------------------------------------------------------------ kernels
    def apply(self, u, Kn, Tv):
        return ops.apply(self.dims, self.h, u, Kn, Tv, bonds=self.bonds)

    def forces(self, u, Kn, Tv):
        o = ops.post(self.dims, self.h, u, Kn, Tv, dl=True, force=True, bonds=self.bonds)
        return o["dl"], o["force"]

    def state(self, u, Kn, Tv):
        """dL (..., N, 18), per-slot own shell sums Si, TSi and neighbour shell sums Sq, TSq."""
        o = ops.post(self.dims, self.h, u, Kn, Tv, dl=True, bonds=self.bonds)
        w = o["shell"]                                                   # (..., N, 4) S0 S1 TS0 TS1
        S, TS = w[..., :2], w[..., 2:]
        Si, TSi = S[..., self.shell], TS[..., self.shell]
        pad = lambda x: torch.cat([x, torch.zeros_like(x[..., :1, :])], -2)
        Sq = pad(Si)[..., self.nbr, self.slot] * self.valid
        TSq = pad(TSi)[..., self.nbr, self.slot] * self.valid
        return o["dl"], Si, TSi, Sq, TSq

    def stress(self, u, Kn, Tv):
        return ops.post(self.dims, self.h, u, Kn, Tv, stress=True, bonds=self.bonds)["stress"]

    def energy(self, u, Kn, Tv):
        return ops.post(self.dims, self.h, u, Kn, Tv, energy=True, bonds=self.bonds)["energy"]

    def block_diag(self, Kn, Tv):
        return ops.block_jacobi(self.dims, self.h, Kn, Tv, raw=True, bonds=self.bonds)

    def precond(self, Kn, Tv, free):
        """Block-Jacobi inverse (..., N, d, d), zero on fixed particles (free: (..., N) float)."""
        return ops.full_from_sym(ops.block_jacobi(self.dims, self.h, Kn, Tv, free=free, bonds=self.bonds))

    # ------------------------------------------------------------------------------------------ PCG
    def _coef(self, Kn, Tv, dtype):
        """Coefficients for the iteration: packed forward planes when the set is symmetric."""
        if self.pack_coef or self.symmetric_only:
            kp, tp, asym = ops.pack(self.dims, Kn, Tv, out_dtype=dtype)
            if int(asym.sum()) == 0:
                return kp, tp, L.PACKED
            if self.symmetric_only:
                raise ValueError("bond parameters are not symmetric (Kn[i,k] != Kn[q,k^1] or Tv likewise)")
        return Kn.to(dtype), Tv.to(dtype), L.DIRECTED

    def pcg(self, b, Kn, Tv, M, free, tol=1e-10, maxit=20000, x0=None, check=10):
        """Solve K_ff x = b for every system of the batch (b zero on fixed, free (..., N)).  Stops each system at
        ||r|| <= tol ||b||.  Returns x, iterations (max over systems), max relative residual."""
        Msym = None if M is None else ops.sym_from_full(M)
        if self.mixed and b.dtype == torch.float64:
            k64, t64, lay64 = self._coef(Kn, Tv, torch.float64)
            k32, t32, lay32 = self._coef(Kn, Tv, torch.float32)
            x, info = ops.pcg_refined(self.dims, self.h, b, k64, t64, k32, t32,
                                      minv32=None if Msym is None else Msym.float(), free=free, tol=tol, maxit=maxit,
                                      x0=x0, check=check, graph=self.graph, layout64=lay64, layout32=lay32,
                                      bonds=self.bonds)
        else:
            kk, tt, lay = self._coef(Kn, Tv, b.dtype)
            x, info = ops.pcg(self.dims, self.h, b, kk, tt, minv=Msym, free=free, tol=tol, maxit=maxit, x0=x0,
                              check=check, graph=self.graph, layout=lay, bonds=self.bonds)
        self.last_info = info
        return x, info.it, info.relmax

    def solve(self, Kn, Tv, up, free, tol=1e-10, maxit=20000, M=None):
        """u with u = up on fixed particles and (K u)_free = 0; up (..., N, 3) zero on free, free (..., N)."""
        b = -self.apply(up, Kn, Tv) * free[..., None]
        if M is None:
            M = self.precond(Kn, Tv, free)
        x, it, rel = self.pcg(b, Kn, Tv, M, free, tol, maxit)
        return up + x, it, rel


class _Bonds:
    """Runs a block with op.bonds set to the bond words a forward pass saw (the op may have been re-masked since)."""

    def __init__(self, op, bonds):
        self.op, self.bonds = op, bonds

    def __enter__(self):
        self.saved = self.op.bonds
        self.op.bonds = self.bonds

    def __exit__(self, *exc):
        self.op.bonds = self.saved


class LinSolve(torch.autograd.Function):
    """u = solve(Kn, Tv) with the exact adjoint gradient w.r.t. Kn, Tv (op: SCT / SCTV; up, free fixed)."""

    @staticmethod
    def forward(ctx, Kn, Tv, op, up, free, tol):
        with torch.no_grad():
            M = op.precond(Kn, Tv, free)
            u, it, rel = op.solve(Kn, Tv, up, free, tol, M=M)
        ctx.save_for_backward(Kn, Tv, u, M)
        ctx.op, ctx.free, ctx.tol, ctx.bonds = op, free, tol, op.bonds
        ctx.info = (it, rel)
        return u

    @staticmethod
    def backward(ctx, g):
        Kn, Tv, u, M = ctx.saved_tensors
        op, free = ctx.op, ctx.free
        with torch.no_grad(), _Bonds(op, ctx.bonds):
            lam, _, _ = op.pcg((g * free[..., None]).contiguous(), Kn, Tv, M, free, ctx.tol)
            gK, gT = ops.bond_grad(op.dims, op.h, Kn, Tv, u, lam, scale=-1.0, bonds=ctx.bonds)
        return gK, gT, None, None, None, None


class SymScatter:
    """Values per undirected bond (..., N, NN/2), one per forward slot 2k, -> symmetric directed (..., N, NN): slot
    2k at p takes (p, k), slot 2k+1 at p takes (q, k) of its neighbour q, for which p sits in slot 2k; zero on
    missing bonds.  Differentiable (a gather), as the reference's scatter of bond parameters into both directed
    slots (coarse3d/train3d.py:77-97, coarse2d/train_cg.py:80-91)."""

    def __init__(self, lat, device):
        N, NN = lat.N, lat.NN
        h = NN // 2
        idx = np.zeros((N, NN), np.int64)
        for k in range(h):
            idx[:, 2 * k] = np.arange(N) * h + k
            q = lat.nbr[:, 2 * k + 1]
            idx[:, 2 * k + 1] = np.where(q >= 0, q * h + k, 0)
        self.idx = torch.as_tensor(idx, device=device)
        self.valid = torch.as_tensor(lat.valid, device=device)

    def __call__(self, v):
        flat = v.reshape(v.shape[:-2] + (-1,))
        return flat[..., self.idx] * self.valid.to(v.dtype)


class LoadSet:
    """Boundary conditions of the load cases on a lattice whose particles belong to blocks IJK."""

    def __init__(self, lat, IJK, nc, s, h, device, dtype, loads=LOADS):
        self.loads = loads
        fr, ups, mov = [], [], []
        for ld in loads:
            fixed, up, moving = prescribed(ld, lat.x, IJK, nc, s, h)
            fr.append(~fixed)
            ups.append(up)
            mov.append(np.zeros(lat.N, bool) if moving is None else moving)
        self.free = torch.as_tensor(np.array(fr), device=device, dtype=dtype)
        self.up = torch.as_tensor(np.array(ups), device=device, dtype=dtype)
        self.moving = torch.as_tensor(np.array(mov), device=device, dtype=dtype)


def _reactions(P, moving, loads, B, like):
    cols = []
    for l, ld in enumerate(loads):
        if ld in PLATES:
            cols.append((P[:, l, :, GDIR[ld]] * moving[l]).sum(-1))
        else:
            cols.append(torch.zeros(B, device=like.device, dtype=like.dtype))
    return torch.stack(cols, 1)


class CoarseT:
    """Coarse SC model: nc^3 particles at the block centres, spacing s h."""

    def __init__(self, nc, s, h=1.0, device="cuda", dtype=torch.float64, loads=LOADS, mixed=False):
        self.nc, self.s, self.h = nc, s, h
        self.lat = LatticeGeom(nc, nc, nc, s * h)
        self.op = SCT(self.lat, device, dtype, mixed=mixed)
        self.N = self.lat.N
        self.loads = loads
        self.bc = LoadSet(self.lat, self.lat.ijk, (nc, nc, nc), s, h, device, dtype, loads)
        self.freecell = self.bc.free.bool()
        surf = [free_surface_cells(ld, (nc, nc, nc)) if ld in PLATES else np.zeros(self.N, bool) for ld in loads]
        self.surf = torch.as_tensor(np.array(surf), device=device)
        self.V = (s * h) ** 3

    def run(self, Kn, Tv, tol=1e-10, differentiable=True):
        """Kn, Tv (B, N, 18) -> dict u (B,L,N,3), sig (B,L,N,6), w (B,L,N), F (B,L)."""
        Kn4, Tv4 = Kn[:, None], Tv[:, None]
        B = Kn.shape[0]
        up, free = self.bc.up[None], self.bc.free[None]
        if differentiable:
            u = LinSolve.apply(Kn4, Tv4, self.op, up.expand(B, -1, -1, -1), free.expand(B, -1, -1), tol)
        else:
            with torch.no_grad():
                u, _, _ = self.op.solve(Kn4, Tv4, up, free, tol)
        o = _PostFn.apply(Kn4, Tv4, u, self.op)
        sig, U, P = o
        F = _reactions(P, self.bc.moving, self.loads, B, Kn)
        return dict(u=u, sig=sig, w=U / self.V, F=F)


class _PostFn(torch.autograd.Function):
    """(stress, energy, K u) of u with gradients w.r.t. Kn, Tv and u from the post-processing VJP kernels
    (replaces autograd through SCT.stress / energy / apply in CoarseT.run, coarse3d/torch_lpm3d.py:212-227)."""

    @staticmethod
    def forward(ctx, Kn, Tv, u, op):
        ctx.set_materialize_grads(False)
        o = ops.post(op.dims, op.h, u, Kn, Tv, ku=True, stress=True, energy=True, bonds=op.bonds)
        ctx.save_for_backward(Kn, Tv, u)
        ctx.op, ctx.bonds = op, op.bonds
        return o["stress"], o["energy"], o["ku"]

    @staticmethod
    def backward(ctx, gs, ge, gp):
        Kn, Tv, u = ctx.saved_tensors
        op = ctx.op
        if gs is None and ge is None and gp is None:
            return None, None, None, None
        with torch.no_grad():
            gu, gK, gT = ops.post_vjp(op.dims, op.h, u, Kn, Tv, g_ku=gp, g_stress=gs, g_energy=ge, bonds=ctx.bonds)
        gu = gu.sum_to_size(u.shape) if ctx.needs_input_grad[2] else None
        return (gK if ctx.needs_input_grad[0] else None, gT if ctx.needs_input_grad[1] else None, gu, None)


class FineRefT:
    """Fine reference (no grad) of an n^3 voxel box and block size s: block means of u, stress, energy density and
    the grip reactions of the load cases, all samples of a batch at once."""

    def __init__(self, n, s, h=1.0, device="cuda", dtype=torch.float64, loads=LOADS, mixed=False):
        self.n, self.s, self.h = n, s, h
        self.nc = n // s
        self.lat = LatticeGeom(n, n, n, h)
        self.op = SCT(self.lat, device, dtype, mixed=mixed)
        IJK = self.lat.ijk // s
        self.bc = LoadSet(self.lat, IJK, (self.nc,) * 3, s, h, device, dtype, loads)
        self.loads = loads
        cid = (IJK[:, 2] * self.nc + IJK[:, 1]) * self.nc + IJK[:, 0]
        self.cid = torch.as_tensor(cid, device=device)
        self.Nc = self.nc ** 3

    def bmean(self, f):
        out = torch.zeros(f.shape[:-2] + (self.Nc, f.shape[-1]), device=f.device, dtype=f.dtype)
        out.index_add_(f.dim() - 2, self.cid, f)
        return out / self.s ** 3

    @torch.no_grad()
    def run(self, Kn, Tv, tol=1e-10):
        Kn4, Tv4 = Kn[:, None], Tv[:, None]
        u, it, rel = self.op.solve(Kn4, Tv4, self.bc.up[None], self.bc.free[None], tol)
        o = ops.post(self.op.dims, self.op.h, u, Kn4, Tv4, ku=True, stress=True, energy=True)
        F = _reactions(o["ku"], self.bc.moving, self.loads, Kn.shape[0], Kn)
        return dict(u=self.bmean(u), sig=self.bmean(o["stress"]), w=self.bmean(o["energy"][..., None])[..., 0] / self.h ** 3,
                    F=F, iters=it, relres=rel)


# =====================================================================================================================
class SCTV(SCT):
    """SC operator with per-specimen removed particles (cement3d/torch_v1.py:18-128): set_mask(solid) with solid
    (B, N) bool (or None) makes a bond exist only if both ends are solid; shell sums, forces, the stress opp-flag
    rule (removed neighbours count as missing bonds), the block diagonal and the gradients all follow.  Removed
    particles have no bonds and must be fixed (free = 0) in solves.  With solid = None it is SCT."""

    def __init__(self, lat, device="cuda", dtype=torch.float64, mixed=False, pack=True, graph=False):
        super().__init__(lat, device, dtype, mixed=mixed, pack=pack, graph=graph)
        self.gvalid = torch.as_tensor(lat.valid, device=self.dev)
        self.l0 = torch.as_tensor(lat.l0, device=self.dev, dtype=dtype)
        self.Vp = lat.V
        self.set_mask(None)

    def set_mask(self, solid):
        """solid: None or (B, N) bool tensor.  Sets the bond words (B, 1, N) and, as the reference exposes them, the
        validity valid (B|1, 1, N, 18) in the operator dtype."""
        if solid is None:
            self.bonds = None
            self.vb = self.gvalid[None, None]
        else:
            solid = torch.as_tensor(solid, device=self.dev)
            if solid.dim() != 2 or solid.shape[1] != self.N:
                raise ValueError("solid must be (B, N) = (B, %d), got %s" % (self.N, tuple(solid.shape)))
            self.bonds = ops.bond_words(self.dims, solid)[:, None]
            bits = torch.arange(self.lat.NN, device=self.dev, dtype=torch.int32)
            self.vb = ((self.bonds[..., None] >> bits) & 1).bool()
        self.valid = self.vb.to(self.dt)
        return self.bonds


class CoarseTV:
    """Coarse SC model (nc^3 grid, spacing s h) with per-specimen cell masks and the 9 load cases
    (cement3d/torch_v1.py:159-199)."""

    def __init__(self, nc, s, h=1.0, device="cuda", dtype=torch.float64, loads=LOADS, mixed=False):
        self.nc, self.s, self.h = nc, s, h
        self.lat = LatticeGeom(nc, nc, nc, s * h)
        self.op = SCTV(self.lat, device, dtype, mixed=mixed)
        self.N = self.lat.N
        self.loads = loads
        self.bc = LoadSet(self.lat, self.lat.ijk, (nc, nc, nc), s, h, device, dtype, loads)
        self.freecell = self.bc.free.bool()
        surf = [free_surface_cells(ld, (nc, nc, nc)) if ld in PLATES else np.zeros(self.N, bool) for ld in loads]
        self.surf = torch.as_tensor(np.array(surf), device=device)
        self.V = (s * h) ** 3
        self.dt = dtype
        self.info = None

    def masks(self, cell):
        """cell (B, N) bool or None -> free (B|1, L, N), up (B|1, L, N, 3); removed cells are fixed at u = 0."""
        if cell is None:
            return self.bc.free[None], self.bc.up[None]
        cm = torch.as_tensor(cell, device=self.bc.free.device).to(self.dt)[:, None]
        return self.bc.free[None] * cm, self.bc.up[None] * cm[..., None]

    def run(self, Kn, Tv, cell=None, tol=1e-10, differentiable=True, maxit=20000):
        """Kn, Tv (B, N, 18), cell (B, N) bool or None -> dict u (B,L,N,3), sig (B,L,N,6), w (B,L,N), F (B,L).
        Kn, Tv are zeroed on missing bonds (differentiably) as in the reference."""
        self.op.set_mask(cell)
        Kn = Kn * self.op.valid[:, 0]
        Tv = Tv * self.op.valid[:, 0]
        Kn4, Tv4 = Kn[:, None], Tv[:, None]
        free, up = self.masks(cell)
        B = Kn.shape[0]
        if differentiable:
            u = LinSolve.apply(Kn4, Tv4, self.op, up.expand(B, -1, -1, -1), free.expand(B, -1, -1), tol)
        else:
            with torch.no_grad():
                u, it, rel = self.op.solve(Kn4, Tv4, up, free, tol, maxit=maxit)
                self.info = (it, rel)
        sig, U, P = _PostFn.apply(Kn4, Tv4, u, self.op)
        F = _reactions(P, self.bc.moving, self.loads, B, Kn)
        return dict(u=u, sig=sig, w=U / self.V, F=F)


class FineRefTV:
    """Fine reference with voids (cement3d/torch_v1.py:202-235): block sums of stress and energy divided by s^3
    (voids count as zero), displacement block means over the solid particles, grip reactions.  No grad."""

    def __init__(self, n, s, h=1.0, device="cuda", dtype=torch.float64, loads=LOADS, mixed=False):
        self.n, self.s, self.h = n, s, h
        self.nc = n // s
        self.lat = LatticeGeom(n, n, n, h)
        self.op = SCTV(self.lat, device, dtype, mixed=mixed)
        IJK = self.lat.ijk // s
        self.bc = LoadSet(self.lat, IJK, (self.nc,) * 3, s, h, device, dtype, loads)
        self.loads = loads
        cid = (IJK[:, 2] * self.nc + IJK[:, 1]) * self.nc + IJK[:, 0]
        self.cid = torch.as_tensor(cid, device=device)
        self.Nc = self.nc ** 3
        self.dt = dtype

    def bsum(self, f):
        out = torch.zeros(f.shape[:-2] + (self.Nc, f.shape[-1]), device=f.device, dtype=f.dtype)
        out.index_add_(f.dim() - 2, self.cid, f)
        return out

    @torch.no_grad()
    def run(self, Kn, Tv, solid=None, tol=1e-10, maxit=40000):
        self.op.set_mask(solid)
        Kn = Kn * self.op.valid[:, 0]
        Tv = Tv * self.op.valid[:, 0]
        Kn4, Tv4 = Kn[:, None], Tv[:, None]
        if solid is None:
            free, up = self.bc.free[None], self.bc.up[None]
            sm = torch.ones(1, 1, self.lat.N, device=Kn.device, dtype=self.dt)
        else:
            sm = torch.as_tensor(solid, device=Kn.device).to(self.dt)[:, None]
            free, up = self.bc.free[None] * sm, self.bc.up[None] * sm[..., None]
        u, it, rel = self.op.solve(Kn4, Tv4, up, free, tol, maxit=maxit)
        u = u * sm[..., None]
        o = ops.post(self.op.dims, self.op.h, u, Kn4, Tv4, ku=True, stress=True, energy=True, bonds=self.op.bonds)
        F = _reactions(o["ku"], self.bc.moving, self.loads, Kn.shape[0], Kn)
        cnt = self.bsum(sm[..., None])
        um = self.bsum(u) / cnt.clamp_min(1.0)
        return dict(u=um, sig=self.bsum(o["stress"]) / self.s ** 3,
                    w=self.bsum(o["energy"][..., None])[..., 0] / self.s ** 3 / self.h ** 3, F=F, iters=it,
                    relres=rel)
