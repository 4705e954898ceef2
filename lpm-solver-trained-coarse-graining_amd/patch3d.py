"""Local patch homogenisation of the 3D SC coarse model on the GPU, with the interface of coarse3d/patch3d.py.

For every coarse cell a window of (3s)^3 voxel particles around it (reflect-padded at the domain boundary), the
window's two-end-mean bonds, the 6 unit strains as affine displacements about the cell centre on the two outer
particle layers, the interior relaxed, and C_ab = (1/s^3) u_a^T K_cell u_b with K_cell the Hessian of the cell
particles' energy (coarse3d/patch3d.py:1-10).  The reference assembles each window's K_ff densely and factorises it
(patch3d.py:69-143, 191-196); here all windows of a chunk and their 6 load cases are one batch of systems for the
matrix-free HIP kernels: the right-hand sides by one batched K u, the interiors by one batched block-Jacobi PCG
(FP64, to rel-res 1e-13), and the bilinear form from the per-particle energies U_i of one batched post-processing
launch, u_a^T K_cell u_b = sum_cell U(u_a + u_b) - U(u_a) - U(u_b) (U is quadratic).
"""
import time

import numpy as np
import torch

import _lpmlib as L
import lpm_ops as ops
from geom3d import SC_OFFS, SC_SHELL, affine_disp, sc_params_from_C, unit_strain

PAIRS = [(a, b) for a in range(6) for b in range(a + 1, 6)]


class PatchGeom:
    """Window geometry of cell size s (patch3d.py:18-38): W = 3s, fixed = the `ring` outer layers, the cell = the
    middle s^3 block, UA (6, W^3, 3) the affine fields of the unit strains about the cell centre (h = 1)."""

    def __init__(self, s, ring=2, device=None):
        self.s = self.m = int(s)
        W = 3 * self.s
        self.W = W
        self.dims = (W, W, W)
        k, j, i = np.meshgrid(np.arange(W), np.arange(W), np.arange(W), indexing="ij")
        ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
        self.N = W ** 3
        self.fixed = np.any((ijk < ring) | (ijk >= W - ring), 1)
        self.free_idx = -np.ones(self.N, np.int64)
        self.free_idx[~self.fixed] = np.arange(int((~self.fixed).sum()))
        self.nfree = int((~self.fixed).sum())
        cell = np.all((ijk >= s) & (ijk < 2 * s), 1)
        self.cell_list = np.nonzero(cell)[0]
        self.xl = (ijk + 0.5 - 1.5 * s).astype(np.float64)
        self.UA = np.stack([affine_disp(self.xl, unit_strain(k_)) for k_ in range(6)], 0)
        nb = ijk[:, None, :] + SC_OFFS[None, :, :]
        ok = np.all((nb >= 0) & (nb < W), 2)
        nbr = (nb[..., 2] * W + nb[..., 1]) * W + nb[..., 0]
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        d = self.device
        self.t_nbr = torch.as_tensor(np.where(ok, nbr, 0), device=d)
        self.t_ok = torch.as_tensor(ok, device=d, dtype=torch.float64)
        self.t_shell = torch.as_tensor(SC_SHELL, device=d)
        self.t_free = torch.as_tensor(~self.fixed, device=d)
        self.t_cell = torch.as_tensor(cell, device=d, dtype=torch.float64)
        self.t_ufix = torch.as_tensor(self.UA * self.fixed[None, :, None], device=d, dtype=torch.float64)

    def bonds(self, PP):
        """Two-end-mean directed Kn, Tv (nw, W^3, 18) of window particle parameters PP (nw, W^3, 3)
        (patch3d.py:44-56); bonds leaving the window are zero."""
        q = self.t_nbr
        col = self.t_shell[None, None, :].expand(PP.shape[0], self.N, 18)
        own_k = torch.gather(PP, 2, col)
        nb_k = torch.gather(PP[:, q.reshape(-1), :].reshape(PP.shape[0], self.N, 18, 3), 3, col[..., None])[..., 0]
        own_t = PP[:, :, 2:3].expand(-1, -1, 18)
        nb_t = PP[:, q.reshape(-1), 2].reshape(PP.shape[0], self.N, 18)
        return 0.5 * (own_k + nb_k) * self.t_ok, 0.5 * (own_t + nb_t) * self.t_ok


_GEOM = {}


def _geom(s, device):
    key = (int(s), str(device))
    if key not in _GEOM:
        _GEOM[key] = PatchGeom(s, device=device)
    return _GEOM[key]


def patch_C(Cvox, n, s, chunk=4096, timing=None, device=None, tol=1e-13):
    """Cvox (n^3, 3) voxel (C11, C12, C44) -> per-cell homogenised C (Nc, 6, 6) (Voigt, engineering shear), cells
    in lattice order (patch3d.py:211-243).  `chunk` windows form one batch of 6 x chunk systems."""
    t0 = time.perf_counter()
    dev = torch.device(device) if device is not None else torch.device("cuda")
    g = _geom(s, dev)
    nc, W, N = n // s, g.W, g.N
    pp = sc_params_from_C(np.asarray(Cvox, float), 0.5).reshape(n, n, n, 3)
    pad = torch.as_tensor(np.pad(pp, ((s, s), (s, s), (s, s), (0, 0)), mode="reflect"), device=dev)
    kk, jj, ii = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij")
    starts = torch.as_tensor(np.stack([kk.ravel(), jj.ravel(), ii.ravel()], 1) * s, device=dev)
    off = torch.as_tensor(np.stack(np.meshgrid(np.arange(W), np.arange(W), np.arange(W), indexing="ij"),
                                   -1).reshape(-1, 3), device=dev)
    Nc = nc ** 3
    C = torch.zeros((Nc, 6, 6), device=dev, dtype=torch.float64)
    ufix = g.t_ufix
    for c0 in range(0, Nc, chunk):
        st = starts[c0:c0 + chunk]
        idx = st[:, None, :] + off[None, :, :]
        PP = pad[idx[..., 0], idx[..., 1], idx[..., 2]]                               # (nw, W^3, 3)
        Kn, Tv = g.bonds(PP)
        Kn, Tv = Kn[:, None], Tv[:, None]                                              # (nw, 1, N, 18)
        b = -ops.apply(g.dims, 1.0, ufix, Kn, Tv) * g.t_free[:, None].to(torch.float64)
        minv = ops.block_jacobi(g.dims, 1.0, Kn, Tv, free=g.t_free)
        x, info = ops.pcg(g.dims, 1.0, b, Kn, Tv, minv=minv, free=g.t_free, tol=tol, maxit=20 * 3 * g.nfree,
                          check=5)
        if not bool(np.all(info.conv)):
            raise FloatingPointError("patch solve did not converge (max relres %.2e)" % info.relmax)
        u = ufix + x                                                                   # (nw, 6, N, 3)
        U = torch.cat([u, u[:, [a for a, _ in PAIRS]] + u[:, [b_ for _, b_ in PAIRS]]], 1)
        E = ops.post(g.dims, 1.0, U, Kn, Tv, energy=True)["energy"]                   # (nw, 21, N)
        Ec = (E * g.t_cell).sum(-1)
        Cc = torch.zeros((Ec.shape[0], 6, 6), device=dev, dtype=torch.float64)
        ar = torch.arange(6, device=dev)
        Cc[:, ar, ar] = 2.0 * Ec[:, :6]
        for m, (a, b_) in enumerate(PAIRS):
            v = Ec[:, 6 + m] - Ec[:, a] - Ec[:, b_]
            Cc[:, a, b_] = v
            Cc[:, b_, a] = v
        C[c0:c0 + st.shape[0]] = Cc / float(s ** 3)
    out = C.cpu().numpy()
    if timing is not None:
        timing["patch"] = time.perf_counter() - t0
    return out
