"""NumPy-facing 3D LPM with the interface of coarse3d/lin3d.py, computed by the HIP kernels of liblpm.so.

Inputs and outputs are host NumPy arrays as in the reference (u (N, 3), Kn / Tv (N, 18) FP64); every call copies its
inputs to the GPU, runs the kernels and copies the result back, so the PCIe transfer is part of each call.  Batched
or resident work goes through lpm_ops / torch_lpm3d instead.  A lattice with an `active` mask (only active sites
numbered, coarse3d/geom3d.py:38-65) runs on its full box with the inactive sites removed by bond words: compact
arrays are scattered into the box on the GPU and the results gathered back.

The assembled K (stiffness) is probed from the batched matrix-free operator into the reference's block pattern
(coarse3d/lin3d.py:306-447); solve_direct keeps the reference's SciPy SuperLU solve (coarse3d/lin3d.py:535-546),
the FP64 oracle path of the reference.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import lpm_ops as ops
from geom3d import (SC_OFFS, SC_SHELL, NSLOT, DIR_OF_SLOT, FWD_SLOTS, LatticeGeom, iso_C, sc_params_from_C,  # noqa: F401
                    bond_params_two_end_mean, affine_disp, unit_strain, VOIGT_PAIRS)


def _dev(device):
    if device is None:
        device = "cuda"
    return torch.device(device)


class Lattice3D(LatticeGeom):
    """SC lattice geometry (geom3d.LatticeGeom) with the operator kernels on the GPU."""

    def __init__(self, nx, ny, nz, h=1.0, origin=(0.0, 0.0, 0.0), active=None, device=None):
        super().__init__(nx, ny, nz, h, origin, active)
        self.device = _dev(device)
        self.Nbox = self.nx * self.ny * self.nz
        if self.full:
            self.words, self._sites = None, None
        else:
            self._sites = torch.as_tensor(self.sites, device=self.device)
            solid = torch.zeros(self.Nbox, dtype=torch.bool, device=self.device)
            solid[self._sites] = True
            self.words = ops.bond_words(self.dims, solid)

    def _t(self, a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.device)

    def _box(self, a, width):
        """Compact (N, width) host array -> (Nbox, width) device tensor, zero on inactive sites."""
        t = self._t(np.reshape(a, (self.N, width)))
        if self.full:
            return t
        out = torch.zeros((self.Nbox, width), device=self.device, dtype=t.dtype)
        out[self._sites] = t
        return out

    def _host(self, t):
        """(Nbox, ...) device tensor -> compact host array."""
        return (t if self.full else t[self._sites]).cpu().numpy()

    def _post(self, u, Kn, Tv, **what):
        return ops.post(self.dims, self.h, self._box(u, 3), self._box(Kn, 18), self._box(Tv, 18), bonds=self.words,
                        **what)

    def apply(self, u, Kn, Tv):
        return self._host(ops.apply(self.dims, self.h, self._box(u, 3), self._box(Kn, 18), self._box(Tv, 18),
                                    bonds=self.words))

    def bond_forces(self, u, Kn, Tv):
        o = self._post(u, Kn, Tv, dl=True, force=True)
        return self._host(o["dl"]), self._host(o["force"])

    def stress(self, u, Kn, Tv):
        """computeStress(): (N, 6) in the order xx, yy, zz, yz, xz, xy."""
        return self._host(self._post(u, Kn, Tv, stress=True)["stress"])

    def energy_per_particle(self, u, Kn, Tv):
        return self._host(self._post(u, Kn, Tv, energy=True)["energy"])

    def block_diag(self, Kn, Tv):
        return self._host(ops.block_jacobi(self.dims, self.h, self._box(Kn, 18), self._box(Tv, 18), raw=True,
                                           bonds=self.words))

    def stiffness(self, Kn, Tv):
        """Assembled K (3N x 3N CSR).  With inactive sites: the rows and columns of the active sites of the box's
        K (entries between two active sites that only a removed site links are stored as zeros)."""
        indptr, indices, data = ops.stiffness_csr(self.dims, self.h, self._box(Kn, 18), self._box(Tv, 18),
                                                  bonds=self.words)
        K = sp.csr_matrix((data, indices, indptr), shape=(3 * self.Nbox, 3 * self.Nbox))
        if self.full:
            return K
        dof = (3 * self.sites[:, None] + np.arange(3)[None, :]).ravel()
        return K[dof][:, dof].tocsr()


def k_block_jacobi(D, free, device=None):
    """Inverse blocks of D (N, 3, 3), zero on fixed particles (coarse3d/lin3d.py:294-301), on the GPU."""
    Dt = torch.as_tensor(np.asarray(D, float), device=_dev(device))
    M = torch.linalg.inv(Dt) * torch.as_tensor(np.asarray(free, bool), device=Dt.device)[:, None, None]
    return M.cpu().numpy()


def solve_pcg(lat, Kn, Tv, fixed, u_presc, tol=1e-10, maxit=20000, M=None, x0=None):
    """u with u = u_presc on fixed particles (fixed (N,) bool) and K_ff u_f = -K_fp u_p, by block-Jacobi PCG on the
    GPU (coarse3d/lin3d.py:521-532).  Returns u (N, 3) and dict(iters, relres)."""
    free = ~np.asarray(fixed, bool)
    up = np.zeros((lat.N, 3))
    up[~free] = np.asarray(u_presc).reshape(lat.N, 3)[~free]
    kn, tv, upt = lat._box(Kn, 18), lat._box(Tv, 18), lat._box(up, 3)
    fr = lat._box(free.astype(float), 1)[:, 0]                              # inactive sites: fixed
    w = lat.words
    b = -ops.apply(lat.dims, lat.h, upt, kn, tv, bonds=w) * fr[:, None]
    if M is None:
        minv = ops.block_jacobi(lat.dims, lat.h, kn, tv, free=fr, bonds=w)
    else:
        minv = ops.sym_from_full(lat._box(np.reshape(M, (lat.N, 9)), 9).reshape(-1, 3, 3))
    xi = None if x0 is None else lat._box(np.where(free[:, None], np.asarray(x0).reshape(lat.N, 3), 0.0), 3)
    x, info = ops.pcg(lat.dims, lat.h, b, kn, tv, minv=minv, free=fr, tol=tol, maxit=maxit, x0=xi, check=1, bonds=w)
    return up + lat._host(x), dict(iters=int(info.iters.max()), relres=float(info.relres.max()))


def solve_direct(K, fixed, u_presc, lu=None):
    """Same problem with the assembled K and SciPy SuperLU (factorisation returned for reuse)."""
    n = K.shape[0]
    fd = np.repeat(np.asarray(fixed, bool), 3)
    free = ~fd
    u = np.zeros(n)
    u[fd] = np.asarray(u_presc).reshape(-1)[fd]
    rhs = -(K[free][:, fd] @ u[fd])
    if lu is None:
        lu = spla.splu(K[free][:, free].tocsc(), permc_spec="COLAMD")
    u[free] = lu.solve(rhs)
    return u.reshape(-1, 3), lu
