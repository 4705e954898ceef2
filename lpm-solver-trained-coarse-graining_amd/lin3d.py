"""NumPy-facing 3D LPM with the interface of coarse3d/lin3d.py, computed by the HIP kernels of liblpm.so.

Inputs and outputs are host NumPy arrays as in the reference (u (N, 3), Kn / Tv (N, 18) FP64); every call copies its
inputs to the GPU, runs the kernels and copies the result back, so the PCIe transfer is part of each call.  Batched
or resident work goes through lpm_ops / torch_lpm3d instead.  Lattices are full boxes.

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
        if not self.full:
            raise NotImplementedError("the MI355X kernels work on full-box lattices (no inactive sites)")
        self.device = _dev(device)

    def _t(self, a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.device)

    def apply(self, u, Kn, Tv):
        return ops.apply(self.dims, self.h, self._t(np.reshape(u, (self.N, 3))), self._t(Kn), self._t(Tv)).cpu().numpy()

    def bond_forces(self, u, Kn, Tv):
        o = ops.post(self.dims, self.h, self._t(np.reshape(u, (self.N, 3))), self._t(Kn), self._t(Tv), dl=True,
                     force=True)
        return o["dl"].cpu().numpy(), o["force"].cpu().numpy()

    def stress(self, u, Kn, Tv):
        """computeStress(): (N, 6) in the order xx, yy, zz, yz, xz, xy."""
        o = ops.post(self.dims, self.h, self._t(np.reshape(u, (self.N, 3))), self._t(Kn), self._t(Tv), stress=True)
        return o["stress"].cpu().numpy()

    def energy_per_particle(self, u, Kn, Tv):
        o = ops.post(self.dims, self.h, self._t(np.reshape(u, (self.N, 3))), self._t(Kn), self._t(Tv), energy=True)
        return o["energy"].cpu().numpy()

    def block_diag(self, Kn, Tv):
        return ops.block_jacobi(self.dims, self.h, self._t(Kn), self._t(Tv), raw=True).cpu().numpy()

    def stiffness(self, Kn, Tv):
        """Assembled K (3N x 3N CSR)."""
        indptr, indices, data = ops.stiffness_csr(self.dims, self.h, self._t(Kn), self._t(Tv))
        return sp.csr_matrix((data, indices, indptr), shape=(3 * self.N, 3 * self.N))


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
    t = lat._t
    kn, tv, upt, fr = t(Kn), t(Tv), t(up), t(free.astype(float))
    b = -ops.apply(lat.dims, lat.h, upt, kn, tv) * fr[:, None]
    if M is None:
        minv = ops.block_jacobi(lat.dims, lat.h, kn, tv, free=fr)
    else:
        minv = ops.sym_from_full(t(M))
    xi = None if x0 is None else t(np.where(free[:, None], np.asarray(x0).reshape(lat.N, 3), 0.0))
    x, info = ops.pcg(lat.dims, lat.h, b, kn, tv, minv=minv, free=fr, tol=tol, maxit=maxit, x0=xi, check=1)
    return up + x.cpu().numpy(), dict(iters=int(info.iters.max()), relres=float(info.relres.max()))


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
