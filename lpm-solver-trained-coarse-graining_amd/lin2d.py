"""NumPy-facing 2D square-lattice LPM with the interface of coarse2d/lin2d.py, computed by the HIP kernels.

Energy of particle i: U_i = 1/2 sum_k Kn_ik dL_ik^2 + 1/2 sum_c (sum_{k in c} Tv_ik dL_ik)(sum_{k in c} dL_ik),
dL_ik = n_k.(u_i - u_q).  Host arrays in and out (u flat (2N,) or (N, 2), Kn / Tv (N, 8)); each call copies to the
GPU and back.  Lattices are full boxes (active=None) and every bond alive (alive=None).

stiffness() returns (K, B, H) like the reference (coarse2d/lin2d.py:113-118): K is probed from the matrix-free
operator on the GPU (equal to the reference's B^T H B for symmetric parameters); B (bond incidence) and H (bond-space
Hessian) are sparse structure matrices built on the host.  solve_disp keeps the reference's SciPy SuperLU solve.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import lpm_ops as ops
from geom2d import (SQ_OFFS, SQ_SHELL, SquareGeom, iso_C, square_params_from_C, bond_params_two_end_mean,  # noqa: F401
                    affine_disp)


# descending tolerance list of the multi-tolerance solves (timing/t2d/solver2d.py:27)
TOLS = np.array(sorted({m * 10.0 ** k for k in range(-6, 0) for m in (1, 1.5, 2, 3, 4, 5, 7)}, reverse=True))


class Lattice2D(SquareGeom):
    """Square lattice of nx x ny sites with spacing h (coarse2d/lin2d.py:29-74)."""

    def __init__(self, kind, nx, ny, h, origin=(0.0, 0.0), active=None, device=None):
        if active is not None and not np.all(active):
            raise NotImplementedError("the MI355X kernels work on full-box lattices (no inactive sites)")
        super().__init__(nx, ny, h, kind, origin)
        self.device = torch.device(device or "cuda")

    def _t(self, a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.device)

    @staticmethod
    def _alive(alive):
        if alive is not None and not np.all(alive):
            raise NotImplementedError("broken bonds (alive masks) are not on the MI355X path")

    def _post(self, u, Kn, Tv, **kw):
        return ops.post(self.dims, self.h, self._t(np.reshape(u, (self.N, 2))), self._t(Kn), self._t(Tv), **kw)

    # ------------------------------------------------------------------------------------------ structure
    def bond_matrix(self, alive=None):
        """B (8N x 2N): row (i, k) maps u to dL_ik = n_k.(u_i - u_q)."""
        self._alive(alive)
        p, k = np.nonzero(self.valid)
        rows = p * self.NN + k
        q = self.nbr[p, k]
        r = np.concatenate([rows, rows, rows, rows])
        c = np.concatenate([2 * p, 2 * p + 1, 2 * q, 2 * q + 1])
        v = np.concatenate([self.ndir[k, 0], self.ndir[k, 1], -self.ndir[k, 0], -self.ndir[k, 1]])
        return sp.csr_matrix((v, (r, c)), shape=(self.N * self.NN, 2 * self.N))

    def bond_hessian(self, Kn, Tv, alive=None):
        """Bond-space Hessian H (8N x 8N) of sum_i U_i: Kn_ik on the diagonal plus 1/2 (Tv_ia + Tv_ib) between the
        bonds a, b of one (particle, shell) group."""
        self._alive(alive)
        Kn, Tv = np.asarray(Kn, float), np.asarray(Tv, float)
        R, C, V = [], [], []
        p, k = np.nonzero(self.valid)
        R.append(p * self.NN + k)
        C.append(p * self.NN + k)
        V.append(Kn[p, k])
        for a in range(self.NN):
            for b in range(self.NN):
                if self.shell[a] != self.shell[b]:
                    continue
                m = np.nonzero(self.valid[:, a] & self.valid[:, b])[0]
                R.append(m * self.NN + a)
                C.append(m * self.NN + b)
                V.append(0.5 * (Tv[m, a] + Tv[m, b]))
        n = self.N * self.NN
        return sp.csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(n, n))

    def stiffness(self, Kn, Tv, alive=None):
        self._alive(alive)
        indptr, indices, data = ops.stiffness_csr(self.dims, self.h, self._t(Kn), self._t(Tv))
        K = sp.csr_matrix((data, indices, indptr), shape=(2 * self.N, 2 * self.N))
        return K, self.bond_matrix(), self.bond_hessian(Kn, Tv)

    # ------------------------------------------------------------------------------------------ kernels
    def bond_forces(self, u, Kn, Tv, alive=None):
        """dL (N, 8) and bond force F (N, 8)."""
        self._alive(alive)
        o = self._post(u, Kn, Tv, dl=True, force=True)
        return o["dl"].cpu().numpy(), o["force"].cpu().numpy()

    def stress(self, u, Kn, Tv, alive=None):
        """computeStress(): (N, 3) xx, yy, xy."""
        self._alive(alive)
        return self._post(u, Kn, Tv, stress=True)["stress"].cpu().numpy()

    def energy_per_particle(self, u, Kn, Tv, alive=None):
        self._alive(alive)
        return self._post(u, Kn, Tv, energy=True)["energy"].cpu().numpy()

    def apply(self, u, Kn, Tv):
        """K u as a flat (2N,) vector."""
        out = ops.apply(self.dims, self.h, self._t(np.reshape(u, (self.N, 2))), self._t(Kn), self._t(Tv))
        return out.cpu().numpy().reshape(-1)

    def solve_mf(self, Kn, Tv, bcs, starts=None, tols=TOLS, maxit=20000, fp32=False, timing=None):
        """Matrix-free 2x2 block-Jacobi PCG with snapshots at every tolerance of `tols` (timing/t2d/solver2d.py:
        577-613).  bcs {load: (fixed (N,) bool, ub (2N,) prescribed displacement)}; starts {load: u0 (2N,)}
        optional initial guesses.  All loads run as one batched launch, each with its own mask and preconditioner.
        fp32=True stores Kn, Tv rounded to FP32 as the reference's mf_arrays(fp32=True) does; the vector arithmetic
        stays FP64, as in the reference.  Returns {load: (snaps (ntol, 2N) with ub added, iters (ntol,))}."""
        import time
        t0 = time.perf_counter()
        loads = list(bcs)
        N = self.N
        kn, tv = np.asarray(Kn, np.float64), np.asarray(Tv, np.float64)
        if fp32:
            kn, tv = kn.astype(np.float32).astype(np.float64), tv.astype(np.float32).astype(np.float64)
        K, T = self._t(kn), self._t(tv)
        free = np.stack([~np.asarray(bcs[ld][0], bool) for ld in loads])                   # (L, N)
        ub = np.stack([np.reshape(bcs[ld][1], (N, 2)) for ld in loads])                   # (L, N, 2)
        fr = torch.as_tensor(free, device=self.device)
        ubt = self._t(ub)
        b = -ops.apply(self.dims, self.h, ubt, K, T) * fr[..., None]
        x0 = None
        if starts:
            x0 = torch.zeros_like(b)
            for l, ld in enumerate(loads):
                if ld in starts:
                    x0[l] = (self._t(np.reshape(starts[ld], (N, 2))) - ubt[l]) * fr[l][..., None]
        minv = ops.block_jacobi(self.dims, self.h, K, T, free=fr)
        t1 = time.perf_counter()
        x, info = ops.pcg(self.dims, self.h, b, K, T, minv=minv, free=fr, maxit=maxit, x0=x0, tols=tols)
        snaps = (info.snaps + ubt[:, None]).cpu().numpy()
        if timing is not None:
            timing.update(setup=t1 - t0, iter=time.perf_counter() - t1)
        return {ld: (snaps[l].reshape(len(tols), 2 * N), info.snap_iters[0, l].astype(np.int64))
                for l, ld in enumerate(loads)}


def solve_disp(K, fixed_dofs, u_fixed, f_ext=None, lu_cache=None):
    """K u = f with u prescribed on fixed_dofs (SciPy SuperLU); returns the full u and the factorisation."""
    n = K.shape[0]
    fixed = np.zeros(n, bool)
    fixed[np.asarray(fixed_dofs)] = True
    free = ~fixed
    u = np.zeros(n)
    u[np.asarray(fixed_dofs)] = u_fixed
    rhs = -(K[free][:, fixed] @ u[fixed])
    if f_ext is not None:
        rhs = rhs + np.asarray(f_ext)[free]
    if lu_cache is None:
        lu_cache = spla.splu(K[free][:, free].tocsc(), permc_spec="COLAMD")
    u[free] = lu_cache.solve(rhs)
    return u, lu_cache
