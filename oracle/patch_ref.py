"""CPU oracle of the local patch homogenisation of the 3D SC coarse model (NumPy / SciPy, FP64).

TEST INFRASTRUCTURE ONLY: only `tests/` (and the cpu_baseline leg of bench.py) use it, as the checker.  Restates
coarse3d/patch3d.py for small lattices: for every coarse cell a (3s)^3 window of the reflect-padded voxel parameters
(patch3d.py:217-232), the window's two-end-mean bonds (patch3d.py:44-56), the 6 unit strains as affine displacements
about the cell centre on the two outer particle layers (patch3d.py:18-38), the interior relaxed by a direct solve of
K_ff x = -(K u_a)_f (patch3d.py:69-143, 191-196) and C_ab = (1/s^3) sum over the cell particles of
sum_k Kn dL_a dL_b + 1/2 sum_c (TS_a S_b + TS_b S_a) (patch3d.py:146-188).
"""
import numpy as np
import scipy.linalg as sla

import lpm_ref as R

VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def unit_strain(k):
    a, b = VOIGT_PAIRS[k]
    e = np.zeros((3, 3))
    e[a, b] = e[b, a] = 1.0 if a == b else 0.5
    return e


def sc_params_from_C(C, r):
    """(Kn axial, Kn diagonal, Tv) = r (C11 - C12 - C44, C44, (C12 - C44) / 18) (coarse3d/geom3d.py:161-164)."""
    C = np.asarray(C, float)
    return r * np.stack([C[..., 0] - C[..., 1] - C[..., 2], C[..., 2], (C[..., 1] - C[..., 2]) / 18.0], -1)


def iso_C66(C3):
    C3 = np.atleast_2d(np.asarray(C3, float))
    C = np.zeros((len(C3), 6, 6))
    C[:, :3, :3] = C3[:, 1, None, None]
    i = np.arange(3)
    C[:, i, i] = C3[:, 0, None]
    C[:, i + 3, i + 3] = C3[:, 2, None]
    return C


class PatchGeom:
    """Window lattice of cell size s (patch3d.py:18-38): W = 3s, fixed = the `ring` outer layers, cell = the middle
    s^3 block, UA (6, W^3, 3) the affine fields of the unit strains about the cell centre."""

    def __init__(self, s, ring=2):
        self.s, self.W = s, 3 * s
        self.lat = R.Box((self.W,) * 3)
        ijk = self.lat.ijk
        self.fixed = np.any((ijk < ring) | (ijk >= self.W - ring), 1)
        self.cell = np.all((ijk >= s) & (ijk < 2 * s), 1)
        xl = ijk + 0.5 - 1.5 * s
        self.UA = np.stack([xl @ unit_strain(k).T for k in range(6)])


def windows(Cvox, n, s):
    """Per-cell window particle parameters (Nc, W^3, 3) from voxel (C11, C12, C44) (n^3, 3), reflect-padded, in the
    reference's cell and window orders (patch3d.py:215-232)."""
    pp = sc_params_from_C(Cvox, 0.5).reshape(n, n, n, 3)
    pad = np.pad(pp, ((s, s), (s, s), (s, s), (0, 0)), mode="reflect")
    nc, W = n // s, 3 * s
    kk, jj, ii = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij")
    starts = np.stack([kk.ravel(), jj.ravel(), ii.ravel()], 1) * s
    off = np.stack(np.meshgrid(np.arange(W), np.arange(W), np.arange(W), indexing="ij"), -1).reshape(-1, 3)
    idx = starts[:, None, :] + off[None, :, :]
    return pad[idx[..., 0], idx[..., 1], idx[..., 2]]


def cell_C(geom, PP):
    """Homogenised C (6, 6) of one window (patch3d.py:69-188, dense solve)."""
    lat = geom.lat
    Kn, Tv = R.two_end_mean(lat, PP)
    K = R.stiffness(lat, Kn, Tv).toarray()
    fd = np.repeat(~geom.fixed, 3)
    us = []
    for a in range(6):
        u = geom.UA[a] * geom.fixed[:, None]
        b = -(K @ u.reshape(-1))[fd]
        x = sla.solve(K[np.ix_(fd, fd)], b, assume_a="pos")
        u = u.reshape(-1).copy()
        u[fd] = x
        us.append(u.reshape(-1, 3))
    C = np.zeros((6, 6))
    cell = geom.cell
    parts = [R.shell_sums(lat, u, Tv) for u in us]
    for a in range(6):
        for b in range(a, 6):
            dla, sa, ta = parts[a]
            dlb, sb, tb = parts[b]
            e = (Kn * dla * dlb)[cell].sum() + 0.5 * (ta * sb + tb * sa)[cell].sum()
            C[a, b] = C[b, a] = e / geom.s ** 3
    return C


def patch_C(Cvox, n, s):
    """Cvox (n^3, 3) -> per-cell C (Nc, 6, 6) (patch3d.py:211-243)."""
    geom = PatchGeom(s)
    PP = windows(Cvox, n, s)
    return np.stack([cell_C(geom, P) for P in PP])
