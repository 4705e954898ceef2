"""Geometry, load cases and parameter maps of the 2D square-lattice LPM (NumPy).

The slot table, shells, particle numbering (p = j nx + i, x fastest), load-case names and the calcKnTv map are data
formats shared with the reference, so they keep its values (coarse2d/lin2d.py:21-22, 190-219,
coarse2d/torch_lpm.py:12-17, 92-111).  Only the square lattice (LPM-C lattice 0) is on the MI355X path.
"""
import numpy as np

SQ_OFFS = np.array([(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)], np.int64)
SQ_SHELL = np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int64)
EPS = 1e-3
LOADS = ["ring_xx", "ring_yy", "ring_xy", "tb_ty", "tb_sx", "lr_tx"]
BCSET = {"ring_xx": "ring", "ring_yy": "ring", "ring_xy": "ring", "tb_ty": "tb", "tb_sx": "tb", "lr_tx": "lr"}
RING_STRAIN = {"ring_xx": ((1.0, 0.0), (0.0, 0.0)), "ring_yy": ((0.0, 0.0), (0.0, 1.0)),
               "ring_xy": ((0.0, 0.5), (0.5, 0.0))}


class SquareGeom:
    """nx x ny square lattice of spacing h: positions at ((i + 1/2) h, (j + 1/2) h) + origin, 8 slots in reverse
    pairs, n_k = -off_k / |off_k|, neighbour table with -1 outside the box, V = h^2 (unit thickness)."""

    def __init__(self, nx, ny, h, kind="square", origin=(0.0, 0.0)):
        if kind != "square":
            raise NotImplementedError("the MI355X kernels implement the square lattice only (kind=%r)" % (kind,))
        self.kind, self.nx, self.ny, self.h = kind, int(nx), int(ny), float(h)
        p = np.arange(self.nx * self.ny)
        self.ij = np.stack([p % self.nx, p // self.nx], 1).astype(np.int64)
        self.x = (self.ij + 0.5) * self.h + np.asarray(origin, float)[None, :]
        self.N = self.nx * self.ny
        self.V = self.h * self.h
        self.offs, self.shell = SQ_OFFS, SQ_SHELL
        self.NN = len(SQ_OFFS)
        self.opp = np.arange(self.NN) ^ 1
        length = np.sqrt((SQ_OFFS ** 2).sum(1))
        self.l0 = length * self.h
        self.ndir = -SQ_OFFS / length[:, None]
        nb = self.ij[:, None, :] + SQ_OFFS[None]
        ok = (nb[..., 0] >= 0) & (nb[..., 0] < self.nx) & (nb[..., 1] >= 0) & (nb[..., 1] < self.ny)
        self.nbr = np.where(ok, nb[..., 1] * self.nx + nb[..., 0], -1)
        self.valid = ok
        self.nb_initial = ok.sum(1)
        self.site_of = p.copy()
        self.sites = p

    @property
    def dims(self):
        return (self.nx, self.ny)


def prescribed(load, x, IJ, ncx, ncy, s, h, L=1):
    """(fixed (N,), prescribed u (N, 2), moving grip (N,) or None) of a 2D load case on blocks IJ.
    ring_*: the outer L block layers follow EPS * strain * (x - centre); tb_ty / tb_sx: bottom L rows clamped, top L
    rows translated by EPS times the grip distance along y / x; lr_tx: left / right columns along x."""
    centre = np.array([ncx * s * h / 2.0, ncy * s * h / 2.0])
    I, J = IJ[:, 0], IJ[:, 1]
    u = np.zeros((len(x), 2))
    if load in RING_STRAIN:
        fixed = (np.minimum(I, J) < L) | (I >= ncx - L) | (J >= ncy - L)
        u[fixed] = (x[fixed] - centre) @ (EPS * np.array(RING_STRAIN[load])).T
        return fixed, u, None
    if load in ("tb_ty", "tb_sx"):
        lo, hi = J < L, J >= ncy - L
        grip = EPS * (ncy - L) * s * h
        u[hi] = (0.0, grip) if load == "tb_ty" else (grip, 0.0)
    elif load == "lr_tx":
        lo, hi = I < L, I >= ncx - L
        u[hi] = (EPS * (ncx - L) * s * h, 0.0)
    else:
        raise ValueError(load)
    return lo | hi, u, hi


def iso_C(E, nu, plane="strain"):
    """(C11, C12, C44) of isotropic (E, nu) in plane strain or plane stress."""
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu)) if plane == "strain" else E * nu / (1 - nu * nu)
    mu = 0.5 * E / (1 + nu)
    return np.stack([lam + 2 * mu, lam, mu], -1)


def square_params_from_C(C):
    """calcKnTv of the square lattice: (Kn axial, Kn diagonal, Tv) = ((C11 - C12) / 2, C44 / 2, (C12 - C44) / 12)."""
    C = np.asarray(C, float)
    return np.stack([(C[..., 0] - C[..., 1]) / 2.0, C[..., 2] / 2.0, (C[..., 1] - C[..., 2]) / 12.0], -1)


def bond_params_two_end_mean(lat, ppar):
    """Per-particle (Kn axial, Kn diagonal, Tv) -> directed Kn, Tv (N, 8): mean over both ends of each bond."""
    ppar = np.asarray(ppar, float)
    q = np.where(lat.valid, lat.nbr, 0)
    col = np.broadcast_to(np.asarray(lat.shell)[None, :], lat.nbr.shape)
    Kn = 0.5 * (np.take_along_axis(ppar[:, :2], col, 1) + ppar[q, col]) * lat.valid
    Tv = 0.5 * (ppar[:, 2:3] + ppar[q, 2]) * lat.valid
    return np.ascontiguousarray(Kn), np.ascontiguousarray(Tv)


def affine_disp(x, eps, x0=(0.0, 0.0)):
    """u = eps (x - x0)."""
    return (x - np.asarray(x0)[None, :]) @ np.asarray(eps, float).T
