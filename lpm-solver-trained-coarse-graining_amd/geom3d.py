"""Geometry, load cases and parameter maps of the 3D simple-cubic LPM (NumPy), interface of coarse3d/geom3d.py.

The slot table, shells, particle numbering, load-case names, Voigt order and the calcKnTv maps are data formats
shared with the reference, so they keep the reference's values (coarse3d/geom3d.py:16-32, 153-227).  Lattice
neighbour tables are built from the box directly (no per-site Python loops).
"""
import numpy as np

SC_OFFS = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0),
                    (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
                    (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)], np.int64)
SC_SHELL = np.repeat(np.array([0, 1], np.int64), [6, 12])
NSLOT = 18
DIR_OF_SLOT = np.arange(NSLOT) // 2
FWD_SLOTS = np.arange(0, NSLOT, 2)

EPS = 1e-3
LOADS = ["k_xx", "k_yy", "k_zz", "k_yz", "k_xz", "k_xy", "tb_tz", "tb_sx", "lr_tx"]
RING = LOADS[:6]
PLATES = LOADS[6:]
BCSET = dict({ld: "ring" for ld in RING}, tb_tz="tb", tb_sx="tb", lr_tx="lr")
GDIR = {"tb_tz": 2, "tb_sx": 0, "lr_tx": 0}
VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


class LatticeGeom:
    """nx x ny x nz simple-cubic lattice of spacing h, particle id p = (k ny + j) nx + i over active sites."""

    def __init__(self, nx, ny, nz, h=1.0, origin=(0.0, 0.0, 0.0), active=None):
        self.nx, self.ny, self.nz, self.h = int(nx), int(ny), int(nz), float(h)
        nsite = self.nx * self.ny * self.nz
        act = np.ones(nsite, bool) if active is None else np.asarray(active, bool).ravel()
        self.full = bool(act.all())
        self.site_of = np.full(nsite, -1, np.int64)
        self.site_of[act] = np.arange(int(act.sum()))
        self.sites = np.flatnonzero(act)
        s = self.sites
        ijk = np.stack([s % self.nx, (s // self.nx) % self.ny, s // (self.nx * self.ny)], 1).astype(np.int64)
        self.ijk = np.ascontiguousarray(ijk)
        self.x = (self.ijk + 0.5) * self.h + np.asarray(origin, float)[None, :]
        self.N = len(s)
        self.V = self.h ** 3
        self.offs, self.shell, self.NN = SC_OFFS, SC_SHELL, NSLOT
        self.opp = np.arange(NSLOT) ^ 1
        length = np.sqrt((SC_OFFS ** 2).sum(1))
        self.l0 = length * self.h
        self.ndir = np.ascontiguousarray(-SC_OFFS / length[:, None])
        nb = self.ijk[:, None, :] + SC_OFFS[None]
        dims = np.array([self.nx, self.ny, self.nz])
        ok = np.all((nb >= 0) & (nb < dims), axis=2)
        lin = (nb[..., 2] * self.ny + nb[..., 1]) * self.nx + nb[..., 0]
        q = self.site_of[np.where(ok, lin, 0)]
        ok &= q >= 0
        self.nbr = np.ascontiguousarray(np.where(ok, q, -1))
        self.valid = ok
        self.nb_initial = ok.sum(1)

    @property
    def dims(self):
        return (self.nx, self.ny, self.nz)


def unit_strain(k):
    """Voigt component k as a tensor strain, engineering shear (gamma = 1 gives eps_ij = 1/2)."""
    a, b = VOIGT_PAIRS[k]
    e = np.zeros((3, 3))
    e[a, b] = e[b, a] = 1.0 if a == b else 0.5
    return e


def affine_disp(x, eps, x0=(0.0, 0.0, 0.0)):
    return (x - np.asarray(x0)[None, :]) @ np.asarray(eps, float).T


def prescribed(load, x, IJK, nc, s, h=1.0):
    """(fixed (N,), prescribed u (N, 3), moving grip (N,) or None) of a load case on blocks IJK, nc blocks per axis.
    KUBC ring: outer block layer follows EPS * strain * (x - box centre); plates: bottom / left layer clamped, the
    opposite layer translated by EPS times the grip distance."""
    ncx, ncy, ncz = nc
    I, J, K = IJK[:, 0], IJK[:, 1], IJK[:, 2]
    u = np.zeros((len(x), 3))
    if load in RING:
        fixed = (np.minimum(np.minimum(I, J), K) < 1) | (I >= ncx - 1) | (J >= ncy - 1) | (K >= ncz - 1)
        centre = 0.5 * np.array([ncx, ncy, ncz], float) * s * h
        u[fixed] = affine_disp(x[fixed], EPS * unit_strain(RING.index(load)), centre)
        return fixed, u, None
    if load in ("tb_tz", "tb_sx"):
        lo, hi = K < 1, K >= ncz - 1
        grip = EPS * (ncz - 1) * s * h
        d = (0.0, 0.0, grip) if load == "tb_tz" else (grip, 0.0, 0.0)
    elif load == "lr_tx":
        lo, hi = I < 1, I >= ncx - 1
        d = (EPS * (ncx - 1) * s * h, 0.0, 0.0)
    else:
        raise ValueError(load)
    u[hi] = d
    return lo | hi, u, hi


def free_surface_cells(load, nc):
    """Free cells of a plate case on a free lateral face."""
    I, J, K = cell_ijk(nc).T
    ncx, ncy, ncz = nc
    if load.startswith("tb"):
        grip = (K < 1) | (K >= ncz - 1)
        face = (I < 1) | (I >= ncx - 1) | (J < 1) | (J >= ncy - 1)
    else:
        grip = (I < 1) | (I >= ncx - 1)
        face = (J < 1) | (J >= ncy - 1) | (K < 1) | (K >= ncz - 1)
    return face & ~grip


def cell_ijk(nc):
    ncx, ncy, ncz = nc
    c = np.arange(ncx * ncy * ncz)
    return np.stack([c % ncx, (c // ncx) % ncy, c // (ncx * ncy)], 1)


def block_index(ijk, s, nc):
    B = ijk // s
    return (B[:, 2] * nc[1] + B[:, 1]) * nc[0] + B[:, 0]


def interface_cells(phase, n, s):
    """Cells (nc^3, x fastest) whose s^3 voxel block holds a voxel with a face neighbour of the other phase
    (coarse3d/geom3d.py:134-147).  The field is padded by repeating its faces, so a box face never counts as an
    interface, and compared with its six unit shifts."""
    p = np.asarray(phase).reshape(n, n, n)
    q = np.pad(p, 1, mode="edge")
    core = (slice(1, -1),) * 3
    hit = np.zeros(p.shape, bool)
    for ax in range(3):
        for sh in (-1, 1):
            hit |= np.roll(q, sh, axis=ax)[core] != p
    nc = n // s
    blocks = hit[:nc * s, :nc * s, :nc * s].reshape(nc, s, nc, s, nc, s)
    return blocks.max(axis=(1, 3, 5)).ravel()


# ---------------------------------------------------------------------------------------------- materials
def iso_C(E, nu):
    """(C11, C12, C44) of isotropic (E, nu)."""
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = 0.5 * E / (1 + nu)
    return np.stack([lam + 2 * mu, lam, mu], -1)


def sc_params_from_C(C, r):
    """calcKnTv of the SC lattice: (Kn axial, Kn diagonal, Tv) = r (C11 - C12 - C44, C44, (C12 - C44) / 18)."""
    C = np.asarray(C, float)
    return r * np.stack([C[..., 0] - C[..., 1] - C[..., 2], C[..., 2], (C[..., 1] - C[..., 2]) / 18.0], -1)


def bond_params_two_end_mean(lat, ppar):
    """Per-particle (Kn axial, Kn diagonal, Tv) -> directed Kn, Tv (N, 18) as the mean over both bond ends."""
    ppar = np.asarray(ppar, float)
    q = np.where(lat.valid, lat.nbr, 0)
    col = np.broadcast_to(SC_SHELL[None, :], lat.nbr.shape)
    own = np.take_along_axis(ppar[:, :2], col, 1)
    other = ppar[q, col]
    Kn = 0.5 * (own + other) * lat.valid
    Tv = 0.5 * (ppar[:, 2:3] + ppar[q, 2]) * lat.valid
    return np.ascontiguousarray(Kn), np.ascontiguousarray(Tv)


def C21_to_mat(C):
    """Cell stiffnesses are (..., 6, 6) Voigt matrices with engineering shear everywhere, so the reference's
    conversion helper is the identity: the input comes back unchanged."""
    return C


def cell_params_from_C66(C, r):
    """Generalised SC calcKnTv of an anisotropic cell C (Ncell, 6, 6), Voigt with engineering shear: nine
    per-direction Kn (x, y, z, xy+, xy-, xz+, xz-, yz+, yz-) and one Tv (coarse3d/geom3d.py:183-201)."""
    c = lambda a, b: C[:, a - 1, b - 1]
    T = (c(1, 2) - c(6, 6) + c(1, 3) - c(5, 5) + c(2, 3) - c(4, 4)) / 3.0
    cols = [c(1, 1) - c(5, 5) - c(6, 6) - T, c(2, 2) - c(4, 4) - c(6, 6) - T, c(3, 3) - c(4, 4) - c(5, 5) - T]
    for diag, a, b in ((6, 1, 2), (5, 1, 3), (4, 2, 3)):
        m = 0.5 * (c(a, diag) + c(b, diag))
        cols += [c(diag, diag) + m, c(diag, diag) - m]
    cols.append(T / 18.0)
    return r * np.stack(cols, 1)


def iso_C66(C3):
    C3 = np.asarray(C3, float)
    C = np.zeros((len(C3), 6, 6))
    C[:, :3, :3] = C3[:, 1, None, None]
    idx = np.arange(3)
    C[:, idx, idx] = C3[:, 0, None]
    C[:, idx + 3, idx + 3] = C3[:, 2, None]
    return C


def bond_params_from_cells(lat_c, cp):
    """Two-end mean of per-cell (9 direction Kn + Tv) -> Kn, Tv (Nc, 18)."""
    cp = np.asarray(cp, float)
    q = np.where(lat_c.valid, lat_c.nbr, 0)
    col = np.broadcast_to(DIR_OF_SLOT[None, :], lat_c.nbr.shape)
    Kn = 0.5 * (cp[:, DIR_OF_SLOT] + cp[q, col]) * lat_c.valid
    Tv = 0.5 * (cp[:, 9:10] + cp[q, 9]) * lat_c.valid
    return np.ascontiguousarray(Kn), np.ascontiguousarray(Tv)


def voigt_cells(Cvox, n, s):
    """Block volume average of voxel (C11, C12, C44) -> (Nc, 6, 6)."""
    nc = n // s
    C = np.asarray(Cvox, float).reshape(nc, s, nc, s, nc, s, 3).mean(axis=(1, 3, 5)).reshape(-1, 3)
    return iso_C66(C)
