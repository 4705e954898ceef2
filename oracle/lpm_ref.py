"""CPU oracle of the LPM elastic solve path (NumPy / SciPy, FP64).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: only `tests/`, `__graft_entry__.smoke()` and
the `cpu_baseline` leg of `bench.py` may use it, and only as the checker.  It restates the reference algorithm in
plain loop-oriented NumPy so that the HIP kernels can be compared against it on the same inputs.

Reference conventions restated here (citations relative to the reference tree):
  * SC lattice, 18 slots in reverse pairs (k, k^1), shell 0 = slots 0-5 (axial), shell 1 = slots 6-17 (face
    diagonals); particle id p = (k ny + j) nx + i; n_k = -off_k / |off_k|; l0 = |off| h; V = h^3
    (coarse3d/geom3d.py:16-20, 35-65).
  * Square lattice, 8 slots, shell 0 = slots 0-3, shell 1 = slots 4-7; p = j nx + i; V = h^2
    (coarse2d/lin2d.py:21-22, 29-74).
  * dL_ik = n_k.(u_i - u_q), S_i^c = sum_{k in c} dL_ik, TS_i^c = sum_{k in c} Tv_ik dL_ik,
    F_ik = 2 Kn_ik dL_ik + 1/2 (TS_i^c + TS_q^c) + 1/2 Tv_ik (S_i^c + S_q^c), (K u)_i = sum_k F_ik n_k
    (coarse3d/lin3d.py:98-149; coarse3d/torch_lpm3d.py:63-83).
  * energy-gradient form g_ik = Kn_ik dL_ik + 1/2 (Tv_ik S_i^c + TS_i^c), f_j = sum_k (g_jk + g_{q,k^1}) n_k
    (timing/t2d/solver2d.py:269-305).
  * computeStress with the opp-flag surface rule (coarse3d/lin3d.py:170-210, coarse2d/lin2d.py:143-154), particle
    energy (coarse3d/lin3d.py:213-236), block diagonal (coarse3d/lin3d.py:239-269), inverse by cofactors
    (coarse3d/lin3d.py:272-291), K = B^T H B (coarse2d/lin2d.py:77-118, coarse3d/lin3d.py:62-92).
  * block-Jacobi PCG from x0 = 0 with ||r|| <= tol ||b|| tested every iteration (coarse3d/lin3d.py:464-532), and
    its multi-tolerance form with snapshots (timing/t2d/solver2d.py:349-421),
    plain CG (timing/t3d/cgplain.py:11-55, timing/t2d/cg2plain.py:11-49), SuperLU direct solve
    (coarse3d/lin3d.py:535-546, coarse2d/lin2d.py:166-180).
  * boundary conditions of the load cases (coarse3d/geom3d.py:71-106, coarse2d/torch_lpm.py:92-111).
  * adjoint bond-parameter gradient: derivative of lam.(K(Kn, Tv) u) w.r.t. every directed slot, which is what
    torch autograd returns in coarse3d/torch_lpm3d.py:165-175 (closed form; pinned by finite differences in tests).
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ------------------------------------------------------------------------------------------------ slot tables
SC_OFF = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                   [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
                   [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
                   [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.int64)
SC_SHL = np.array([0] * 6 + [1] * 12, dtype=np.int64)
SQ_OFF = np.array([[1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [1, -1], [-1, 1]], dtype=np.int64)
SQ_SHL = np.array([0] * 4 + [1] * 4, dtype=np.int64)
EPS = 1e-3
LOADS3 = ["k_xx", "k_yy", "k_zz", "k_yz", "k_xz", "k_xy", "tb_tz", "tb_sx", "lr_tx"]
LOADS2 = ["ring_xx", "ring_yy", "ring_xy", "tb_ty", "tb_sx", "lr_tx"]
VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


class Box:
    """Full box lattice (2D square or 3D simple cubic).  Attributes mirror the reference lattice objects."""

    def __init__(self, dims, h=1.0):
        self.dims = tuple(int(d) for d in dims)
        self.dim = len(self.dims)
        self.h = float(h)
        off = SC_OFF if self.dim == 3 else SQ_OFF
        self.off, self.shell = off, (SC_SHL if self.dim == 3 else SQ_SHL)
        self.NN = len(off)
        self.N = int(np.prod(self.dims))
        # particle id: x fastest, then y, then z
        idx = np.indices(self.dims[::-1]).reshape(self.dim, -1)[::-1].T       # (N, dim) = (i, j[, k])
        self.ijk = np.ascontiguousarray(idx)
        self.x = (self.ijk + 0.5) * self.h
        length = np.sqrt((off * off).sum(1))
        self.l0 = length * self.h
        self.ndir = -off / length[:, None]
        self.V = self.h ** self.dim
        nb = self.ijk[:, None, :] + off[None, :, :]
        ok = np.all((nb >= 0) & (nb < np.array(self.dims)[None, None, :]), axis=2)
        strides = np.cumprod([1] + list(self.dims[:-1]))
        q = (nb * strides[None, None, :]).sum(2)
        self.nbr = np.where(ok, q, -1)
        self.valid = ok
        self.opp = np.arange(self.NN) ^ 1


class ActiveBox(Box):
    """Lattice of the active (solid) sites of a box, numbered compactly in box order: the reference's lattice with
    removed particles (coarse3d/geom3d.py:38-65, `active` mask; only active sites are numbered and a slot is a bond
    iff its neighbour is an active site).  Every operator function of this module accepts it."""

    def __init__(self, dims, active, h=1.0):
        super().__init__(dims, h)
        act = np.asarray(active, bool).ravel()
        full_nbr, full_ok = self.nbr, self.valid
        self.sites = np.flatnonzero(act)
        site_of = np.full(len(act), -1, np.int64)
        site_of[self.sites] = np.arange(len(self.sites))
        self.site_of = site_of
        self.N = len(self.sites)
        self.ijk = self.ijk[self.sites]
        self.x = self.x[self.sites]
        q = np.where(full_ok, full_nbr, 0)[self.sites]
        ok = full_ok[self.sites] & act[q]
        self.nbr = np.where(ok, site_of[q], -1)
        self.valid = ok


# ------------------------------------------------------------------------------------------------ operator
def bond_dl(lat, u):
    """dL (N, NN) = n_k.(u_i - u_q), zero on missing bonds."""
    u = u.reshape(lat.N, lat.dim)
    q = np.where(lat.valid, lat.nbr, 0)
    d = np.einsum("ikd,kd->ik", u[:, None, :] - u[q], lat.ndir)
    return d * lat.valid


def shell_sums(lat, u, Tv):
    dL = bond_dl(lat, u)
    S = np.zeros((lat.N, 2))
    TS = np.zeros((lat.N, 2))
    for c in (0, 1):
        m = lat.shell == c
        S[:, c] = dL[:, m].sum(1)
        TS[:, c] = (Tv[:, m] * dL[:, m]).sum(1)
    return dL, S, TS


def bond_forces(lat, u, Kn, Tv):
    """(dL, F) per directed slot in the force form (coarse3d/lin3d.py:152-167)."""
    dL, S, TS = shell_sums(lat, u, Tv)
    q = np.where(lat.valid, lat.nbr, 0)
    c = lat.shell[None, :]
    Si = np.take_along_axis(S, np.broadcast_to(c, dL.shape), 1)
    TSi = np.take_along_axis(TS, np.broadcast_to(c, dL.shape), 1)
    Sq = S[q, np.broadcast_to(c, dL.shape)]
    TSq = TS[q, np.broadcast_to(c, dL.shape)]
    F = (2.0 * Kn * dL + 0.5 * (TSi + TSq) + 0.5 * Tv * (Si + Sq)) * lat.valid
    return dL, F


def apply(lat, u, Kn, Tv):
    """K u in the force form, (N, dim)."""
    _, F = bond_forces(lat, u, Kn, Tv)
    return F @ lat.ndir


def apply_loops(lat, u, Kn, Tv):
    """K u by explicit per-particle loops, line by line after coarse3d/lin3d.py:98-149 (small lattices only)."""
    u = u.reshape(lat.N, lat.dim)
    N, NN = lat.N, lat.NN
    S = np.zeros((N, 2))
    TS = np.zeros((N, 2))
    for i in range(N):
        for k in range(NN):
            q = lat.nbr[i, k]
            if q < 0:
                continue
            dl = float(np.dot(lat.ndir[k], u[i] - u[q]))
            S[i, lat.shell[k]] += dl
            TS[i, lat.shell[k]] += Tv[i, k] * dl
    out = np.zeros((N, lat.dim))
    for i in range(N):
        for k in range(NN):
            q = lat.nbr[i, k]
            if q < 0:
                continue
            c = lat.shell[k]
            dl = float(np.dot(lat.ndir[k], u[i] - u[q]))
            f = 2.0 * Kn[i, k] * dl + 0.5 * (TS[i, c] + TS[q, c]) + 0.5 * Tv[i, k] * (S[i, c] + S[q, c])
            out[i] += f * lat.ndir[k]
    return out


def apply_gradform(lat, u, Kn, Tv):
    """K u in the energy-gradient form of timing/t2d/solver2d.py:269-305 (equal to the force form for symmetric
    parameters)."""
    dL, S, TS = shell_sums(lat, u, Tv)
    c = np.broadcast_to(lat.shell[None, :], dL.shape)
    g = (Kn * dL + 0.5 * (Tv * np.take_along_axis(S, c, 1) + np.take_along_axis(TS, c, 1))) * lat.valid
    q = np.where(lat.valid, lat.nbr, 0)
    t = (g + g[q, np.broadcast_to(lat.opp[None, :], dL.shape)]) * lat.valid
    return t @ lat.ndir


def energy(lat, u, Kn, Tv):
    """Particle energy U_i (coarse3d/lin3d.py:213-236)."""
    dL, S, TS = shell_sums(lat, u, Tv)
    return 0.5 * (Kn * dL * dL).sum(1) + 0.5 * (S * TS).sum(1)


def stress_weights(lat):
    """opp-flag / V * l0 per slot (coarse3d/lin3d.py:189-194)."""
    full = lat.valid.sum(1) == lat.NN
    of = np.where(full[:, None], 0.5, np.where(lat.valid[:, lat.opp], 0.5, 1.0))
    return of / lat.V * lat.l0[None, :] * lat.valid


def stress(lat, u, Kn, Tv):
    """computeStress: 3D (N, 6) xx yy zz yz xz xy; 2D (N, 3) xx yy xy."""
    _, F = bond_forces(lat, u, Kn, Tv)
    a = stress_weights(lat) * F
    n = lat.ndir
    if lat.dim == 3:
        comps = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    else:
        comps = [(0, 0), (1, 1), (0, 1)]
    return np.stack([(a * n[None, :, p] * n[None, :, r]).sum(1) for p, r in comps], 1)


def block_diag(lat, Kn, Tv):
    """dim x dim diagonal blocks of K (coarse3d/lin3d.py:239-269)."""
    d = lat.dim
    D = np.zeros((lat.N, d, d))
    q = np.where(lat.valid, lat.nbr, 0)
    for c in (0, 1):
        ks = np.nonzero(lat.shell == c)[0]
        T = (Tv[:, ks] * lat.valid[:, ks]) @ lat.ndir[ks]
        s = lat.valid[:, ks].astype(float) @ lat.ndir[ks]
        D += 0.5 * (T[:, :, None] * s[:, None, :] + s[:, :, None] * T[:, None, :])
        for a in ks:
            coef = (Kn[:, a] + Kn[q[:, a], a ^ 1] + Tv[q[:, a], a ^ 1]) * lat.valid[:, a]
            D += coef[:, None, None] * np.outer(lat.ndir[a], lat.ndir[a])[None]
    return D


def inv_blocks(D):
    return np.linalg.inv(D)


def precond(lat, Kn, Tv, free):
    """Block-Jacobi inverse, zero on fixed particles."""
    return inv_blocks(block_diag(lat, Kn, Tv)) * np.asarray(free, float)[:, None, None]


# ------------------------------------------------------------------------------------------------ assembly
def bond_matrix(lat):
    N, NN, d = lat.N, lat.NN, lat.dim
    rows, cols, vals = [], [], []
    for i in range(N):
        for k in range(NN):
            q = lat.nbr[i, k]
            if q < 0:
                continue
            r = i * NN + k
            for a in range(d):
                rows += [r, r]
                cols += [d * i + a, d * q + a]
                vals += [lat.ndir[k, a], -lat.ndir[k, a]]
    return sp.csr_matrix((vals, (rows, cols)), shape=(N * NN, d * N))


def bond_hessian(lat, Kn, Tv):
    N, NN = lat.N, lat.NN
    rows, cols, vals = [], [], []
    for i in range(N):
        for a in range(NN):
            if not lat.valid[i, a]:
                continue
            rows.append(i * NN + a)
            cols.append(i * NN + a)
            vals.append(Kn[i, a])
            for b in range(NN):
                if lat.valid[i, b] and lat.shell[b] == lat.shell[a]:
                    rows.append(i * NN + a)
                    cols.append(i * NN + b)
                    vals.append(0.5 * (Tv[i, a] + Tv[i, b]))
    return sp.csr_matrix((vals, (rows, cols)), shape=(N * NN, N * NN))


def stiffness(lat, Kn, Tv):
    """Assembled K = B^T H B (exact Hessian of sum_i U_i for symmetric parameters)."""
    B = bond_matrix(lat)
    H = bond_hessian(lat, Kn, Tv)
    return (B.T @ H @ B).tocsr()


def dense_from_apply(lat, Kn, Tv):
    """Dense K by applying the operator to unit vectors (brute force, tiny lattices)."""
    n = lat.dim * lat.N
    K = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        K[:, j] = apply(lat, e, Kn, Tv).ravel()
    return K


def dense_from_energy(lat, Kn, Tv):
    """Dense Hessian of E(u) = sum_i U_i by second differences (E is quadratic, so h = 1 is exact up to
    round-off)."""
    n = lat.dim * lat.N
    E = lambda v: float(energy(lat, v, Kn, Tv).sum())
    H = np.zeros((n, n))
    z = np.zeros(n)
    e0 = E(z)
    Ei = np.zeros(n)
    for a in range(n):
        v = z.copy()
        v[a] = 1.0
        Ei[a] = E(v)
    for a in range(n):
        for b in range(a, n):
            v = z.copy()
            v[a] += 1.0
            v[b] += 1.0
            # a != b: E(e_a + e_b) - E(e_a) - E(e_b) + E(0) = H_ab;  a == b: E(2 e_a) - 2 E(e_a) + E(0) = H_aa
            H[a, b] = H[b, a] = E(v) - Ei[a] - Ei[b] + e0
    return H


# ------------------------------------------------------------------------------------------------ parameters
def iso_C(E, nu, plane=None):
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    if plane == "stress":
        lam = E * nu / (1 - nu * nu)
    else:
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return np.stack([lam + 2 * mu, lam, mu], -1)


def sc_params_from_C(C, r):
    """(Kn axial, Kn diagonal, Tv) of coarse3d/geom3d.py:161-164."""
    C11, C12, C44 = C[..., 0], C[..., 1], C[..., 2]
    return np.stack([r * (C11 - C12 - C44), r * C44, r * (C12 - C44) / 18.0], -1)


def square_params_from_C(C):
    """(Kn axial, Kn diagonal, Tv) of coarse2d/lin2d.py:197-200."""
    C11, C12, C44 = C[..., 0], C[..., 1], C[..., 2]
    return np.stack([(C11 - C12) / 2, C44 / 2, (C12 - C44) / 12], -1)


def two_end_mean(lat, ppar):
    """Per-particle (Kn shell 0, Kn shell 1, Tv) -> directed Kn, Tv as the mean of both ends."""
    Kn = np.zeros((lat.N, lat.NN))
    Tv = np.zeros((lat.N, lat.NN))
    for k in range(lat.NN):
        q = lat.nbr[:, k]
        ok = q >= 0
        qq = np.where(ok, q, 0)
        col = lat.shell[k]
        Kn[:, k] = np.where(ok, 0.5 * (ppar[:, col] + ppar[qq, col]), 0.0)
        Tv[:, k] = np.where(ok, 0.5 * (ppar[:, 2] + ppar[qq, 2]), 0.0)
    return Kn, Tv


def random_symmetric_params(lat, rng, kn=(0.8, 1.2), tv=(-0.03, 0.03)):
    """Per undirected bond random Kn, Tv written to both directed slots (as coarse3d/check_torch3d.py:54-61)."""
    Kn = np.zeros((lat.N, lat.NN))
    Tv = np.zeros((lat.N, lat.NN))
    for k in range(0, lat.NN, 2):
        q = lat.nbr[:, k]
        ok = q >= 0
        a = rng.uniform(*kn, size=lat.N)
        b = rng.uniform(*tv, size=lat.N)
        Kn[ok, k] = a[ok]
        Tv[ok, k] = b[ok]
        Kn[q[ok], k ^ 1] = a[ok]
        Tv[q[ok], k ^ 1] = b[ok]
    return Kn, Tv


# ------------------------------------------------------------------------------------------------ loads
def unit_strain(k):
    e = np.zeros((3, 3))
    a, b = VOIGT[k]
    if a == b:
        e[a, a] = 1.0
    else:
        e[a, b] = e[b, a] = 0.5
    return e


def prescribed3(load, x, IJK, nc, s, h=1.0):
    """(fixed (N,), u_presc (N,3), moving (N,) or None) of coarse3d/geom3d.py:86-106."""
    ncx, ncy, ncz = nc
    I, J, K = IJK[:, 0], IJK[:, 1], IJK[:, 2]
    u = np.zeros((len(x), 3))
    if load in LOADS3[:6]:
        fixed = (I < 1) | (I >= ncx - 1) | (J < 1) | (J >= ncy - 1) | (K < 1) | (K >= ncz - 1)
        xc = 0.5 * np.array([ncx, ncy, ncz]) * s * h
        u[fixed] = (x[fixed] - xc) @ (EPS * unit_strain(LOADS3.index(load))).T
        return fixed, u, None
    if load.startswith("tb"):
        a, b = K < 1, K >= ncz - 1
        dist = (ncz - 1) * s * h
        d = [0.0, 0.0, EPS * dist] if load == "tb_tz" else [EPS * dist, 0.0, 0.0]
    else:
        a, b = I < 1, I >= ncx - 1
        d = [EPS * (ncx - 1) * s * h, 0.0, 0.0]
    u[b] = d
    return a | b, u, b


def prescribed2(load, x, IJ, ncx, ncy, s, h, L=1):
    """(fixed, u_presc (N,2), moving or None) of coarse2d/torch_lpm.py:92-111."""
    xc = np.array([ncx * s * h / 2, ncy * s * h / 2])
    I, J = IJ[:, 0], IJ[:, 1]
    u = np.zeros((len(x), 2))
    if load.startswith("ring"):
        e = {"ring_xx": [[1, 0], [0, 0]], "ring_yy": [[0, 0], [0, 1]], "ring_xy": [[0, .5], [.5, 0]]}[load]
        fixed = (I < L) | (I >= ncx - L) | (J < L) | (J >= ncy - L)
        u[fixed] = (x[fixed] - xc) @ (EPS * np.array(e)).T
        return fixed, u, None
    if load.startswith("tb"):
        a, b = J < L, J >= ncy - L
        dist = (ncy - L) * s * h
        d = [0.0, EPS * dist] if load == "tb_ty" else [EPS * dist, 0.0]
    else:
        a, b = I < L, I >= ncx - L
        d = [EPS * (ncx - L) * s * h, 0.0]
    u[b] = d
    return a | b, u, b


# ------------------------------------------------------------------------------------------------ solvers
def pcg(lat, b, Kn, Tv, M, free, tol=1e-10, maxit=20000):
    """Block-Jacobi PCG of coarse3d/lin3d.py:464-518 (x0 = 0).  free (N,) bool.  Returns x, iters, relres."""
    d = lat.dim
    fr = np.asarray(free, float)[:, None]
    b = b.reshape(lat.N, d)
    x = np.zeros_like(b)
    r = b.copy() * fr
    bn = math.sqrt(float((b * b).sum()))
    if bn == 0.0:
        return x, 0, 0.0
    z = np.einsum("nij,nj->ni", M, r)
    p = z.copy()
    rz = float((r * z).sum())
    it, rn = 0, 0.0
    for it in range(1, maxit + 1):
        Ap = apply(lat, p, Kn, Tv) * fr
        pAp = float((p * Ap).sum())
        alpha = rz / pAp
        x += alpha * p * fr
        r -= alpha * Ap
        rn = math.sqrt(float((r * r).sum()))
        if rn <= tol * bn:
            break
        z = np.einsum("nij,nj->ni", M, r)
        rz2 = float((r * z).sum())
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, it, rn / bn


def pcg_snapshots(lat, b, Kn, Tv, M, mask, tols, maxit=20000, x0=None):
    """Multi-tolerance (P)CG of timing/t2d/solver2d.py:349-421 (pcg_mf): x0 (zero on fixed DOFs) or 0, per-DOF mask
    (N, dim) float, block-Jacobi M (N, dim, dim) or None (plain CG); tols descending.  The snapshot of level k is the
    iterate of the first iteration whose recursive residual meets ||r|| <= tols[k] ||b|| (iteration 0 included);
    the loop stops once every level is met; levels never met get the final iterate and -1.  Returns (snaps
    (ntol, N, dim), iters (ntol,), iterations run)."""
    d = lat.dim
    m = np.asarray(mask, float).reshape(lat.N, d)
    b = np.asarray(b, float).reshape(lat.N, d)
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, float).reshape(lat.N, d).copy()
    prec = (lambda r: np.einsum("nij,nj->ni", M, r)) if M is not None else (lambda r: r.copy())
    r = (b - apply(lat, x, Kn, Tv)) * m
    bn = math.sqrt(float((b * b).sum()))
    tols = np.asarray(tols, float)
    nt = len(tols)
    snaps = np.empty((nt, lat.N, d))
    its = np.zeros(nt, np.int64)
    k = 0
    rn = math.sqrt(float((r * r).sum()))
    while k < nt and rn <= tols[k] * bn:
        snaps[k], its[k] = x, 0
        k += 1
    if k == nt:
        return snaps, its, 0
    z = prec(r) * m
    p = z.copy()
    rz = float((r * z).sum())
    it = 0
    while it < maxit:
        Ap = apply(lat, p, Kn, Tv) * m
        alpha = rz / float((p * Ap).sum())
        x = x + alpha * p
        r = r - alpha * Ap
        it += 1
        rn = math.sqrt(float((r * r).sum()))
        while k < nt and rn <= tols[k] * bn:
            snaps[k], its[k] = x, it
            k += 1
        if k == nt:
            break
        z = prec(r) * m
        rzn = float((r * z).sum())
        p = z + (rzn / rz) * p
        rz = rzn
    while k < nt:
        snaps[k], its[k] = x, -1
        k += 1
    return snaps, its, it


def cg_plain(lat, b, Kn, Tv, mask, tol=1e-8, maxit=200000):
    """Unpreconditioned CG from zero with a per-DOF mask (timing/t2d/cg2plain.py:11-49, timing/t3d/cgplain.py:11-55).
    mask (N, dim) float (1 free / 0 fixed).  Returns x, iters, relres."""
    d = lat.dim
    m = np.asarray(mask, float).reshape(lat.N, d)
    b = b.reshape(lat.N, d)
    x = np.zeros_like(b)
    r = b * m
    bn = math.sqrt(float((b * b).sum()))
    rr = float((r * r).sum())
    if bn == 0.0 or math.sqrt(rr) <= tol * bn:
        return x, 0, (math.sqrt(rr) / bn if bn > 0 else 0.0)
    p = r.copy()
    it = 0
    while it < maxit:
        Ap = apply(lat, p, Kn, Tv) * m
        alpha = rr / float((p * Ap).sum())
        x += alpha * p
        r -= alpha * Ap
        rn = float((r * r).sum())
        it += 1
        if math.sqrt(rn) <= tol * bn:
            rr = rn
            break
        p = r + (rn / rr) * p
        rr = rn
    return x, it, math.sqrt(rr) / bn


def solve_direct(K, fixed_dofs_mask, u_presc):
    """SuperLU solve of K_ff u_f = -K_fp u_p (coarse3d/lin3d.py:535-546).  fixed_dofs_mask (n,) bool."""
    fd = np.asarray(fixed_dofs_mask, bool).ravel()
    free = ~fd
    u = np.zeros(K.shape[0])
    u[fd] = np.asarray(u_presc).ravel()[fd]
    rhs = -(K[free][:, fd] @ u[fd])
    lu = spla.splu(K[free][:, free].tocsc(), permc_spec="COLAMD")
    u[free] = lu.solve(rhs)
    return u


# ------------------------------------------------------------------------------------------------ adjoint
def bond_grad(lat, u, lam, Kn, Tv):
    """d/dKn_ik and d/dTv_ik of E = lam . (K(Kn, Tv) u) for every directed slot (force form).
    dE/dKn_ik = 2 dL_ik(u) mu_ik, dE/dTv_ik = 1/2 mu_ik (S_i^c(u) + S_q^c(u)) + 1/2 dL_ik(u) S_i^c(lam),
    mu_ik = n_k . lam_i."""
    dL, S, _ = shell_sums(lat, u, Tv)
    _, Sl, _ = shell_sums(lat, lam, Tv)
    lam = lam.reshape(lat.N, lat.dim)
    mu = lam @ lat.ndir.T
    q = np.where(lat.valid, lat.nbr, 0)
    c = np.broadcast_to(lat.shell[None, :], dL.shape)
    Si = np.take_along_axis(S, c, 1)
    Sq = S[q, c]
    Li = np.take_along_axis(Sl, c, 1)
    gK = 2.0 * dL * mu * lat.valid
    gT = (0.5 * mu * (Si + Sq) + 0.5 * dL * Li) * lat.valid
    return gK, gT
