"""CPU oracle of the aggregation AMG preconditioner of the 3D SC LPM (NumPy / SciPy, FP64).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: only `tests/` (and the `cpu_baseline` leg of
`bench.py`) may use it, and only as the checker.  It restates, for small lattices, the reference's aggregation AMG
(timing/t3d/amg3d.py) with the same conventions:

  * aggregation of the free particles (a box) into chunks of `agg` particles per axis, the remainder merged into the
    last chunk (chunk_index, amg3d.py:42-48; AMGGeom, amg3d.py:438-481);
  * rigid-body modes about the lattice centroid, rows interleaved by DOF (gmg._rbm, gmg.py:145-162), tentative
    prolongator by per-aggregate modified Gram-Schmidt with rank detection (gmg._tentative, gmg.py:70-143), coarse
    near-null space = the R factors, levels >= 1 aggregated the same way on the node grid (amg3d.py:466-481);
  * Galerkin coarse operators A_{l+1} = P^T A_l P stored as 27-point 6x6-block stencils (d = (z+1)9+(y+1)3+(x+1),
    amg3d.py:39, 151-196, 256-300); the oracle forms them by sparse matrix products of the assembled K, an
    independent route from the reference's bond-level loops and from the GPU's probing;
  * damped block-Jacobi smoothing with omega / rho(D^-1 A), rho from a power iteration started from
    default_rng(seed).standard_normal (amg3d.py:493-505), V(nu, nu) on the finest level, W-cycles (gamma) on levels
    >= w_from, dense Cholesky on the coarsest level (amg3d.py:511-646);
  * PCG with ||r|| <= tol ||b|| tested every iteration, x0 = 0 (amg3d.py:649-672).
"""
import math

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

import lpm_ref as R

OFF27 = np.array([(d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1) for d in range(27)], np.int64)


def chunk_index(L, a, rule="merge"):
    """chunk id of positions 0..L-1 along one axis and the number of chunks (amg3d.py:42-48)."""
    if rule == "floor":
        return np.arange(L) // a, -(-L // a)
    nch = max(1, L // a)
    return np.minimum(np.arange(L) // a, nch - 1), nch


def grid_ijk(g):
    gx, gy, gz = g
    k, j, i = np.meshgrid(np.arange(gz), np.arange(gy), np.arange(gx), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.int64)


def rbm(X):
    """6 rigid-body modes of the points X (n, 3) about their centroid, rows interleaved by DOF (gmg.py:145-162)."""
    n = X.shape[0]
    c = X - X.mean(0)
    B = np.zeros((3 * n, 6))
    for d in range(3):
        B[d::3, d] = 1.0
    B[0::3, 3] = -c[:, 1]
    B[1::3, 3] = c[:, 0]
    B[1::3, 4] = -c[:, 2]
    B[2::3, 4] = c[:, 1]
    B[0::3, 5] = c[:, 2]
    B[2::3, 5] = -c[:, 0]
    return B


def tentative(agg_of_row, nagg, B, m=6, tol=1e-8):
    """Per-aggregate modified Gram-Schmidt of the near-null space rows B (gmg.py:70-143).

    Returns Q (rows, m) with the orthonormal columns of each row's aggregate (zero columns where a mode was dropped),
    Bc (nagg, m, m) = the R factors, and keep (nagg, m)."""
    n = B.shape[0]
    Q = np.zeros((n, m))
    Bc = np.zeros((nagg, m, m))
    keep_all = np.zeros((nagg, m), bool)
    for a in range(nagg):
        rows = np.nonzero(agg_of_row == a)[0]      # rows in increasing order, as the reference's stable grouping
        if len(rows) == 0:
            continue
        blk = B[rows]
        scale = max(math.sqrt(float((blk[:, c] ** 2).sum())) for c in range(m))
        Qa = np.zeros((len(rows), m))
        Ra = np.zeros((m, m))
        keep = np.zeros(m, bool)
        for c in range(m):
            v = blk[:, c].copy()
            for d in range(c):
                if keep[d]:
                    pr = float(Qa[:, d] @ v)
                    Ra[d, c] = pr
                    v -= pr * Qa[:, d]
            nrm = math.sqrt(float(v @ v))
            if nrm > tol * scale and nrm > 0.0:
                keep[c] = True
                Ra[c, c] = nrm
                Qa[:, c] = v / nrm
        Q[rows] = Qa
        Bc[a] = Ra
        keep_all[a] = keep
    return Q, Bc, keep_all


class AMGGeom:
    """Aggregation and tentative prolongators of all levels (amg3d.py:435-481).  lat: lpm_ref.Box (3D)."""

    def __init__(self, lat, free, agg=3, rule="merge", coarse_max=3000, max_levels=8):
        self.lat = lat
        self.free = np.asarray(free, bool)
        self.agg, self.rule, self.coarse_max = agg, rule, coarse_max
        fi = np.nonzero(self.free)[0]
        ijk = lat.ijk
        lo, hi = ijk[fi].min(0), ijk[fi].max(0)
        Lx = hi - lo + 1
        assert int(np.prod(Lx)) == len(fi), "free set must be a box"
        self.lo, self.L = lo, Lx
        ch = [chunk_index(int(Lx[d]), agg, rule) for d in range(3)]
        g = tuple(int(c[1]) for c in ch)
        rel = ijk[fi] - lo
        agg_of = -np.ones(lat.N, np.int64)
        agg_of[fi] = (ch[2][0][rel[:, 2]] * g[1] + ch[1][0][rel[:, 1]]) * g[0] + ch[0][0][rel[:, 0]]
        nagg = g[0] * g[1] * g[2]
        dofs = (3 * fi[:, None] + np.arange(3)[None, :]).ravel()
        B = rbm(lat.x)[dofs]
        Qr, Bc, keep = tentative(np.repeat(agg_of[fi], 3), nagg, B, 6, 1e-8)
        assert keep.all(), "rank-deficient aggregate"
        P = np.zeros((lat.N, 3, 6))
        P[fi] = Qr.reshape(len(fi), 3, 6)
        self.l0 = dict(agg_of=agg_of, P=P, g=g, Bc=Bc)
        self.levels = []
        Bn = Bc
        while 6 * nagg > coarse_max and len(self.levels) < max_levels:
            chn = [chunk_index(g[d], agg, rule) for d in range(3)]
            gc = tuple(int(c[1]) for c in chn)
            nij = grid_ijk(g)
            a_of = (chn[2][0][nij[:, 2]] * gc[1] + chn[1][0][nij[:, 1]]) * gc[0] + chn[0][0][nij[:, 0]]
            nc_ = gc[0] * gc[1] * gc[2]
            Qn, Bc2, keep = tentative(np.repeat(a_of, 6), nc_, Bn.reshape(-1, 6), 6, 1e-8)
            assert keep.all()
            self.levels.append(dict(g=g, agg_of=a_of, Q=Qn.reshape(nagg, 6, 6), gc=gc, Bc=Bc2))
            g, nagg, Bn = gc, nc_, Bc2
        self.g_coarsest = g


# ------------------------------------------------------------------------------------------------ stencils
def p_matrix(agg_of, P, nagg):
    """sparse (k n, 6 nagg) prolongator from per-item blocks P (n, k, 6); rows of items with agg < 0 are zero."""
    n, k, _ = P.shape
    it = np.nonzero(agg_of >= 0)[0]
    rows = (k * it[:, None, None] + np.arange(k)[None, :, None]) + np.zeros((1, 1, 6), np.int64)
    cols = (6 * agg_of[it][:, None, None] + np.arange(6)[None, None, :]) + np.zeros((1, k, 1), np.int64)
    return sp.csr_matrix((P[it].ravel(), (rows.ravel(), cols.ravel())), shape=(k * n, 6 * nagg))


def stencil_from_matrix(M, g):
    """(n, 27, 6, 6) blocks of the 6-per-node matrix M on the node grid g; asserts nothing lies outside the
    27-point pattern."""
    M = sp.csr_matrix(M)
    gx, gy, gz = g
    n = gx * gy * gz
    A = np.zeros((n, 27, 6, 6))
    ij = grid_ijk(g)
    Md = M.toarray()
    covered = np.zeros_like(Md, bool)
    for I in range(n):
        for d in range(27):
            j = ij[I] + OFF27[d]
            if np.any(j < 0) or np.any(j >= np.array(g)):
                continue
            J = (j[2] * gy + j[1]) * gx + j[0]
            A[I, d] = Md[6 * I:6 * I + 6, 6 * J:6 * J + 6]
            covered[6 * I:6 * I + 6, 6 * J:6 * J + 6] = True
    assert np.abs(Md[~covered]).max(initial=0.0) == 0.0, "coupling outside the 27-point stencil"
    return A


def dense_from_stencil(A, g):
    """dense (6n, 6n) matrix of a 27-point stencil (amg3d.py:418-432)."""
    gx, gy, gz = g
    n = A.shape[0]
    D = np.zeros((6 * n, 6 * n))
    ij = grid_ijk(g)
    for I in range(n):
        for d in range(27):
            j = ij[I] + OFF27[d]
            if np.any(j < 0) or np.any(j >= np.array(g)):
                continue
            J = (j[2] * gy + j[1]) * gx + j[0]
            D[6 * I:6 * I + 6, 6 * J:6 * J + 6] = A[I, d]
    return D


def st_mv(A, x, g):
    """y = A x on the 27-point 6x6-block stencil (amg3d.py:230-253), written as a loop over offsets."""
    gx, gy, gz = g
    n = A.shape[0]
    y = np.zeros((n, 6))
    ij = grid_ijk(g)
    for d in range(27):
        j = ij + OFF27[d]
        ok = np.all((j >= 0) & (j < np.array(g)), 1)
        J = (j[:, 2] * gy + j[:, 1]) * gx + j[:, 0]
        y[ok] += np.einsum("nrc,nc->nr", A[ok, d], x[J[ok]])
    return y


def power_rho(mv, Minv, n, k, its, seed=0, start=None):
    """rho(Minv A) by power iteration (amg3d.py:493-505); `start` replaces the seeded normal start vectors."""
    x = np.random.default_rng(seed).standard_normal((n, k)) if start is None else np.array(start, float)
    y = np.einsum("nab,nb->na", Minv, x)
    x, y = y, x
    lam = 1.0
    for _ in range(its):
        x = x / np.linalg.norm(x)
        Ax = mv(x)
        y = np.einsum("nab,nb->na", Minv, Ax)
        lam = float(np.linalg.norm(y))
        x, y = y, x
    return lam


class AMG3D:
    """Numeric hierarchy, V/W-cycle and PCG (amg3d.py:508-672) on dense / sparse NumPy operators.

    rho: optional list of the spectral radii per level (finest first) to use instead of power iterations."""

    def __init__(self, geom, Kn, Tv, nu=1, omega=1.32, power_its=10, nu_c=None, gamma=1, w_from=1, rho=None,
                 starts=None):
        lat = geom.lat
        self.geom, self.nu, self.omega = geom, nu, omega
        self.nu_c = nu if nu_c is None else nu_c
        self.gamma, self.w_from = gamma, w_from
        self.lat, self.Kn, self.Tv = lat, Kn, Tv
        self.free = geom.free
        fr = self.free.astype(float)
        self.M0 = R.precond(lat, Kn, Tv, self.free)
        K = R.stiffness(lat, Kn, Tv)
        Dm = sp.diags(np.repeat(fr, 3))
        self.Kff = (Dm @ K @ Dm).tocsr()
        starts = starts or {}
        self.rho0 = rho[0] if rho is not None else power_rho(self.apply0, self.M0, lat.N, 3, power_its,
                                                             start=starts.get(0))
        g0 = geom.l0
        nagg = int(np.prod(g0["g"]))
        P0 = p_matrix(g0["agg_of"], g0["P"], nagg)
        self.P0 = P0
        Afull = (P0.T @ self.Kff @ P0).tocsr()
        self.A = [stencil_from_matrix(Afull, g0["g"])]
        self.Minv, self.rho = [], []
        gs = [g0["g"]]
        for k, Lg in enumerate(geom.levels):
            A = self.A[-1]
            Mi = np.linalg.inv(A[:, 13])
            self.Minv.append(Mi)
            if rho is not None:
                self.rho.append(rho[k + 1])
            else:
                self.rho.append(power_rho(lambda x, A=A, g=Lg["g"]: st_mv(A, x, g), Mi, A.shape[0], 6, power_its,
                                          start=starts.get(k + 1)))
            nc_ = int(np.prod(Lg["gc"]))
            Q = p_matrix(Lg["agg_of"], Lg["Q"], nc_)
            Ac_full = (Q.T @ sp.csr_matrix(dense_from_stencil(A, Lg["g"])) @ Q).tocsr()
            self.A.append(stencil_from_matrix(Ac_full, Lg["gc"]))
            gs.append(Lg["gc"])
        self.gs = gs
        Dn = dense_from_stencil(self.A[-1], gs[-1])
        Dn = 0.5 * (Dn + Dn.T)
        self.cho = sla.cho_factor(Dn, lower=False, check_finite=False)

    def apply0(self, x):
        return (self.Kff @ x.reshape(-1)).reshape(-1, 3)

    def _coarse(self, lvl, rc):
        ec = self._vc(lvl, rc)
        k = lvl - 1
        if k == len(self.A) - 1:
            return ec
        for _ in range(self.gamma - 1 if lvl >= self.w_from else 0):
            ec = ec + self._vc(lvl, rc - st_mv(self.A[k], ec, self.gs[k]))
        return ec

    def _vc(self, lvl, b):
        k = lvl - 1
        A = self.A[k]
        if k == len(self.A) - 1:
            return sla.cho_solve(self.cho, b.reshape(-1), check_finite=False).reshape(-1, 6)
        g = self.gs[k]
        Lg = self.geom.levels[k]
        w = self.omega / self.rho[k]
        Mi = self.Minv[k]
        bj = lambda r: w * np.einsum("nab,nb->na", Mi, r)
        x = bj(b)
        for _ in range(self.nu_c - 1):
            x = x + bj(b - st_mv(A, x, g))
        r = b - st_mv(A, x, g)
        nc_ = int(np.prod(Lg["gc"]))
        Q = p_matrix(Lg["agg_of"], Lg["Q"], nc_)
        ec = self._coarse(lvl + 1, (Q.T @ r.reshape(-1)).reshape(-1, 6))
        x = x + (Q @ ec.reshape(-1)).reshape(-1, 6)
        for _ in range(self.nu_c):
            x = x + bj(b - st_mv(A, x, g))
        return x

    def vcycle(self, b):
        """finest level: b (N, 3), zero on fixed particles -> z (N, 3)."""
        w = self.omega / self.rho0
        bj = lambda r: w * np.einsum("nab,nb->na", self.M0, r)
        x = bj(b)
        for _ in range(self.nu - 1):
            x = x + bj(b - self.apply0(x))
        r = b - self.apply0(x)
        ec = self._coarse(1, (self.P0.T @ r.reshape(-1)).reshape(-1, 6))
        x = x + (self.P0 @ ec.reshape(-1)).reshape(-1, 3)
        for _ in range(self.nu):
            x = x + bj(b - self.apply0(x))
        return x

    def pcg(self, b, tol, maxit=2000):
        """K_ff x = b (b zero on fixed), x0 = 0, ||r|| <= tol ||b|| (amg3d.py:649-672).  Returns x, iters, relres."""
        x = np.zeros_like(b)
        r = b.copy()
        bn = math.sqrt(float((b * b).sum()))
        if bn == 0.0:
            return x, 0, 0.0
        z = self.vcycle(r)
        p = z.copy()
        rz = float((r * z).sum())
        it, rn = 0, bn
        for it in range(1, maxit + 1):
            Ap = self.apply0(p)
            alpha = rz / float((p * Ap).sum())
            x += alpha * p
            r -= alpha * Ap
            rn = math.sqrt(float((r * r).sum()))
            if rn <= tol * bn:
                break
            z = self.vcycle(r)
            rz2 = float((r * z).sum())
            p = z + (rz2 / rz) * p
            rz = rz2
        return x, it, rn / bn


# ------------------------------------------------------------------------------------------------ large lattices
def _mgs_batched(B, tol=1e-8):
    """modified Gram-Schmidt with rank detection (gmg.py:101-130) of a batch B (G, rows, 6) of equal-size blocks."""
    G, n, m = B.shape
    Q = np.zeros_like(B)
    Rf = np.zeros((G, m, m))
    keep = np.zeros((G, m), bool)
    scale = np.sqrt((B ** 2).sum(1)).max(1)
    for c in range(m):
        v = B[:, :, c].copy()
        for d in range(c):
            pr = np.where(keep[:, d], np.einsum("gi,gi->g", Q[:, :, d], v), 0.0)
            Rf[:, d, c] = pr
            v -= pr[:, None] * Q[:, :, d]
        nrm = np.sqrt((v * v).sum(1))
        ok = (nrm > tol * scale) & (nrm > 0.0)
        keep[:, c] = ok
        Rf[:, c, c] = np.where(ok, nrm, 0.0)
        Q[:, :, c] = np.where(ok[:, None], v / np.where(ok, nrm, 1.0)[:, None], 0.0)
    assert keep.all(), "rank-deficient aggregate"
    return Q, Rf


def box_hierarchy(dims, lo, L, h=1.0, agg=3, coarse_max=1000, max_levels=8):
    """AMGGeom of a free box (lo, L) on a full dims lattice without per-aggregate loops, for the CPU baseline on
    large lattices: level-0 blocks from the Gram-Schmidt of one aggregate per shape (at most 8 shapes; the blocks
    depend on the shape only, pinned against AMGGeom in tests), the coarse near-null space of every aggregate as the
    shape's R factor times the centroid shift, and the node levels by batched Gram-Schmidt of equal-size aggregates.
    Returns dict(g0, tmpl (concatenated (s, 3, 6) blocks, x fastest), toff (8), levels [dict(g, gc, Q (n, 6, 6))])."""
    dims, lo, L = np.array(dims), np.array(lo), np.array(L)
    g = [max(1, int(L[d]) // agg) for d in range(3)]
    lens = [(agg if g[d] > 1 else int(L[d]), int(L[d]) - agg * (g[d] - 1)) for d in range(3)]
    tm, toff, Rs, off = [], [0] * 8, [], 0
    for s in range(8):
        shape = [lens[d][(s >> d) & 1] for d in range(3)]
        k, j, i = np.meshgrid(np.arange(shape[2]), np.arange(shape[1]), np.arange(shape[0]), indexing="ij")
        X = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(float) * h
        Q, Rf = _mgs_batched(rbm(X)[None])
        tm.append(Q[0].reshape(-1))
        toff[s] = off
        off += Q[0].size
        Rs.append(Rf[0])
    gi = grid_ijk(tuple(g))
    last = gi == np.array(g) - 1
    sid = last[:, 0] * 1 + last[:, 1] * 2 + last[:, 2] * 4
    size = np.where(last, np.array([lens[d][1] for d in range(3)]), np.array([lens[d][0] for d in range(3)]))
    D = (lo + agg * gi + 0.5 * (size - 1)) * h - 0.5 * (dims - 1) * h
    T = np.tile(np.eye(6), (len(gi), 1, 1))
    T[:, 0, 3], T[:, 1, 3] = -D[:, 1], D[:, 0]
    T[:, 1, 4], T[:, 2, 4] = -D[:, 2], D[:, 1]
    T[:, 0, 5], T[:, 2, 5] = D[:, 2], -D[:, 0]
    Bn = np.einsum("gij,gjk->gik", np.stack(Rs)[sid], T)
    levels, nagg = [], int(np.prod(g))
    g = tuple(g)
    while 6 * nagg > coarse_max and len(levels) < max_levels:
        chn = [chunk_index(g[d], agg) for d in range(3)]
        gc = tuple(int(c[1]) for c in chn)
        ni = grid_ijk(g)
        a_of = (chn[2][0][ni[:, 2]] * gc[1] + chn[1][0][ni[:, 1]]) * gc[0] + chn[0][0][ni[:, 0]]
        nc_ = int(np.prod(gc))
        Qn = np.zeros((nagg, 6, 6))
        Bc2 = np.zeros((nc_, 6, 6))
        order = np.argsort(a_of, kind="stable")
        counts = np.bincount(a_of, minlength=nc_)
        starts = np.concatenate([[0], np.cumsum(counts)])
        for cnt in np.unique(counts):
            ks = np.nonzero(counts == cnt)[0]
            mem = np.stack([order[starts[k_]:starts[k_] + cnt] for k_ in ks])
            Qg, Rg = _mgs_batched(Bn[mem].reshape(len(ks), cnt * 6, 6))
            Qn[mem.reshape(-1)] = Qg.reshape(-1, 6, 6)
            Bc2[ks] = Rg
        levels.append(dict(g=g, gc=gc, Q=Qn))
        g, nagg, Bn = gc, nc_, Bc2
    return dict(g0=tuple(int(v) for v in (L // agg).clip(min=1)), tmpl=np.concatenate(tm), toff=toff,
                levels=levels, Bc0=None)


# ------------------------------------------------------------------------------------------------ C restatement
class CAmg:
    """The C/OpenMP restatement of the AMG-PCG (oracle/lpm_oracle_amg.c) on a full box lattice with a free box:
    hierarchy geometry from box_hierarchy, numeric setup (rho by power iteration, level-1 stencil by probing with the
    C K u, coarse stencils by the direct Galerkin product, 6x6 inverses, dense coarsest inverse) and the PCG.
    nbr (N, 18) int64 (-1 = no bond), Kn / Tv (N, 18) directed, M0 (N, 3, 3) block-Jacobi inverse (zero on fixed
    particles), free (N,) bool.  starts: {level: start vector} of the power iterations (default_rng(0) normals)."""

    def __init__(self, lib, dims, lo, L, nbr, Kn, Tv, M0, free, h=1.0, agg=3, coarse_max=1000, nu=1, gamma=2,
                 power_its=5, omega=1.32, w_from=1, starts=None):
        import ctypes
        self.lib, self.ct = lib, ctypes
        P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        lib.orc_apply3.argtypes = [I64, P, P, P, P, P, P]
        for f in ("orc_amg0_restrict", "orc_amg0_prolong"):
            getattr(lib, f).argtypes = [P, P, P]
        lib.orc_amg0_probe.argtypes = [P, ctypes.c_int, P]
        lib.orc_amg0_gather.argtypes = [P, ctypes.c_int, P, P]
        lib.orc_q_galerkin.argtypes = [P, P, ctypes.c_int, P, P, P]
        lib.orc_inv6.argtypes = [I64, P, P]
        lib.orc_inv6.restype = ctypes.c_int
        lib.orc_power_rho0.argtypes = [I64, P, P, P, P, P, P, ctypes.c_int]
        lib.orc_power_rho0.restype = D
        lib.orc_power_rho_st.argtypes = [P, P, P, P, ctypes.c_int]
        lib.orc_power_rho_st.restype = D
        lib.orc_amg_pcg.argtypes = [P, I64, P, P, P, P, P, P, D, ctypes.c_int, P]
        lib.orc_amg_pcg.restype = ctypes.c_int
        ptr = lambda a: a.ctypes.data_as(P)
        self.ptr = ptr
        starts = starts or {}
        self.N = int(np.prod(dims))
        self.keep = [nbr, Kn, Tv, M0]
        self.nbr, self.Kn, self.Tv = nbr, Kn, Tv
        self.free = np.ascontiguousarray(np.asarray(free, np.uint8))
        hb = box_hierarchy(dims, lo, L, h, agg, coarse_max)
        self.hb = hb
        self.tmpl = np.ascontiguousarray(hb["tmpl"])

        class A0(ctypes.Structure):
            _fields_ = [("nx", I64), ("ny", I64), ("nz", I64), ("lo", I32 * 3), ("L", I32 * 3), ("g0", I32 * 3),
                        ("agg", I32), ("tmpl", P), ("toff", I64 * 8)]

        class AM(ctypes.Structure):
            _fields_ = [("l0", A0), ("M0", P), ("w0", D), ("nlev", I32), ("nu", I32), ("nu_c", I32),
                        ("gamma", I32), ("w_from", I32), ("pad_", I32), ("g", (I32 * 3) * 12), ("A", P * 12),
                        ("Q", P * 12), ("Minv", P * 12), ("w", D * 12), ("dinv", P)]

        s = AM()
        s.l0.nx, s.l0.ny, s.l0.nz = dims
        for k in range(3):
            s.l0.lo[k], s.l0.L[k], s.l0.g0[k] = int(lo[k]), int(L[k]), int(hb["g0"][k])
        s.l0.agg = agg
        s.l0.tmpl = ptr(self.tmpl).value
        for k in range(8):
            s.l0.toff[k] = int(hb["toff"][k])
        self.M0 = np.ascontiguousarray(M0, dtype=np.float64)
        s.M0 = ptr(self.M0).value
        s.nu, s.nu_c, s.gamma, s.w_from = nu, nu, gamma, w_from
        grids = [tuple(hb["g0"])] + [Lg["gc"] for Lg in hb["levels"]]
        s.nlev = len(grids)
        for k, g in enumerate(grids):
            for d in range(3):
                s.g[k][d] = int(g[d])
        N = self.N
        st0 = starts.get(0)
        if st0 is None:
            st0 = np.random.default_rng(0).standard_normal((N, 3))
        st0 = np.ascontiguousarray(st0, dtype=np.float64)
        self.rho0 = lib.orc_power_rho0(N, ptr(nbr), ptr(Kn), ptr(Tv), ptr(self.M0), ptr(self.free), ptr(st0),
                                       power_its)
        s.w0 = omega / self.rho0
        # level-1 stencil by probing
        g0 = grids[0]
        nagg = int(np.prod(g0))
        A = np.zeros((nagg, 27, 6, 6))
        v = np.empty((6, N, 3))
        kv = np.empty((6, N, 3))
        S = np.empty((N, 4))
        for col in range(27):
            lib.orc_amg0_probe(ctypes.byref(s.l0), col, ptr(v))
            for j in range(6):
                lib.orc_apply3(N, ptr(nbr), ptr(v[j]), ptr(Kn), ptr(Tv), ptr(kv[j]), ptr(S))
            lib.orc_amg0_gather(ctypes.byref(s.l0), col, ptr(kv), ptr(A))
        del v, kv, S
        self.A, self.Minv, self.rho = [A], [], []
        for k, Lg in enumerate(hb["levels"]):
            g = np.array(Lg["g"], np.int32)
            gc = np.array(Lg["gc"], np.int32)
            n = int(np.prod(g))
            Mi = np.empty((n, 6, 6))
            assert lib.orc_inv6(n, ptr(A), ptr(Mi)) == 0
            stk = starts.get(k + 1)
            if stk is None:
                stk = np.random.default_rng(0).standard_normal((n, 6))
            stk = np.ascontiguousarray(stk, dtype=np.float64)
            r_k = lib.orc_power_rho_st(ptr(g), ptr(A), ptr(Mi), ptr(stk), power_its)
            Q = np.ascontiguousarray(Lg["Q"])
            Ac = np.empty((int(np.prod(gc)), 27, 6, 6))
            lib.orc_q_galerkin(ptr(g), ptr(gc), agg, ptr(A), ptr(Q), ptr(Ac))
            self.keep += [Q]
            self.Minv.append(Mi)
            self.rho.append(r_k)
            s.A[k] = ptr(A).value
            s.Q[k] = ptr(Q).value
            s.Minv[k] = ptr(Mi).value
            s.w[k] = omega / r_k
            self.A.append(Ac)
            A = Ac
        s.A[len(grids) - 1] = ptr(A).value
        Dn = dense_from_stencil(A, grids[-1])
        self.dinv = np.ascontiguousarray(np.linalg.inv(0.5 * (Dn + Dn.T)))
        s.dinv = ptr(self.dinv).value
        self.s = s

    def pcg(self, b, tol, maxit=2000):
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(self.N, 3)
        x = np.zeros_like(b)
        rel = np.zeros(1)
        p = self.ptr
        it = self.lib.orc_amg_pcg(self.ct.byref(self.s), self.N, p(self.nbr), p(self.Kn), p(self.Tv), p(self.free),
                                  p(b), p(x), float(tol), int(maxit), p(rel))
        return x, int(it), float(rel[0])
