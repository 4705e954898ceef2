"""Pins of the oracle (oracle/lpm_ref.py, oracle/lpm_oracle.c) by what the mathematics fixes: brute-force Hessians
of the particle energy, symmetry, the energy identity, rigid-body null space, the affine patch test and
C = calcKnTv^-1 for homogeneous material, the SuperLU direct solve, finite differences of the adjoint, and the
two independent restatements (NumPy, C) against each other.  The reference ships no golden vectors for this path
(its check scripts print and read an npz that is not shipped), so these are the pins."""
import ctypes
import math
import os

import numpy as np
import pytest

import lpm_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SMALL = [(3, 3, 3), (4, 3, 2), (2, 2, 2), (4, 4), (5, 3), (3, 2)]


def _sym(dims, seed):
    rng = np.random.default_rng(seed)
    lat = R.Box(dims)
    Kn, Tv = R.random_symmetric_params(lat, rng, kn=(0.5, 1.5), tv=(-0.1, 0.1))
    return lat, Kn, Tv, rng


@pytest.mark.parametrize("dims", SMALL)
def test_stiffness_equals_energy_hessian(dims):
    lat, Kn, Tv, _ = _sym(dims, 1)
    H = R.dense_from_energy(lat, Kn, Tv)
    assert np.abs(R.stiffness(lat, Kn, Tv).toarray() - H).max() < 1e-12 * max(1, np.abs(H).max())
    assert np.abs(R.dense_from_apply(lat, Kn, Tv) - H).max() < 1e-12 * max(1, np.abs(H).max())


@pytest.mark.parametrize("dims", SMALL)
def test_loop_and_gradient_forms_agree(dims):
    lat, Kn, Tv, rng = _sym(dims, 2)
    u = rng.standard_normal((lat.N, lat.dim))
    a = R.apply(lat, u, Kn, Tv)
    assert np.abs(a - R.apply_loops(lat, u, Kn, Tv)).max() < 1e-13 * np.abs(a).max()
    assert np.abs(a - R.apply_gradform(lat, u, Kn, Tv)).max() < 1e-13 * np.abs(a).max()


def test_forms_differ_for_asymmetric_params():
    """The force form uses the directed slot of each particle: a dropped or swapped index shows here."""
    lat, Kn, Tv, rng = _sym((3, 3, 3), 3)
    Kn = Kn.copy()
    Kn[13, 0] += 0.3
    u = rng.standard_normal((lat.N, 3))
    assert np.abs(R.apply(lat, u, Kn, Tv) - R.apply_gradform(lat, u, Kn, Tv)).max() > 1e-3
    assert np.abs(R.apply(lat, u, Kn, Tv) - R.apply_loops(lat, u, Kn, Tv)).max() < 1e-13


@pytest.mark.parametrize("dims", [(6, 5, 4), (9, 7)])
def test_symmetry_energy_identity_null_space(dims):
    lat, Kn, Tv, rng = _sym(dims, 4)
    u = rng.standard_normal((lat.N, lat.dim))
    v = rng.standard_normal((lat.N, lat.dim))
    Ku, Kv = R.apply(lat, u, Kn, Tv), R.apply(lat, v, Kn, Tv)
    assert abs((v * Ku).sum() - (u * Kv).sum()) < 1e-11 * abs((v * Ku).sum())
    assert abs((u * Ku).sum() - 2 * R.energy(lat, u, Kn, Tv).sum()) < 1e-11 * (u * Ku).sum()
    for a in range(lat.dim):
        t = np.zeros((lat.N, lat.dim))
        t[:, a] = 1.0
        assert np.abs(R.apply(lat, t, Kn, Tv)).max() < 1e-12
    W = np.zeros((lat.dim, lat.dim))
    W[0, 1], W[1, 0] = 0.7, -0.7
    rot = lat.x @ W.T
    assert np.abs(R.apply(lat, rot, Kn, Tv)).max() < 1e-12


def _homog(dims, E=3.0, nu=0.25):
    lat = R.Box(dims)
    C = R.iso_C(np.full(lat.N, E), np.full(lat.N, nu))
    pp = R.sc_params_from_C(C, 0.5 * lat.h) if lat.dim == 3 else R.square_params_from_C(C)
    Kn, Tv = R.two_end_mean(lat, pp)
    return lat, Kn, Tv, C[0]


@pytest.mark.parametrize("dims", [(7, 7, 7), (9, 9)])
def test_affine_patch_and_stress_equals_C_eps(dims):
    lat, Kn, Tv, C3 = _homog(dims)
    d = lat.dim
    rng = np.random.default_rng(5)
    e = rng.standard_normal((d, d))
    e = 0.5 * (e + e.T)
    u = lat.x @ e.T
    f = R.apply(lat, u, Kn, Tv)
    lo = 2
    inner = np.all((lat.ijk >= lo) & (lat.ijk < np.array(dims) - lo), axis=1)
    assert np.abs(f[inner]).max() < 1e-12
    sig = R.stress(lat, u, Kn, Tv)
    C11, C12, C44 = C3
    tr = np.trace(e)
    if d == 3:
        expect = [C12 * tr + (C11 - C12) * e[0, 0], C12 * tr + (C11 - C12) * e[1, 1],
                  C12 * tr + (C11 - C12) * e[2, 2], 2 * C44 * e[1, 2], 2 * C44 * e[0, 2], 2 * C44 * e[0, 1]]
    else:
        expect = [C12 * tr + (C11 - C12) * e[0, 0], C12 * tr + (C11 - C12) * e[1, 1], 2 * C44 * e[0, 1]]
    full = lat.valid.all(1)
    assert np.abs(sig[full] - np.array(expect)[None]).max() < 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("dims", [(4, 3, 3), (5, 4)])
def test_block_diag_is_diagonal_of_K(dims):
    lat, Kn, Tv, _ = _sym(dims, 6)
    K = R.stiffness(lat, Kn, Tv).toarray()
    d = lat.dim
    D = R.block_diag(lat, Kn, Tv)
    for i in range(lat.N):
        assert np.abs(D[i] - K[d * i:d * i + d, d * i:d * i + d]).max() < 1e-13


def _bvp(dims, seed, load):
    lat, Kn, Tv, rng = _sym(dims, seed)
    if lat.dim == 3:
        fixed, up, mov = R.prescribed3(load, lat.x, lat.ijk, dims, 1)
    else:
        fixed, up, mov = R.prescribed2(load, lat.x, lat.ijk, dims[0], dims[1], 1, 1.0)
    b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * (~fixed)[:, None]
    return lat, Kn, Tv, fixed, up, b


@pytest.mark.parametrize("dims,load", [((8, 7, 6), "k_xz"), ((8, 7, 6), "tb_tz"), ((12, 10), "ring_xy"),
                                       ((12, 10), "lr_tx")])
def test_pcg_and_cg_match_superlu(dims, load):
    lat, Kn, Tv, fixed, up, b = _bvp(dims, 7, load)
    K = R.stiffness(lat, Kn, Tv)
    ud = R.solve_direct(K, np.repeat(fixed, lat.dim), up.ravel()).reshape(-1, lat.dim)
    x, it, rr = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, ~fixed), ~fixed, tol=1e-12)
    assert rr <= 1e-12 and np.abs(x + up * fixed[:, None] - ud).max() < 1e-9 * np.abs(ud).max()
    x2, it2, rr2 = R.cg_plain(lat, b, Kn, Tv, np.repeat((~fixed)[:, None], lat.dim, 1), tol=1e-12)
    assert rr2 <= 1e-12 and np.abs(x2 + up * fixed[:, None] - ud).max() < 1e-9 * np.abs(ud).max()
    assert it <= it2


def test_load_cases_geometry():
    lat = R.Box((6, 6, 6))
    IJK = lat.ijk // 2
    fixed, up, mov = R.prescribed3("k_xx", lat.x, IJK, (3, 3, 3), 2)
    assert fixed.sum() == 216 - 8 and mov is None
    i = np.flatnonzero(fixed)[0]
    assert abs(up[i, 0] - 1e-3 * (lat.x[i, 0] - 3.0)) < 1e-15 and up[i, 1] == 0 and up[i, 2] == 0
    fixed, up, mov = R.prescribed3("tb_tz", lat.x, IJK, (3, 3, 3), 2)
    assert fixed.sum() == 2 * 72 and mov.sum() == 72 and np.allclose(up[mov], [0, 0, 1e-3 * 4])
    fixed, up, mov = R.prescribed3("k_yz", lat.x, IJK, (3, 3, 3), 2)
    j = np.flatnonzero(fixed)[5]
    assert abs(up[j, 1] - 0.5e-3 * (lat.x[j, 2] - 3.0)) < 1e-15


@pytest.mark.parametrize("dims", [(4, 4, 3), (6, 5)])
def test_adjoint_closed_form_per_directed_slot(dims):
    """d/dKn_ik and d/dTv_ik of lam.(K u) by central differences of the operator, slot by slot (also for
    non-symmetric perturbations: each directed value enters the force form once)."""
    lat, Kn, Tv, rng = _sym(dims, 8)
    u = rng.standard_normal((lat.N, lat.dim))
    lam = rng.standard_normal((lat.N, lat.dim))
    gK, gT = R.bond_grad(lat, u, lam, Kn, Tv)
    E = lambda k, t: float((lam * R.apply(lat, u, k, t)).sum())
    h = 1e-3
    for i in range(lat.N):
        for k in range(lat.NN):
            if not lat.valid[i, k]:
                continue
            kp, km = Kn.copy(), Kn.copy()
            kp[i, k] += h
            km[i, k] -= h
            assert abs((E(kp, Tv) - E(km, Tv)) / (2 * h) - gK[i, k]) < 1e-9
            tp, tm = Tv.copy(), Tv.copy()
            tp[i, k] += h
            tm[i, k] -= h
            assert abs((E(Kn, tp) - E(Kn, tm)) / (2 * h) - gT[i, k]) < 1e-9


def test_adjoint_gradient_of_a_solve_by_finite_differences():
    """Loss of the solved displacement: adjoint (lam = K_ff^-1 g, grad = -lam.dK u) vs central differences of
    symmetric bond pairs (coarse3d/check_torch3d.py:48-95 recipe)."""
    dims = (5, 5, 5)
    lat, Kn, Tv, fixed, up, b = _bvp(dims, 9, "k_xy")
    fr = ~fixed
    W = np.random.default_rng(1).standard_normal((lat.N, 3))
    K = R.stiffness

    def solve(kn, tv):
        bb = -R.apply(lat, up * fixed[:, None], kn, tv) * fr[:, None]
        x, _, _ = R.pcg(lat, bb, kn, tv, R.precond(lat, kn, tv, fr), fr, tol=1e-14, maxit=10000)
        return x + up * fixed[:, None]

    u = solve(Kn, Tv)
    lam, _, _ = R.pcg(lat, W * fr[:, None], Kn, Tv, R.precond(lat, Kn, Tv, fr), fr, tol=1e-14, maxit=10000)
    gK, gT = R.bond_grad(lat, u, lam, Kn, Tv)
    gK, gT = -gK, -gT
    rng = np.random.default_rng(4)
    h = 1e-4
    for _ in range(4):
        i = int(rng.integers(lat.N))
        k = int(rng.choice(np.flatnonzero(lat.valid[i])))
        q = lat.nbr[i, k]
        for which in (0, 1):
            A = [Kn.copy(), Tv.copy()]
            Bm = [Kn.copy(), Tv.copy()]
            A[which][i, k] += h
            A[which][q, k ^ 1] += h
            Bm[which][i, k] -= h
            Bm[which][q, k ^ 1] -= h
            fd = ((W * solve(*A)).sum() - (W * solve(*Bm)).sum()) / (2 * h)
            g = (gK if which == 0 else gT)
            ad = g[i, k] + g[q, k ^ 1]
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-8)
    del K


# ------------------------------------------------------------------------------------------------ C oracle
def _clib():
    import build_oracle
    path = build_oracle.LIB
    if not os.path.exists(path):
        build_oracle.build()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.orc_apply3.argtypes = [ctypes.c_int64, P, P, P, P, P, P]
    L.orc_kmv2.argtypes = [ctypes.c_int64, P, P, P, P, P, P]
    L.orc_cg3.argtypes = [ctypes.c_int64, P, P, P, P, P, P, ctypes.c_double, ctypes.c_int, P]
    L.orc_pcg3.argtypes = [ctypes.c_int64, P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_int, P]
    L.orc_cg2.argtypes = [ctypes.c_int64, P, P, P, P, P, P, ctypes.c_double, ctypes.c_int, P]
    L.orc_cg3.restype = L.orc_pcg3.restype = L.orc_cg2.restype = ctypes.c_int
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_c_oracle_operators_match_numpy():
    L = _clib()
    for dims in [(7, 6, 5), (11, 9)]:
        lat, Kn, Tv, rng = _sym(dims, 10)
        u = np.ascontiguousarray(rng.standard_normal((lat.N, lat.dim)))
        nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
        out = np.zeros_like(u)
        if lat.dim == 3:
            L.orc_apply3(lat.N, _p(nbr), _p(u), _p(Kn), _p(Tv), _p(out), _p(np.zeros((lat.N, 4))))
        else:
            L.orc_kmv2(lat.N, _p(nbr), _p(u), _p(Kn), _p(Tv), _p(np.zeros((lat.N, 8))), _p(out))
        ref = R.apply(lat, u, Kn, Tv)
        assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


def test_c_oracle_solvers_match_numpy():
    L = _clib()
    lat, Kn, Tv, fixed, up, b = _bvp((9, 8, 7), 11, "k_yy")
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    fr = np.ascontiguousarray((~fixed).astype(np.uint8))
    b = np.ascontiguousarray(b)
    x = np.zeros_like(b)
    rel = np.zeros(1)
    it = L.orc_cg3(lat.N, _p(nbr), _p(Kn), _p(Tv), _p(fr), _p(b), _p(x), 1e-10, 100000, _p(rel))
    xr, itr, rr = R.cg_plain(lat, b, Kn, Tv, np.repeat(fr[:, None], 3, 1).astype(float), tol=1e-10)
    assert abs(it - itr) <= 1 and np.abs(x - xr).max() < 1e-8 * np.abs(xr).max() and rel[0] <= 1e-10
    M = np.ascontiguousarray(R.precond(lat, Kn, Tv, ~fixed))
    x = np.zeros_like(b)
    it = L.orc_pcg3(lat.N, _p(nbr), _p(Kn), _p(Tv), _p(M), _p(fr), _p(b), _p(x), 1e-10, 100000, _p(rel))
    xr, itr, rr = R.pcg(lat, b, Kn, Tv, M, ~fixed, tol=1e-10)
    assert abs(it - itr) <= 1 and np.abs(x - xr).max() < 1e-8 * np.abs(xr).max() and rel[0] <= 1e-10
    lat, Kn, Tv, fixed, up, b = _bvp((14, 11), 12, "tb_ty")
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    mask = np.ascontiguousarray(np.repeat((~fixed).astype(float), 2))
    b = np.ascontiguousarray(b.ravel())
    x = np.zeros_like(b)
    it = L.orc_cg2(lat.N, _p(nbr), _p(Kn), _p(Tv), _p(mask), _p(b), _p(x), 1e-10, 100000, _p(rel))
    xr, itr, rr = R.cg_plain(lat, b.reshape(-1, 2), Kn, Tv, mask.reshape(-1, 2), tol=1e-10)
    assert abs(it - itr) <= 1 and np.abs(x.reshape(-1, 2) - xr).max() < 1e-8 * np.abs(xr).max()
    assert math.isfinite(rel[0])


def test_pcg_snapshots_are_the_first_iterates_meeting_each_tolerance():
    """The multi-tolerance restatement (timing/t2d/solver2d.py:349-421): snapshot k is the plain PCG iterate after
    its[k] iterations, the recursive residual meets tols[k] there and not one iteration earlier; with a small maxit
    the levels never met hold the final iterate and -1."""
    lat = R.Box((11, 9))
    Kn, Tv = R.random_symmetric_params(lat, np.random.default_rng(3))
    fixed = lat.ijk[:, 1] == 0
    rng = np.random.default_rng(4)
    b = rng.standard_normal((lat.N, 2)) * (~fixed)[:, None]
    M = R.precond(lat, Kn, Tv, ~fixed)
    mask = np.repeat((~fixed)[:, None], 2, 1).astype(float)
    tols = np.array([0.5, 0.3, 0.1, 1e-2, 5e-3, 1e-4, 1e-7, 1e-9])
    snaps, its, n = R.pcg_snapshots(lat, b, Kn, Tv, M, mask, tols, maxit=5000)
    assert (its >= 0).all() and np.all(np.diff(its) >= 0) and n == its[-1]
    for k, t in enumerate(tols):
        xk, itk, rk = R.pcg(lat, b, Kn, Tv, M, ~fixed, tol=0.0, maxit=int(its[k])) if its[k] > 0 else (
            np.zeros_like(b), 0, 1.0)
        assert np.abs(snaps[k] - xk).max() <= 1e-12 * max(np.abs(xk).max(), 1e-300)
        assert its[k] == 0 or rk <= t
        if its[k] > 1:
            _, _, rprev = R.pcg(lat, b, Kn, Tv, M, ~fixed, tol=0.0, maxit=int(its[k]) - 1)
            assert rprev > t
    s2, i2, n2 = R.pcg_snapshots(lat, b, Kn, Tv, M, mask, tols, maxit=int(its[3]) + 1)
    assert (i2[:4] == its[:4]).all() and (i2[5:] == -1).all() and n2 == its[3] + 1
    xf, _, _ = R.pcg(lat, b, Kn, Tv, M, ~fixed, tol=0.0, maxit=n2)
    assert np.abs(s2[-1] - xf).max() <= 1e-12 * np.abs(xf).max()
