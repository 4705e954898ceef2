"""Pins the AMG oracle (oracle/amg_ref.py) by what the mathematics fixes: the tentative prolongator reproduces the
rigid-body modes with orthonormal columns, the level-0 blocks depend on the aggregate's shape only, the stencils equal
the sparse Galerkin products, the cycle is a symmetric operator, and AMG-PCG reaches the SuperLU solution in fewer
iterations than block-Jacobi PCG."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import amg_ref as A  # noqa: E402
import lpm_ref as R  # noqa: E402


def _problem(dims, seed=0, contrast=True):
    rng = np.random.default_rng(seed)
    lat = R.Box(dims)
    Kn, Tv = R.random_symmetric_params(lat, rng, kn=(0.5, 5.0) if contrast else (0.8, 1.2), tv=(-0.05, 0.3))
    fixed, up, _ = R.prescribed3("k_xy", lat.x, lat.ijk, dims, 1)
    return lat, Kn, Tv, ~fixed, up


def test_chunk_index_merges_the_remainder():
    c, n = A.chunk_index(11, 3)
    assert n == 3 and list(c) == [0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2]
    c, n = A.chunk_index(2, 3)
    assert n == 1 and list(c) == [0, 0]
    c, n = A.chunk_index(7, 3, "floor")
    assert n == 3 and list(c) == [0, 0, 0, 1, 1, 1, 2]


def test_rbm_are_rigid_motions():
    lat = R.Box((4, 3, 5))
    B = A.rbm(lat.x)
    rng = np.random.default_rng(1)
    Kn, Tv = R.random_symmetric_params(lat, rng)
    for c in range(6):
        ku = R.apply(lat, B[:, c].reshape(-1, 3), Kn, Tv)
        assert np.abs(ku).max() < 1e-11 * max(1.0, np.abs(B[:, c]).max())


def test_tentative_prolongator_reproduces_rigid_modes():
    lat, Kn, Tv, free, _ = _problem((11, 10, 9))
    geom = A.AMGGeom(lat, free, agg=3, coarse_max=10)
    g0 = geom.l0
    nagg = int(np.prod(g0["g"]))
    P = A.p_matrix(g0["agg_of"], g0["P"], nagg)
    PtP = (P.T @ P).toarray()
    assert np.allclose(PtP, np.eye(6 * nagg), atol=1e-12)
    fi = np.nonzero(free)[0]
    B = A.rbm(lat.x)
    Bf = np.zeros_like(B)
    dofs = (3 * fi[:, None] + np.arange(3)).ravel()
    Bf[dofs] = B[dofs]
    assert np.allclose(P @ g0["Bc"].reshape(-1, 6), Bf, atol=1e-11)
    # coarse level: Q orthonormal per aggregate and reproduces the coarse near-null space
    for Lg in geom.levels:
        nc_ = int(np.prod(Lg["gc"]))
        Q = A.p_matrix(Lg["agg_of"], Lg["Q"], nc_)
        assert np.allclose((Q.T @ Q).toarray(), np.eye(6 * nc_), atol=1e-12)


def test_level0_blocks_depend_on_the_aggregate_shape_only():
    """Gram-Schmidt removes the translation part of the global rotations, so an aggregate's block equals the one of
    an aggregate of the same shape anywhere: the GPU evaluates P from at most 8 shape templates."""
    lat, Kn, Tv, free, _ = _problem((12, 11, 10))
    geom = A.AMGGeom(lat, free, agg=3)
    g0 = geom.l0
    fi = np.nonzero(free)[0]
    rel = lat.ijk[fi] - geom.lo
    for I in np.unique(g0["agg_of"][fi]):
        sel = fi[g0["agg_of"][fi] == I]
        r = lat.ijk[sel]
        shape = r.max(0) - r.min(0) + 1
        sub = R.Box(tuple(int(s) for s in shape))
        Qt, _, keep = A.tentative(np.zeros(3 * sub.N, np.int64), 1, A.rbm(sub.x), 6, 1e-8)
        assert keep.all()
        loc = r - r.min(0)
        lid = (loc[:, 2] * shape[1] + loc[:, 1]) * shape[0] + loc[:, 0]
        assert np.allclose(g0["P"][sel], Qt.reshape(-1, 3, 6)[lid], atol=1e-12)
    assert rel.min() == 0


def test_galerkin_stencils_equal_the_products():
    lat, Kn, Tv, free, _ = _problem((14, 14, 14))
    geom = A.AMGGeom(lat, free, agg=3, coarse_max=30)
    assert len(geom.levels) >= 1
    amg = A.AMG3D(geom, Kn, Tv, rho=[2.0] * (len(geom.levels) + 1))
    rng = np.random.default_rng(3)
    g0 = geom.l0
    nagg = int(np.prod(g0["g"]))
    x = rng.standard_normal((nagg, 6))
    ref = (amg.P0.T @ (amg.Kff @ (amg.P0 @ x.reshape(-1)))).reshape(-1, 6)
    assert np.allclose(A.st_mv(amg.A[0], x, g0["g"]), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # symmetric stencil: A[I, d] = A[J, 26 - d]^T
    D = A.dense_from_stencil(amg.A[0], g0["g"])
    assert np.allclose(D, D.T, atol=1e-12 * np.abs(D).max())
    for k, Lg in enumerate(geom.levels):
        nc_ = int(np.prod(Lg["gc"]))
        Q = A.p_matrix(Lg["agg_of"], Lg["Q"], nc_)
        xc = rng.standard_normal((nc_, 6))
        Dk = A.dense_from_stencil(amg.A[k], Lg["g"])
        refc = Q.T @ (Dk @ (Q @ xc.reshape(-1)))
        got = A.st_mv(amg.A[k + 1], xc, Lg["gc"]).reshape(-1)
        assert np.allclose(got, refc, rtol=1e-12, atol=1e-12 * np.abs(refc).max())


@pytest.mark.parametrize("gamma", [1, 2])
def test_cycle_is_symmetric_positive(gamma):
    lat, Kn, Tv, free, _ = _problem((14, 13, 12))
    geom = A.AMGGeom(lat, free, agg=3, coarse_max=30)
    amg = A.AMG3D(geom, Kn, Tv, nu=1, gamma=gamma, power_its=5)
    rng = np.random.default_rng(4)
    fr = free.astype(float)[:, None]
    a = rng.standard_normal((lat.N, 3)) * fr
    b = rng.standard_normal((lat.N, 3)) * fr
    va, vb = amg.vcycle(a), amg.vcycle(b)
    assert abs(float((va * b).sum()) - float((a * vb).sum())) < 1e-10 * float(np.abs(va).sum())
    assert float((va * a).sum()) > 0


def test_amg_pcg_matches_superlu_and_beats_block_jacobi():
    lat, Kn, Tv, free, up = _problem((15, 15, 15), seed=2)
    fixed = ~free
    b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * free[:, None]
    geom = A.AMGGeom(lat, free, agg=3, coarse_max=100)
    amg = A.AMG3D(geom, Kn, Tv, nu=1, gamma=2, power_its=5)
    x, it, rr = amg.pcg(b, 1e-10, 500)
    assert rr <= 1e-10
    K = R.stiffness(lat, Kn, Tv)
    fd = np.repeat(free, 3)
    xs = np.zeros(3 * lat.N)
    xs[fd] = spla.splu(sp.csc_matrix(K[fd][:, fd])).solve(b.reshape(-1)[fd])
    assert np.abs(x.reshape(-1) - xs).max() <= 1e-8 * np.abs(xs).max()
    _, itj, _ = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, free), free, tol=1e-10)
    assert it < itj


def test_box_hierarchy_equals_per_aggregate_gram_schmidt():
    """The loop-free hierarchy of the CPU baseline equals AMGGeom's per-aggregate Gram-Schmidt."""
    dims = (17, 11, 16)
    lat, Kn, Tv, free, _ = _problem(dims)
    geom = A.AMGGeom(lat, free, agg=3, coarse_max=50)
    hb = A.box_hierarchy(dims, geom.lo, geom.L, agg=3, coarse_max=50)
    assert tuple(hb["g0"]) == tuple(geom.l0["g"]) and len(hb["levels"]) == len(geom.levels)
    fi = np.nonzero(free)[0]
    g0 = np.array(hb["g0"])
    for p in fi:
        rel = lat.ijk[p] - geom.lo
        cc = np.minimum(rel // 3, g0 - 1)
        loc = rel - 3 * cc
        s = sum(1 << d for d in range(3) if cc[d] == g0[d] - 1)
        ln = [geom.L[d] - 3 * (g0[d] - 1) if cc[d] == g0[d] - 1 else 3 for d in range(3)]
        row = (loc[2] * ln[1] + loc[1]) * ln[0] + loc[0]
        blk = hb["tmpl"][hb["toff"][s] + 18 * row: hb["toff"][s] + 18 * row + 18].reshape(3, 6)
        assert np.abs(blk - geom.l0["P"][p]).max() < 1e-12
    for Lh, Lr in zip(hb["levels"], geom.levels):
        assert tuple(Lh["gc"]) == tuple(Lr["gc"]) and np.abs(Lh["Q"] - Lr["Q"]).max() < 1e-10


def test_c_amg_pcg_matches_numpy_oracle():
    """The C/OpenMP AMG-PCG (CPU baseline) reproduces the NumPy oracle: same rho, stencils, iterations, solution."""
    import ctypes
    import build_oracle
    build_oracle.build()
    lib = ctypes.CDLL(build_oracle.LIB)
    dims = (20, 19, 21)
    lat, Kn, Tv, free, up = _problem(dims, seed=2)
    fixed = ~free
    b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * free[:, None]
    geom = A.AMGGeom(lat, free, agg=3, coarse_max=100)
    ref = A.AMG3D(geom, Kn, Tv, nu=1, gamma=2, power_its=5)
    M0 = R.precond(lat, Kn, Tv, free)
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    c = A.CAmg(lib, dims, geom.lo, geom.L, nbr, np.ascontiguousarray(Kn), np.ascontiguousarray(Tv), M0, free,
               agg=3, coarse_max=100, nu=1, gamma=2, power_its=5)
    assert abs(c.rho0 - ref.rho0) <= 1e-10 * ref.rho0
    for a_, b_ in zip(c.rho, ref.rho):
        assert abs(a_ - b_) <= 1e-10 * b_
    for k in range(len(ref.A)):
        assert np.abs(c.A[k] - ref.A[k]).max() <= 1e-10 * np.abs(ref.A[k]).max()
    x, it, rr = c.pcg(b, 1e-10, 500)
    xr, itr, _ = ref.pcg(b, 1e-10, 500)
    assert rr <= 1e-10 and abs(it - itr) <= 1
    assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max()
