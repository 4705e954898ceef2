"""The 2D drop-in (torch_lpm, lin2d) and the CSR assembly (lin2d / lin3d stiffness) on the GPU against the oracle.

Tolerances: kernels 1e-12 relative (FP64); matrix-free solves at tol 1e-12 vs SuperLU 1e-9 relative; assembled K
vs the oracle's B^T H B 1e-12 relative; gradients vs central differences (step 1e-4 on O(1) parameters) 1e-6."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import lpm_ref as R
import lpm_ops as ops
import geom2d
import torch_lpm as T2
import lin2d
import lin3d

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    b = np.asarray(b.detach().cpu() if torch.is_tensor(b) else b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def params2(lat, seed, contrast=10.0):
    rng = np.random.default_rng(seed)
    ph = rng.random(lat.N) < 0.35
    C = R.iso_C(np.where(ph, contrast, 1.0), np.full(lat.N, 0.2))
    return R.two_end_mean(lat, R.square_params_from_C(C))


@pytest.mark.parametrize("dims", [(7, 6, 5), (9, 8), (3, 1, 2), (1, 4)])
def test_stiffness_csr_matches_oracle(dev, dims):
    print("csr", dims, flush=True)
    lat = R.Box(dims, 0.7)
    Kn, Tv = R.random_symmetric_params(lat, np.random.default_rng(1))
    Kref = R.stiffness(lat, Kn, Tv)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    indptr, indices, data = ops.stiffness_csr(dims, 0.7, t(Kn), t(Tv))
    n = lat.dim * lat.N
    K = sp.csr_matrix((data, indices, indptr), shape=(n, n))
    assert rel(K.toarray(), Kref.toarray()) < 1e-12
    # pattern: sorted columns, and every structural entry of the oracle's product is present
    for r in range(n):
        cols = indices[indptr[r]:indptr[r + 1]]
        assert np.all(np.diff(cols) > 0)
    Kr = Kref.copy()
    Kr.eliminate_zeros()
    pattern = sp.csr_matrix((np.ones(len(indices)), indices, indptr), shape=(n, n))
    assert (Kr != 0).multiply(pattern).nnz == Kr.nnz


def test_lin3d_surface(dev):
    dims = (6, 5, 4)
    lat = lin3d.Lattice3D(*dims, h=1.0, device=dev)
    ref = R.Box(dims)
    Kn, Tv = R.random_symmetric_params(ref, np.random.default_rng(2))
    u = np.random.default_rng(3).standard_normal((lat.N, 3))
    assert rel(lat.apply(u, Kn, Tv), R.apply(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.stress(u, Kn, Tv), R.stress(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.energy_per_particle(u, Kn, Tv), R.energy(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.block_diag(Kn, Tv), R.block_diag(ref, Kn, Tv)) < 1e-12
    K = lat.stiffness(Kn, Tv)
    fixed, up, _ = R.prescribed3("k_yz", ref.x, ref.ijk, dims, 1)
    ud, _ = lin3d.solve_direct(K, fixed, up)
    uo = R.solve_direct(R.stiffness(ref, Kn, Tv), np.repeat(fixed, 3), up).reshape(-1, 3)
    assert rel(ud, uo) < 1e-10
    up_, info = lin3d.solve_pcg(lat, Kn, Tv, fixed, up, tol=1e-12)
    assert info["relres"] <= 1e-12 and rel(up_, uo) < 1e-9
    M = lin3d.k_block_jacobi(lat.block_diag(Kn, Tv), ~fixed, device=dev)
    assert rel(M, R.precond(ref, Kn, Tv, ~fixed)) < 1e-10


def test_lin2d_surface(dev):
    nx, ny, h = 9, 7, 0.5
    lat = lin2d.Lattice2D("square", nx, ny, h, device=dev)
    ref = R.Box((nx, ny), h)
    Kn, Tv = params2(ref, 4)
    u = np.random.default_rng(5).standard_normal(2 * lat.N)
    dl, F = lat.bond_forces(u, Kn, Tv)
    rdl, rF = R.bond_forces(ref, u, Kn, Tv)
    assert rel(dl, rdl) < 1e-12 and rel(F, rF) < 1e-12
    assert rel(lat.stress(u, Kn, Tv), R.stress(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.energy_per_particle(u, Kn, Tv), R.energy(ref, u, Kn, Tv)) < 1e-12
    K, B, H = lat.stiffness(Kn, Tv)
    assert rel(K.toarray(), R.stiffness(ref, Kn, Tv).toarray()) < 1e-12
    assert rel(B.toarray(), R.bond_matrix(ref).toarray()) == 0.0
    assert rel(H.toarray(), R.bond_hessian(ref, Kn, Tv).toarray()) < 1e-15
    fixed, up, _ = R.prescribed2("tb_sx", ref.x, ref.ijk, nx, ny, 1, h)
    fd = np.nonzero(np.repeat(fixed, 2))[0]
    u2, lu = lin2d.solve_disp(K, fd, up.ravel()[fd])
    uo = R.solve_direct(R.stiffness(ref, Kn, Tv), np.repeat(fixed, 2), up)
    assert rel(u2, uo) < 1e-10


def _coarse_case(dev, ncx=8, ncy=6, s=2, B=2):
    model = T2.CoarseT(ncx, ncy, s, h=1.0, device=dev)
    ref = R.Box((ncx, ncy), s * 1.0)
    Ks, Ts = zip(*[params2(ref, 10 + b) for b in range(B)])
    return model, ref, np.array(Ks), np.array(Ts)


def test_coarse2d_run_matches_superlu(dev):
    model, ref, Kn, Tv = _coarse_case(dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    model.tol = 1e-12
    out = model.run(t(Kn), t(Tv))
    for b in range(Kn.shape[0]):
        K = R.stiffness(ref, Kn[b], Tv[b])
        for ld in model.loads:
            fixed, up, moving = R.prescribed2(ld, ref.x, ref.ijk, model.ncx, model.ncy, model.s, model.h)
            uo = R.solve_direct(K, np.repeat(fixed, 2), up).reshape(-1, 2)
            o = out[ld]
            assert rel(o["u"][b], uo) < 1e-9
            assert rel(o["sig"][b], R.stress(ref, uo, Kn[b], Tv[b])) < 1e-8
            assert rel(o["w"][b], R.energy(ref, uo, Kn[b], Tv[b]) / (model.s * model.h) ** 2) < 1e-8
            if moving is None:
                assert o["F"] is None
            else:
                Fo = R.apply(ref, uo, Kn[b], Tv[b])[moving].sum(0)
                assert rel(o["F"][b], Fo) < 1e-8


def test_coarse2d_dense_compat_matches_matrix_free(dev):
    model, ref, Kn, Tv = _coarse_case(dev, 6, 5, 1, 2)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    Kd = model.assemble(t(Kn), t(Tv))
    for b in range(Kn.shape[0]):
        assert rel(Kd[b], R.stiffness(ref, Kn[b], Tv[b]).toarray()) < 1e-12
    model.tol = 1e-12
    out = model.run(t(Kn), t(Tv))
    for ld in model.loads:
        ud = model.solve(Kd, ld)
        assert rel(out[ld]["u"].reshape(Kn.shape[0], -1), ud) < 1e-9


def test_coarse2d_gradient_matches_fd_and_dense_autograd(dev):
    model, ref, Kn, Tv = _coarse_case(dev, 6, 6, 2, 2)
    model.tol = 1e-13
    rng = np.random.default_rng(7)
    W = {ld: (torch.as_tensor(rng.standard_normal((2, model.N, 2)), device=dev),
              torch.as_tensor(rng.standard_normal((2, model.N, 3)), device=dev),
              torch.as_tensor(rng.standard_normal((2, model.N)), device=dev)) for ld in model.loads}
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)

    def objective(out):
        val = 0.0
        for ld in model.loads:
            wu, ws, ww = W[ld]
            o = out[ld]
            val = val + (wu * o["u"]).sum() + 1e3 * (ws * o["sig"]).sum() + 1e3 * (ww * o["w"]).sum()
            if o["F"] is not None:
                val = val + 1e2 * o["F"].sum()
        return val

    def loss(k, v, grad=False):
        k, v = t(k).requires_grad_(grad), t(v).requires_grad_(grad)
        val = objective(model.run(k, v))
        if not grad:
            return float(val)
        val.backward()
        return k.grad.cpu().numpy(), v.grad.cpu().numpy()

    gk, gt = loss(Kn, Tv, True)
    dk = np.array([R.random_symmetric_params(ref, np.random.default_rng(20 + b), (-1, 1), (-1, 1))[0] for b in range(2)])
    dt = np.array([R.random_symmetric_params(ref, np.random.default_rng(30 + b), (-1, 1), (-1, 1))[1] for b in range(2)])
    eps = 1e-4
    fd_k = (loss(Kn + eps * dk, Tv) - loss(Kn - eps * dk, Tv)) / (2 * eps)
    fd_t = (loss(Kn, Tv + eps * dt) - loss(Kn, Tv - eps * dt)) / (2 * eps)
    assert abs((gk * dk).sum() - fd_k) < 1e-6 * abs(fd_k)
    assert abs((gt * dt).sum() - fd_t) < 1e-6 * abs(fd_t)

    # the dense reference formulation (assemble + Cholesky + torch autograd) gives the same directional derivatives
    kd, vd = t(Kn).requires_grad_(True), t(Tv).requires_grad_(True)
    K = model.assemble(kd, vd)
    out = {}
    for ld in model.loads:
        u = model.solve(K, ld)
        mv = model.bc[ld]["moving"]
        P = (K @ u.unsqueeze(-1)).squeeze(-1).reshape(2, model.N, 2)
        out[ld] = dict(u=u.reshape(2, model.N, 2), sig=model.stress(u, kd, vd), w=model.energy_density(u, kd, vd),
                       F=None if mv is None else P[:, mv].sum(1))
    objective(out).backward()
    assert abs((kd.grad.cpu().numpy() * dk).sum() - (gk * dk).sum()) < 1e-7 * abs(fd_k)
    assert abs((vd.grad.cpu().numpy() * dt).sum() - (gt * dt).sum()) < 1e-7 * abs(fd_t)


def test_coarse2d_refuses_asymmetric(dev):
    model, ref, Kn, Tv = _coarse_case(dev, 5, 5, 1, 1)
    Kn = Kn.copy()
    Kn[0, 7, 0] *= 1.5
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    with pytest.raises(ValueError):
        model.run(t(Kn), t(Tv))


def test_stability_penalty(dev):
    model, ref, Kn, Tv = _coarse_case(dev, 5, 4, 1, 1)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    Tv = Tv.copy()
    Tv[0, 3, :4] = -2.0 * Kn[0, 3, :4]                                # one unstable axial group
    pen, lam = model.stability_penalty(t(Kn), t(Tv))
    H = model.group_hessians(t(Kn), t(Tv)).cpu().numpy()
    lmin = np.linalg.eigvalsh(H)[..., 0]
    sc = np.abs(Kn).mean()
    assert rel(lam, lmin / sc) < 1e-12
    assert float(pen[0]) > 0 and lmin[0, 3, 0] < 0
    # H of a group is the Hessian of the group's bond energy 1/2 sum Kn e^2 + 1/2 (sum Tv e)(sum e) over its live
    # bonds (the particle energy of coarse3d/lin3d.py:213-236 split by shell): compare the quadratic forms
    rng = np.random.default_rng(11)
    shell = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    for p in (0, 3, 7, 12):
        for c in (0, 1):
            ks = np.nonzero(shell == c)[0]
            live = ref.valid[p, ks]
            e = rng.standard_normal(4) * live
            kn, tv = Kn[0, p, ks] * live, Tv[0, p, ks] * live
            U = 0.5 * (kn * e * e).sum() + 0.5 * (tv * e).sum() * e.sum()
            assert abs(0.5 * e @ H[0, p, c] @ e - U) <= 1e-12 * max(1.0, abs(U))
    # the penalty is the sum of squared shortfalls of lambda_min / mean|Kn| below frac
    frac = 0.05
    assert abs(float(pen[0]) - (np.maximum(frac - lmin[0] / sc, 0.0) ** 2).sum()) <= 1e-12 * float(pen[0])


@pytest.mark.parametrize("dims,precond", [((24, 20), True), ((24, 20), False), ((9, 8, 7), True)])
def test_pcg_snapshots_match_oracle(dev, dims, precond):
    """Multi-tolerance solve (timing/t2d/solver2d.py:349-421): three systems with different masks in one launch;
    per level the first iteration meeting it and that iterate equal the oracle's restatement; a small maxit leaves
    the last levels at -1 with the final iterate."""
    print("snapshots", dims, precond, flush=True)
    lat = R.Box(dims)
    d = lat.dim
    Kn, Tv = R.random_symmetric_params(lat, np.random.default_rng(6))
    rng = np.random.default_rng(7)
    fixeds = [lat.ijk[:, 0] == 0, lat.ijk[:, 1] == 0, (lat.ijk[:, 0] == 0) | (lat.ijk[:, 1] == lat.dims[1] - 1)]
    bs = [rng.standard_normal((lat.N, d)) * (~f)[:, None] for f in fixeds]
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    free = torch.as_tensor(np.stack([~f for f in fixeds]), device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=free) if precond else None
    tols = lin2d.TOLS
    for maxit in (20000, 25):
        x, info = ops.pcg(dims, 1.0, t(np.stack(bs)), t(Kn), t(Tv), minv=minv, free=free, maxit=maxit, tols=tols)
        for s, (f, b) in enumerate(zip(fixeds, bs)):
            M = R.precond(lat, Kn, Tv, ~f) if precond else None
            snaps, its, _ = R.pcg_snapshots(lat, b, Kn, Tv, M, np.repeat((~f)[:, None], d, 1), tols, maxit=maxit)
            assert (info.snap_iters[0, s] == its).all(), (s, info.snap_iters[0, s], its)
            assert rel(info.snaps[s], snaps) < 1e-9
            if maxit == 25:
                assert (its == -1).any() and not info.conv[0, s]
            else:
                assert info.conv[0, s] and info.iters[0, s] == its[-1]


@pytest.mark.parametrize("fp32", [False, True])
def test_lin2d_solve_mf_matches_oracle(dev, fp32):
    """Lattice2D.solve_mf keeps solve_mf's shape (snapshots with ub added, per-level iterations per load);
    fp32=True rounds the stored coefficients to FP32 (the reference's mf_arrays(fp32=True)), arithmetic FP64."""
    nx, ny = 16, 12
    lat = lin2d.Lattice2D("square", nx, ny, 1.0, device=dev)
    ref = R.Box((nx, ny))
    Kn, Tv = params2(ref, 8)
    bcs = {}
    for ld in ("tb_ty", "ring_xy"):
        fixed, up, _ = R.prescribed2(ld, ref.x, ref.ijk, nx, ny, 1, 1.0)
        bcs[ld] = (fixed, up.ravel())
    out = lat.solve_mf(Kn, Tv, bcs, tols=lin2d.TOLS[::4], fp32=fp32)
    kn, tv = (Kn.astype(np.float32).astype(np.float64), Tv.astype(np.float32).astype(np.float64)) if fp32 else (Kn, Tv)
    for ld, (fixed, ub) in bcs.items():
        ubm = ub.reshape(-1, 2)
        b = -R.apply(ref, ubm, kn, tv) * (~fixed)[:, None]
        snaps, its, _ = R.pcg_snapshots(ref, b, kn, tv, R.precond(ref, kn, tv, ~fixed),
                                        np.repeat((~fixed)[:, None], 2, 1), lin2d.TOLS[::4])
        S, I = out[ld]
        assert S.shape == (len(its), 2 * lat.N) and (I == its).all()
        assert rel(S, (snaps + ubm[None]).reshape(len(its), -1)) < 1e-9
    # the FP32-rounded set is a different operator: its final snapshot differs from the FP64 one by ~1e-7
    if fp32:
        o64 = lat.solve_mf(Kn, Tv, bcs, tols=lin2d.TOLS[::4])
        d = rel(out["tb_ty"][0][-1], o64["tb_ty"][0][-1])
        assert 1e-10 < d < 1e-5
