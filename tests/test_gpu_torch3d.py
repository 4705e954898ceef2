"""The 3D drop-in surface (torch_lpm3d) on the GPU against the oracle: SCT kernels, solves vs SuperLU, the
adjoint gradient of LinSolve and the differentiable CoarseT.run vs central finite differences, FineRefT block means.

Tolerances: FP64 kernels 1e-12 relative; solves at tol 1e-12 agree with SuperLU to 1e-9; gradients agree with
central differences (step 1e-4 relative to the parameter scale, O(step^2) truncation) to 1e-6 relative."""
import numpy as np
import pytest
import torch

import lpm_ref as R
import geom3d
import torch_lpm3d as T3

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    b = np.asarray(b.detach().cpu() if torch.is_tensor(b) else b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def two_phase(dims, h, seed):
    rng = np.random.default_rng(seed)
    lat = R.Box(dims, h)
    ph = rng.random(lat.N) < 0.3
    C = R.iso_C(np.where(ph, 10.0, 1.0), np.full(lat.N, 0.25))
    Kn, Tv = R.two_end_mean(lat, R.sc_params_from_C(C, 0.5))
    return lat, Kn, Tv


def sym_direction(lat, seed):
    """Random symmetric (Kn, Tv) direction (the same value on both slots of a bond)."""
    dk, dt = R.random_symmetric_params(lat, np.random.default_rng(seed), kn=(-1.0, 1.0), tv=(-1.0, 1.0))
    return dk, dt


def test_sct_kernels_match_oracle(dev):
    dims, h = (6, 5, 4), 0.5
    lat, Kn, Tv = two_phase(dims, h, 1)
    u = np.random.default_rng(2).standard_normal((lat.N, 3))
    op = T3.SCT(geom3d.LatticeGeom(*dims, h), device=dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    assert rel(op.apply(t(u), t(Kn), t(Tv)), R.apply(lat, u, Kn, Tv)) < 1e-12
    assert rel(op.stress(t(u), t(Kn), t(Tv)), R.stress(lat, u, Kn, Tv)) < 1e-12
    assert rel(op.energy(t(u), t(Kn), t(Tv)), R.energy(lat, u, Kn, Tv)) < 1e-12
    dl, F = op.forces(t(u), t(Kn), t(Tv))
    rdl, rF = R.bond_forces(lat, u, Kn, Tv)
    assert rel(dl, rdl) < 1e-12 and rel(F, rF) < 1e-12
    D = op.block_diag(t(Kn), t(Tv))
    assert rel(D, R.block_diag(lat, Kn, Tv)) < 1e-12
    free = np.ones(lat.N, bool)
    free[:7] = False
    M = op.precond(t(Kn), t(Tv), t(free.astype(float)))
    assert rel(M, R.precond(lat, Kn, Tv, free)) < 1e-10


@pytest.mark.parametrize("load", ["k_xy", "tb_tz"])
@pytest.mark.parametrize("mixed", [False, True])
def test_sct_solve_matches_superlu(dev, load, mixed):
    print("solve", load, mixed, flush=True)
    dims, h = (8, 7, 6), 1.0
    lat, Kn, Tv = two_phase(dims, h, 3)
    fixed, up, _ = R.prescribed3(load, lat.x, lat.ijk, dims, 1, h)
    ud = R.solve_direct(R.stiffness(lat, Kn, Tv), np.repeat(fixed, 3), up).reshape(lat.N, 3)
    op = T3.SCT(geom3d.LatticeGeom(*dims, h), device=dev, mixed=mixed)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    u, it, relres = op.solve(t(Kn), t(Tv), t(up), t((~fixed).astype(float)), tol=1e-12)
    assert relres <= 1e-12
    assert rel(u, ud) < 1e-9


def test_linsolve_gradient_matches_fd(dev):
    """d/dp of L = sum w * u(p) along symmetric directions, adjoint vs central differences (LinSolve contract:
    coarse3d/torch_lpm3d.py:151-175)."""
    dims, h = (5, 5, 5), 1.0
    lat, Kn, Tv = two_phase(dims, h, 4)
    op = T3.SCT(geom3d.LatticeGeom(*dims, h), device=dev)
    fixed, up, _ = R.prescribed3("k_xz", lat.x, lat.ijk, dims, 1, h)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    upt, fr = t(up), t((~fixed).astype(float))
    w = t(np.random.default_rng(5).standard_normal((lat.N, 3)))
    dk, dt = sym_direction(lat, 6)

    def loss(k, v, grad=False):
        k, v = t(k).requires_grad_(grad), t(v).requires_grad_(grad)
        u = T3.LinSolve.apply(k, v, op, upt, fr, 1e-13)
        val = (w * u).sum()
        if not grad:
            return float(val)
        val.backward()
        return k.grad.cpu().numpy(), v.grad.cpu().numpy()

    gk, gt = loss(Kn, Tv, True)
    eps = 1e-4
    fd_k = (loss(Kn + eps * dk, Tv) - loss(Kn - eps * dk, Tv)) / (2 * eps)
    fd_t = (loss(Kn, Tv + eps * dt) - loss(Kn, Tv - eps * dt)) / (2 * eps)
    assert abs((gk * dk).sum() - fd_k) < 1e-6 * abs(fd_k) + 1e-12
    assert abs((gt * dt).sum() - fd_t) < 1e-6 * abs(fd_t) + 1e-12


def test_coarse_run_gradient_matches_fd(dev):
    """Gradient of a loss over CoarseT.run's u, sig, w and F (all 9 loads, 2 specimens) vs central differences."""
    nc, s = 4, 2
    model = T3.CoarseT(nc, s, device=dev)
    lat = R.Box((nc, nc, nc), s * 1.0)
    rng = np.random.default_rng(7)
    Ks, Ts = [], []
    for b in range(2):
        _, k, v = two_phase((nc, nc, nc), s * 1.0, 10 + b)
        Ks.append(k)
        Ts.append(v)
    Kn, Tv = np.array(Ks), np.array(Ts)
    wu = torch.as_tensor(rng.standard_normal((2, 9, lat.N, 3)), device=dev)
    ws = torch.as_tensor(rng.standard_normal((2, 9, lat.N, 6)), device=dev)
    ww = torch.as_tensor(rng.standard_normal((2, 9, lat.N)), device=dev)
    wf = torch.as_tensor(rng.standard_normal((2, 9)), device=dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)

    def loss(k, v, grad=False):
        k, v = t(k).requires_grad_(grad), t(v).requires_grad_(grad)
        o = model.run(k, v, tol=1e-13)
        val = ((wu * o["u"]).sum() + 1e3 * (ws * o["sig"]).sum() + 1e3 * (ww * o["w"]).sum() +
               1e2 * (wf * o["F"]).sum())
        if not grad:
            return float(val)
        val.backward()
        return k.grad.cpu().numpy(), v.grad.cpu().numpy()

    gk, gt = loss(Kn, Tv, True)
    dk = np.array([sym_direction(lat, 20 + b)[0] for b in range(2)])
    dt = np.array([sym_direction(lat, 30 + b)[1] for b in range(2)])
    eps = 1e-4
    fd_k = (loss(Kn + eps * dk, Tv) - loss(Kn - eps * dk, Tv)) / (2 * eps)
    fd_t = (loss(Kn, Tv + eps * dt) - loss(Kn, Tv - eps * dt)) / (2 * eps)
    assert abs((gk * dk).sum() - fd_k) < 1e-6 * abs(fd_k) + 1e-12
    assert abs((gt * dt).sum() - fd_t) < 1e-6 * abs(fd_t) + 1e-12


def test_fineref_block_means_match_superlu(dev):
    n, s = 8, 2
    fine = T3.FineRefT(n, s, device=dev)
    lat, Kn, Tv = two_phase((n, n, n), 1.0, 8)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    out = fine.run(t(Kn)[None], t(Tv)[None], tol=1e-12)
    K = R.stiffness(lat, Kn, Tv)
    nc = n // s
    cid = ((lat.ijk[:, 2] // s) * nc + lat.ijk[:, 1] // s) * nc + lat.ijk[:, 0] // s
    bmean = lambda f: np.stack([np.bincount(cid, f[:, a], nc ** 3) for a in range(f.shape[1])], 1) / s ** 3
    for l, ld in enumerate(geom3d.LOADS):
        fixed, up, moving = R.prescribed3(ld, lat.x, lat.ijk // s, (nc,) * 3, s)
        ud = R.solve_direct(K, np.repeat(fixed, 3), up).reshape(lat.N, 3)
        assert rel(out["u"][0, l], bmean(ud)) < 1e-9
        assert rel(out["sig"][0, l], bmean(R.stress(lat, ud, Kn, Tv))) < 1e-8
        assert rel(out["w"][0, l], bmean(R.energy(lat, ud, Kn, Tv)[:, None])[:, 0]) < 1e-8
        if moving is not None:
            Fr = (R.apply(lat, ud, Kn, Tv)[moving, geom3d.GDIR[ld]]).sum()
            assert abs(float(out["F"][0, l]) - Fr) < 1e-8 * abs(Fr)
