"""Accuracy of the mixed-precision solve against an FP64 oracle at scale, and of the training gradient.

* 64^3 sub-volume of the 256^3 cement-like benchmark microstructure (tension z): the FP32 PCG + FP64 refinement
  solve to rel-res 1e-8 against the C oracle's FP64 block-Jacobi PCG to 1e-13 (a sparse direct factorisation of
  786k DOF does not fit the test budget).  Bars: relative displacement error 1e-6, relative stress and reaction
  (apparent stiffness) error 1e-5.
* Coarse 3D training gradient (FP32 parameters, 9 loads, mixed-precision solves): adjoint gradient vs central
  finite differences with step 1e-3 (relative to the O(1) parameter scale) and vs torch autograd through a dense
  FP64 restatement of the operator (K probed column by column from a torch force-form K.u, torch.linalg.solve).
"""
import ctypes
import os
import sys

import numpy as np
import pytest
import torch

import lpm_ref as R
import lpm_ops as ops
import _lpmlib as L
import synth
import geom3d
import torch_lpm3d as T3

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle_lib():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_oracle
    if not os.path.exists(build_oracle.LIB):
        build_oracle.build()
    lib = ctypes.CDLL(build_oracle.LIB)
    P = ctypes.c_void_p
    lib.orc_pcg3.argtypes = [ctypes.c_int64, P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_int, P]
    lib.orc_pcg3.restype = ctypes.c_int
    return lib


def test_subvolume_64_mixed_vs_fp64_oracle(dev):
    n = 64
    lab = synth.cement_phases(256, 2024).reshape(256, 256, 256)[:n, :n, :n].ravel()
    kn = np.array([0.1, 1.0, 2.0, 5.0])[lab]
    ppar = np.stack([kn, kn, 0.3 * kn], 1)
    dims = (n, n, n)
    lat = R.Box(dims)
    Kn, Tv = R.two_end_mean(lat, ppar)
    z = lat.ijk[:, 2]
    fixed = (z == 0) | (z == n - 1)
    up = np.zeros((lat.N, 3))
    up[z == n - 1, 2] = 1e-3 * (n - 1)
    # GPU: the benchmark's solve path (packed FP64 / FP32 operators, FP32 block Jacobi, refinement to 1e-8)
    kp64, tp64 = synth.packed_from_particles(dims, ppar, dev, torch.float64)
    upt = torch.as_tensor(up, device=dev)
    free = torch.as_tensor((~fixed).astype(float), device=dev)
    b = -ops.apply(dims, 1.0, upt, kp64, tp64, layout=L.PACKED) * free[:, None]
    M = ops.block_jacobi(dims, 1.0, kp64, tp64, free=free, out_dtype=torch.float32, layout=L.PACKED)
    x, info = ops.pcg_refined(dims, 1.0, b, kp64, tp64, kp64.float(), tp64.float(), minv32=M, free=free, tol=1e-8,
                              maxit=100000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
    assert info.relmax <= 1e-8
    ug = up + x.cpu().numpy()
    # oracle: FP64 block-Jacobi PCG to 1e-13 (C restatement of coarse3d/lin3d.py:464-518)
    lib = _oracle_lib()
    bo = np.ascontiguousarray(-R.apply(lat, up, Kn, Tv) * (~fixed)[:, None])
    Mo = np.ascontiguousarray(R.precond(lat, Kn, Tv, ~fixed))
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    fr = np.ascontiguousarray((~fixed).astype(np.uint8))
    xo = np.zeros_like(bo)
    rel = np.zeros(1)
    pt = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    it = lib.orc_pcg3(lat.N, pt(nbr), pt(np.ascontiguousarray(Kn)), pt(np.ascontiguousarray(Tv)), pt(Mo), pt(fr),
                      pt(bo), pt(xo), 1e-13, 100000, pt(rel))
    assert rel[0] <= 1e-13
    uo = up + xo
    eu = np.abs(ug - uo).max() / np.abs(uo).max()
    so, sg = R.stress(lat, uo, Kn, Tv), R.stress(lat, ug, Kn, Tv)
    es = np.abs(sg - so).max() / np.abs(so).max()
    top = z == n - 1
    Fo, Fg = R.apply(lat, uo, Kn, Tv)[top, 2].sum(), R.apply(lat, ug, Kn, Tv)[top, 2].sum()
    eF = abs(Fg - Fo) / abs(Fo)
    print("64^3: %d inner its, %d outer; oracle %d its; rel err u %.2e, stress %.2e, reaction %.2e"
          % (info.it, int(np.max(info.outer)), it, eu, es, eF), flush=True)
    assert eu < 1e-6 and es < 1e-5 and eF < 1e-5


def _torch_apply(lat, u, Kn, Tv):
    """Force-form K u restated with torch ops (autograd through Kn, Tv)."""
    nbr = torch.as_tensor(np.where(lat.valid, lat.nbr, 0), device=u.device)
    valid = torch.as_tensor(lat.valid, device=u.device, dtype=u.dtype)
    nd = torch.as_tensor(lat.ndir, device=u.device, dtype=u.dtype)
    sh = torch.as_tensor(lat.shell, device=u.device)
    dL = ((u[:, None, :] - u[nbr]) * nd[None]).sum(-1) * valid
    S = torch.stack([dL[:, sh == c].sum(1) for c in (0, 1)], 1)
    TS = torch.stack([(Tv * dL)[:, sh == c].sum(1) for c in (0, 1)], 1)
    Si, TSi = S[:, sh], TS[:, sh]
    Sq, TSq = S[nbr, sh[None, :].expand_as(nbr)], TS[nbr, sh[None, :].expand_as(nbr)]
    F = (2 * Kn * dL + 0.5 * (TSi + TSq) + 0.5 * Tv * (Si + Sq)) * valid
    return F @ nd


def test_training_gradient_fd_h1e3_and_dense_autograd(dev):
    """Coarse 4^3 x 2 specimens x 9 loads, FP32 parameters, mixed-precision solves (the training-step path)."""
    nc, s = 4, 2
    model = T3.CoarseT(nc, s, device=dev, dtype=torch.float64, mixed=True)
    lat = R.Box((nc, nc, nc), s * 1.0)
    rng = np.random.default_rng(11)
    Ks, Ts = zip(*[R.random_symmetric_params(lat, np.random.default_rng(40 + b), (0.5, 1.5), (-0.05, 0.05))
                   for b in range(2)])
    Kn = np.array(Ks, np.float32).astype(np.float64)
    Tv = np.array(Ts, np.float32).astype(np.float64)
    wu = torch.as_tensor(rng.standard_normal((2, 9, lat.N, 3)), device=dev)
    ws = torch.as_tensor(rng.standard_normal((2, 9, lat.N, 6)), device=dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)

    def loss(k, v, grad=False, tol=1e-12):
        k, v = t(k).requires_grad_(grad), t(v).requires_grad_(grad)
        o = model.run(k, v, tol=tol)
        val = (wu * o["u"]).sum() + 1e3 * (ws * o["sig"]).sum()
        if not grad:
            return float(val)
        val.backward()
        return k.grad.cpu().numpy(), v.grad.cpu().numpy()

    gk, gt = loss(Kn, Tv, True)
    h = 1e-3
    for seed in (1, 2):
        dk = np.array([R.random_symmetric_params(lat, np.random.default_rng(seed * 10 + b), (-1, 1), (-1, 1))[0]
                       for b in range(2)])
        fd = (loss(Kn + h * dk, Tv) - loss(Kn - h * dk, Tv)) / (2 * h)
        assert abs((gk * dk).sum() - fd) < 1e-5 * abs(fd)
        dt = np.array([R.random_symmetric_params(lat, np.random.default_rng(seed * 20 + b), (-1, 1), (-1, 1))[1]
                       for b in range(2)])
        fd = (loss(Kn, Tv + h * dt) - loss(Kn, Tv - h * dt)) / (2 * h)
        assert abs((gt * dt).sum() - fd) < 1e-5 * abs(fd)

    # torch autograd through a dense FP64 restatement: u = up + K_ff^-1 (-K_fp up), for the specimen 0 displacement
    # loss only (u part), compared with the adjoint gradient of the same loss
    geo = geom3d.LatticeGeom(nc, nc, nc, s * 1.0)
    bcs = model.bc
    k0 = t(Kn[0]).requires_grad_(True)
    v0 = t(Tv[0]).requires_grad_(True)
    n3 = 3 * lat.N
    eye = torch.eye(n3, device=dev, dtype=torch.float64)
    cols = [_torch_apply(lat, eye[j].reshape(lat.N, 3), k0, v0).reshape(-1) for j in range(n3)]
    K = torch.stack(cols, 1)
    val = 0.0
    for l in range(9):
        fr = bcs.free[l].repeat_interleave(3).bool()
        upl = bcs.up[l].reshape(-1)
        Kff = K[fr][:, fr]
        rhs = -(K[fr][:, ~fr] @ upl[~fr])
        u = upl.clone()
        u = u.masked_scatter(fr, torch.linalg.solve(Kff, rhs))
        val = val + (wu[0, l].reshape(-1) * u).sum()
    val.backward()
    k1, v1 = t(Kn).requires_grad_(True), t(Tv).requires_grad_(True)
    o = model.run(k1, v1, tol=1e-12)
    (wu[0:1] * o["u"][0:1]).sum().backward()
    ga, gb = k0.grad.cpu().numpy(), k1.grad[0].cpu().numpy()
    assert np.abs(ga - gb).max() < 1e-7 * np.abs(ga).max()
    ga, gb = v0.grad.cpu().numpy(), v1.grad[0].cpu().numpy()
    assert np.abs(ga - gb).max() < 1e-7 * np.abs(ga).max()
    del geo
