"""The two fixed-size workloads of the benchmark set that the operator tests do not reach, at their full sizes.

C1: plain 2D CG, 128 x 128 square lattice (two-phase disks, contrast 10, Tv = 0.3 Kn, two-end mean), per-DOF mask
    (bottom row roller in y, one pinned node, top row displaced by 1e-3 of the height), FP64 unpreconditioned CG to
    rel-res 1e-8 (timing/t2d/cg2plain.py:11-68).  Checked against the C oracle's plain CG with the same mask and
    tolerance (the reference's solver: solution and iteration count) and against SciPy SuperLU on the oracle's
    assembled K (32 768 DOF).
C2: 48^3 KUBC, nine right-hand sides (six unit strains, three combined) in one batched FP32-inner / FP64-refined
    block-Jacobi PCG launch to rel-res 1e-8 (coarse3d/geom3d.py:71-106, coarse3d/lin3d.py:464-532).  Checked by
    superposition (each combined load equals the sum of its two unit-strain solutions: the ring displacements are
    linear in the strain) and, for the first load, against the C oracle's FP64 PCG to 1e-13.

Bars: displacement 1e-6 relative, stress and reactions 1e-5 relative."""
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
import workloads

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
    lib.orc_cg2.argtypes = [ctypes.c_int64, P, P, P, P, P, P, ctypes.c_double, ctypes.c_int, P]
    lib.orc_cg2.restype = ctypes.c_int
    return lib


def _pt(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("path", ["batched", "persistent", "pipelined"])
def test_c1_plain2d_128_matches_superlu(dev, path):
    """path: the batched launch-per-iteration CG (lpm_pcg_solve), the single-launch persistent CG
    (lpm_cg2_persistent) or its pipelined form (lpm_cg2_pipelined)."""
    print("C1 plain2d 128", path, flush=True)
    W = workloads.plain2d(128, device=dev)
    dims = W["dims"]
    b = -ops.apply(dims, 1.0, W["up"], W["Kn"], W["Tv"]) * W["mask"]
    if path == "batched":
        x, info = ops.pcg(dims, 1.0, b, W["Kn"], W["Tv"], minv=None, free=W["mask"], tol=1e-8, maxit=200000)
    else:
        x, info = ops.Cg2Persistent(dims, W["Kn"], W["Tv"], W["mask"],
                                    pipelined=path == "pipelined").solve(b, tol=1e-8, maxit=200000)
    assert info.conv.all() and info.relmax <= 1e-8
    ug = W["up_np"] + x.reshape(-1, 2).cpu().numpy()
    lat = R.Box(dims)
    Kn, Tv, mask, up = W["Kn_np"], W["Tv_np"], W["mask_np"], W["up_np"]
    uo = R.solve_direct(R.stiffness(lat, Kn, Tv), mask.ravel() == 0, up).reshape(-1, 2)
    top = lat.ijk[:, 1] == dims[1] - 1
    Fo, Fg = R.apply(lat, uo, Kn, Tv)[top, 1].sum(), R.apply(lat, ug, Kn, Tv)[top, 1].sum()
    eu, es = rel(ug, uo), rel(R.stress(lat, ug, Kn, Tv), R.stress(lat, uo, Kn, Tv))
    # the C oracle's plain CG (cg2plain.cg_mf restated) with the same mask and tolerance
    lib = _oracle_lib()
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    bo = np.ascontiguousarray(-R.apply(lat, up, Kn, Tv).ravel() * mask.ravel())
    xo = np.zeros_like(bo)
    rr = np.zeros(1)
    it = lib.orc_cg2(lat.N, _pt(nbr), _pt(np.ascontiguousarray(Kn)), _pt(np.ascontiguousarray(Tv)),
                     _pt(np.ascontiguousarray(mask.ravel().astype(np.float64))), _pt(bo), _pt(xo), 1e-8, 200000,
                     _pt(rr))
    uc = up + xo.reshape(-1, 2)
    ec, ecs = rel(ug, uc), rel(R.stress(lat, ug, Kn, Tv), R.stress(lat, uc, Kn, Tv))
    print("C1: %d its (oracle CG %d); vs the oracle's CG: u %.2e, stress %.2e; vs SuperLU: u %.2e, stress %.2e, "
          "reaction %.2e" % (info.it, it, ec, ecs, eu, es, abs(Fg - Fo) / abs(Fo)), flush=True)
    if path == "pipelined":
        # the pipelined recurrences drift from the CG iterates on this lattice (replacements keep its true residual
        # under the tolerance); at the same residual its error is about 4x the classic form's, so it is held to the
        # bars a rel-res 1e-8 solve meets against the exact solution, not to the reference's iterate
        res = R.apply(lat, ug, Kn, Tv) * mask
        assert np.linalg.norm(res.ravel()) <= 1e-8 * np.linalg.norm(bo) * 1.0001
        assert eu < 1e-5 and es < 1e-4 and abs(Fg - Fo) <= 1e-6 * abs(Fo) and info.it <= 1.1 * it
        return
    # against the reference's solver at the same tolerance (the parity bar) ...
    assert ec < 1e-6 and ecs < 1e-5 and abs(info.it - it) <= max(5, it // 100)
    # ... and against the exact solution, which a solve stopped at rel-res 1e-8 reaches to about 1e-6 on this
    # lattice (the oracle's CG is 1.1e-6 from SuperLU)
    assert eu < 1e-5 and es < 1e-4 and abs(Fg - Fo) <= 1e-5 * abs(Fo)


def test_c2_kubc3d_48_nine_rhs_one_launch(dev):
    print("C2 kubc3d 48", flush=True)
    W = workloads.kubc3d(48, device=dev)
    dims = W["dims"]
    up, free = W["up"], W["free"]
    b = -ops.apply(dims, 1.0, up, W["kp64"], W["tp64"], layout=L.PACKED) * free[:, None]
    M = ops.block_jacobi(dims, 1.0, W["kp64"], W["tp64"], free=free, out_dtype=torch.float32, layout=L.PACKED)
    x, info = ops.pcg_refined(dims, 1.0, b, W["kp64"], W["tp64"], W["kp32"], W["tp32"], minv32=M, free=free,
                              tol=1e-8, maxit=100000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
    assert b.shape[0] == 9 and info.conv.shape == (1, 9)
    assert info.conv.all() and info.relmax <= 1e-8
    u = (up + x).cpu().numpy()
    # superposition: load 6 + c = unit strains a + b
    for c, (a, bb) in enumerate(workloads.COMBINED):
        assert rel(u[6 + c], u[a] + u[bb]) < 1e-6, (c, rel(u[6 + c], u[a] + u[bb]))
    # load 0 against the C oracle's FP64 block-Jacobi PCG to 1e-13
    lat = R.Box(dims)
    Kn, Tv = synth.directed_from_packed(dims, W["kp64"], W["tp64"])
    Kn = np.ascontiguousarray(Kn.cpu().numpy())
    Tv = np.ascontiguousarray(Tv.cpu().numpy())
    fixed = W["fixed"]
    up0 = up[0].cpu().numpy()
    bo = np.ascontiguousarray(-R.apply(lat, up0, Kn, Tv) * (~fixed)[:, None])
    Mo = np.ascontiguousarray(R.precond(lat, Kn, Tv, ~fixed))
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    fr = np.ascontiguousarray((~fixed).astype(np.uint8))
    xo = np.zeros_like(bo)
    rr = np.zeros(1)
    it = _oracle_lib().orc_pcg3(lat.N, _pt(nbr), _pt(Kn), _pt(Tv), _pt(Mo), _pt(fr), _pt(bo), _pt(xo), 1e-13, 100000,
                                _pt(rr))
    assert rr[0] <= 1e-13
    uo = up0 + xo
    eu, es = rel(u[0], uo), rel(R.stress(lat, u[0], Kn, Tv), R.stress(lat, uo, Kn, Tv))
    print("C2: %d inner its (max), %d outer; oracle %d its; load 0 rel err u %.2e, stress %.2e"
          % (info.it, info.outer, it, eu, es), flush=True)
    assert eu < 1e-6 and es < 1e-5
