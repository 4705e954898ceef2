"""Failure paths of the solve path on the GPU (the reference's conventions):

* a coefficient set whose K is not positive definite raises FloatingPointError, as the reference's Cholesky does
  (coarse2d/torch_lpm.py:20-25); the batched PCG detects it on the device (r.z < 0 or p.Ap <= 0) and the host maps
  LPM_E_NOT_SPD to FloatingPointError;
* NaN inputs end in RuntimeError (LPM_E_NONFINITE) right away instead of iterating or passing as a zero right-hand side;
* an iteration cap is not an error: the system is reported unconverged with its relative residual
  (coarse3d/torch_lpm3d.py:140, coarse3d/lin3d.py:518), which equals the oracle's after the same iterations;
* a zero right-hand side is done at iteration 0 (coarse3d/lin3d.py:479-480).

Every kernel here exits normally: the conditions are produced by the inputs, never by a fault."""
import numpy as np
import pytest
import torch

import lpm_ref as R
import lpm_ops as ops
import torch_lpm as T2
import torch_lpm3d as T3

pytestmark = pytest.mark.gpu


def _case3(dev, dims=(7, 6, 5), seed=0):
    lat = R.Box(dims)
    Kn, Tv = R.random_symmetric_params(lat, np.random.default_rng(seed))
    fixed, up, _ = R.prescribed3("k_xy", lat.x, lat.ijk, dims, 1)
    b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * (~fixed)[:, None]
    return lat, Kn, Tv, fixed, b


def _t(dev):
    return lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)


def test_coarse2d_unstable_parameters_raise_floating_point_error(dev):
    """Negative bond stiffness makes K negative definite: the dense reference path (Cholesky) and the matrix-free run
    both raise FloatingPointError."""
    print("coarse2d not SPD", flush=True)
    model = T2.CoarseT(6, 5, 2, h=1.0, device=dev)
    ref = R.Box((6, 5), 2.0)
    C = R.iso_C(np.ones(ref.N), np.full(ref.N, 0.2))
    Kn, Tv = R.two_end_mean(ref, R.square_params_from_C(C))
    t = _t(dev)
    bad_kn, bad_tv = t(-Kn)[None], t(-Tv)[None]
    with pytest.raises(FloatingPointError):
        model.run(bad_kn, bad_tv)
    with pytest.raises(FloatingPointError):
        model.solve(model.assemble(bad_kn, bad_tv), model.loads[0])
    # the same model still solves a stable set afterwards
    out = model.run(t(Kn)[None], t(Tv)[None])
    assert all(torch.isfinite(o["u"]).all() for o in out.values())


def test_pcg_detects_indefinite_operator_during_the_iteration(dev):
    """Positive diagonal blocks but an indefinite K (one bond pair strongly negative): block Jacobi is SPD at the
    start, and the first search direction with p.Ap <= 0 stops the solve with FloatingPointError.  The oracle's K is
    checked to be indefinite on the free DOFs and its diagonal blocks positive definite, so the case is what it
    claims to be."""
    print("pcg indefinite", flush=True)
    dims = (5, 4, 4)
    lat = R.Box(dims)
    Kn = np.where(lat.valid, 1.0, 0.0)
    Tv = np.zeros_like(Kn)
    # make the x bond between two interior neighbours strongly negative (both directed slots)
    p = (2 * dims[1] + 2) * dims[0] + 1
    q = p + 1
    Kn[p, 0] = Kn[q, 1] = -4.0
    fixed = (lat.ijk[:, 2] == 0)
    free = ~fixed
    K = R.stiffness(lat, Kn, Tv).toarray()
    fd = np.repeat(free, 3)
    Kff = K[np.ix_(fd, fd)]
    assert np.linalg.eigvalsh(Kff)[0] < 0
    D = R.block_diag(lat, Kn, Tv)[free]
    assert (np.linalg.eigvalsh(D)[:, 0] > 0).all()
    # the first search direction is z = M b (PCG) or b (plain CG); choose b so that it is the eigenvector v of the
    # negative eigenvalue: p.Ap = lambda_min |v|^2 < 0 in the first iteration
    w, V = np.linalg.eigh(Kff)
    v = np.zeros(3 * lat.N)
    v[fd] = V[:, 0]
    v = v.reshape(-1, 3)
    Dv = np.einsum("nij,nj->ni", R.block_diag(lat, Kn, Tv), v) * free[:, None]
    t = _t(dev)
    fr = torch.as_tensor(free, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=fr)
    with pytest.raises(FloatingPointError):
        ops.pcg(dims, 1.0, t(Dv), t(Kn), t(Tv), minv=minv, free=fr, tol=1e-12, maxit=500)
    with pytest.raises(FloatingPointError):
        ops.pcg(dims, 1.0, t(v), t(Kn), t(Tv), minv=None, free=fr, tol=1e-12, maxit=500)
    # the mixed-precision driver reports it from its FP32 inner solve
    with pytest.raises(FloatingPointError):
        ops.pcg_refined(dims, 1.0, t(Dv), t(Kn), t(Tv), t(Kn).float(), t(Tv).float(), minv32=minv.float(), free=fr,
                        tol=1e-8, maxit=500)


@pytest.mark.parametrize("where", ["kn", "tv", "b"])
def test_nan_inputs_raise_nonfinite(dev, where):
    print("nan", where, flush=True)
    dims = (7, 6, 5)
    lat, Kn, Tv, fixed, b = _case3(dev, dims)
    Kn, Tv, b = Kn.copy(), Tv.copy(), b.copy()
    p = (2 * dims[1] + 3) * dims[0] + 3                  # interior (free) particle (3, 3, 2)
    assert not fixed[p]
    if where == "kn":
        Kn[p, 0] = np.nan
    elif where == "tv":
        Tv[p, 6] = np.nan
    else:
        b[p, 1] = np.nan
    t = _t(dev)
    fr = torch.as_tensor(~fixed, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=fr)
    for precond in (True, False):
        with pytest.raises(RuntimeError, match="non-finite"):
            ops.pcg(dims, 1.0, t(b), t(Kn), t(Tv), minv=minv if precond else None, free=fr, tol=1e-10,
                    maxit=100000)
    kp32, tp32 = t(Kn).float(), t(Tv).float()
    with pytest.raises(RuntimeError, match="non-finite"):
        ops.pcg_refined(dims, 1.0, t(b), t(Kn), t(Tv), kp32, tp32, minv32=minv.float(), free=fr, tol=1e-8,
                        maxit=100000)


def test_maxit_reports_not_converged_with_the_oracle_residual(dev):
    print("maxit", flush=True)
    dims = (7, 6, 5)
    lat, Kn, Tv, fixed, b = _case3(dev, dims)
    t = _t(dev)
    fr = torch.as_tensor(~fixed, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=fr)
    for maxit in (1, 5):
        x, info = ops.pcg(dims, 1.0, t(b), t(Kn), t(Tv), minv=minv, free=fr, tol=1e-10, maxit=maxit, check=3)
        xr, it, rr = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, ~fixed), ~fixed, tol=1e-10, maxit=maxit)
        assert not info.conv.any() and info.it == maxit == it
        assert np.isfinite(info.relmax) and info.relmax > 1e-10
        assert abs(info.relmax - rr) <= 1e-10 * rr
        assert np.abs(x.cpu().numpy() - xr).max() <= 1e-10 * np.abs(xr).max()
    # the drop-in surface reports the same without raising (the reference returns it and relres)
    op = T3.SCT(T3.LatticeGeom(*dims), device=dev)
    _, it, rel = op.pcg(t(b), t(Kn), t(Tv), op.precond(t(Kn), t(Tv), fr.double()), fr.double(), tol=1e-10, maxit=5)
    assert it == 5 and rel > 1e-10 and not op.last_info.conv.any()
    # refined mode: the FP64 refinement stops at max_outer with the true residual reported
    x, info = ops.pcg_refined(dims, 1.0, t(b), t(Kn), t(Tv), t(Kn).float(), t(Tv).float(), minv32=minv.float(),
                              free=fr, tol=1e-12, maxit=3, inner_tol=1e-6, max_outer=2)
    assert not info.conv.any() and info.outer == 2 and np.isfinite(info.relmax) and info.relmax > 1e-12


def test_zero_rhs_is_done_at_iteration_zero(dev):
    print("zero rhs", flush=True)
    dims = (6, 5, 4)
    lat, Kn, Tv, fixed, b = _case3(dev, dims)
    t = _t(dev)
    fr = torch.as_tensor(~fixed, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=fr)
    bb = torch.stack([t(b), torch.zeros_like(t(b))])
    x, info = ops.pcg(dims, 1.0, bb, t(Kn), t(Tv), minv=minv, free=fr, tol=1e-12, maxit=1000)
    assert info.conv.all() and info.iters[0, 1] == 0 and float(x[1].abs().max()) == 0.0 and info.iters[0, 0] > 0
