"""lpm_pcg_solve_noalloc (workspace in the library's static device scratch) against lpm_pcg_solve (caller
workspace), both called through the C ABI: two calls in a row, the second on a smaller lattice, in FP64 (3D block
Jacobi PCG) and FP32 (2D plain CG with a per-DOF mask).  Same kernels, same fixed-order reductions: the solutions,
iteration counts and residuals must be identical.  A solve larger than the scratch is refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

import lpm_ref as R
import lpm_ops as ops
import _lpmlib as L

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _problem(dims, dev, dt, precond, seed):
    lat = R.Box(dims)
    rng = np.random.default_rng(seed)
    Kn, Tv = R.random_symmetric_params(lat, rng)
    d = lat.dim
    fixed = lat.ijk[:, 0] == 0
    b = rng.standard_normal((lat.N, d)) * (~fixed)[:, None]
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    free = torch.as_tensor(~fixed, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=free) if precond else None
    mask = torch.as_tensor(np.repeat((~fixed)[:, None], d, 1).astype(np.uint8), device=dev).contiguous()
    return t(Kn).contiguous(), t(Tv).contiguous(), t(b).contiguous(), mask, minv


def _solve(dims, Kn, Tv, b, mask, minv, noalloc, tol=1e-10):
    lib = L.lib()
    g = L.grid_of(dims, 1.0)
    nsys = 1
    dt = L.dtype_code(b.dtype)
    o = ops._opts(minv is not None, tol, 20000, 10, False)
    x = torch.empty_like(b)
    iters = np.zeros(nsys, np.int32)
    rel = np.zeros(nsys, np.float64)
    conv = np.zeros(nsys, np.int32)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    z = L.Bcast(0, 0)
    st = L.stream_of(b)
    if noalloc:
        rc = lib.lpm_pcg_solve_noalloc(ctypes.byref(g), 1, 1, dt, L.DIRECTED, _ptr(Kn), _ptr(Tv), z, _ptr(mask), z,
                                       _ptr(minv), z, _ptr(b), _ptr(x), 0, ctypes.byref(o), vp(iters), vp(rel),
                                       vp(conv), st)
    else:
        work = torch.empty(int(lib.lpm_pcg_workspace(ctypes.byref(g), nsys, dt)), device=b.device, dtype=torch.uint8)
        rc = lib.lpm_pcg_solve(ctypes.byref(g), 1, 1, dt, L.DIRECTED, _ptr(Kn), _ptr(Tv), z, _ptr(mask), z,
                               _ptr(minv), z, _ptr(b), _ptr(x), 0, ctypes.byref(o), _ptr(work), vp(iters), vp(rel),
                               vp(conv), st)
    L.check(rc)
    torch.cuda.synchronize()
    return x, int(iters[0]), float(rel[0]), int(conv[0])


@pytest.mark.parametrize("case", [((24, 20, 16), (9, 7, 5), torch.float64, True),
                                  ((96, 80), (21, 13), torch.float32, False)])
def test_noalloc_solve_equals_workspace_solve_twice(dev, case):
    big, small, dt, precond = case
    print("noalloc", big, small, dt, flush=True)
    for dims, seed in ((big, 1), (small, 2)):
        Kn, Tv, b, mask, minv = _problem(dims, dev, dt, precond, seed)
        xa, ia, ra, ca = _solve(dims, Kn, Tv, b, mask, minv, noalloc=False)
        xb, ib, rb, cb = _solve(dims, Kn, Tv, b, mask, minv, noalloc=True)
        assert ca == cb == 1 and ia == ib and ra == rb
        assert torch.equal(xa, xb)
        # and the solution is the oracle's (the existing entry point's parity bar)
        lat = R.Box(dims)
        if precond:
            fixed = lat.ijk[:, 0] == 0
            xr, _, _ = R.pcg(lat, b.cpu().numpy(), Kn.cpu().numpy(), Tv.cpu().numpy(),
                             R.precond(lat, Kn.cpu().numpy(), Tv.cpu().numpy(), ~fixed), ~fixed, tol=1e-12)
            assert np.abs(xb.cpu().numpy() - xr).max() <= 1e-8 * np.abs(xr).max()


def test_noalloc_refuses_a_solve_beyond_its_scratch(dev):
    lib = L.lib()
    cap = int(lib.lpm_noalloc_scratch_bytes())
    g = L.grid_of((256, 256, 256), 1.0)
    assert int(lib.lpm_pcg_workspace(ctypes.byref(g), 1, L.F64)) > cap
    o = ops._opts(True, 1e-8, 10, 10, False)
    z = L.Bcast(0, 0)
    dummy = ctypes.c_void_p(16)
    rc = lib.lpm_pcg_solve_noalloc(ctypes.byref(g), 1, 1, L.F64, L.PACKED, dummy, dummy, z, None, z, dummy, z, dummy,
                                   dummy, 0, ctypes.byref(o), None, None, None, None)
    assert rc == L.LPM_E_BADARG and b"static scratch" in lib.lpm_last_error()
