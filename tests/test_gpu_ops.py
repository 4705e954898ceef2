"""HIP kernels of liblpm.so against the oracle (oracle/lpm_ref.py) on the same seeded inputs.

Tolerances: FP64 kernels agree with the FP64 oracle to 1e-12 relative (max-norm); FP32 kernels to 2e-6 relative
(FP32 round-off of sums of 18 terms with O(1) coefficients)."""
import numpy as np
import pytest
import torch

import lpm_ref as R
import lpm_ops as ops
import _lpmlib as L

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 2e-6}


def rel(a, b):
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def case(dims, seed=0, two_phase=False):
    rng = np.random.default_rng(seed)
    lat = R.Box(dims)
    if two_phase:
        ph = rng.random(lat.N) < 0.3
        E = np.where(ph, 10.0, 1.0)
        C = R.iso_C(E, np.full(lat.N, 0.25))
        pp = R.sc_params_from_C(C, 0.5) if lat.dim == 3 else R.square_params_from_C(C)
        Kn, Tv = R.two_end_mean(lat, pp)
    else:
        Kn, Tv = R.random_symmetric_params(lat, rng)
    u = rng.standard_normal((lat.N, lat.dim))
    return lat, Kn, Tv, u


DIMS = [(12, 12, 12), (7, 5, 9), (50, 9, 6), (3, 2, 1), (1, 1, 1), (16, 16), (37, 5), (1, 3)]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_apply_directed(dev, dims, dt):
    print("apply", dims, dt, flush=True)
    lat, Kn, Tv, u = case(dims, seed=len(dims) * 7 + dims[0])
    ref = R.apply(lat, u, Kn, Tv)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    out = ops.apply(dims, 1.0, t(u), t(Kn), t(Tv))
    assert rel(out, ref) < TOL[dt]


@pytest.mark.parametrize("dims", [(12, 12, 12), (64, 10, 5), (70, 9, 4), (1, 1, 1), (3, 2, 1), (20, 14), (2, 3)])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_apply_packed_matches_directed(dev, dims, dt):
    print("packed", dims, dt, flush=True)
    lat, Kn, Tv, u = case(dims, seed=3, two_phase=True)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    kp, tp, asym = ops.pack(dims, t(Kn), t(Tv))
    assert int(asym.sum()) == 0
    out = ops.apply(dims, 1.0, t(u), kp, tp, layout=L.PACKED)
    assert rel(out, R.apply(lat, u, Kn, Tv)) < TOL[dt]


@pytest.mark.parametrize("dims", [(9, 7, 6), (70, 9, 5), (11, 8)])
def test_slots_leaving_the_box_are_ignored_and_packed_as_zero(dev, dims):
    """Entries of slots whose bond leaves the box are no bonds: garbage there changes neither K u (directed layout)
    nor the packed planes, which hold zeros on such slots (the fused kernels rely on it)."""
    lat, Kn, Tv, u = case(dims, seed=5, two_phase=True)
    rng = np.random.default_rng(8)
    out_of_box = ~lat.valid
    Kg, Tg = Kn.copy(), Tv.copy()
    Kg[out_of_box] = rng.uniform(5.0, 9.0, out_of_box.sum())
    Tg[out_of_box] = rng.uniform(-3.0, 3.0, out_of_box.sum())
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    ref = R.apply(lat, u, Kn, Tv)
    assert rel(ops.apply(dims, 1.0, t(u), t(Kg), t(Tg)), ref) < 1e-12
    kp, tp, asym = ops.pack(dims, t(Kg), t(Tg))
    kp0, tp0, _ = ops.pack(dims, t(Kn), t(Tv))
    assert int(asym.sum()) == 0
    assert torch.equal(kp, kp0) and torch.equal(tp, tp0)
    for f32 in (False, True):
        k, tt_, uu = (kp.float(), tp.float(), t(u).float()) if f32 else (kp, tp, t(u))
        assert rel(ops.apply(dims, 1.0, uu, k, tt_, layout=L.PACKED), ref) < (2e-6 if f32 else 1e-12)


def test_pack_flags_asymmetry(dev):
    lat, Kn, Tv, u = case((6, 5, 4), seed=1)
    Kn = Kn.copy()
    Kn[7, 0] += 0.5
    _, _, asym = ops.pack((6, 5, 4), torch.as_tensor(Kn, device=dev), torch.as_tensor(Tv, device=dev))
    assert int(asym.sum()) == 1


def test_apply_broadcast_batch(dev):
    dims = (10, 9, 8)
    rng = np.random.default_rng(11)
    lat = R.Box(dims)
    B, Lc = 2, 3
    Ks = [R.random_symmetric_params(lat, rng) for _ in range(B)]
    U = rng.standard_normal((B, Lc, lat.N, 3))
    t = lambda a: torch.as_tensor(np.asarray(a), device=dev, dtype=torch.float64)
    Kn = t([k for k, _ in Ks])[:, None]
    Tv = t([v for _, v in Ks])[:, None]
    out = ops.apply(dims, 1.0, t(U), Kn, Tv)
    for b in range(B):
        for l in range(Lc):
            assert rel(out[b, l], R.apply(lat, U[b, l], *Ks[b])) < 1e-12


@pytest.mark.parametrize("dims", [(9, 8, 7), (13, 11)])
def test_post_outputs(dev, dims):
    lat, Kn, Tv, u = case(dims, seed=5, two_phase=True)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    o = ops.post(dims, 1.0, t(u), t(Kn), t(Tv), ku=True, stress=True, energy=True, dl=True, force=True)
    dL, F = R.bond_forces(lat, u, Kn, Tv)
    assert rel(o["ku"], R.apply(lat, u, Kn, Tv)) < 1e-12
    assert rel(o["stress"], R.stress(lat, u, Kn, Tv)) < 1e-12
    assert rel(o["energy"], R.energy(lat, u, Kn, Tv)) < 1e-12
    assert rel(o["dl"], dL) < 1e-12
    assert rel(o["force"], F) < 1e-12


@pytest.mark.parametrize("dims", [(6, 7, 5), (9, 6)])
def test_block_jacobi(dev, dims):
    lat, Kn, Tv, u = case(dims, seed=9, two_phase=True)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    D = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), raw=True)
    assert rel(D, R.block_diag(lat, Kn, Tv)) < 1e-12
    rng = np.random.default_rng(2)
    free = rng.random(lat.N) > 0.2
    M = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=torch.as_tensor(free, device=dev))
    assert rel(ops.full_from_sym(M), R.precond(lat, Kn, Tv, free)) < 1e-11


def _bc_problem(dims, seed):
    lat, Kn, Tv, _ = case(dims, seed=seed, two_phase=True)
    if lat.dim == 3:
        nc = dims
        fixed, up, _ = R.prescribed3("k_xy", lat.x, lat.ijk, nc, 1)
    else:
        fixed, up, _ = R.prescribed2("ring_xx", lat.x, lat.ijk, dims[0], dims[1], 1, 1.0)
    up = up * fixed[:, None]
    b = -R.apply(lat, up, Kn, Tv) * (~fixed)[:, None]
    return lat, Kn, Tv, fixed, up, b


@pytest.mark.parametrize("dims", [(10, 9, 8), (14, 12)])
@pytest.mark.parametrize("precond", [True, False])
def test_pcg_fp64_matches_oracle(dev, dims, precond):
    print("pcg", dims, precond, flush=True)
    lat, Kn, Tv, fixed, up, b = _bc_problem(dims, 4)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    free = torch.as_tensor(~fixed, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=free) if precond else None
    x, info = ops.pcg(dims, 1.0, t(b), t(Kn), t(Tv), minv=minv, free=free, tol=1e-12, maxit=5000)
    if precond:
        xr, it, rr = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, ~fixed), ~fixed, tol=1e-12, maxit=5000)
    else:
        xr, it, rr = R.cg_plain(lat, b, Kn, Tv, np.repeat((~fixed)[:, None], lat.dim, 1), tol=1e-12)
    assert info.conv.all()
    assert abs(info.it - it) <= 2
    K = R.stiffness(lat, Kn, Tv)
    ud = R.solve_direct(K, np.repeat(fixed, lat.dim), up.ravel()).reshape(-1, lat.dim)
    assert rel(x.cpu().numpy() + up, ud) < 1e-9
    assert rel(x, xr) < 1e-9


@pytest.mark.parametrize("dims", [(12, 10, 9), (20, 18)])
def test_refined_reaches_fp64_accuracy(dev, dims):
    print("refined", dims, flush=True)
    lat, Kn, Tv, fixed, up, b = _bc_problem(dims, 8)
    t64 = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    free = torch.as_tensor(~fixed, device=dev)
    k64, t64_ = t64(Kn), t64(Tv)
    k32, t32 = k64.float(), t64_.float()
    minv = ops.block_jacobi(dims, 1.0, k64, t64_, free=free, out_dtype=torch.float32)
    x, info = ops.pcg_refined(dims, 1.0, t64(b), k64, t64_, k32, t32, minv32=minv, free=free, tol=1e-10)
    assert info.conv.all() and info.relmax <= 1e-10
    K = R.stiffness(lat, Kn, Tv)
    ud = R.solve_direct(K, np.repeat(fixed, lat.dim), up.ravel()).reshape(-1, lat.dim)
    assert rel(x.cpu().numpy() + up, ud) < 1e-7


@pytest.mark.parametrize("dims", [(7, 6, 5), (9, 8)])
def test_bond_grad_matches_oracle(dev, dims):
    lat, Kn, Tv, u = case(dims, seed=12, two_phase=True)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(u.shape)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    gk, gt = ops.bond_grad(dims, 1.0, t(Kn), t(Tv), t(u), t(lam))
    rk, rt = R.bond_grad(lat, u, lam, Kn, Tv)
    assert rel(gk, rk) < 1e-12
    assert rel(gt, rt) < 1e-12


@pytest.mark.parametrize("dims", [(7, 6, 5), (9, 8), (1, 1, 1)])
@pytest.mark.parametrize("nb,nl", [(1, 1), (2, 3)])
def test_post_vjp_matches_linearisation(dev, dims, nb, nl):
    """<gu, du> and <gKn, dKn> + <gTv, dTv> equal the oracle's directional derivatives of
    gP.Ku + gS:stress + gU.energy; Ku and stress are linear in u, energy is quadratic (central difference exact), and
    all three are linear in (Kn, Tv) for fixed u."""
    print("post_vjp", dims, nb, nl, flush=True)
    lat = R.Box(dims)
    rng = np.random.default_rng(21)
    d, NN, N = lat.dim, lat.NN, lat.N
    ns = 6 if d == 3 else 3
    Kn = rng.uniform(0.5, 1.5, (nb, 1, N, NN))         # directed, not symmetric: per-slot gradients are defined
    Tv = rng.uniform(-0.2, 0.2, (nb, 1, N, NN))
    u = rng.standard_normal((nb, nl, N, d))
    gP, gS, gU = rng.standard_normal((nb, nl, N, d)), rng.standard_normal((nb, nl, N, ns)), rng.standard_normal((nb, nl, N))
    du = rng.standard_normal(u.shape)
    dK, dT = rng.standard_normal(Kn.shape), rng.standard_normal(Tv.shape)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    gu, gk, gt = ops.post_vjp(dims, 1.0, t(u), t(Kn), t(Tv), g_ku=t(gP), g_stress=t(gS), g_energy=t(gU))
    assert gu.shape == u.shape and gk.shape == Kn.shape and gt.shape == Tv.shape

    def obj(uu, kk, tt, b, l):
        return (np.sum(gP[b, l] * R.apply(lat, uu, kk, tt)) + np.sum(gS[b, l] * R.stress(lat, uu, kk, tt)) +
                np.sum(gU[b, l] * R.energy(lat, uu, kk, tt)))

    ref_u, ref_p = 0.0, 0.0
    for b in range(nb):
        for l in range(nl):
            k, v = Kn[b, 0], Tv[b, 0]
            ref_u += 0.5 * (obj(u[b, l] + du[b, l], k, v, b, l) - obj(u[b, l] - du[b, l], k, v, b, l))
            ref_p += obj(u[b, l], dK[b, 0], dT[b, 0], b, l)
    got_u = float((gu.cpu().numpy() * du).sum())
    got_p = float((gk.cpu().numpy() * dK).sum() + (gt.cpu().numpy() * dT).sum())
    assert abs(got_u - ref_u) < 1e-10 * max(1.0, abs(ref_u))
    assert abs(got_p - ref_p) < 1e-10 * max(1.0, abs(ref_p))


@pytest.mark.parametrize("dims", [(6, 5, 4), (8, 7)])
def test_post_vjp_packed_and_fp32(dev, dims):
    """Packed symmetric coefficients give the same gu and directed gradients as the directed layout; FP32 agrees
    with FP64 to FP32 round-off."""
    lat, Kn, Tv, u = case(dims, seed=4, two_phase=True)
    rng = np.random.default_rng(5)
    ns = 6 if lat.dim == 3 else 3
    gP, gS, gU = rng.standard_normal(u.shape), rng.standard_normal((lat.N, ns)), rng.standard_normal(lat.N)
    t = lambda a, dt=torch.float64: torch.as_tensor(a, device=dev, dtype=dt)
    ref = ops.post_vjp(dims, 1.0, t(u), t(Kn), t(Tv), g_ku=t(gP), g_stress=t(gS), g_energy=t(gU))
    kp, tp, asym = ops.pack(dims, t(Kn), t(Tv))
    assert int(asym.sum()) == 0
    pk = ops.post_vjp(dims, 1.0, t(u), kp, tp, g_ku=t(gP), g_stress=t(gS), g_energy=t(gU), layout=L.PACKED)
    for a, b in zip(pk, ref):
        assert rel(a, b.cpu().numpy()) < 1e-12
    f32 = ops.post_vjp(dims, 1.0, t(u, torch.float32), t(Kn, torch.float32), t(Tv, torch.float32),
                       g_ku=t(gP, torch.float32), g_stress=t(gS, torch.float32), g_energy=t(gU, torch.float32))
    for a, b in zip(f32, ref):
        assert rel(a, b.cpu().numpy()) < 2e-5


@pytest.mark.parametrize("dims", [(10, 9, 8), (14, 12)])
def test_pcg_graph_replay_matches_eager(dev, dims):
    """The check-iteration block replayed as a captured hipGraph (many replays per solve, with the host's flag
    polls between them) gives the eager iteration bit for bit, twice in a row."""
    lat, Kn, Tv, _ = case(dims, seed=31, two_phase=True)
    fixed, up, _ = (R.prescribed3("k_xy", lat.x, lat.ijk, dims, 1) if lat.dim == 3 else
                    R.prescribed2("ring_xy", lat.x, lat.ijk, dims[0], dims[1], 1, 1.0))
    b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * (~fixed)[:, None]
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    free = torch.as_tensor(~fixed, device=dev)
    minv = ops.block_jacobi(dims, 1.0, t(Kn), t(Tv), free=free)
    xe, ie = ops.pcg(dims, 1.0, t(b), t(Kn), t(Tv), minv=minv, free=free, tol=1e-12, maxit=5000, check=5)
    for _ in range(2):
        xg, ig = ops.pcg(dims, 1.0, t(b), t(Kn), t(Tv), minv=minv, free=free, tol=1e-12, maxit=5000, check=5,
                         graph=True)
        assert ig.it == ie.it and ie.it > 20
        assert torch.equal(xg, xe)


_TILING_REF = {}


def _tiling_case():
    """The 70 x 11 x 40 case of the tiling test with its oracle K u and PCG solution (computed once)."""
    if not _TILING_REF:
        dims = (70, 11, 40)
        lat, Kn, Tv, u = case(dims, seed=17, two_phase=True)
        fixed, up, _ = R.prescribed3("tb_sx", lat.x, lat.ijk, dims, 1)
        b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * (~fixed)[:, None]
        xr, _, _ = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, ~fixed), ~fixed, tol=1e-11, maxit=5000)
        _TILING_REF.update(dims=dims, lat=lat, Kn=Kn, Tv=Tv, u=u, ref=R.apply(lat, u, Kn, Tv), fixed=fixed, b=b, xr=xr)
    return _TILING_REF


@pytest.mark.parametrize("wide,zc", [(1, 7), (1, 13), (-1, 5), (2, 11), (1, 0)])
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_apply_tilings_and_z_chunks(dev, wide, zc, dt):
    """Every tile shape and z-chunk length (chunks that end inside the 6-step unrolled group, a last chunk shorter
    than the others) gives the oracle's K u, packed and directed, plain and inside PCG."""
    print("tiling", wide, zc, dt, flush=True)
    c = _tiling_case()
    dims, Kn, Tv, u, ref = c["dims"], c["Kn"], c["Tv"], c["u"], c["ref"]
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    kp, tp, _ = ops.pack(dims, t(Kn), t(Tv))
    try:
        L.lib().lpm_set_ku_tiling(wide, zc)
        assert rel(ops.apply(dims, 1.0, t(u), kp, tp, layout=L.PACKED), ref) < TOL[dt]
        assert rel(ops.apply(dims, 1.0, t(u), t(Kn), t(Tv)), ref) < TOL[dt]
        free = torch.as_tensor(~c["fixed"], device=dev)
        if dt == torch.float64:
            x, info = ops.pcg(dims, 1.0, t(c["b"]), kp, tp, minv=ops.block_jacobi(dims, 1.0, kp, tp, free=free,
                              layout=L.PACKED), free=free, tol=1e-11, maxit=5000, layout=L.PACKED)
            assert info.conv.all() and rel(x, c["xr"]) < 1e-8
    finally:
        L.lib().lpm_set_ku_tiling(0, 0)


_BOX_DIMS = [(1, 1, 1), (3, 2, 1), (17, 5, 9), (33, 7, 6), (48, 9, 13)]


@pytest.mark.parametrize("dims", _BOX_DIMS)
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_box_kernel_matches_oracle_and_marching_kernel(dev, dims, dt):
    """The box K u kernel (one block per 16 x 4 x 4 sites: partial boxes on every axis, single sites) and the
    marching kernel, each forced, give the oracle's K u, packed and directed; inside the FP64 PCG (masked-dot form)
    and the refined solver (FP64 residual form) the box kernel reaches the oracle's PCG solution."""
    print("box", dims, dt, flush=True)
    lat, Kn, Tv, u = case(dims, seed=23, two_phase=True)
    ref = R.apply(lat, u, Kn, Tv)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    kp, tp, _ = ops.pack(dims, t(Kn), t(Tv))
    try:
        outs = {}
        for box in (1, 0):
            L.lib().lpm_set_ku_box(box)
            outs[box] = ops.apply(dims, 1.0, t(u), kp, tp, layout=L.PACKED)
            assert rel(outs[box], ref) < TOL[dt], box
            assert rel(ops.apply(dims, 1.0, t(u), t(Kn), t(Tv)), ref) < TOL[dt], box
        assert rel(outs[1], outs[0].double().cpu().numpy()) < TOL[dt]
        L.lib().lpm_set_ku_box(1)
        if dt == torch.float64 and min(dims) >= 2:
            fixed, up, _ = R.prescribed3("tb_tz", lat.x, lat.ijk, dims, 1)
            b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * (~fixed)[:, None]
            xr, _, _ = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, ~fixed), ~fixed, tol=1e-12, maxit=20000)
            free = torch.as_tensor(~fixed, device=dev)
            minv = ops.block_jacobi(dims, 1.0, kp, tp, free=free, layout=L.PACKED)
            x, info = ops.pcg(dims, 1.0, t(b), kp, tp, minv=minv, free=free, tol=1e-12, maxit=20000,
                              layout=L.PACKED)
            assert info.conv.all() and rel(x, xr) < 1e-9
            k32, t32, _ = ops.pack(dims, t(Kn), t(Tv), out_dtype=torch.float32)
            m32 = ops.block_jacobi(dims, 1.0, kp, tp, free=free, layout=L.PACKED, out_dtype=torch.float32)
            x, info = ops.pcg_refined(dims, 1.0, t(b), kp, tp, k32, t32, minv32=m32, free=free, tol=1e-10,
                                      maxit=20000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
            assert info.conv.all() and rel(x, xr) < 1e-8
    finally:
        L.lib().lpm_set_ku_box(-1)


_LEAN_DIMS = [(64, 4, 1), (64, 1, 3), (128, 5, 7), (130, 9, 13), (64, 8, 20), (96, 6, 15)]
_LEAN_REF = {}


@pytest.mark.parametrize("dims", _LEAN_DIMS)
@pytest.mark.parametrize("zc", [0, 7, 9])
def test_lean_kernel_matches_oracle_and_general_kernel(dev, dims, zc):
    """The lean FP32 packed K u (64-wide tiles, used when nx >= 48 without bond words): single planes, partial x and
    y tiles, z chunks that end inside the unrolled group; plain K u against the oracle (FP32 bar) and against the
    general kernel, and its masked-dot form inside the refined solver against the oracle's FP64 PCG solution."""
    print("lean", dims, zc, flush=True)
    lat, Kn, Tv, u = case(dims, seed=21, two_phase=True)
    t64 = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    kp, tp, _ = ops.pack(dims, t64(Kn), t64(Tv), out_dtype=torch.float32)
    u32 = t64(u).float()
    ref = R.apply(lat, u, Kn, Tv)
    try:
        L.lib().lpm_set_ku_tiling(1, zc)
        L.lib().lpm_set_ku_kernel(1)
        lean = ops.apply(dims, 1.0, u32, kp, tp, layout=L.PACKED)
        assert rel(lean, ref) < TOL[torch.float32]
        L.lib().lpm_set_ku_kernel(0)
        general = ops.apply(dims, 1.0, u32, kp, tp, layout=L.PACKED)
        L.lib().lpm_set_ku_kernel(1)
        assert rel(lean, general.double().cpu().numpy()) < 1e-6
        if dims[2] >= 3 and dims[1] >= 3:
            if dims not in _LEAN_REF:
                fixed, up, _ = R.prescribed3("tb_tz", lat.x, lat.ijk, dims, 1)
                b = -R.apply(lat, up * fixed[:, None], Kn, Tv) * (~fixed)[:, None]
                xr, _, _ = R.pcg(lat, b, Kn, Tv, R.precond(lat, Kn, Tv, ~fixed), ~fixed, tol=1e-12, maxit=20000)
                _LEAN_REF[dims] = (fixed, b, xr)
            fixed, b, xr = _LEAN_REF[dims]
            free = torch.as_tensor(~fixed, device=dev)
            k64, tp64, _ = ops.pack(dims, t64(Kn), t64(Tv))
            minv = ops.block_jacobi(dims, 1.0, k64, tp64, free=free, layout=L.PACKED, out_dtype=torch.float32)
            x, info = ops.pcg_refined(dims, 1.0, t64(b), k64, tp64, kp, tp, minv32=minv, free=free, tol=1e-10,
                                      maxit=20000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
            assert info.conv.all() and rel(x, xr) < 1e-8
    finally:
        L.lib().lpm_set_ku_kernel(0)
        L.lib().lpm_set_ku_tiling(0, 0)
