"""Removed particles (voids) on the GPU against the oracle's reduced lattice (oracle/lpm_ref.py ActiveBox, the
reference's lattice with only active sites numbered).

The GPU runs the full box with bond words; its results on the solid particles must equal the oracle's on the
compact lattice, and removed particles get zero K u, stress and energy whatever their u holds.  Tolerances as in
test_gpu_ops: FP64 1e-12 relative (max-norm), FP32 2e-6; solves against SuperLU on the oracle's assembled K."""
import numpy as np
import pytest
import torch

import lpm_ref as R
import lpm_ops as ops
import _lpmlib as L
import synth

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 2e-6}


def rel(a, b):
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def porous(dims, seed, frac=0.2, clean=True):
    """Random solid set with spherical-ish pores (3D cleaned like the reference's materials); 2D random holes."""
    rng = np.random.default_rng(seed)
    if len(dims) == 3:
        img = np.ones(tuple(reversed(dims)), bool)
        n_p = max(1, int(frac * img.size / 7))
        for _ in range(n_p):
            c = [rng.integers(0, d) for d in dims]
            img[max(c[2] - (rng.random() < 0.5), 0):c[2] + 1, c[1], max(c[0] - 1, 0):c[0] + 1] = False
        if clean:
            img, _ = synth.clean_solid(img)
        return img.ravel()
    return rng.random(int(np.prod(dims))) >= frac


def case(dims, seed, frac=0.2):
    rng = np.random.default_rng(seed)
    full = R.Box(dims)
    solid = porous(dims, seed, frac)
    lat = R.ActiveBox(dims, solid)
    Kf, Tf = R.random_symmetric_params(full, rng, kn=(0.5, 1.5), tv=(-0.05, 0.05))
    u = rng.standard_normal((full.N, full.dim))
    return full, lat, solid, Kf, Tf, u


DIMS = [(9, 8, 7), (70, 9, 13), (33, 17, 20), (13, 11), (40, 9)]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_apply_with_voids_matches_reduced_lattice(dev, dims, dt):
    print("void apply", dims, dt, flush=True)
    full, lat, solid, Kf, Tf, u = case(dims, seed=sum(dims))
    assert 0 < lat.N < full.N
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev))
    ref = R.apply(lat, u[lat.sites], Kf[lat.sites], Tf[lat.sites])
    ug = t(u)
    ug[torch.as_tensor(~solid, device=dev)] = 1e3                 # removed particles' u must not matter
    out = ops.apply(dims, 1.0, ug, t(Kf), t(Tf), bonds=w)
    sites = torch.as_tensor(lat.sites, device=dev)
    assert rel(out[sites], ref) < TOL[dt]
    assert float(out[torch.as_tensor(~solid, device=dev)].abs().max()) == 0.0
    kp, tp, asym = ops.pack(dims, t(Kf), t(Tf))
    assert int(asym.sum()) == 0
    outp = ops.apply(dims, 1.0, ug, kp, tp, layout=L.PACKED, bonds=w)
    assert rel(outp[sites], ref) < TOL[dt]


def test_bond_words_bits(dev):
    dims = (6, 5, 4)
    full, lat, solid, _, _, _ = case(dims, seed=3, frac=0.3)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev)).cpu().numpy().view(np.uint32)
    bits = (w[:, None] >> np.arange(18)) & 1
    expect = np.zeros((full.N, 18), bool)
    expect[lat.sites] = lat.valid
    assert np.array_equal(bits.astype(bool), expect)


@pytest.mark.parametrize("dims", [(9, 8, 7), (70, 9, 13), (13, 11)])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_post_and_block_diag_with_voids(dev, dims, dt):
    print("void post", dims, dt, flush=True)
    full, lat, solid, Kf, Tf, u = case(dims, seed=5 + dims[0])
    t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev))
    sites = torch.as_tensor(lat.sites, device=dev)
    kc, tc, uc = Kf[lat.sites], Tf[lat.sites], u[lat.sites]
    o = ops.post(dims, 1.0, t(u), t(Kf), t(Tf), ku=True, stress=True, energy=True, dl=True, force=True, bonds=w)
    dL, F = R.bond_forces(lat, uc, kc, tc)
    assert rel(o["ku"][sites], R.apply(lat, uc, kc, tc)) < TOL[dt]
    assert rel(o["stress"][sites], R.stress(lat, uc, kc, tc)) < TOL[dt]
    assert rel(o["energy"][sites], R.energy(lat, uc, kc, tc)) < TOL[dt]
    assert rel(o["dl"][sites], dL) < TOL[dt] and rel(o["force"][sites], F) < TOL[dt]
    void = torch.as_tensor(~solid, device=dev)
    for key in ("ku", "stress", "energy", "force"):
        assert float(o[key][void].abs().max()) == 0.0
    D = ops.block_jacobi(dims, 1.0, t(Kf), t(Tf), raw=True, bonds=w)
    assert rel(D[sites], R.block_diag(lat, kc, tc)) < TOL[dt]


def test_batched_masks_broadcast_over_loads(dev):
    """Two specimens with their own masks and coefficients, three right-hand sides each: words (B, 1, N)."""
    dims = (20, 9, 11)
    cases = [case(dims, seed=40 + b, frac=0.15 + 0.1 * b) for b in range(2)]
    rng = np.random.default_rng(2)
    u = rng.standard_normal((2, 3) + (cases[0][0].N, 3))
    t = lambda a: torch.as_tensor(np.asarray(a), device=dev, dtype=torch.float64)
    solid = torch.as_tensor(np.array([c[2] for c in cases]), device=dev)
    w = ops.bond_words(dims, solid)[:, None]
    Kn = t([c[3] for c in cases])[:, None]
    Tv = t([c[4] for c in cases])[:, None]
    out = ops.apply(dims, 1.0, t(u), Kn, Tv, bonds=w)
    for b, (full, lat, _, Kf, Tf, _) in enumerate(cases):
        for l in range(3):
            ref = R.apply(lat, u[b, l][lat.sites], Kf[lat.sites], Tf[lat.sites])
            assert rel(out[b, l][torch.as_tensor(lat.sites, device=dev)], ref) < 1e-12


def _void_bvp(dims, seed, load):
    full, lat, solid, Kf, Tf, _ = case(dims, seed)
    fixed_f, up_f, _ = R.prescribed3(load, full.x, full.ijk, dims, 1)
    fixed = fixed_f[lat.sites]
    up = up_f[lat.sites] * fixed[:, None]
    kc, tc = Kf[lat.sites], Tf[lat.sites]
    K = R.stiffness(lat, kc, tc)
    ud = R.solve_direct(K, np.repeat(fixed, 3), up.ravel()).reshape(-1, 3)
    return full, lat, solid, Kf, Tf, fixed_f, up_f, ud


@pytest.mark.parametrize("dims,load", [((14, 12, 10), "tb_tz"), ((70, 9, 13), "k_xz")])
def test_pcg_with_voids_matches_superlu(dev, dims, load):
    print("void pcg", dims, load, flush=True)
    full, lat, solid, Kf, Tf, fixed_f, up_f, ud = _void_bvp(dims, 7, load)
    t = lambda a, dt=torch.float64: torch.as_tensor(a, device=dev, dtype=dt)
    sd = torch.as_tensor(solid, device=dev)
    w = ops.bond_words(dims, sd)
    free = t((~fixed_f) & solid)
    up = t(up_f) * t(solid)[:, None]
    sites = torch.as_tensor(lat.sites, device=dev)
    for layout in (L.DIRECTED, L.PACKED):
        if layout == L.PACKED:
            kk, tt, _ = ops.pack(dims, t(Kf), t(Tf))
        else:
            kk, tt = t(Kf), t(Tf)
        b = -ops.apply(dims, 1.0, up, kk, tt, layout=layout, bonds=w) * free[:, None]
        M = ops.block_jacobi(dims, 1.0, kk, tt, free=free, layout=layout, bonds=w)
        x, info = ops.pcg(dims, 1.0, b, kk, tt, minv=M, free=free, tol=1e-12, maxit=20000, layout=layout, bonds=w)
        assert info.conv.all()
        u = up + x
        assert rel(u[sites], ud) < 1e-9
        assert float(x[~sd].abs().max()) == 0.0
    # FP32 inner iterations in FP64 refinement, packed
    k64, t64, _ = ops.pack(dims, t(Kf), t(Tf))
    M32 = ops.block_jacobi(dims, 1.0, k64, t64, free=free, out_dtype=torch.float32, layout=L.PACKED, bonds=w)
    b = -ops.apply(dims, 1.0, up, k64, t64, layout=L.PACKED, bonds=w) * free[:, None]
    x, info = ops.pcg_refined(dims, 1.0, b, k64, t64, k64.float(), t64.float(), minv32=M32, free=free, tol=1e-10,
                              maxit=20000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED, bonds=w)
    assert info.relmax <= 1e-10
    assert rel((up + x)[sites], ud) < 1e-7


def test_bond_grad_with_voids_matches_oracle(dev):
    dims = (11, 9, 8)
    full, lat, solid, Kf, Tf, u = case(dims, seed=21)
    lam = np.random.default_rng(22).standard_normal(u.shape)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev))
    gk, gt = ops.bond_grad(dims, 1.0, t(Kf), t(Tf), t(u), t(lam), bonds=w)
    rk, rt = R.bond_grad(lat, u[lat.sites], lam[lat.sites], Kf[lat.sites], Tf[lat.sites])
    ek, et = np.zeros_like(Kf), np.zeros_like(Tf)
    ek[lat.sites], et[lat.sites] = rk, rt
    assert rel(gk, ek) < 1e-12 and rel(gt, et) < 1e-12


def test_post_vjp_with_voids_matches_linearisation(dev):
    """<gP, dP> + <gS, dS> + <gU, dU> for a perturbation (du, dKn, dTv) equals the VJP's inner product (the outputs
    are linear in u for fixed parameters, linear in (Kn, Tv) for fixed u, and quadratic only through the energy)."""
    dims = (10, 9, 7)
    full, lat, solid, Kf, Tf, u = case(dims, seed=31)
    rng = np.random.default_rng(32)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev))[None, None]
    gP, gS, gU = t(rng.standard_normal((1, 2, full.N, 3))), t(rng.standard_normal((1, 2, full.N, 6))), \
        t(rng.standard_normal((1, 2, full.N)))
    uu = t(rng.standard_normal((1, 2, full.N, 3)))
    kk, tt = t(Kf)[None, None], t(Tf)[None, None]
    gu, gK, gT = ops.post_vjp(dims, 1.0, uu, kk, tt, g_ku=gP, g_stress=gS, g_energy=gU, bonds=w)

    def obj(u_, k_, t_):
        o = ops.post(dims, 1.0, u_, k_, t_, ku=True, stress=True, energy=True, bonds=w)
        return float((gP * o["ku"]).sum() + (gS * o["stress"]).sum() + (gU * o["energy"]).sum())

    du = t(rng.standard_normal(uu.shape))
    dk = t(R.random_symmetric_params(full, rng, (-1, 1), (-1, 1))[0])[None, None]
    dtv = t(R.random_symmetric_params(full, rng, (-1, 1), (-1, 1))[1])[None, None]
    e = 1e-4
    fd = (obj(uu + e * du, kk + e * dk, tt + e * dtv) - obj(uu - e * du, kk - e * dk, tt - e * dtv)) / (2 * e)
    ad = float((gu * du).sum() + (gK * dk).sum() + (gT * dtv).sum())
    assert abs(fd - ad) < 1e-8 * abs(fd)


@pytest.mark.parametrize("wide,zc", [(1, 7), (-1, 5), (2, 11), (1, 0)])
def test_void_tilings_and_z_chunks(dev, wide, zc):
    print("void tiling", wide, zc, flush=True)
    dims = (70, 11, 40)
    full, lat, solid, Kf, Tf, u = case(dims, seed=17, frac=0.25)
    ref = R.apply(lat, u[lat.sites], Kf[lat.sites], Tf[lat.sites])
    sites = torch.as_tensor(lat.sites, device=dev)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev))
    try:
        L.lib().lpm_set_ku_tiling(wide, zc)
        for dt in (torch.float32, torch.float64):
            t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
            kp, tp, _ = ops.pack(dims, t(Kf), t(Tf))
            assert rel(ops.apply(dims, 1.0, t(u), kp, tp, layout=L.PACKED, bonds=w)[sites], ref) < TOL[dt]
            assert rel(ops.apply(dims, 1.0, t(u), t(Kf), t(Tf), bonds=w)[sites], ref) < TOL[dt]
    finally:
        L.lib().lpm_set_ku_tiling(0, 0)


def test_void_box_kernel(dev):
    """The box K u kernel with bond words (forced on), both value types and layouts, against the oracle."""
    print("void box", flush=True)
    dims = (37, 11, 14)
    full, lat, solid, Kf, Tf, u = case(dims, seed=19, frac=0.25)
    ref = R.apply(lat, u[lat.sites], Kf[lat.sites], Tf[lat.sites])
    sites = torch.as_tensor(lat.sites, device=dev)
    w = ops.bond_words(dims, torch.as_tensor(solid, device=dev))
    try:
        L.lib().lpm_set_ku_box(1)
        for dt in (torch.float32, torch.float64):
            t = lambda a: torch.as_tensor(a, device=dev, dtype=dt)
            kp, tp, _ = ops.pack(dims, t(Kf), t(Tf))
            assert rel(ops.apply(dims, 1.0, t(u), kp, tp, layout=L.PACKED, bonds=w)[sites], ref) < TOL[dt]
            assert rel(ops.apply(dims, 1.0, t(u), t(Kf), t(Tf), bonds=w)[sites], ref) < TOL[dt]
    finally:
        L.lib().lpm_set_ku_box(-1)


# ------------------------------------------------------------------------------------------------ torch layer
import torch_lpm3d as T3  # noqa: E402
import lin3d  # noqa: E402
from geom3d import LOADS, PLATES, GDIR  # noqa: E402


def _cells(nc, B, seed, frac=0.15):
    return np.array([porous((nc, nc, nc), seed + b, frac) for b in range(B)])


def _oracle_run(nc, s, cell, Kf, Tf):
    """Oracle u, stress, energy density and reactions of the nine load cases on the reduced coarse lattice."""
    full = R.Box((nc, nc, nc), s * 1.0)
    lat = R.ActiveBox((nc, nc, nc), cell, s * 1.0)
    kc, tc = Kf[lat.sites], Tf[lat.sites]
    K = R.stiffness(lat, kc, tc)
    out = dict(u=[], sig=[], w=[], F=[])
    for ld in LOADS:
        fixed_f, up_f, mov_f = R.prescribed3(ld, full.x, full.ijk, (nc,) * 3, s, 1.0)
        fixed = fixed_f[lat.sites]
        uc = R.solve_direct(K, np.repeat(fixed, 3), (up_f[lat.sites] * fixed[:, None]).ravel()).reshape(-1, 3)
        u = np.zeros((full.N, 3))
        u[lat.sites] = uc
        sig = np.zeros((full.N, 6))
        sig[lat.sites] = R.stress(lat, uc, kc, tc)
        w = np.zeros(full.N)
        w[lat.sites] = R.energy(lat, uc, kc, tc) / (s ** 3)
        P = R.apply(lat, uc, kc, tc)
        F = (P[:, GDIR[ld]] * mov_f[lat.sites]).sum() if ld in PLATES else 0.0
        for k, v in zip(("u", "sig", "w", "F"), (u, sig, w, F)):
            out[k].append(v)
    return {k: np.array(v) for k, v in out.items()}


def test_sctv_without_mask_is_sct(dev):
    nc = 6
    lat = T3.LatticeGeom(nc, nc, nc, 2.0)
    rng = np.random.default_rng(3)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    Kf, Tf = R.random_symmetric_params(R.Box((nc,) * 3, 2.0), rng)
    u = t(rng.standard_normal((lat.N, 3)))
    a, b = T3.SCT(lat, dev), T3.SCTV(lat, dev)
    b.set_mask(None)
    assert torch.equal(a.apply(u, t(Kf), t(Tf)), b.apply(u, t(Kf), t(Tf)))
    assert torch.equal(a.stress(u, t(Kf), t(Tf)), b.stress(u, t(Kf), t(Tf)))


@pytest.mark.parametrize("mixed", [False, True])
def test_coarsetv_run_matches_oracle(dev, mixed):
    nc, s, B = 6, 2, 2
    cells = _cells(nc, B, 50)
    rng = np.random.default_rng(51)
    full = R.Box((nc,) * 3, s * 1.0)
    KT = [R.random_symmetric_params(full, rng, (0.5, 1.5), (-0.05, 0.05)) for _ in range(B)]
    Kn = torch.as_tensor(np.array([k for k, _ in KT]), device=dev)
    Tv = torch.as_tensor(np.array([t for _, t in KT]), device=dev)
    model = T3.CoarseTV(nc, s, device=dev, mixed=mixed)
    o = model.run(Kn, Tv, cell=torch.as_tensor(cells, device=dev), tol=1e-12, differentiable=False)
    for b in range(B):
        ref = _oracle_run(nc, s, cells[b], KT[b][0], KT[b][1])
        tol = 1e-9 if not mixed else 1e-8
        assert rel(o["u"][b], ref["u"]) < tol
        assert rel(o["sig"][b], ref["sig"]) < tol and rel(o["w"][b], ref["w"]) < tol
        assert rel(o["F"][b], ref["F"]) < tol


def test_coarsetv_gradient_by_finite_differences(dev):
    """Adjoint gradient of a loss of u and sigma through LinSolve + the post VJP with voids, against central
    differences (h = 1e-3 on O(1) parameters) of symmetric perturbations; removed bonds get zero gradient."""
    nc, s, B = 5, 2, 2
    cells = torch.as_tensor(_cells(nc, B, 60), device=dev)
    rng = np.random.default_rng(61)
    full = R.Box((nc,) * 3, s * 1.0)
    KT = [R.random_symmetric_params(full, rng, (0.5, 1.5), (-0.05, 0.05)) for _ in range(B)]
    t = lambda a: torch.as_tensor(np.asarray(a), device=dev, dtype=torch.float64)
    Kn, Tv = t([k for k, _ in KT]), t([v for _, v in KT])
    model = T3.CoarseTV(nc, s, device=dev)
    wu = t(rng.standard_normal((B, 9, full.N, 3)))
    ws = t(rng.standard_normal((B, 9, full.N, 6)))

    def loss(k, v, grad=False):
        k, v = k.clone().requires_grad_(grad), v.clone().requires_grad_(grad)
        o = model.run(k, v, cell=cells, tol=1e-13)
        val = (wu * o["u"]).sum() + 1e3 * (ws * o["sig"]).sum()
        if not grad:
            return float(val)
        val.backward()
        return k.grad, v.grad

    gk, gt = loss(Kn, Tv, True)
    valid = model.op.valid[:, 0]
    assert float((gk * (1 - valid)).abs().max()) == 0.0 and float((gt * (1 - valid)).abs().max()) == 0.0
    h = 1e-3
    for seed in (1, 2):
        r2 = np.random.default_rng(seed)
        dk = t([R.random_symmetric_params(full, r2, (-1, 1), (-1, 1))[0] for _ in range(B)])
        dv = t([R.random_symmetric_params(full, r2, (-1, 1), (-1, 1))[1] for _ in range(B)])
        fd = (loss(Kn + h * dk, Tv) - loss(Kn - h * dk, Tv)) / (2 * h)
        assert abs(float((gk * dk).sum()) - fd) < 1e-5 * abs(fd)
        fd = (loss(Kn, Tv + h * dv) - loss(Kn, Tv - h * dv)) / (2 * h)
        assert abs(float((gt * dv).sum()) - fd) < 1e-5 * abs(fd)


def test_finereftv_matches_oracle(dev):
    n, s = 8, 2
    solid = porous((n, n, n), 70, 0.15)
    full = R.Box((n,) * 3)
    rng = np.random.default_rng(71)
    Kf, Tf = R.random_symmetric_params(full, rng, (0.5, 1.5), (-0.05, 0.05))
    t = lambda a: torch.as_tensor(np.asarray(a), device=dev, dtype=torch.float64)
    model = T3.FineRefTV(n, s, device=dev)
    o = model.run(t(Kf)[None], t(Tf)[None], solid=torch.as_tensor(solid[None], device=dev), tol=1e-12)
    lat = R.ActiveBox((n,) * 3, solid)
    kc, tc = Kf[lat.sites], Tf[lat.sites]
    K = R.stiffness(lat, kc, tc)
    nc = n // s
    cid = ((full.ijk[:, 2] // s) * nc + full.ijk[:, 1] // s) * nc + full.ijk[:, 0] // s
    cnt = np.bincount(cid, weights=solid.astype(float), minlength=nc ** 3)
    for l, ld in enumerate(LOADS):
        fixed_f, up_f, mov_f = R.prescribed3(ld, full.x, full.ijk // s, (nc,) * 3, s, 1.0)
        fixed = fixed_f[lat.sites]
        uc = R.solve_direct(K, np.repeat(fixed, 3), (up_f[lat.sites] * fixed[:, None]).ravel()).reshape(-1, 3)
        u = np.zeros((full.N, 3))
        u[lat.sites] = uc
        um = np.stack([np.bincount(cid, weights=u[:, a], minlength=nc ** 3) for a in range(3)], 1)
        um /= np.maximum(cnt, 1.0)[:, None]
        sg = np.zeros((full.N, 6))
        sg[lat.sites] = R.stress(lat, uc, kc, tc)
        sm = np.stack([np.bincount(cid, weights=sg[:, a], minlength=nc ** 3) for a in range(6)], 1) / s ** 3
        U = np.zeros(full.N)
        U[lat.sites] = R.energy(lat, uc, kc, tc)
        wm = np.bincount(cid, weights=U, minlength=nc ** 3) / s ** 3
        assert rel(o["u"][0, l], um) < 1e-9 and rel(o["sig"][0, l], sm) < 1e-9 and rel(o["w"][0, l], wm) < 1e-9
        if ld in PLATES:
            F = (R.apply(lat, uc, kc, tc)[:, GDIR[ld]] * mov_f[lat.sites]).sum()
            assert abs(float(o["F"][0, l]) - F) < 1e-9 * abs(F)


def test_lin3d_active_lattice(dev):
    dims = (9, 8, 7)
    act = porous(dims, 80, 0.2)
    lat = lin3d.Lattice3D(*dims, active=act, device=dev)
    ref = R.ActiveBox(dims, act)
    assert lat.N == ref.N and np.array_equal(lat.nbr, ref.nbr)
    rng = np.random.default_rng(81)
    Kf, Tf = R.random_symmetric_params(R.Box(dims), rng, (0.5, 1.5), (-0.05, 0.05))
    Kn, Tv = Kf[ref.sites], Tf[ref.sites]
    u = rng.standard_normal((ref.N, 3))
    assert rel(lat.apply(u, Kn, Tv), R.apply(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.stress(u, Kn, Tv), R.stress(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.energy_per_particle(u, Kn, Tv), R.energy(ref, u, Kn, Tv)) < 1e-12
    assert rel(lat.block_diag(Kn, Tv), R.block_diag(ref, Kn, Tv)) < 1e-12
    K = lat.stiffness(Kn, Tv)
    Kr = R.stiffness(ref, Kn, Tv)
    assert abs(K - Kr).max() < 1e-12 * abs(Kr).max()
    fixed = (ref.ijk[:, 2] == 0) | (ref.ijk[:, 2] == dims[2] - 1)
    up = np.zeros((ref.N, 3))
    up[ref.ijk[:, 2] == dims[2] - 1, 2] = 1e-2
    ug, info = lin3d.solve_pcg(lat, Kn, Tv, fixed, up, tol=1e-12)
    ud = R.solve_direct(Kr, np.repeat(fixed, 3), up.ravel()).reshape(-1, 3)
    assert info["relres"] <= 1e-12 and rel(ug, ud) < 1e-9
