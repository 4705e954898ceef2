"""The drop-in surface against vectors written by the reference's own torch modules (tests/golden/*.npz from
tests/golden/make_surface_golden.py: coarse3d/torch_lpm3d.py, coarse2d/torch_lpm.py, cement3d/torch_v1.py on the CPU
in FP64).

CPU tests pin the oracle (oracle/lpm_ref.py) with them: operator, computeStress with its surface (opp-flag) rule at
box faces and at removed particles, particle energy, block diagonal, the fine-reference block means and the coarse
runs (SuperLU / dense solves), and the reference's autograd gradients (checked by central differences of the
oracle's loss: per directed slot in 3D, per bond pair in 2D, where the reference's dense energy-form assembly makes
only the pair sums representation-independent).  GPU tests run the HIP path through the same surface against the
same vectors: FP64 kernels to 1e-12, solves at tol 1e-12 / 1e-13 to 1e-8, gradients to 1e-7 of the largest."""
import os

import numpy as np
import pytest
import torch

import lpm_ref as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")
GDIR = {"tb_tz": 2, "tb_sx": 0, "lr_tx": 0}


def _f(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------------------------------------------ oracle helpers
def _solve3(lat, K, fixed, up):
    """Dense or sparse solve of K_ff u_f = -K_fp u_p (u = up on fixed particles)."""
    fd = np.repeat(fixed, 3)
    if isinstance(K, np.ndarray):
        u = up.ravel().copy()
        f = ~fd
        u[f] = np.linalg.solve(K[np.ix_(f, f)], -K[np.ix_(f, fd)] @ up.ravel()[fd])
        return u.reshape(-1, 3)
    return R.solve_direct(K, fd, up).reshape(-1, 3)


def _run3(lat, Kn, Tv, IJK, nc, s, h, dense=False):
    """Oracle coarse / fine run of the 9 load cases: per-particle u, stress, energy and the reactions."""
    K = R.dense_from_apply(lat, Kn, Tv) if dense else R.stiffness(lat, Kn, Tv)
    out = {k: [] for k in ("u", "sig", "U", "F")}
    for ld in R.LOADS3:
        fixed, up, mov = R.prescribed3(ld, lat.x, IJK, nc, s, h)
        u = _solve3(lat, K, fixed, up)
        out["u"].append(u)
        out["sig"].append(R.stress(lat, u, Kn, Tv))
        out["U"].append(R.energy(lat, u, Kn, Tv))
        out["F"].append(0.0 if mov is None else R.apply(lat, u, Kn, Tv)[mov, GDIR[ld]].sum())
    return {k: np.array(v) for k, v in out.items()}


def _block_means(n, s, f, count=None):
    """(L, N, k) per particle of an n^3 box -> (L, nc^3, k) block sums / s^3 (or / count)."""
    nc = n // s
    ijk = np.indices((n, n, n)).reshape(3, -1)[::-1].T // s
    cid = (ijk[:, 2] * nc + ijk[:, 1]) * nc + ijk[:, 0]
    out = np.zeros(f.shape[:-2] + (nc ** 3, f.shape[-1]))
    np.add.at(out, (slice(None), cid), f)
    return out / (s ** 3 if count is None else np.maximum(count, 1.0))


def _coarse3_loss(lat, Kn, Tv, g):
    r = _run3(lat, Kn, Tv, lat.ijk, (g["nc"],) * 3, int(g["s"]), 1.0, dense=True)
    V = (float(g["s"]) * 1.0) ** 3
    return ((r["sig"] * g["Wsig"][0]).sum() + 1e3 * (r["u"] * g["Wu"][0]).sum() + (r["U"] / V * g["Ww"][0]).sum()
            + (r["F"] * g["WF"][0]).sum())


# ------------------------------------------------------------------------------------------------ CPU: the oracle
def test_oracle_operator_matches_reference_surface3d():
    """K u, computeStress (opp-flag weights at every face, edge and corner of the box), particle energy and the
    block diagonal of the reference's SCT on a two-phase n = 10 box."""
    g = _f("surface3d_n10")
    n = int(g["n"])
    lat = R.Box((n, n, n))
    Kn, Tv, u = g["Kn"], g["Tv"], g["u"]
    assert rel(R.stress_weights(lat), g["stw"]) < 1e-15
    assert rel(R.apply(lat, u, Kn, Tv), g["Ku"]) < 1e-13
    assert rel(R.stress(lat, u, Kn, Tv), g["stress"]) < 1e-13
    assert rel(R.energy(lat, u, Kn, Tv), g["energy"]) < 1e-13
    assert rel(R.block_diag(lat, Kn, Tv), g["block_diag"]) < 1e-13


def test_oracle_fine_reference_block_means_match_reference():
    """FineRefT.run at s = 2 (reference PCG to 1e-12) against the oracle's SuperLU solves: block means of u,
    stress and energy density, grip reactions of the plate cases."""
    g = _f("surface3d_n10")
    n, s = int(g["n"]), 2
    lat = R.Box((n, n, n))
    r = _run3(lat, g["Kn"], g["Tv"], lat.ijk // s, (n // s,) * 3, s, 1.0)
    assert rel(_block_means(n, s, r["u"]), g["fine_u"]) < 1e-9
    assert rel(_block_means(n, s, r["sig"]), g["fine_sig"]) < 1e-9
    assert rel(_block_means(n, s, r["U"][..., None])[..., 0], g["fine_w"]) < 1e-9
    assert rel(r["F"], g["fine_F"]) < 1e-9


def test_oracle_coarse_run_and_autograd_gradient_match_reference():
    """CoarseT.run on 5^3 (s = 2, tol 1e-13) against dense oracle solves, and the reference's autograd gradient of
    a linear functional of (u, sig, w, F) against central differences of the oracle's loss, per directed slot."""
    g = _f("coarse3d_nc5")
    nc, s = int(g["nc"]), int(g["s"])
    lat = R.Box((nc,) * 3, float(s))
    Kn, Tv = g["Kn"], g["Tv"]
    r = _run3(lat, Kn, Tv, lat.ijk, (nc,) * 3, s, 1.0, dense=True)
    assert rel(r["u"], g["u"]) < 1e-10 and rel(r["sig"], g["sig"]) < 1e-10
    assert rel(r["U"] / s ** 3, g["w"]) < 1e-10 and rel(r["F"], g["F"]) < 1e-10
    assert abs(_coarse3_loss(lat, Kn, Tv, g) - float(g["loss"])) <= 1e-10 * abs(float(g["loss"]))
    rng = np.random.default_rng(2)
    gmax = max(np.abs(g["gKn"]).max(), np.abs(g["gTv"]).max())
    for _ in range(4):
        i = int(rng.integers(lat.N))
        k = int(rng.choice(np.nonzero(lat.valid[i])[0]))
        for which in ("gKn", "gTv"):
            hs = 1e-5
            vals = []
            for sg in (1, -1):
                A = [Kn.copy(), Tv.copy()]
                A[0 if which == "gKn" else 1][i, k] += sg * hs
                vals.append(_coarse3_loss(lat, A[0], A[1], g))
            fd = (vals[0] - vals[1]) / (2 * hs)
            assert abs(fd - g[which][i, k]) <= 1e-6 * gmax, (which, i, k, fd, g[which][i, k])


def _run2(lat, Kn, Tv, ncx, ncy, s, loads):
    K = R.stiffness(lat, Kn, Tv)
    out = {}
    for ld in loads:
        fixed, up, mov = R.prescribed2(ld, lat.x, lat.ijk, ncx, ncy, s, 1.0)
        u = R.solve_direct(K, np.repeat(fixed, 2), up).reshape(-1, 2)
        out[ld] = dict(u=u, sig=R.stress(lat, u, Kn, Tv), w=R.energy(lat, u, Kn, Tv) / (s * 1.0) ** 2,
                       F=None if mov is None else R.apply(lat, u, Kn, Tv)[mov].sum(0))
    return out


def _loss2(lat, Kn, Tv, g, b):
    loads = [str(v) for v in g["loads"]]
    o = _run2(lat, Kn, Tv, int(g["ncx"]), int(g["ncy"]), int(g["s"]), loads)
    L = 0.0
    for ld in loads:
        for k, v in o[ld].items():
            if v is not None:
                L += (v * g["%s/W%s" % (ld, k)][b]).sum() * (1e3 if k == "u" else 1.0)
    return L


def test_oracle_coarse2d_matches_reference_dense_solve():
    """The 2D CoarseT.run (dense assembly + Cholesky) on 8 x 6 at s = 2 for two specimens and all loads, and the
    bond-pair sums of its gradient against central differences of the oracle's loss (symmetric perturbation)."""
    g = _f("coarse2d_8x6")
    ncx, ncy, s = int(g["ncx"]), int(g["ncy"]), int(g["s"])
    lat = R.Box((ncx, ncy), float(s))
    loads = [str(v) for v in g["loads"]]
    for b in range(g["Kn"].shape[0]):
        o = _run2(lat, g["Kn"][b], g["Tv"][b], ncx, ncy, s, loads)
        for ld in loads:
            assert rel(o[ld]["u"], g["%s/u" % ld][b]) < 1e-10
            assert rel(o[ld]["sig"], g["%s/sig" % ld][b]) < 1e-10
            assert rel(o[ld]["w"], g["%s/w" % ld][b]) < 1e-10
            if o[ld]["F"] is not None:
                assert rel(o[ld]["F"], g["%s/F" % ld][b]) < 1e-10
    total = _loss2(lat, g["Kn"][0], g["Tv"][0], g, 0) + _loss2(lat, g["Kn"][1], g["Tv"][1], g, 1)
    assert abs(total - float(g["loss"])) <= 1e-10 * abs(float(g["loss"]))
    rng = np.random.default_rng(4)
    gmax = max(np.abs(g["gKn"]).max(), np.abs(g["gTv"]).max())
    for _ in range(4):
        i = int(rng.integers(lat.N))
        k = int(rng.choice([kk for kk in range(0, 8, 2) if lat.valid[i, kk]]))
        q = lat.nbr[i, k]
        for which in ("gKn", "gTv"):
            hs = 1e-5
            vals = []
            for sg in (1, -1):
                A = [g["Kn"][0].copy(), g["Tv"][0].copy()]
                M = A[0 if which == "gKn" else 1]
                M[i, k] += sg * hs
                M[q, k ^ 1] += sg * hs
                vals.append(_loss2(lat, A[0], A[1], g, 0))
            fd = (vals[0] - vals[1]) / (2 * hs)
            pair = g[which][0, i, k] + g[which][0, q, k ^ 1]
            assert abs(fd - pair) <= 1e-6 * gmax, (which, i, k, fd, pair)


def _void_fine_oracle(g):
    """FineRefTV at s = 2 restated on the reduced lattice of the solid particles (SuperLU), embedded back into the
    box: (u block means over solid particles, stress and energy block sums / s^3, reactions)."""
    n, s = int(g["n"]), 2
    solid = g["solid"]
    full = R.Box((n, n, n))
    lat = R.ActiveBox((n, n, n), solid)
    Kn, Tv = g["Kn"][lat.sites], g["Tv"][lat.sites]
    Kn = np.where(lat.valid, Kn, 0.0)
    Tv = np.where(lat.valid, Tv, 0.0)
    K = R.stiffness(lat, Kn, Tv)
    U, SG, EN, F = [], [], [], []
    for ld in R.LOADS3:
        fixed, up, mov = R.prescribed3(ld, full.x, full.ijk // s, (n // s,) * 3, s, 1.0)
        u = _solve3(lat, K, fixed[lat.sites], up[lat.sites])
        ue = np.zeros((full.N, 3))
        ue[lat.sites] = u
        se = np.zeros((full.N, 6))
        se[lat.sites] = R.stress(lat, u, Kn, Tv)
        ee = np.zeros(full.N)
        ee[lat.sites] = R.energy(lat, u, Kn, Tv)
        U.append(ue)
        SG.append(se)
        EN.append(ee)
        F.append(0.0 if mov is None else R.apply(lat, u, Kn, Tv)[mov[lat.sites], GDIR[ld]].sum())
    cnt = _block_means(n, s, np.broadcast_to(solid.astype(float)[None, :, None], (1, full.N, 1)), count=np.ones(1))
    return (_block_means(n, s, np.array(U), count=cnt), _block_means(n, s, np.array(SG)),
            _block_means(n, s, np.array(EN)[..., None])[..., 0], np.array(F))


def test_oracle_removed_particles_match_reference():
    """SCTV with removed particles: the opp-flag stress rule at removed neighbours, energy, block diagonal (zero at
    removed particles) and K u of the reduced lattice; FineRefTV block means; CoarseTV with removed cells."""
    g = _f("voids3d_n8")
    n = int(g["n"])
    solid = g["solid"]
    lat = R.ActiveBox((n, n, n), solid)
    Kn = np.where(lat.valid, g["Kn"][lat.sites], 0.0)
    Tv = np.where(lat.valid, g["Tv"][lat.sites], 0.0)
    u = g["u"][lat.sites]
    for name, fn in (("Ku", R.apply), ("stress", R.stress), ("energy", R.energy)):
        ref = g[name]
        assert rel(fn(lat, u, Kn, Tv), ref[lat.sites]) < 1e-13, name
        assert np.abs(ref[~solid]).max() == 0.0
    assert rel(R.block_diag(lat, Kn, Tv), g["block_diag"][lat.sites]) < 1e-13
    um, sm, wm, F = _void_fine_oracle(g)
    assert rel(um, g["fine_u"]) < 1e-9 and rel(sm, g["fine_sig"]) < 1e-9
    assert rel(wm, g["fine_w"]) < 1e-9 and rel(F, g["fine_F"]) < 1e-9
    nc, s = int(g["nc"]), int(g["s"])
    cell = g["cell"]
    clat = R.ActiveBox((nc,) * 3, cell, float(s))
    Kc = np.where(clat.valid, g["Kc"][clat.sites], 0.0)
    Tc = np.where(clat.valid, g["Tc"][clat.sites], 0.0)
    fullc = R.Box((nc,) * 3, float(s))
    K = R.stiffness(clat, Kc, Tc)
    for li, ld in enumerate(R.LOADS3):
        fixed, up, mov = R.prescribed3(ld, fullc.x, fullc.ijk, (nc,) * 3, s, 1.0)
        u = _solve3(clat, K, fixed[clat.sites], up[clat.sites])
        assert rel(u, g["coarse_u"][li][clat.sites]) < 1e-9
        assert np.abs(g["coarse_u"][li][~cell]).max() == 0.0
        assert rel(R.stress(clat, u, Kc, Tc), g["coarse_sig"][li][clat.sites]) < 1e-9
        assert rel(R.energy(clat, u, Kc, Tc) / s ** 3, g["coarse_w"][li][clat.sites]) < 1e-9
        if mov is not None:
            Fo = R.apply(clat, u, Kc, Tc)[mov[clat.sites], GDIR[ld]].sum()
            assert abs(Fo - g["coarse_F"][li]) <= 1e-9 * abs(g["coarse_F"][li])


# ------------------------------------------------------------------------------------------------ GPU: the HIP path
@pytest.mark.gpu
def test_gpu_sct_surface_matches_reference(dev):
    import torch_lpm3d as T3
    g = _f("surface3d_n10")
    n = int(g["n"])
    op = T3.SCT(T3.LatticeGeom(n, n, n), device=dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    Kn, Tv, u = t(g["Kn"]), t(g["Tv"]), t(g["u"])
    assert rel(op.apply(u, Kn, Tv), g["Ku"]) < 1e-12
    assert rel(op.stress(u, Kn, Tv), g["stress"]) < 1e-12
    assert rel(op.energy(u, Kn, Tv), g["energy"]) < 1e-12
    assert rel(op.block_diag(Kn, Tv), g["block_diag"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_gpu_fineref_matches_reference(dev, mixed):
    import torch_lpm3d as T3
    g = _f("surface3d_n10")
    n = int(g["n"])
    fr = T3.FineRefT(n, 2, device=dev, mixed=mixed)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    r = fr.run(t(g["Kn"])[None], t(g["Tv"])[None], tol=1e-12)
    for k in ("u", "sig", "w", "F"):
        assert rel(r[k][0], g["fine_" + k]) < 1e-8, k


@pytest.mark.gpu
def test_gpu_coarse3d_run_and_gradient_match_reference_autograd(dev):
    """CoarseT.run outputs and the gradient of the reference's loss per directed slot: LinSolve's adjoint PCG and
    bond-gradient kernel plus the post-processing VJP kernels against the reference's autograd."""
    import torch_lpm3d as T3
    g = _f("coarse3d_nc5")
    nc, s = int(g["nc"]), int(g["s"])
    cm = T3.CoarseT(nc, s, device=dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    kn = t(g["Kn"])[None].requires_grad_(True)
    tv = t(g["Tv"])[None].requires_grad_(True)
    o = cm.run(kn, tv, tol=1e-13)
    for k in ("u", "sig", "w", "F"):
        assert rel(o[k][0], g[k]) < 1e-9, k
    loss = ((o["sig"] * t(g["Wsig"])).sum() + 1e3 * (o["u"] * t(g["Wu"])).sum() + (o["w"] * t(g["Ww"])).sum()
            + (o["F"] * t(g["WF"])).sum())
    assert abs(float(loss.detach()) - float(g["loss"])) <= 1e-9 * abs(float(g["loss"]))
    loss.backward()
    gmax = max(np.abs(g["gKn"]).max(), np.abs(g["gTv"]).max())
    valid = R.Box((nc,) * 3).valid
    for mine, ref in ((kn.grad[0], g["gKn"]), (tv.grad[0], g["gTv"])):
        m = mine.cpu().numpy() * valid
        assert np.abs(m - ref * valid).max() <= 1e-7 * gmax


@pytest.mark.gpu
def test_gpu_coarse2d_run_and_gradient_pairs_match_reference(dev):
    import torch_lpm as T2
    g = _f("coarse2d_8x6")
    ncx, ncy, s = int(g["ncx"]), int(g["ncy"]), int(g["s"])
    m = T2.CoarseT(ncx, ncy, s, device=dev, tol=1e-13)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    kn = t(g["Kn"]).requires_grad_(True)
    tv = t(g["Tv"]).requires_grad_(True)
    o = m.run(kn, tv)
    loss = 0.0
    for ld in [str(v) for v in g["loads"]]:
        for k in ("u", "sig", "w", "F"):
            v = o[ld][k]
            if v is None:
                assert ("%s/%s" % (ld, k)) not in g.files
                continue
            assert rel(v, g["%s/%s" % (ld, k)]) < 1e-9, (ld, k)
            loss = loss + (v * t(g["%s/W%s" % (ld, k)])).sum() * (1e3 if k == "u" else 1.0)
    loss.backward()
    lat = R.Box((ncx, ncy), float(s))
    gmax = max(np.abs(g["gKn"]).max(), np.abs(g["gTv"]).max())
    for mine, ref in ((kn.grad, g["gKn"]), (tv.grad, g["gTv"])):
        mm = mine.cpu().numpy()
        for k in range(0, 8, 2):
            ok = lat.valid[:, k]
            q = lat.nbr[ok, k]
            pm = mm[:, ok, k] + mm[:, q, k ^ 1]
            pr = ref[:, ok, k] + ref[:, q, k ^ 1]
            assert np.abs(pm - pr).max() <= 1e-7 * gmax


@pytest.mark.gpu
def test_gpu_removed_particles_match_reference(dev):
    import torch_lpm3d as T3
    g = _f("voids3d_n8")
    n = int(g["n"])
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    solid = torch.as_tensor(g["solid"][None], device=dev)
    op = T3.SCTV(T3.LatticeGeom(n, n, n), device=dev)
    op.set_mask(solid)
    Kn, Tv, u = t(g["Kn"])[None], t(g["Tv"])[None], t(g["u"])[None]
    assert rel(op.apply(u, Kn, Tv)[0], g["Ku"]) < 1e-12
    assert rel(op.stress(u, Kn, Tv)[0], g["stress"]) < 1e-12
    assert rel(op.energy(u, Kn, Tv)[0], g["energy"]) < 1e-12
    assert rel(op.block_diag(Kn, Tv)[0], g["block_diag"]) < 1e-12
    fr = T3.FineRefTV(n, 2, device=dev)
    r = fr.run(Kn, Tv, solid, tol=1e-12)
    for k in ("u", "sig", "w", "F"):
        assert rel(r[k][0], g["fine_" + k]) < 1e-8, k
    cm = T3.CoarseTV(int(g["nc"]), int(g["s"]), device=dev)
    r = cm.run(t(g["Kc"])[None], t(g["Tc"])[None], torch.as_tensor(g["cell"][None], device=dev), tol=1e-12,
               differentiable=False)
    for k in ("u", "sig", "w", "F"):
        assert rel(r[k][0], g["coarse_" + k]) < 1e-8, k
