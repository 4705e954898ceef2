"""Bond-network inference (coarse3d/net3d.py): the oracle (oracle/bondnet_ref.py) pinned by the reference's own outputs
(tests/golden/bondnet_*.npz, written by tests/golden/make_bondnet_golden.py), the class-sum convolution against
torch's conv3d with the expanded kernel, and the cube-symmetry equivariance that the reference guarantees; the
product's window tables against the reference's BondWindows; the HIP path (net3d.py of the package) against the oracle
and the fixtures.  GPU results are FP32 against FP64 references: tolerances are relative to the largest output."""
import ctypes
import os

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import bondnet_ref as BR
import net3d

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["bondnet_s2_small", "bondnet_s3_small", "bondnet_s2_full"]
TOL32 = 2e-5          # FP32 kernels against the FP64 oracle, relative to the largest magnitude


def _fixture(name):
    f = np.load(os.path.join(GOLD, name + ".npz"))
    P = {k[2:]: f[k].astype(np.float64) for k in f.files if k.startswith("w/")}
    return f, P


def _net_from(P, s, ch, D, hid, device):
    net = net3d.BondNet(s, ch=ch, D=D, hid=hid)
    net.load_state_dict({k: torch.as_tensor(v, dtype=torch.float32) for k, v in P.items()}, strict=False)
    return net.to(device)


def _random_net(s, ch, D, hid, seed, head_scale=0.05):
    torch.manual_seed(seed)
    net = net3d.BondNet(s, ch=ch, D=D, hid=hid)
    with torch.no_grad():
        for l1, l2, l3 in net.heads:
            l3.weight.normal_(0.0, head_scale)
            l3.bias.normal_(0.0, head_scale)
        for m in (net.trunk[0], net.trunk[2], net.trunk[4]):
            m.b.normal_(0.0, 0.1)
    P = {k: v.detach().double().numpy() for k, v in net.state_dict().items() if not k.endswith("masks")}
    return net, P


def _micro(n, s, seed, contrast=4.0):
    rng = np.random.default_rng(seed)
    phase = (rng.random(n ** 3) < 0.4).astype(np.int64)
    return phase, BR.voxel_features(phase, contrast, 0.3, n, s)


# ------------------------------------------------------------------------------------------------- symmetry helpers
def _rotate(Fld, g):
    """Field over the last three axes (z, y, x) of a cube, rotated about its centre: out[g v] = Fld[v]."""
    n = Fld.shape[-1]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v2 = 2 * np.stack([x, y, z], -1) - (n - 1)                   # doubled coordinates of the output voxel
    src = ((v2 @ g) + (n - 1)) // 2                               # g^T v' (g orthogonal)
    return Fld[..., src[..., 2], src[..., 1], src[..., 0]]


def _rotate_bonds(A, g, nc):
    """Directed per-slot field (nc^3, 18): out[g p, k'] = A[p, k] with offset(k') = g offset(k)."""
    off = BR.SC_OFF
    kmap = np.array([next(j for j in range(18) if np.array_equal(off[j], g @ off[k])) for k in range(18)])
    ijk = np.stack(np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    d2 = (2 * ijk - (nc - 1)) @ g.T + (nc - 1)
    q = ((d2[:, 2] // 2) * nc + d2[:, 1] // 2) * nc + d2[:, 0] // 2
    out = np.zeros_like(A)
    out[q[:, None], kmap[None, :]] = A
    return out


# -------------------------------------------------------------------------------------------------------- CPU
@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_fixture(name):
    f, P = _fixture(name)
    outs, groups = BR.forward(f["X"], int(f["nc"]), int(f["s"]), P)
    Fv = BR.trunk(f["X"][0].astype(float), P)
    assert np.abs(Fv - f["Fv"][0]).max() <= 1e-10 * np.abs(f["Fv"][0]).max()
    for gi in range(2):
        for key in ("iv", "ic", "bonds", "perms"):
            assert np.array_equal(groups[gi][key], f["g%d/%s" % (gi, key)])
        ref = f["g%d/out" % gi]
        assert np.abs(outs[gi] - ref).max() <= 1e-10 * np.abs(ref).max()


def test_oracle_iso_conv_is_conv3d_with_the_class_kernel():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 7, 6, 9))
    w = rng.standard_normal((4, 5, 4))
    b = rng.standard_normal(4)
    off = np.array(np.meshgrid(*(np.arange(-1, 2),) * 3, indexing="ij")).reshape(3, -1).T
    cls = np.abs(off).sum(1)
    K = w[:, :, cls].reshape(4, 5, 3, 3, 3)
    ref = F.conv3d(torch.as_tensor(x)[None], torch.as_tensor(K), torch.as_tensor(b), padding=1)[0].numpy()
    assert np.abs(BR.iso_conv(x, w, b) - ref).max() < 1e-12


@pytest.mark.parametrize("gidx", [9, 20, 47])
def test_oracle_is_cube_equivariant(gidx):
    net, P = _random_net(2, 6, 4, 16, seed=5, head_scale=0.3)
    nc, s = 3, 2
    _, X = _micro(nc * s, s, 7)
    g = BR.OH[gidx]
    cp = np.ones((nc ** 3, 10))
    outs, groups = BR.forward(X[None], nc, s, P)
    Kn, Tv = BR.post([o[0] for o in outs], groups, cp, nc)
    outs_r, _ = BR.forward(_rotate(X, g)[None], nc, s, P)
    Kn_r, Tv_r = BR.post([o[0] for o in outs_r], groups, cp, nc)
    assert np.abs(Kn_r - _rotate_bonds(Kn, g, nc)).max() < 1e-10 * np.abs(Kn).max()
    assert np.abs(Tv_r - _rotate_bonds(Tv, g, nc)).max() < 1e-10 * np.abs(Tv).max()
    # the rotation is not the identity on this field: the check has something to detect
    assert np.abs(Kn_r - Kn).max() > 1e-6


@pytest.mark.parametrize("name", CASES)
def test_product_windows_match_reference(name):
    f, _ = _fixture(name)
    bw = net3d.BondWindows(int(f["nc"]), int(f["s"]))
    for gi, g in enumerate(bw.groups):
        for key in ("iv", "ic", "bonds", "perms"):
            assert np.array_equal(g[key], f["g%d/%s" % (gi, key)])


def test_checkpoint_names_load():
    f, P = _fixture("bondnet_s2_full")
    net = net3d.BondNet(2)
    res = net.load_state_dict({k: torch.as_tensor(v, dtype=torch.float32) for k, v in P.items()}, strict=False)
    assert not res.unexpected_keys and all(k.endswith("masks") for k in res.missing_keys)


def test_voigt_cell_params_match_general_calcKnTv():
    import geom3d as G
    rng = np.random.default_rng(2)
    n, s = 6, 2
    Cv = G.iso_C(rng.uniform(1, 5, n ** 3), rng.uniform(0.1, 0.35, n ** 3))
    ref = G.cell_params_from_C66(G.voigt_cells(Cv, n, s), s / 2.0)
    assert np.abs(net3d.voigt_cell_params(Cv, n, s).numpy() - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(BR.voigt_cell_params(Cv, n, s) - ref).max() < 1e-12 * np.abs(ref).max()


def test_zero_last_layer_gives_voigt():
    net, P = _random_net(2, 6, 4, 16, seed=1, head_scale=0.0)
    nc, s = 3, 2
    _, X = _micro(nc * s, s, 3)
    outs, groups = BR.forward(X[None], nc, s, P)
    cp = np.random.default_rng(0).uniform(0.5, 1.5, (nc ** 3, 10))
    Kn, Tv = BR.post([o[0] for o in outs], groups, cp, nc)
    import geom3d as G
    KnV, TvV = G.bond_params_from_cells(G.LatticeGeom(nc, nc, nc, s), cp)
    assert np.abs(Kn - KnV).max() == 0.0 and np.abs(Tv - TvV).max() < 1e-15


# -------------------------------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("cin,cout,dims,gelu,proj", [(3, 24, (9, 70, 5), True, 0), (24, 24, (13, 5, 66), True, 0),
                                                     (8, 6, (4, 4, 4), False, 0), (24, 24, (7, 9, 65), True, 12),
                                                     (6, 30, (5, 6, 7), True, 5)])
def test_gpu_iso_conv(dev, cin, cout, dims, gelu, proj):
    rng = np.random.default_rng(cin * 100 + cout)
    nz, ny, nx = dims
    x = rng.standard_normal((2, cin, nz, ny, nx)).astype(np.float32)
    w = (rng.standard_normal((cout, cin, 4)) * 0.2).astype(np.float32)
    b = rng.standard_normal(cout).astype(np.float32)
    refs = []
    for xb in x.astype(float):
        r = BR.iso_conv(xb, w.astype(float), b.astype(float))
        refs.append(BR.gelu(r) if gelu else r)
    ref = np.stack(refs)
    wt, bt = torch.as_tensor(w, device=dev), torch.as_tensor(b, device=dev)
    if proj:
        pw = (rng.standard_normal((proj, cout)) * 0.3).astype(np.float32)
        pb = rng.standard_normal(proj).astype(np.float32)
        ref = np.einsum("dc,bczyx->bzyxd", pw, ref) + pb
        y = net3d.iso_conv(torch.as_tensor(x, device=dev), wt, bt, gelu=gelu,
                           proj=(torch.as_tensor(pw, device=dev), torch.as_tensor(pb, device=dev)), dstride=16)
        y = y.cpu().numpy().reshape(2, nz, ny, nx, 16)
        assert np.all(y[..., proj:] == 0)
        y = y[..., :proj]
    else:
        y = net3d.iso_conv(torch.as_tensor(x, device=dev), wt, bt, gelu=gelu).cpu().numpy()
    assert np.abs(y - ref).max() <= TOL32 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_forward_matches_reference_fixture(dev, name):
    f, P = _fixture(name)
    net = _net_from(P, int(f["s"]), int(f["ch"]), int(f["D"]), int(f["hid"]), dev)
    outs = net(torch.as_tensor(f["X"], device=dev), int(f["nc"]))
    for gi in range(2):
        ref = f["g%d/out" % gi]
        got = outs[gi].cpu().numpy()
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max(), (gi, np.abs(got - ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("nc,s,B", [(7, 2, 2), (5, 3, 1)])
def test_gpu_bond_params_match_oracle(dev, nc, s, B):
    """Kn / Tv of the fused heads + post against oracle forward + post; bond counts that are not multiples of the
    16-bond block and blocks that straddle two directions."""
    net, P = _random_net(s, 8, 4, 32, seed=nc)
    n = nc * s
    Xs, cps = [], []
    import geom3d as G
    for b in range(B):
        phase, X = _micro(n, s, 10 + b, contrast=3.0 + b)
        Xs.append(X)
        Cv = G.iso_C(np.where(phase == 1, 3.0 + b, 1.0), np.where(phase == 1, 0.3, 0.2))
        cps.append(net3d.voigt_cell_params(Cv, n, s).numpy())
    X = np.stack(Xs)
    Kn, Tv = net.to(dev).bond_params(torch.as_tensor(X, device=dev), nc, torch.as_tensor(np.stack(cps), device=dev))
    outs, groups = BR.forward(X, nc, s, P)
    for b in range(B):
        KnR, TvR = BR.post([o[b] for o in outs], groups, cps[b], nc)
        assert np.abs(Kn[b].cpu().numpy() - KnR).max() <= TOL32 * 10 * np.abs(KnR).max()
        assert np.abs(Tv[b].cpu().numpy() - TvR).max() <= TOL32 * 10 * np.abs(KnR).max()
        # slots leaving the box stay zero, every in-box slot is set (Kn > 0)
        lat = G.LatticeGeom(nc, nc, nc, s)
        assert np.all((Kn[b].cpu().numpy() > 0) == lat.valid)


@pytest.mark.gpu
def test_gpu_forward_is_cube_equivariant(dev):
    net, _ = _random_net(2, 24, 12, 64, seed=3, head_scale=0.3)
    net = net.to(dev)
    nc, s = 6, 2
    _, X = _micro(nc * s, s, 9)
    cp = torch.ones((nc ** 3, 10), device=dev, dtype=torch.float64)
    Kn, _ = net.bond_params(torch.as_tensor(X[None], device=dev), nc, cp)
    g = BR.OH[21]
    Kn_r, _ = net.bond_params(torch.as_tensor(_rotate(X, g)[None].copy(), device=dev), nc, cp)
    A = Kn[0].cpu().numpy()
    assert np.abs(Kn_r[0].cpu().numpy() - _rotate_bonds(A, g, nc)).max() <= 1e-4 * np.abs(A).max()


@pytest.mark.gpu
def test_gpu_inferred_parameters_feed_the_solve(dev):
    """The network's Kn / Tv of a 6^3 coarse lattice (h = s) go straight into the GPU solve (train3d.py:100-108 feeds
    CoarseT the same way), which matches SuperLU on the oracle's K built from the same parameters."""
    import geom3d
    import lpm_ref as R
    import torch_lpm3d as T3
    nc, s = 6, 2
    net, _ = _random_net(s, 8, 4, 32, seed=4)
    phase, X = _micro(nc * s, s, 21)
    Cv = geom3d.iso_C(np.where(phase == 1, 4.0, 1.0), np.where(phase == 1, 0.3, 0.2))
    cp = net3d.voigt_cell_params(Cv, nc * s, s)
    Kn, Tv = net.to(dev).bond_params(torch.as_tensor(X[None], device=dev), nc, cp.to(dev))
    Kn, Tv = Kn[0], Tv[0]
    dims, h = (nc, nc, nc), float(s)
    lat = R.Box(dims, h)
    fixed, up, _ = R.prescribed3("k_xx", lat.x, lat.ijk, dims, 1, h)
    ud = R.solve_direct(R.stiffness(lat, Kn.cpu().numpy(), Tv.cpu().numpy()), np.repeat(fixed, 3), up)
    op = T3.SCT(geom3d.LatticeGeom(*dims, h), device=dev)
    t = lambda a: torch.as_tensor(a, device=dev, dtype=torch.float64)
    u, it, relres = op.solve(Kn, Tv, t(up), t((~fixed).astype(float)), tol=1e-12)
    ud = ud.reshape(lat.N, 3)
    assert relres <= 1e-12
    assert np.abs(u.cpu().numpy() - ud).max() <= 1e-9 * np.abs(ud).max()


@pytest.mark.gpu
def test_gpu_post_matches_fused_bond_params(dev):
    """infer_fast.Post's two-step interface (bases, scatter of the raw outputs) equals the fused heads + post."""
    import geom3d
    nc, s = 5, 2
    net, _ = _random_net(s, 8, 4, 32, seed=8)
    net = net.to(dev)
    phase, X = _micro(nc * s, s, 31)
    Cv = geom3d.iso_C(np.where(phase == 1, 4.0, 1.0), np.where(phase == 1, 0.3, 0.2))
    Xd = torch.as_tensor(X[None], device=dev)
    post = net3d.Post(nc * s, s, net, device=dev)
    Kn, Tv = post.scatter(net(Xd, nc), post.bases(Cv))
    Kf, Tf = net.bond_params(Xd, nc, net3d.voigt_cell_params(Cv, nc * s, s).to(dev))
    assert torch.allclose(Kn, Kf[0], rtol=1e-12, atol=0) and torch.allclose(Tv, Tf[0], rtol=1e-12, atol=1e-15)


def test_post_glue_matches_oracle_on_host_tensors():
    """net3d.Post (torch glue around the kernels) scatters given raw outputs exactly as the oracle's post."""
    import geom3d
    nc, s = 4, 2
    net = net3d.BondNet(s, ch=6, D=4, hid=16)
    rng = np.random.default_rng(12)
    Cv = geom3d.iso_C(rng.uniform(1, 4, (nc * s) ** 3), rng.uniform(0.15, 0.3, (nc * s) ** 3))
    groups = BR.bond_windows(nc, s)
    outs = [rng.normal(0, 0.1, (1, len(g["bonds"]), 2)) for g in groups]
    post = net3d.Post(nc * s, s, net, device="cpu")
    Kn, Tv = post.scatter([torch.as_tensor(o) for o in outs], post.bases(Cv))
    KnR, TvR = BR.post([o[0] for o in outs], groups, BR.voigt_cell_params(Cv, nc * s, s), nc)
    assert np.abs(Kn.numpy() - KnR).max() <= 1e-14 * np.abs(KnR).max()
    assert np.abs(Tv.numpy() - TvR).max() <= 1e-14 * np.abs(KnR).max()


def test_bond_abi_rejects_bad_arguments_without_a_gpu():
    """The bond-network entry points validate shapes and tables before any HIP call."""
    import ctypes
    import _lpmlib as L
    lib = net3d._lib()
    assert lib.lpm_iso_conv(1, 3, 0, 8, 8, 8, None, None, None, 1, 0, 0, None, None, None, None) == L.LPM_E_BADARG
    assert lib.lpm_iso_conv(1, 40, 8, 8, 8, 8, None, None, None, 1, 0, 0, None, None, None, None) == L.LPM_E_BADARG
    assert lib.lpm_bond_heads(None, 1, 4, 2, *([None] * 13)) == L.LPM_E_BADARG
    grp = net3d.BondGroup()
    grp.ndir, grp.nv, grp.ncell, grp.H, grp.D, grp.dstride, grp.hid = 3, 16, 34, 8, 12, 12, 64
    for d, k in enumerate((0, 2, 4)):
        grp.slot[d] = k
    assert lib.lpm_bond_w1_layout(ctypes.byref(grp), None, None, None) == L.LPM_E_BADARG   # axial with H = 8
    assert b"stabiliser" in lib.lpm_last_error() or b"directions" in lib.lpm_last_error()
    grp.H = 16
    grp.voff = grp.coff = grp.inv = 16                                   # non-null dummies: never dereferenced
    grp.slot[1] = 7                                                      # odd slot
    assert lib.lpm_bond_w1_layout(ctypes.byref(grp), None, None, None) == L.LPM_E_BADARG
    assert b"forward slot" in lib.lpm_last_error()
    grp.slot[1] = 2
    grp.dstride = 10                                                     # not a multiple of 4
    assert lib.lpm_bond_w1_layout(ctypes.byref(grp), None, None, None) == L.LPM_E_BADARG


def test_iso_conv_rejects_mismatched_weights():
    """Shape checks run before any device call: wrong weight shapes raise instead of reading out of bounds."""
    x = torch.zeros((1, 3, 4, 4, 4))
    with pytest.raises(RuntimeError):                      # CPU tensors are refused first
        net3d.iso_conv(x, torch.zeros((8, 3, 4)), torch.zeros(8))


@pytest.mark.gpu
def test_gpu_shape_checks_raise_before_launch(dev):
    x = torch.zeros((1, 3, 6, 6, 6), device=dev)
    with pytest.raises(ValueError):
        net3d.iso_conv(x, torch.zeros((8, 4, 4), device=dev), torch.zeros(8, device=dev))
    net = net3d.BondNet(2, ch=6, D=4, hid=16).to(dev)
    nc, s = 3, 2
    _, X = _micro(nc * s, s, 1)
    Fv, Fc = net.features(torch.as_tensor(X[None], device=dev))
    with pytest.raises(ValueError):
        net.heads_run(Fv, Fc, nc + 1)                     # features of a 3^3 box read as a 4^3 one
    with pytest.raises(ValueError):
        net.heads_run(Fv, Fc, nc, cp=torch.zeros((1, nc ** 3, 10), device=dev, dtype=torch.float64))
    assert [o.shape[1] for o in net.heads_run(Fv, Fc, nc)] == [54, 72]


def test_void_features_reduce_to_the_two_phase_features():
    """voxel_features_v (infer3d.py:38-50) with every voxel solid equals voxel_features of the same phases; void voxels
    carry the padding's values."""
    n, s = 6, 2
    rng = np.random.default_rng(4)
    phase = (rng.random(n ** 3) < 0.4).astype(np.int64)
    E = np.where(phase == 1, 3.0, 1.0)
    nu = np.where(phase == 1, 0.3, 0.2)
    X = net3d.voxel_features_v(E, nu, np.ones(n ** 3, bool), n, s)
    assert np.array_equal(X, net3d.voxel_features(phase, 3.0, 0.3, n, s))
    solid = rng.random(n ** 3) > 0.2
    Xv = net3d.voxel_features_v(E, nu, solid, n, s)
    core = Xv[:, 2 * s:2 * s + n, 2 * s:2 * s + n, 2 * s:2 * s + n].reshape(3, -1)
    assert np.all(core[:, ~solid] == np.array([[0.0], [0.0], [1.0]]))
    mu = E / (2 * (1 + nu))
    assert abs(np.log(mu[solid] / mu[solid].mean()).astype(np.float32) - core[0, solid]).max() == 0


def test_mask_cells_matches_oracle_and_its_rules():
    import geom3d
    nc = 5
    rng = np.random.default_rng(6)
    lat = geom3d.LatticeGeom(nc, nc, nc, 2)
    Kn = rng.uniform(0.0, 2.0, (nc ** 3, 18)) * lat.valid
    Kn[rng.random(Kn.shape) < 0.1] = 1e-9                       # a few weak bonds that the floor must raise
    Tv = rng.normal(0.0, 0.1, (nc ** 3, 18)) * lat.valid
    cell = rng.random(nc ** 3) > 0.25
    K, T = net3d.mask_cells(torch.as_tensor(Kn), torch.as_tensor(Tv), cell, nc)
    KR, TR = BR.mask_cells(Kn, Tv, cell, nc)
    assert np.array_equal(K.numpy(), KR) and np.array_equal(T.numpy(), TR)
    q = np.where(lat.valid, lat.nbr, 0)
    keep = lat.valid & cell[:, None] & cell[q]                  # the rule, from the lattice tables
    km = 1e-3 * (Kn * keep).sum() / keep.sum()                   # floor from the kept values before flooring
    assert np.all(KR[~keep] == 0) and np.all(TR[~keep] == 0)
    assert np.array_equal(KR[keep], np.maximum(Kn[keep], km)) and np.array_equal(TR[keep], Tv[keep])
    assert np.sum(Kn[keep] < km) > 0                             # the floor acted on some bonds


@pytest.mark.gpu
def test_gpu_mask_cells_matches_oracle(dev):
    import geom3d
    nc = 6
    rng = np.random.default_rng(9)
    lat = geom3d.LatticeGeom(nc, nc, nc, 2)
    Kn = rng.uniform(0.0, 2.0, (nc ** 3, 18)) * lat.valid
    Tv = rng.normal(0.0, 0.1, (nc ** 3, 18)) * lat.valid
    cell = rng.random(nc ** 3) > 0.3
    K, T = net3d.mask_cells(torch.as_tensor(Kn, device=dev), torch.as_tensor(Tv, device=dev), cell, nc)
    KR, TR = BR.mask_cells(Kn, Tv, cell, nc)
    assert np.abs(K.cpu().numpy() - KR).max() <= 1e-15 * np.abs(KR).max()
    assert np.abs(T.cpu().numpy() - TR).max() <= 1e-15 * np.abs(TR).max()


def test_split_precision_needs_twelve_channels():
    with pytest.raises(NotImplementedError):
        net3d.BondNet(2, D=4, precision="bf16x3")
    with pytest.raises(ValueError):
        net3d.BondNet(2, precision="fp16")


@pytest.mark.gpu
def test_gpu_split_bf16_forward_matches_reference_fixture_and_fp32(dev):
    """The opt-in split-BF16 first layer against the reference's outputs (same bar as the FP32 path) and against the
    FP32 kernel (within 1e-5 of the largest output)."""
    f, P = _fixture("bondnet_s2_full")
    net = _net_from(P, 2, 24, 12, 64, dev)
    split = net3d.BondNet(2, precision="bf16x3")
    split.load_state_dict(net.state_dict())
    split = split.to(dev)
    X = torch.as_tensor(f["X"], device=dev)
    o32 = net(X, int(f["nc"]))
    osp = split(X, int(f["nc"]))
    for gi in range(2):
        ref = f["g%d/out" % gi]
        got = osp[gi].cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
        assert np.abs(got - o32[gi].cpu().numpy()).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("nslot", [50, 51, 62, 88, 100, 162, 174])
@pytest.mark.parametrize("hid", [16, 32, 64])
def test_split_w1_buffer_covers_every_fragment_load(nslot, hid):
    """The split-BF16 heads loop (bond_heads_split_kernel) loads W1 fragments two slot pairs ahead: two loads before
    the loop and two per pass over pairs (pr, pr + 1).  Replay that schedule and check that the largest pair index
    it reads lies inside the buffer that lpm_bond_w1_split_size reports for the group (no GPU needed)."""
    lib = net3d._lib()
    grp = net3d.BondGroup()
    grp.ndir, grp.nv, grp.ncell, grp.H, grp.D, grp.dstride, grp.hid = 3, nslot - 8, 8, 16, 12, 12, hid
    nbytes = lib.lpm_bond_w1_split_size(ctypes.byref(grp))
    nt = 1 if hid <= 16 else (2 if hid <= 32 else 4)
    pairs_held = nbytes // (nt * 64 * 16)                      # one bf16x8 (16 bytes) per lane, column tile and pair
    assert pairs_held * nt * 64 * 16 == nbytes
    npair = (nslot + 1) // 2
    lp, reads = 0, []
    for _ in range(2):
        reads.append(lp)
        lp += 1
    for pr in range(0, npair, 2):
        for _ in range(2):
            reads.append(lp)
            lp += 1
    assert max(reads) < pairs_held
    # every pair a step consumes holds real slots; the pairs read past npair are the zero padding
    assert pairs_held >= npair


@pytest.mark.gpu
def test_gpu_split_bf16_refuses_windows_beyond_its_lds(dev):
    """At s = 3 a bond's window (88 / 100 slots of 12 channels, hi and lo planes) does not fit the split kernel's LDS
    budget: the call is refused before any launch (the FP32 heads serve s = 3)."""
    net, P = _random_net(3, 8, 12, 16, seed=5)
    split = net3d.BondNet(3, ch=8, D=12, hid=16, precision="bf16x3")
    split.load_state_dict(net.state_dict())
    split = split.to(dev)
    phase, X = _micro(15, 3, 41)
    with pytest.raises(ValueError, match="LDS"):
        split(torch.as_tensor(X[None], device=dev), 5)


@pytest.mark.gpu
@pytest.mark.parametrize("hid", [16, 32, 64])
def test_gpu_split_bf16_bond_params_match_oracle(dev, hid):
    nc, s, B = 7, 2, 2
    net, P = _random_net(s, 8, 12, hid, seed=hid)
    split = net3d.BondNet(s, ch=8, D=12, hid=hid, precision="bf16x3")
    split.load_state_dict(net.state_dict())
    split = split.to(dev)
    import geom3d as G
    n = nc * s
    Xs, cps = [], []
    for b in range(B):
        phase, X = _micro(n, s, 40 + b, contrast=2.0 + b)
        Xs.append(X)
        Cv = G.iso_C(np.where(phase == 1, 2.0 + b, 1.0), np.where(phase == 1, 0.3, 0.2))
        cps.append(net3d.voigt_cell_params(Cv, n, s).numpy())
    X = np.stack(Xs)
    Kn, Tv = split.bond_params(torch.as_tensor(X, device=dev), nc, torch.as_tensor(np.stack(cps), device=dev))
    outs, groups = BR.forward(X, nc, s, P)
    for b in range(B):
        KnR, TvR = BR.post([o[b] for o in outs], groups, cps[b], nc)
        assert np.abs(Kn[b].cpu().numpy() - KnR).max() <= TOL32 * 10 * np.abs(KnR).max()
        assert np.abs(Tv[b].cpu().numpy() - TvR).max() <= TOL32 * 10 * np.abs(KnR).max()
