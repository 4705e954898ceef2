"""Pins of the oracle's lattice with removed particles (oracle/lpm_ref.py ActiveBox) and of the solid-set cleaning
(synth.clean_solid), by what the mathematics fixes: the Hessian of the particle energy of the reduced lattice,
symmetry, the energy identity, rigid-body null space, the affine patch test, the opp-flag stress rule at a pore
(a particle next to removed sites has the stress weights of a box-face particle with the same missing slots), and
the full mask giving the full box."""
import numpy as np
import pytest

import lpm_ref as R
import synth


def _mask(dims, seed, frac):
    rng = np.random.default_rng(seed)
    act = rng.random(int(np.prod(dims))) >= frac
    return act


def _params(lat, seed):
    return R.random_symmetric_params(lat, np.random.default_rng(seed), kn=(0.5, 1.5), tv=(-0.1, 0.1))


@pytest.mark.parametrize("dims", [(4, 4, 3), (3, 3, 3), (6, 5), (5, 4)])
def test_full_mask_is_the_box(dims):
    box = R.Box(dims)
    act = R.ActiveBox(dims, np.ones(box.N, bool))
    Kn, Tv = _params(box, 1)
    u = np.random.default_rng(2).standard_normal((box.N, box.dim))
    assert np.array_equal(act.nbr, box.nbr) and np.array_equal(act.valid, box.valid)
    assert np.abs(R.apply(act, u, Kn, Tv) - R.apply(box, u, Kn, Tv)).max() == 0.0
    assert np.abs(R.stress(act, u, Kn, Tv) - R.stress(box, u, Kn, Tv)).max() == 0.0


@pytest.mark.parametrize("dims,frac,seed", [((4, 4, 3), 0.15, 3), ((5, 3, 3), 0.25, 4), ((6, 5), 0.2, 5),
                                            ((7, 4), 0.3, 6)])
def test_reduced_lattice_stiffness_is_energy_hessian(dims, frac, seed):
    act = _mask(dims, seed, frac)
    lat = R.ActiveBox(dims, act)
    full = R.Box(dims)
    Kf, Tf = _params(full, seed)
    Kn, Tv = Kf[lat.sites], Tf[lat.sites]          # compact rows; slots to removed sites are ignored
    H = R.dense_from_energy(lat, Kn, Tv)
    sc = max(1.0, np.abs(H).max())
    assert np.abs(R.stiffness(lat, Kn, Tv).toarray() - H).max() < 1e-12 * sc
    assert np.abs(R.dense_from_apply(lat, Kn, Tv) - H).max() < 1e-12 * sc
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((lat.N, lat.dim))
    a = R.apply(lat, u, Kn, Tv)
    assert np.abs(a - R.apply_loops(lat, u, Kn, Tv)).max() < 1e-13 * np.abs(a).max()
    D = R.block_diag(lat, Kn, Tv)
    K = H
    d = lat.dim
    for i in range(lat.N):
        assert np.abs(D[i] - K[d * i:d * i + d, d * i:d * i + d]).max() < 1e-12 * sc


def test_removed_neighbours_change_the_operator():
    """A removed site drops every bond to it: K of the reduced lattice differs from the box's K restricted to the
    remaining sites (the shell sums of the neighbours lose the removed bonds)."""
    dims = (4, 4, 4)
    act = np.ones(64, bool)
    act[21] = False
    lat, box = R.ActiveBox(dims, act), R.Box(dims)
    Kf, Tf = _params(box, 7)
    Kr = R.stiffness(lat, Kf[lat.sites], Tf[lat.sites]).toarray()
    dof = (3 * lat.sites[:, None] + np.arange(3)).ravel()
    Kb = R.stiffness(box, Kf, Tf).toarray()[np.ix_(dof, dof)]
    assert np.abs(Kr - Kb).max() > 1e-3


@pytest.mark.parametrize("dims", [(7, 6, 5), (9, 8)])
def test_reduced_lattice_symmetry_energy_null_space(dims):
    full = R.Box(dims)
    # one pore in the middle: the remaining sites stay connected
    c = np.array(dims) // 2
    act = ~np.all(np.abs(full.ijk - c) <= 1, axis=1)
    lat = R.ActiveBox(dims, act)
    Kf, Tf = _params(full, 8)
    Kn, Tv = Kf[lat.sites], Tf[lat.sites]
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal((2, lat.N, lat.dim))
    Ku, Kv = R.apply(lat, u, Kn, Tv), R.apply(lat, v, Kn, Tv)
    assert abs((v * Ku).sum() - (u * Kv).sum()) < 1e-11 * abs((v * Ku).sum())
    assert abs((u * Ku).sum() - 2 * R.energy(lat, u, Kn, Tv).sum()) < 1e-11 * (u * Ku).sum()
    for a in range(lat.dim):
        t = np.zeros((lat.N, lat.dim))
        t[:, a] = 1.0
        assert np.abs(R.apply(lat, t, Kn, Tv)).max() < 1e-12
    W = np.zeros((lat.dim, lat.dim))
    W[0, 1], W[1, 0] = 0.7, -0.7
    assert np.abs(R.apply(lat, lat.x @ W.T, Kn, Tv)).max() < 1e-12


def test_affine_patch_away_from_the_pore_and_stress_of_full_particles():
    dims = (13, 13, 13)
    full = R.Box(dims)
    act = ~np.all(full.ijk == 6, axis=1)                                  # one removed site at the centre
    lat = R.ActiveBox(dims, act)
    C = R.iso_C(np.full(full.N, 3.0), np.full(full.N, 0.25))
    Kf, Tf = R.two_end_mean(full, R.sc_params_from_C(C, 0.5))
    Kn, Tv = Kf[lat.sites], Tf[lat.sites]
    rng = np.random.default_rng(10)
    e = rng.standard_normal((3, 3))
    e = 0.5 * (e + e.T)
    u = lat.x @ e.T
    f = R.apply(lat, u, Kn, Tv)
    # two hops away from the pore and from the box faces: every bond and every neighbour shell sum intact
    far = np.all((lat.ijk >= 2) & (lat.ijk < 11), axis=1) & (np.abs(lat.ijk - 6).max(1) > 2)
    assert far.any() and np.abs(f[far]).max() < 1e-12
    near = np.abs(lat.ijk - 6).max(1) == 1
    assert np.abs(f[near]).max() > 1e-6                                   # the pore surface carries forces
    sig = R.stress(lat, u, Kn, Tv)
    C11, C12, C44 = C[0]
    tr = np.trace(e)
    expect = [C12 * tr + (C11 - C12) * e[0, 0], C12 * tr + (C11 - C12) * e[1, 1], C12 * tr + (C11 - C12) * e[2, 2],
              2 * C44 * e[1, 2], 2 * C44 * e[0, 2], 2 * C44 * e[0, 1]]
    fullp = lat.valid.all(1)
    assert np.abs(sig[fullp] - np.array(expect)[None]).max() < 1e-12 * np.abs(expect).max()


def test_opp_flag_rule_at_a_pore_equals_the_box_face_rule():
    """The stress weights of a particle depend only on which of its slots are bonds: a particle whose five -x
    neighbours are removed has the weights of a particle on the x = 0 face."""
    dims = (5, 5, 5)
    box = R.Box(dims)
    ijk = box.ijk
    p = int(np.flatnonzero(np.all(ijk == (2, 2, 2), axis=1))[0])
    act = np.ones(box.N, bool)
    for k in range(18):
        if R.SC_OFF[k, 0] == -1:
            act[box.nbr[p, k]] = False
    lat = R.ActiveBox(dims, act)
    face = int(np.flatnonzero(np.all(ijk == (0, 2, 2), axis=1))[0])
    assert np.array_equal(lat.valid[lat.site_of[p]], box.valid[face])
    w_pore = R.stress_weights(lat)[lat.site_of[p]]
    w_face = R.stress_weights(box)[face]
    assert np.array_equal(w_pore, w_face)
    of = (w_face * box.V / box.l0)[box.valid[face]]
    assert set(np.round(of, 12)) == {0.5, 1.0}


def _clean_plain(s, kmin=4, max_it=50):
    """Direct restatement of the cleaning (spanning tested on every solid site) for the pin below."""
    from scipy import ndimage as ndi
    s = s.copy()
    st = ndi.generate_binary_structure(3, 2)
    nd = R.SC_OFF / np.linalg.norm(R.SC_OFF, axis=1)[:, None]
    for _ in range(max_it):
        lab, nl = ndi.label(s, structure=st)
        if nl > 1:
            sz = np.bincount(lab.ravel())
            sz[0] = 0
            s = lab == sz.argmax()
        nb = np.zeros(s.shape, int)
        M = np.zeros(s.shape + (3, 3))
        for k, o in enumerate(R.SC_OFF):
            sh = synth._shifted(s, tuple(o))
            nb += sh
            M += sh[..., None, None] * np.outer(nd[k], nd[k])
        deg = np.zeros_like(s)
        deg[s] = np.linalg.eigvalsh(M[s])[:, 0] < 1e-8
        d = deg | (s & (nb < kmin))
        if not d.any() and nl <= 1:
            break
        s &= ~d
    return s


@pytest.mark.parametrize("seed,frac", [(1, 0.3), (2, 0.45), (3, 0.6)])
def test_clean_solid_matches_plain_restatement(seed, frac):
    rng = np.random.default_rng(seed)
    s = rng.random((12, 11, 10)) >= frac
    got, nrm = synth.clean_solid(s)
    ref = _clean_plain(s)
    assert np.array_equal(got, ref) and nrm == int(s.sum() - ref.sum())
    # every kept site has >= 4 bonds spanning R^3 and the kept set is one 18-connected component
    lat = R.ActiveBox((10, 11, 12), got.ravel())
    assert lat.N == 0 or (lat.valid.sum(1) >= 4).all()
