"""CPU oracle of the coarse bond network's inference (NumPy / SciPy, FP64).

TEST INFRASTRUCTURE ONLY: only `tests/` and the cpu_baseline leg of bench.py use it, as the checker.  Restates
coarse3d/net3d.py and the post-processing of coarse3d/train3d.py:100-108 / timing/t3d/infer_fast.py:150-177:

  * IsoConv3d (net3d.py:35-49): a 3x3x3 convolution with zero padding whose kernel has one weight per offset class
    (centre, 6 faces, 12 edges, 8 corners), computed here as four class sums of the input followed by a channel mix;
  * trunk (net3d.py:125-126): IsoConv3d, GELU (erf form), IsoConv3d, GELU, IsoConv3d, GELU, 1x1 convolution to D;
  * bond windows (net3d.py:55-118): the canonical window of a bond (end-cell voxels, then the other cells of a 4x3x3 /
    4x4x3 cell box), the stabiliser of the bond direction, every forward bond of the nc^3 box reached by the first
    cube symmetry that maps the canonical direction onto it; positions are handled in doubled integer units;
  * heads (net3d.py:151-166): h = mean_g GELU(W1 P_g x + b1), out = W3 GELU(W2 h + b2) + b3;
  * post (train3d.py:100-108, infer_fast.py:160-177): Voigt bases by the two-end mean of the block-averaged cell
    parameters, Kn = Kn_V exp(o1), Tv = Tv_V + Kn_V o2, scattered to both directed slots of every bond.
Pins: the reference's own outputs on the committed fixtures (tests/golden/bondnet_*.npz), the class-sum convolution
against scipy's direct correlation with the expanded kernel, and the cube-symmetry equivariance of the whole map.
"""
import itertools

import numpy as np
from scipy.special import erf

SC_OFF = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                   [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
                   [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
                   [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.int64)


def cube_group():
    """The 48 signed permutation matrices, rows M[r, p_r] = sign_r, in the order (permutation, signs) of
    itertools (net3d.py:23-32); the order decides which symmetry carries the canonical bond onto each slot."""
    out = []
    for p in itertools.permutations(range(3)):
        for sg in itertools.product((1, -1), repeat=3):
            M = np.zeros((3, 3), np.int64)
            M[np.arange(3), list(p)] = sg
            out.append(M)
    return out


OH = cube_group()


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


# ------------------------------------------------------------------------------------------------------ trunk
def class_sums(x):
    """x (C, Z, Y, X) -> (4, C, Z, Y, X): sums over the centre, face, edge and corner neighbours (zero outside)."""
    p = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1)))
    Z, Y, X = x.shape[1:]
    S = np.zeros((4,) + x.shape)
    for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
        m = abs(dz) + abs(dy) + abs(dx)
        S[m] += p[:, 1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X]
    return S


def iso_conv(x, w, b):
    """IsoConv3d: x (C, Z, Y, X), w (Cout, C, 4), b (Cout,)."""
    S = class_sums(x)
    return np.einsum("oim,mizyx->ozyx", w, S) + b[:, None, None, None]


def trunk(X, P):
    """X (cin, npd^3 as Z, Y, X) -> voxel features (D, Z, Y, X); P = dict of the reference's state-dict names."""
    h = gelu(iso_conv(X, P["trunk.0.w"], P["trunk.0.b"]))
    h = gelu(iso_conv(h, P["trunk.2.w"], P["trunk.2.b"]))
    h = gelu(iso_conv(h, P["trunk.4.w"], P["trunk.4.b"]))
    W = P["trunk.6.weight"].reshape(P["trunk.6.weight"].shape[0], -1)
    return np.einsum("dc,czyx->dzyx", W, h) + P["trunk.6.bias"][:, None, None, None]


def cell_pool(Fv, s):
    """Mean over the s^3 voxels of every cell (avg_pool3d(s, s)); Fv (D, Z, Y, X) with Z, Y, X multiples of s."""
    D, Z, Y, X = Fv.shape
    return Fv.reshape(D, Z // s, s, Y // s, s, X // s, s).mean(axis=(2, 4, 6))


# ---------------------------------------------------------------------------------------------------- windows
def canonical_slots(s, axial):
    """Doubled-unit positions (relative to the bond midpoint) of the voxel slots (2 s^3: the first end cell, then
    the second, each in (a, b, c) order with c fastest) and of the other cells of the window (x slowest, z
    fastest), the canonical direction c and its stabiliser (symmetries with g c = +-c) (net3d.py:55-71)."""
    c = np.array([1, 0, 0]) if axial else np.array([1, 1, 0])
    t2 = 2 * np.arange(s) - (s - 1)
    tv = np.array(list(itertools.product(t2, t2, t2)))
    pv = np.concatenate([tv - c * s, tv + c * s])
    ys = (-1, 0, 1) if axial else (-1, 0, 1, 2)
    cells = [np.array(e) for e in itertools.product((-1, 0, 1, 2), ys, (-1, 0, 1))]
    cells = [e for e in cells if e.any() and not np.array_equal(e, c)]
    pc = np.array([2 * s * e - s * c for e in cells])
    stab = [g for g in OH if np.array_equal(g @ c, c) or np.array_equal(g @ c, -c)]
    return pv, pc, c, stab


def slot_perm(P, g):
    """j -> j' with P[j'] = g P[j]."""
    where = {tuple(p): i for i, p in enumerate(P)}
    return np.array([where[tuple(q)] for q in P @ g.T])


def bond_windows(nc, s):
    """Per group (axial, diagonal): dict(iv (nb, 2 s^3) padded-voxel ids, ic (nb, ncell) padded-cell ids,
    bonds (nb, 2) = (cell, forward slot), perms (H, nslot)) (net3d.py:74-118)."""
    npd, ncp = nc * s + 4 * s, nc + 4
    groups = []
    for axial in (True, False):
        pv, pc, c, stab = canonical_slots(s, axial)
        perms = np.array([np.concatenate([slot_perm(pv, g), len(pv) + slot_perm(pc, g)]) for g in stab])
        iv, ic, bonds = [], [], []
        for k in (range(0, 6, 2) if axial else range(6, 18, 2)):
            o = SC_OFF[k]
            R = next(g for g in OH if np.array_equal(g @ c, o))
            z, y, x = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij")
            I = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
            I = I[np.all((I + o >= 0) & (I + o < nc), 1)]
            m2 = 2 * (I + 2) * s + s + o * s                          # doubled midpoint, padded voxel units
            vox = (m2[:, None, :] + (pv @ R.T)[None] - 1) // 2
            cel = (m2[:, None, :] + (pc @ R.T)[None] - s) // (2 * s)
            iv.append((vox[..., 2] * npd + vox[..., 1]) * npd + vox[..., 0])
            ic.append((cel[..., 2] * ncp + cel[..., 1]) * ncp + cel[..., 0])
            bonds.append(np.stack([(I[:, 2] * nc + I[:, 1]) * nc + I[:, 0], np.full(len(I), k)], 1))
        groups.append(dict(iv=np.concatenate(iv), ic=np.concatenate(ic), bonds=np.concatenate(bonds), perms=perms))
    return groups


# ------------------------------------------------------------------------------------------------------ heads
def heads(Fv, Fc, groups, P):
    """Raw outputs per group (nb, 2) of one specimen; Fv (D, npd^3), Fc (D, ncp^3) flattened (net3d.py:151-166)."""
    D = Fv.shape[0]
    out = []
    for gi, g in enumerate(groups):
        W1, b1 = P["heads.%d.0.weight" % gi], P["heads.%d.0.bias" % gi]
        x = np.concatenate([Fv.T[g["iv"]], Fc.T[g["ic"]]], 1)           # (nb, nslot, D)
        nb, nslot = x.shape[:2]
        W1s = W1.reshape(W1.shape[0], nslot, D)
        xf = x.reshape(nb, nslot * D)
        h = np.zeros((nb, W1.shape[0]))
        for perm in g["perms"]:
            h += gelu(xf @ W1s[:, perm, :].reshape(W1.shape[0], -1).T + b1)
        h /= len(g["perms"])
        h2 = gelu(h @ P["heads.%d.1.weight" % gi].T + P["heads.%d.1.bias" % gi])
        out.append(h2 @ P["heads.%d.2.weight" % gi].T + P["heads.%d.2.bias" % gi])
    return out


def forward(X, nc, s, P):
    """X (B, cin, npd, npd, npd) -> list over groups of (B, nb, 2) (net3d.py:151-166)."""
    groups = bond_windows(nc, s)
    res = [[], []]
    for xb in np.asarray(X, float):
        Fv = trunk(xb, P)
        Fc = cell_pool(Fv, s)
        for gi, o in enumerate(heads(Fv.reshape(Fv.shape[0], -1), Fc.reshape(Fc.shape[0], -1), groups, P)):
            res[gi].append(o)
    return [np.stack(r) for r in res], groups


# ---------------------------------------------------------------------------------------------------- features
def voxel_features(phase, contrast, nu_i, n, s):
    """(3, n + 4s, ...) float32: log(mu / mean mu), (lam - mu) / (18 mu), void flag; 2s void voxels on every side
    (net3d.py:169-181)."""
    E = np.where(phase == 1, contrast, 1.0)
    nu = np.where(phase == 1, nu_i, 0.2)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    P = 2 * s
    X = np.zeros((3,) + (n + 2 * P,) * 3, np.float32)
    X[2] = 1.0
    inner = (slice(P, P + n),) * 3
    X[(0,) + inner] = np.log(mu / mu.mean()).reshape(n, n, n)
    X[(1,) + inner] = ((lam - mu) / (18 * mu)).reshape(n, n, n)
    X[(2,) + inner] = 0.0
    return X


# -------------------------------------------------------------------------------------------------------- post
def voigt_cell_params(Cvox, n, s):
    """Block-averaged voxel (C11, C12, C44) -> per-cell (9 direction Kn, Tv) at r = s/2; for an isotropic cell the
    generalised calcKnTv reduces to Kn = r (C11 - C12 - C44) on the axial and r C44 on the diagonal directions
    and Tv = r (C12 - C44) / 18 (coarse3d/geom3d.py:183-227)."""
    nc = n // s
    C = np.asarray(Cvox, float).reshape(nc, s, nc, s, nc, s, 3).mean(axis=(1, 3, 5)).reshape(-1, 3)
    r = s / 2.0
    T = C[:, 1] - C[:, 2]
    kax = r * (C[:, 0] - 2 * C[:, 2] - T)
    kdg = r * C[:, 2]
    return np.concatenate([np.repeat(kax[:, None], 3, 1), np.repeat(kdg[:, None], 6, 1), (r * T / 18.0)[:, None]], 1)


def post(outs, groups, cp, nc):
    """Raw outputs -> directed Kn, Tv (nc^3, 18) (infer_fast.py:160-177)."""
    N = nc ** 3
    Kn = np.zeros((N, 18))
    Tv = np.zeros((N, 18))
    for g, o in zip(groups, outs):
        p, k = g["bonds"][:, 0], g["bonds"][:, 1]
        ijk = np.stack([p % nc, (p // nc) % nc, p // (nc * nc)], 1) + SC_OFF[k]
        q = (ijk[:, 2] * nc + ijk[:, 1]) * nc + ijk[:, 0]
        kv = 0.5 * (cp[p, k // 2] + cp[q, k // 2])
        tv = 0.5 * (cp[p, 9] + cp[q, 9])
        kn = kv * np.exp(o[:, 0])
        tt = tv + kv * o[:, 1]
        Kn[p, k] = kn
        Kn[q, k ^ 1] = kn
        Tv[p, k] = tt
        Tv[q, k ^ 1] = tt
    return Kn, Tv


def mask_cells(Kn, Tv, cell, nc, kmin=1e-3):
    """Void-cell masking of the coarse parameters with the minimum-stiffness floor (timing/t3d/infer3d.py:95-103),
    written as explicit loops over particles and slots."""
    Kn = np.array(Kn, float)
    Tv = np.array(Tv, float)
    keep = np.zeros(Kn.shape, bool)
    for p in range(nc ** 3):
        x, y, z = p % nc, (p // nc) % nc, p // (nc * nc)
        for k in range(18):
            ox, oy, oz = SC_OFF[k]
            X, Y, Z = x + ox, y + oy, z + oz
            if 0 <= X < nc and 0 <= Y < nc and 0 <= Z < nc:
                keep[p, k] = cell[p] and cell[(Z * nc + Y) * nc + X]
    Kn[~keep] = 0.0
    Tv[~keep] = 0.0
    km = kmin * Kn.sum() / max(keep.sum(), 1)
    Kn[keep] = np.maximum(Kn[keep], km)
    return Kn, Tv
