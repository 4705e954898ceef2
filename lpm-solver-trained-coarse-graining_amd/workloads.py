"""Seeded synthetic workloads of the solve path at their stated sizes (built on the GPU, outside any timed region).

plain2d     128^2 square lattice, disks r = 6 +- 2 at fraction 0.30, Kn 1 / 10, Tv = 0.3 Kn, two-end (Voigt) mean;
            bottom row fixed in y, one pinned node, top row displaced by 1e-3 of the height; FP64 plain CG.
kubc3d      48^3 simple cubic, spheres r = 4 +- 1 at fraction 0.25, contrast 10, Tv = 0.3 Kn; nine KUBC load cases
            (6 unit strains, eps = 1e-3, plus 3 combined) as nine right-hand sides of one batched refined PCG.
train3d     16 specimens of 48^3 fine voxels with ellipsoidal inclusions (aspect 1-3, random orientation, fraction
            0.3, contrast 5 or 10, nu = 0.2) coarsened at s = 2 to 24^3 coarse lattices with per-bond Kn, Tv; nine loads.
porous3d    n^3 four-phase cement-like labels (synth.cement_phases) with the pore voxels removed (the void treatment
            of cement3d/materials.py), the solid set cleaned (synth.clean_solid); hydrate 1 / portlandite 2 /
            clinker 5, Tv = 0.3 Kn, two-end mean; tension z on the solid particles of the bottom / top layers.
"""
import numpy as np
import torch

import geom2d
import geom3d
import synth


# ------------------------------------------------------------------------------------------------ plain 2D
def plain2d(n=128, seed=1, device="cuda", radius=6.0, rjit=2.0 / 6.0, phi=0.30, contrast=10.0):
    img = synth.disks_two_phase(n, n, radius, phi, seed, rjit=rjit)
    kn = np.where(img > 0, contrast, 1.0)
    lat = geom2d.SquareGeom(n, n, 1.0)
    Kn, Tv = geom2d.bond_params_two_end_mean(lat, np.stack([kn, kn, 0.3 * kn], 1))
    j = lat.ij[:, 1]
    mask = np.ones((lat.N, 2), np.uint8)
    mask[j == 0, 1] = 0
    mask[j == n - 1, 1] = 0
    mask[0, 0] = 0                                    # pinned node: bottom-left corner fixed in x as well
    up = np.zeros((lat.N, 2))
    up[j == n - 1, 1] = 1e-3 * (n - 1)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), device=device, dtype=dt)
    return dict(dims=(n, n), lat=lat, Kn=t(Kn), Tv=t(Tv), mask=torch.as_tensor(mask, device=device), up=t(up),
                phase_fraction=float(img.mean()), Kn_np=Kn, Tv_np=Tv, mask_np=mask, up_np=up)


# ------------------------------------------------------------------------------------------------ KUBC 3D
COMBINED = [(0, 5), (1, 3), (2, 4)]      # xx + xy, yy + yz, zz + xz


def kubc_strains():
    eps = [geom3d.unit_strain(k) for k in range(6)]
    eps += [geom3d.unit_strain(a) + geom3d.unit_strain(b) for a, b in COMBINED]
    return eps


def kubc3d(n=48, seed=2, device="cuda", radius=4.0, phi=0.25, contrast=10.0):
    rng = np.random.default_rng(seed)
    img = synth.spheres3d(n, phi, radius, rng, rjit=0.25, overlap=False)
    kn = np.where(img.ravel(), contrast, 1.0)
    dims = (n, n, n)
    ppar = np.stack([kn, kn, 0.3 * kn], 1)
    kp64, tp64 = synth.packed_from_particles(dims, ppar, device, torch.float64)
    lat = geom3d.LatticeGeom(n, n, n, 1.0)
    fixed, _, _ = geom3d.prescribed("k_xx", lat.x, lat.ijk, dims, 1, 1.0)
    centre = 0.5 * n * np.ones(3)
    ups = []
    for e in kubc_strains():
        u = np.zeros((lat.N, 3))
        u[fixed] = geom3d.affine_disp(lat.x[fixed], geom3d.EPS * e, centre)
        ups.append(u)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device, dtype=torch.float64)
    return dict(dims=dims, lat=lat, kp64=kp64, tp64=tp64, kp32=kp64.float(), tp32=tp64.float(), up=t(np.array(ups)),
                free=t((~fixed).astype(float)), fixed=fixed, phase_fraction=float(img.mean()), ppar=ppar)


# ------------------------------------------------------------------------------------------------ training 3D
def ellipsoids3d(n, phi, rng, rmean=4.0, aspect=(1.0, 3.0), max_tries=20000):
    """Random ellipsoids (semi-axes r, r, a r with a uniform in `aspect`, random orientation) until the volume
    fraction phi; overlaps allowed."""
    img = np.zeros((n, n, n), bool)
    g = np.arange(n) + 0.5
    tries = 0
    while img.mean() < phi and tries < max_tries:
        tries += 1
        a = rng.uniform(*aspect)
        r = rmean * (0.75 + 0.5 * rng.random())
        semi = np.array([r, r, a * r])
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        c = rng.uniform(0, n, 3)
        ext = int(np.ceil(semi.max())) + 1
        lo = np.maximum(np.floor(c - ext).astype(int), 0)
        hi = np.minimum(np.ceil(c + ext).astype(int), n)
        if np.any(hi <= lo):
            continue
        Z, Y, X = np.meshgrid(g[lo[2]:hi[2]], g[lo[1]:hi[1]], g[lo[0]:hi[0]], indexing="ij")
        d = np.stack([X - c[0], Y - c[1], Z - c[2]], -1) @ q          # body-frame coordinates
        m = ((d / semi) ** 2).sum(-1) <= 1.0
        img[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] |= m
    return img


def train3d(B=16, nf=48, s=2, seed=3, device="cuda", phi=0.3):
    """Per-specimen coarse parameters (B, Nc, 18) FP32: Voigt mean of the fine stiffness over each s^3 block, the
    generalised calcKnTv of the cell (geom3d.cell_params_from_C66, radius s/2) and the two-end mean per bond."""
    rng = np.random.default_rng(seed)
    nc = nf // s
    lat = geom3d.LatticeGeom(nc, nc, nc, s * 1.0)
    Ks, Ts = [], []
    for b in range(B):
        img = ellipsoids3d(nf, phi, rng)
        contrast = 5.0 if rng.random() < 0.5 else 10.0
        E = np.where(img, contrast, 1.0).ravel()
        C66 = geom3d.iso_C66(geom3d.iso_C(E, np.full(E.shape, 0.2)))     # (nf^3, 6, 6), z, y, x order
        Cc = C66.reshape(nc, s, nc, s, nc, s, 6, 6).mean(axis=(1, 3, 5)).reshape(-1, 6, 6)  # Voigt mean per block
        Kn, Tv = geom3d.bond_params_from_cells(lat, geom3d.cell_params_from_C66(Cc, s / 2.0))
        Ks.append(Kn)
        Ts.append(Tv)
    f32 = lambda a: torch.as_tensor(np.array(a), device=device, dtype=torch.float32)
    return dict(nc=nc, s=s, lat=lat, Kn=f32(Ks), Tv=f32(Ts))


# ------------------------------------------------------------------------------------------------ voids 3D
PHASE_KN = np.array([0.1, 1.0, 2.0, 5.0])


def porous3d(n=256, seed=2024, device="cuda", lab=None):
    """Removed pore particles on the cement-like labels; returns packed FP64 / FP32 coefficients, the bond words,
    tension-z boundary data restricted to the solid particles and the solid mask (host and device)."""
    import lpm_ops as ops
    if lab is None:
        lab = synth.cement_phases(n, seed)
    solid, nrm = synth.clean_solid((lab != 0).reshape(n, n, n))
    solid = solid.ravel()
    kn = PHASE_KN[lab]
    ppar = np.stack([kn, kn, 0.3 * kn], 1)
    dims = (n, n, n)
    kp64, tp64 = synth.packed_from_particles(dims, ppar, device, torch.float64)
    sd = torch.as_tensor(solid, device=device)
    words = ops.bond_words(dims, sd)
    z = torch.arange(n, device=device).repeat_interleave(n * n)
    fixed = ((z == 0) | (z == n - 1)) & sd
    up = torch.zeros((n ** 3, 3), device=device, dtype=torch.float64)
    up[(z == n - 1) & sd, 2] = 1e-3 * (n - 1)
    free = (sd & ~fixed).to(torch.float64)
    return dict(dims=dims, lab=lab, solid=solid, removed_by_cleaning=int(nrm), kp64=kp64, tp64=tp64,
                kp32=kp64.float(), tp32=tp64.float(), words=words, up=up, free=free, fixed=fixed, ppar=ppar,
                solid_fraction=float(solid.mean()))
