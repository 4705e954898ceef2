"""Seeded synthetic inputs of the solve path: microstructures and per-bond coefficients.

Generators follow the semantics of the reference's micro.disks_two_phase / grf_two_phase (coarse2d/micro.py:13-38)
and micro3d spheres (coarse3d/micro3d.py:84-132); the bond coefficients are the two-end (Voigt) mean of per-particle
(Kn axial, Kn diagonal, Tv) (coarse3d/geom3d.py:167-175, coarse2d/lin2d.py:210-219).  Large lattices build the
packed forward-slot planes directly on the GPU with tensor shifts (set-up work, outside any timed region).
"""
import numpy as np
import torch

SC_FWD = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1)]
SQ_FWD = [(1, 0), (0, 1), (1, 1), (1, -1)]


# ------------------------------------------------------------------------------------------------ 2D fields
def disks_two_phase(nx, ny, radius, phi, seed, rjit=0.3, max_tries=20000):
    """Random disks of radius radius*(1 +- rjit) until area fraction phi; disks keep a half-pixel gap from earlier
    ones while fewer than half the attempts are used (non-overlapping where possible)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:ny, 0:nx] + 0.5
    img = np.zeros((ny, nx), bool)
    cen = []
    tries = 0
    while img.mean() < phi and tries < max_tries:
        tries += 1
        r = radius * (1 + rjit * (2 * rng.random() - 1))
        c = rng.uniform([-r, -r], [nx + r, ny + r])
        if cen and tries < max_tries // 2:
            a = np.array(cen)
            gap = np.hypot(c[0] - a[:, 0], c[1] - a[:, 1]) - r - a[:, 2]
            if gap.min() < 0.5:
                continue
        cen.append((c[0], c[1], r))
        img |= (xx - c[0]) ** 2 + (yy - c[1]) ** 2 <= r * r
    return img.astype(np.int64).ravel()


def grf_two_phase(nx, ny, corr, phi, seed, pad=None):
    """Thresholded Gaussian random field (Gaussian filter in Fourier space, padded and cropped, so not
    periodic): phase 1 on the top phi quantile."""
    rng = np.random.default_rng(seed)
    p = int(4 * corr) + 4 if pad is None else pad
    noise = rng.standard_normal((ny + 2 * p, nx + 2 * p))
    kx = np.fft.fftfreq(noise.shape[1])[None, :]
    ky = np.fft.fftfreq(noise.shape[0])[:, None]
    filt = np.exp(-2 * (np.pi * corr) ** 2 * (kx * kx + ky * ky))
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))[p:p + ny, p:p + nx]
    return (f > np.quantile(f, 1 - phi)).astype(np.int64).ravel()


# ------------------------------------------------------------------------------------------------ 3D fields
def spheres3d(n, phi, rmean, rng, rjit=0.25, overlap=False, max_tries=200000, exclude=None):
    """Random spheres in an n^3 box until volume fraction phi (of the whole box).  Without overlap a sphere must not
    touch earlier ones (6-neighbour dilation) for the first 70 % of the attempts, as in micro3d._rsa; `exclude`
    marks voxels that may not be taken (overlapping multi-phase generation)."""
    img = np.zeros((n, n, n), bool)
    target = phi * img.size
    occ = 0
    tries = 0
    g = np.arange(n) + 0.5
    while occ < target and tries < max_tries:
        tries += 1
        r = rmean * (1 + rjit * (2 * rng.random() - 1))
        c = rng.uniform(-0.5 * r, n + 0.5 * r, size=3)            # (x, y, z)
        lo = np.maximum(np.floor(c - r - 1).astype(int), 0)
        hi = np.minimum(np.ceil(c + r + 1).astype(int), n)
        if np.any(hi <= lo):
            continue
        dz = (g[lo[2]:hi[2]] - c[2])[:, None, None]
        dy = (g[lo[1]:hi[1]] - c[1])[None, :, None]
        dx = (g[lo[0]:hi[0]] - c[0])[None, None, :]
        m = dx * dx + dy * dy + dz * dz <= r * r
        if not m.any():
            continue
        cur = img[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        if exclude is not None:
            m &= ~exclude[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        if not overlap and tries < 0.7 * max_tries:
            md = m.copy()
            md[1:] |= m[:-1]
            md[:-1] |= m[1:]
            md[:, 1:] |= m[:, :-1]
            md[:, :-1] |= m[:, 1:]
            md[:, :, 1:] |= m[:, :, :-1]
            md[:, :, :-1] |= m[:, :, 1:]
            if (md & cur).any():
                continue
        new = m & ~cur
        k = int(new.sum())
        if k == 0:
            continue
        cur |= m
        occ += k
    return img


def cement_phases(n, seed, fractions=(0.15, 0.55, 0.15, 0.15), rmean=6.0):
    """Four-phase labels (0 pore, 1 hydrate, 2 portlandite, 3 clinker) of random overlapping spheres at the given
    volume fractions: clinker, portlandite and pore spheres are placed in turn into the space still hydrate."""
    rng = np.random.default_rng(seed)
    lab = np.ones((n, n, n), np.int8)
    taken = np.zeros((n, n, n), bool)
    for ph, fr in ((3, fractions[3]), (2, fractions[2]), (0, fractions[0])):
        m = spheres3d(n, fr, rmean, rng, overlap=True, exclude=taken)
        lab[m & ~taken] = ph
        taken |= m
    return lab.ravel()


# ------------------------------------------------------------------------------------------------ removed particles
_SC18 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0), (1, -1, 0),
         (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]


def _shifted(a, o):
    """b[z, y, x] = a[(z, y, x) + o] with o = (dx, dy, dz), False outside the box."""
    b = np.zeros_like(a)
    n = a.shape
    src = [slice(max(0, d), n[ax] + min(0, d)) for ax, d in zip((2, 1, 0), o)]
    dst = [slice(max(0, -d), n[ax] - max(0, d)) for ax, d in zip((2, 1, 0), o)]
    b[dst[2], dst[1], dst[0]] = a[src[2], src[1], src[0]]
    return b


def clean_solid(solid, kmin=4, max_it=50):
    """Connectivity cleaning of a solid set (nz, ny, nx) bool, as cement3d/materials.py:51-84: keep the largest
    18-connected component, then remove, until nothing changes, sites with fewer than kmin bonds to solid
    neighbours or whose bond directions do not span R^3 (a singular diagonal block of K).  At most 8 of the 18
    directions are coplanar, so only sites with <= 8 bonds are tested for spanning.  Returns (mask, #removed)."""
    from scipy import ndimage as ndi
    s = np.asarray(solid, bool).copy()
    n0 = int(s.sum())
    st = ndi.generate_binary_structure(3, 2)
    nd = np.array(_SC18, float)
    nd /= np.linalg.norm(nd, axis=1)[:, None]
    for _ in range(max_it):
        lab, nl = ndi.label(s, structure=st)
        if nl > 1:
            sz = np.bincount(lab.ravel())
            sz[0] = 0
            s = lab == sz.argmax()
        shifted = [_shifted(s, o) for o in _SC18]
        nb = np.zeros(s.shape, np.int16)
        for sh in shifted:
            nb += sh
        bad = s & (nb < kmin)
        cand = s & ~bad & (nb <= 8)
        if cand.any():
            idx = np.nonzero(cand)
            M = np.zeros((len(idx[0]), 3, 3))
            for k, sh in enumerate(shifted):
                M += sh[idx][:, None, None] * np.outer(nd[k], nd[k])[None]
            deg = np.linalg.eigvalsh(M)[:, 0] < 1e-8
            bad[tuple(i[deg] for i in idx)] = True
        if not bad.any() and nl <= 1:
            break
        s &= ~bad
    return s, n0 - int(s.sum())


# ------------------------------------------------------------------------------------------------ coefficients
def guard_len(dims):
    """Leading zero guard values of a packed plane (one z-plane + one row + one site; 2D: one row + one), rounded
    up to 64 values: the layout of lpm_ops.packed_guard / the C ABI."""
    g = dims[0] * dims[1] + dims[0] + 1 if len(dims) == 3 else dims[0] + 1
    return (g + 63) // 64 * 64


def plane_len(dims):
    """Plane stride of the packed layout (lpm_ops.packed_len)."""
    n = int(np.prod(dims)) + 2 * guard_len(dims)
    return (n + 63) // 64 * 64


def packed_from_particles(dims, ppar, device, dtype):
    """Per-particle (Kn axial, Kn diagonal, Tv) (N, 3) -> packed forward planes Kn, Tv (NN/2, N + 2 G) of the
    two-end mean with zero guards, zero on bonds leaving the box.  ppar may be a NumPy array or a tensor."""
    d = len(dims)
    pp = torch.as_tensor(ppar, device=device, dtype=torch.float64)
    shape = tuple(reversed(dims))                                 # (nz, ny, nx) or (ny, nx)
    fwd = SC_FWD if d == 3 else SQ_FWD
    N = int(np.prod(dims))
    G = guard_len(dims)
    kn = torch.zeros((len(fwd), plane_len(dims)), device=device, dtype=dtype)
    tv = torch.zeros_like(kn)
    cols = [pp[:, c].reshape(shape) for c in range(3)]
    for k, off in enumerate(fwd):
        shell = 0 if sum(abs(o) for o in off) == 1 else 1
        a = cols[shell]
        t = cols[2]
        ka = torch.zeros(shape, device=device, dtype=torch.float64)
        ta = torch.zeros_like(ka)
        src, dst = [], []
        for ax in range(d):                       # array axis order is reversed: axis d-1-ax is coordinate ax
            o = off[ax]
            n_ = dims[ax]
            src.append(slice(max(o, 0), n_ + min(o, 0)))
            dst.append(slice(max(-o, 0), n_ - max(o, 0)))
        src_t = tuple(reversed(src))
        dst_t = tuple(reversed(dst))
        ka[dst_t] = 0.5 * (a[dst_t] + a[src_t])
        ta[dst_t] = 0.5 * (t[dst_t] + t[src_t])
        kn[k, G:G + N] = ka.reshape(-1).to(dtype)
        tv[k, G:G + N] = ta.reshape(-1).to(dtype)
    return kn, tv


def directed_from_packed(dims, kp, tp):
    """Packed forward planes (NN/2, N + 2 G) -> directed (N, NN): slot 2k at p, slot 2k+1 at p from the plane value
    of the neighbour p - off_k (zero outside the box)."""
    d = len(dims)
    fwd = SC_FWD if d == 3 else SQ_FWD
    N = int(np.prod(dims))
    G = guard_len(dims)
    kp, tp = kp[..., G:G + N], tp[..., G:G + N]
    shape = tuple(reversed(dims))
    Kn = torch.zeros((N, 2 * len(fwd)), device=kp.device, dtype=kp.dtype)
    Tv = torch.zeros_like(Kn)
    for k, off in enumerate(fwd):
        Kn[:, 2 * k] = kp[k]
        Tv[:, 2 * k] = tp[k]
        kb = torch.zeros(shape, device=kp.device, dtype=kp.dtype)
        tb = torch.zeros_like(kb)
        src, dst = [], []
        for ax in range(d):
            o = off[ax]
            n_ = dims[ax]
            # value at p for the backward slot comes from q = p - off
            src.append(slice(max(-o, 0), n_ - max(o, 0)))
            dst.append(slice(max(o, 0), n_ + min(o, 0)))
        src_t = tuple(reversed(src))
        dst_t = tuple(reversed(dst))
        kb[dst_t] = kp[k].reshape(shape)[src_t]
        tb[dst_t] = tp[k].reshape(shape)[src_t]
        Kn[:, 2 * k + 1] = kb.reshape(-1)
        Tv[:, 2 * k + 1] = tb.reshape(-1)
    return Kn, Tv
