"""Bond-network inference on the GPU with the interface of coarse3d/net3d.py: the producer of the coarse Kn / Tv.

`BondNet` keeps the reference's module tree and parameter names (trunk.0.w ... heads.1.2.bias), so a checkpoint of
the reference loads with `load_state_dict`; its forward runs the HIP kernels of liblpm.so:
  * trunk: three O_h-invariant 3x3x3 convolutions (lpm_iso_conv; the class sums of each input channel are staged in
    LDS and mixed into the output channels in registers), GELU fused, the closing 1x1 convolution fused into the
    third launch, which writes the D features voxel-major;
  * cell means (lpm_cell_pool);
  * per direction group one fused launch of the heads (lpm_bond_heads): window gather, the first layer for every
    element of the bond's stabiliser on the FP32 matrix cores, GELU, mean, second and third layer, and optionally
    Kn = Kn_V exp(o1), Tv = Tv_V + Kn_V o2 scattered to both directed slots (coarse3d/train3d.py:100-108,
    timing/t3d/infer_fast.py:160-177).
Inference only: the forward builds no autograd graph (training the network is not part of this library).
"""
import ctypes
import itertools

import numpy as np
import torch
import torch.nn as nn

import _lpmlib as L
from geom3d import SC_OFFS

P = ctypes.c_void_p
I32 = ctypes.c_int32


class BondGroup(ctypes.Structure):
    _fields_ = [("ndir", I32), ("nv", I32), ("ncell", I32), ("H", I32), ("D", I32), ("dstride", I32), ("hid", I32),
                ("slot", I32 * 6), ("voff", P), ("coff", P), ("inv", P)]


_SIGS = {
    "lpm_iso_conv": (I32, [I32, I32, I32, I32, I32, I32, P, P, P, I32, I32, I32, P, P, P, P]),
    "lpm_cell_pool": (I32, [I32, I32, I32, I32, P, P, P]),
    "lpm_bond_w1_size": (ctypes.c_size_t, [ctypes.POINTER(BondGroup)]),
    "lpm_bond_w1_layout": (I32, [ctypes.POINTER(BondGroup), P, P, P]),
    "lpm_bond_heads": (I32, [ctypes.POINTER(BondGroup), I32, I32, I32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "lpm_bond_w1_split_size": (ctypes.c_size_t, [ctypes.POINTER(BondGroup)]),
    "lpm_bond_w1_split_layout": (I32, [ctypes.POINTER(BondGroup), P, P, P, P]),
    "lpm_bond_heads_split": (I32, [ctypes.POINTER(BondGroup), I32, I32, I32] + [P] * 14),
}
L.SIGS.update(_SIGS)


def _lib():
    lib = L.lib()
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


# ------------------------------------------------------------------------------------------------- geometry
def signed_perms():
    """The 48 cube symmetries as signed permutation matrices M[r, p_r] = sign_r, permutations outer and signs inner
    in itertools order (net3d.py:23-32)."""
    out = []
    for p in itertools.permutations(range(3)):
        for sg in itertools.product((1, -1), repeat=3):
            M = np.zeros((3, 3), np.int64)
            M[np.arange(3), list(p)] = sg
            out.append(M)
    return out


OH = signed_perms()


def canonical_window(s, axial):
    """Window slots of a bond in its canonical frame (net3d.py:55-71), in doubled voxel units about the bond
    midpoint: pv (2 s^3, 3) the voxels of the two end cells, pc the other cells of the 4x3x3 (axial) / 4x4x3
    (diagonal) cell box, c the canonical direction and the stabiliser (g c = +-c)."""
    c = np.array([1, 0, 0]) if axial else np.array([1, 1, 0])
    t2 = 2 * np.arange(s) - (s - 1)
    tv = np.array(list(itertools.product(t2, t2, t2)))
    pv = np.concatenate([tv - c * s, tv + c * s])
    ys = (-1, 0, 1) if axial else (-1, 0, 1, 2)
    box = [np.array(e) for e in itertools.product((-1, 0, 1, 2), ys, (-1, 0, 1))]
    pc = np.array([2 * s * e - s * c for e in box if e.any() and not np.array_equal(e, c)])
    stab = [g for g in OH if np.array_equal(g @ c, c) or np.array_equal(g @ c, -c)]
    return pv, pc, c, stab


def _perm(P2, g):
    idx = {tuple(p): i for i, p in enumerate(P2)}
    return np.array([idx[tuple(q)] for q in P2 @ g.T])


class WindowTables:
    """Per group: forward slots, window offsets per direction (voxel (dx, dy, dz) relative to (I + 2) s, cell
    offsets relative to I + 2) and the stabiliser perms (H, nslot) of net3d.BondWindows (net3d.py:74-118)."""

    def __init__(self, s):
        self.s = s
        self.groups = []
        for axial in (True, False):
            pv, pc, c, stab = canonical_window(s, axial)
            perms = np.array([np.concatenate([_perm(pv, g), len(pv) + _perm(pc, g)]) for g in stab])
            slots = list(range(0, 6, 2) if axial else range(6, 18, 2))
            vo, co = [], []
            for k in slots:
                o = SC_OFFS[k]
                R = next(g for g in OH if np.array_equal(g @ c, o))
                vo.append((s + o * s + pv @ R.T - 1) // 2)
                co.append((pc @ R.T + o * s) // (2 * s))
            self.groups.append(dict(axial=axial, slots=slots, voff=np.stack(vo), coff=np.stack(co), perms=perms,
                                    nv=len(pv), ncell=len(pc)))


class BondWindows:
    """Gather indices of every forward coarse bond of an nc^3 box (net3d.py:74-118): per group iv (nb, nv) padded
    voxel ids, ic (nb, ncell) padded cell ids, bonds (nb, 2) = (cell, slot) and perms (H, nslot)."""

    def __init__(self, nc, s):
        self.nc, self.s = nc, s
        self.npd, self.ncp = nc * s + 4 * s, nc + 4
        self.groups = []
        for g in WindowTables(s).groups:
            iv, ic, bonds = [], [], []
            for d, k in enumerate(g["slots"]):
                I = _cells(nc, SC_OFFS[k])
                vb = (I + 2) * s
                v = vb[:, None, :] + g["voff"][d][None]
                cc = (I + 2)[:, None, :] + g["coff"][d][None]
                iv.append((v[..., 2] * self.npd + v[..., 1]) * self.npd + v[..., 0])
                ic.append((cc[..., 2] * self.ncp + cc[..., 1]) * self.ncp + cc[..., 0])
                bonds.append(np.stack([(I[:, 2] * nc + I[:, 1]) * nc + I[:, 0], np.full(len(I), k)], 1))
            self.groups.append(dict(iv=np.concatenate(iv), ic=np.concatenate(ic), bonds=np.concatenate(bonds),
                                    perms=g["perms"], nv=g["nv"], nc=g["ncell"]))


def _cells(nc, o):
    """(x, y, z) of the cells whose neighbour I + o stays in the nc^3 box, x fastest."""
    rng = [np.arange(max(0, -int(o[a])), nc - max(0, int(o[a]))) for a in range(3)]
    z, y, x = np.meshgrid(rng[2], rng[1], rng[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


# ---------------------------------------------------------------------------------------------------- modules
class IsoConv3d(nn.Module):
    """3x3x3 convolution with an O_h-invariant kernel: 4 weights per channel pair (centre, faces, edges, corners),
    zero padding 1 (net3d.py:35-49)."""

    def __init__(self, cin, cout):
        super().__init__()
        off = np.array(list(itertools.product((-1, 0, 1), repeat=3)))
        m = np.zeros((4, 27))
        m[(off != 0).sum(1), np.arange(27)] = 1.0
        self.register_buffer("masks", torch.as_tensor(m.reshape(4, 3, 3, 3), dtype=torch.float32))
        self.w = nn.Parameter(torch.randn(cout, cin, 4) * np.sqrt(2.0 / (cin * 27)))
        self.b = nn.Parameter(torch.zeros(cout))

    def forward(self, x, gelu=False, proj=None, dstride=0):
        return iso_conv(x, self.w, self.b, gelu=gelu, proj=proj, dstride=dstride)


def iso_conv(x, w, b, gelu=False, proj=None, dstride=0):
    """x (B, cin, nz, ny, nx) float32 on the GPU -> (B, cout, nz, ny, nx); with proj = (pw (D, cout), pb (D,)) the
    1x1 convolution is applied too and the result is voxel-major (B, nz*ny*nx, dstride), channels >= D zero."""
    L.require_gpu(x)
    if x.dtype != torch.float32 or x.dim() != 5:
        raise TypeError("iso_conv takes (B, C, nz, ny, nx) float32")
    x = x.contiguous()
    B, cin, nz, ny, nx = x.shape
    cout = w.shape[0]
    if tuple(w.shape) != (cout, cin, 4) or tuple(b.shape) != (cout,):
        raise ValueError("iso_conv: weights %s / bias %s do not match %d input channels"
                         % (tuple(w.shape), tuple(b.shape), cin))
    w = w.detach().to(x.device, torch.float32).contiguous()
    b = b.detach().to(x.device, torch.float32).contiguous()
    if proj is None:
        y = torch.empty((B, cout, nz, ny, nx), device=x.device, dtype=torch.float32)
        pw = pb = None
        dproj = dstride = 0
    else:
        pw = proj[0].detach().to(x.device, torch.float32).reshape(proj[0].shape[0], -1).contiguous()
        pb = proj[1].detach().to(x.device, torch.float32).contiguous()
        dproj = pw.shape[0]
        if pw.shape[1] != cout or tuple(pb.shape) != (dproj,):
            raise ValueError("iso_conv: projection %s / %s does not match %d channels"
                             % (tuple(pw.shape), tuple(pb.shape), cout))
        dstride = dstride or 4 * ((dproj + 3) // 4)
        y = torch.empty((B, nz * ny * nx, dstride), device=x.device, dtype=torch.float32)
    L.check(_lib().lpm_iso_conv(B, cin, cout, nx, ny, nz, _ptr(x), _ptr(w), _ptr(b), int(bool(gelu)), dproj, dstride,
                                _ptr(pw), _ptr(pb), _ptr(y), L.stream_of(x)))
    return y


class BondNet(nn.Module):
    """Exactly cubic-symmetric bond-parameter network (net3d.py:121-166); same parameters and names.

    precision: "fp32" (default; the first head layer on the FP32 matrix cores, as the reference computes it) or
    "bf16x3" (opt-in, D = 12 only: that layer as three BF16 products of the split operands hi + lo, accumulated in
    FP32, about twice as fast; the raw outputs stay within a few 1e-6 of the FP32 ones, relative to their largest)."""

    def __init__(self, s, cin=3, ch=24, D=12, hid=64, precision="fp32"):
        super().__init__()
        if precision not in ("fp32", "bf16x3"):
            raise ValueError("precision must be 'fp32' or 'bf16x3'")
        if precision == "bf16x3" and D != 12:
            raise NotImplementedError("the split BF16 heads take D = 12 feature channels")
        self.s, self.D, self.precision = s, D, precision
        self.trunk = nn.Sequential(IsoConv3d(cin, ch), nn.GELU(), IsoConv3d(ch, ch), nn.GELU(), IsoConv3d(ch, ch),
                                   nn.GELU(), nn.Conv3d(ch, D, 1))
        self.heads = nn.ModuleList()
        self.tables = WindowTables(s)
        for g in self.tables.groups:
            nin = (g["nv"] + g["ncell"]) * D
            l1 = nn.Linear(nin, hid)
            l3 = nn.Linear(hid, 2)
            nn.init.zeros_(l3.weight)
            nn.init.zeros_(l3.bias)
            self.heads.append(nn.ModuleList([l1, nn.Linear(hid, hid), l3]))
        self.dstride = 4 * ((D + 3) // 4)
        self._dev = {}

    def _device_tables(self, nc, device):
        key = (nc, str(device))
        if key not in self._dev:
            s = self.s
            npd, ncp = nc * s + 4 * s, nc + 4
            out = []
            for g in self.tables.groups:
                vo, co = g["voff"], g["coff"]
                voff = (vo[..., 2] * npd + vo[..., 1]) * npd + vo[..., 0]
                coff = (co[..., 2] * ncp + co[..., 1]) * ncp + co[..., 0]
                inv = np.argsort(g["perms"], axis=1)
                t = dict(voff=torch.as_tensor(voff, dtype=torch.int32, device=device).contiguous(),
                         coff=torch.as_tensor(coff, dtype=torch.int32, device=device).contiguous(),
                         inv=torch.as_tensor(inv, dtype=torch.int32, device=device).contiguous())
                grp = BondGroup()
                grp.ndir = len(g["slots"])
                grp.nv, grp.ncell = g["nv"], g["ncell"]
                grp.H = len(g["perms"])
                grp.D, grp.dstride = self.D, self.dstride
                grp.hid = self.heads[0][0].weight.shape[0]
                for d, k in enumerate(g["slots"]):
                    grp.slot[d] = k
                grp.voff, grp.coff, grp.inv = t["voff"].data_ptr(), t["coff"].data_ptr(), t["inv"].data_ptr()
                t["grp"] = grp
                nb = sum(len(_cells(nc, SC_OFFS[k])) for k in g["slots"])
                t["nb"] = nb
                out.append(t)
            self._dev[key] = out
        return self._dev[key]

    def windows(self, nc, device=None):
        """BondWindows of an nc^3 box (index lists, numpy)."""
        return BondWindows(nc, self.s).groups

    @torch.no_grad()
    def features(self, X):
        """Trunk and cell means: X (B, cin, npd, npd, npd) -> Fv (B, npd^3, dstride), Fc (B, ncp^3, dstride)."""
        t = self.trunk
        h = t[0](X, gelu=True)
        h = t[2](h, gelu=True)
        Fv = t[4](h, gelu=True, proj=(t[6].weight, t[6].bias), dstride=self.dstride)
        del h
        npd = X.shape[-1]
        ncp = npd // self.s
        Fc = torch.empty((X.shape[0], ncp ** 3, self.dstride), device=X.device, dtype=torch.float32)
        L.check(_lib().lpm_cell_pool(X.shape[0], ncp, self.s, self.dstride, _ptr(Fv), _ptr(Fc), L.stream_of(X)))
        return Fv, Fc

    def _check_input(self, X, nc):
        L.require_gpu(X)
        npd = nc * self.s + 4 * self.s
        if X.dim() != 5 or tuple(X.shape[2:]) != (npd,) * 3:
            raise ValueError("X must be (B, cin, %d, %d, %d) for nc = %d, s = %d" % (npd, npd, npd, nc, self.s))
        return X.to(torch.float32).contiguous()

    @torch.no_grad()
    def heads_run(self, Fv, Fc, nc, cp=None, Kn=None, Tv=None, events=None):
        """Raw outputs per group from the features; with cp, Kn / Tv are filled as in bond_params.  `events`: an
        optional list that receives a (start, end) pair of CUDA events around every heads launch."""
        B = Fv.shape[0]
        s, ds = self.s, self.dstride
        npd, ncp, N = (nc + 4) * s, nc + 4, nc ** 3
        L.require_gpu(Fv, Fc, cp, Kn, Tv)
        if (tuple(Fv.shape) != (B, npd ** 3, ds) or tuple(Fc.shape) != (B, ncp ** 3, ds)
                or Fv.dtype != torch.float32 or Fc.dtype != torch.float32
                or not Fv.is_contiguous() or not Fc.is_contiguous()):
            raise ValueError("heads_run: features must be float32 (B, %d, %d) and (B, %d, %d) for nc = %d"
                             % (npd ** 3, ds, ncp ** 3, ds, nc))
        if cp is not None:
            for t_, shp in ((cp, (B, N, 10)), (Kn, (B, N, 18)), (Tv, (B, N, 18))):
                if t_ is None or tuple(t_.shape) != shp or t_.dtype != torch.float64 or not t_.is_contiguous():
                    raise ValueError("heads_run: cp / Kn / Tv must be contiguous float64 (B, %d, 10 | 18)" % N)
        outs = []
        lib = _lib()
        st = L.stream_of(Fv)
        split = self.precision == "bf16x3"
        for (l1, l2, l3), t in zip(self.heads, self._device_tables(nc, Fv.device)):
            grp = t["grp"]
            w1 = l1.weight.detach().to(Fv.device, torch.float32).contiguous()
            if split:
                nbytes = lib.lpm_bond_w1_split_size(ctypes.byref(grp))
                w1h = torch.empty(nbytes, device=Fv.device, dtype=torch.uint8)
                w1l = torch.empty(nbytes, device=Fv.device, dtype=torch.uint8)
                L.check(lib.lpm_bond_w1_split_layout(ctypes.byref(grp), _ptr(w1), _ptr(w1h), _ptr(w1l), st))
            else:
                w1f = torch.empty(lib.lpm_bond_w1_size(ctypes.byref(grp)) // 4, device=Fv.device,
                                  dtype=torch.float32)
                L.check(lib.lpm_bond_w1_layout(ctypes.byref(grp), _ptr(w1), _ptr(w1f), st))
            par = [p.detach().to(Fv.device, torch.float32).contiguous() for p in
                   (l1.bias, l2.weight, l2.bias, l3.weight, l3.bias)]
            out = torch.empty((B, t["nb"], 2), device=Fv.device, dtype=torch.float32)
            if events is not None:
                events.append((torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)))
                events[-1][0].record()
            if split:
                L.check(lib.lpm_bond_heads_split(ctypes.byref(grp), B, nc, self.s, _ptr(Fv), _ptr(Fc), _ptr(w1h),
                                                 _ptr(w1l), *[_ptr(p) for p in par], _ptr(out), _ptr(cp), _ptr(Kn),
                                                 _ptr(Tv), st))
            else:
                L.check(lib.lpm_bond_heads(ctypes.byref(grp), B, nc, self.s, _ptr(Fv), _ptr(Fc), _ptr(w1f),
                                           *[_ptr(p) for p in par], _ptr(out), _ptr(cp), _ptr(Kn), _ptr(Tv), st))
            if events is not None:
                events[-1][1].record()
            outs.append(out)
        return outs

    @torch.no_grad()
    def forward(self, X, nc):
        """X (B, cin, npd, npd, npd) -> list over (axial, diagonal) of raw outputs (B, nbonds, 2)
        (net3d.py:151-166)."""
        X = self._check_input(X, nc)
        Fv, Fc = self.features(X)
        return self.heads_run(Fv, Fc, nc)

    @torch.no_grad()
    def bond_params(self, X, nc, cp):
        """Kn, Tv (B, nc^3, 18) float64 of the coarse lattice: Kn = Kn_V exp(o1), Tv = Tv_V + Kn_V o2 on both directed
        slots of every bond, zero on slots that leave the box; cp (B | 1, nc^3, 10) per-cell Voigt parameters
        (9 direction Kn and Tv, voigt_cell_params) (train3d.py:100-108, infer_fast.py:160-177)."""
        X = self._check_input(X, nc)
        B = X.shape[0]
        cp = cp.to(X.device, torch.float64)
        if cp.dim() == 2:
            cp = cp[None]
        cp = cp.expand(B, nc ** 3, 10).contiguous()
        Fv, Fc = self.features(X)
        Kn = torch.zeros((B, nc ** 3, 18), device=X.device, dtype=torch.float64)
        Tv = torch.zeros_like(Kn)
        self.heads_run(Fv, Fc, nc, cp=cp, Kn=Kn, Tv=Tv)
        return Kn, Tv


# --------------------------------------------------------------------------------------------------- host prep
def voxel_features(phase, contrast, nu_i, n, s):
    """(3, n + 4s, ...) float32 voxel features of a two-phase microstructure: log(mu / mean mu), (lam - mu) /
    (18 mu), void flag, with 2s void voxels on every side (net3d.py:169-181)."""
    E = np.where(phase == 1, contrast, 1.0)
    nu = np.where(phase == 1, nu_i, 0.2)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    pad = 2 * s
    X = np.zeros((3,) + (n + 2 * pad,) * 3, np.float32)
    X[2] = 1.0
    core = (slice(pad, pad + n),) * 3
    X[(0,) + core] = np.log(mu / mu.mean()).reshape(n, n, n)
    X[(1,) + core] = ((lam - mu) / (18 * mu)).reshape(n, n, n)
    X[(2,) + core] = 0.0
    return X


def voigt_cell_params(Cvox, n, s):
    """Per-cell (9 direction Kn, Tv) (nc^3, 10) float64 of the block means of the voxel (C11, C12, C44) (n^3, 3)
    at r = s/2, on the device of Cvox: the generalised SC calcKnTv of an isotropic-form cell stiffness, axial
    Kn = r (C11 - C12 - C44), diagonal Kn = r C44, Tv = r (C12 - C44) / 18 (coarse3d/geom3d.py:183-227)."""
    nc = n // s
    C = torch.as_tensor(Cvox, dtype=torch.float64).reshape(nc, s, nc, s, nc, s, 3).mean(dim=(1, 3, 5)).reshape(-1, 3)
    r = s / 2.0
    T = C[:, 1] - C[:, 2]
    kax = (r * (C[:, 0] - 2 * C[:, 2] - T))[:, None].expand(-1, 3)
    kdg = (r * C[:, 2])[:, None].expand(-1, 6)
    return torch.cat([kax, kdg, (r * T / 18.0)[:, None]], 1).contiguous()


class Post:
    """Voigt bond bases of the bond list and the scatter of raw outputs to per-slot Kn / Tv, on the GPU, with the
    interface of timing/t3d/infer_fast.py:150-177: `bases(Cv)` -> per group (Kn_V, Tv_V) of every listed bond,
    `scatter(outs, base)` -> Kn, Tv (nc^3, 18) float64 with Kn = Kn_V exp(o1), Tv = Tv_V + Kn_V o2 on both directed
    slots (raw outputs of specimen 0, as the reference takes).  `BondNet.bond_params` does the same inside the heads
    kernel."""

    def __init__(self, n, s, net, device=None):
        self.n, self.s, self.nc = n, s, n // s
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        nc = self.nc
        self.bonds, self.q = [], []
        for g in net.windows(nc):
            b = g["bonds"]
            ijk = np.stack([b[:, 0] % nc, (b[:, 0] // nc) % nc, b[:, 0] // (nc * nc)], 1) + SC_OFFS[b[:, 1]]
            q = (ijk[:, 2] * nc + ijk[:, 1]) * nc + ijk[:, 0]
            self.bonds.append(torch.as_tensor(b, device=self.device))
            self.q.append(torch.as_tensor(q, device=self.device))

    def bases(self, Cv):
        cp = voigt_cell_params(Cv, self.n, self.s).to(self.device)
        out = []
        for b, q in zip(self.bonds, self.q):
            p, d = b[:, 0], b[:, 1] // 2
            out.append((0.5 * (cp[p, d] + cp[q, d]), 0.5 * (cp[p, 9] + cp[q, 9])))
        return out

    def scatter(self, outs, base):
        N = self.nc ** 3
        Kn = torch.zeros((N, 18), device=self.device, dtype=torch.float64)
        Tv = torch.zeros_like(Kn)
        for b, q, o, (kv, tv) in zip(self.bonds, self.q, outs, base):
            o = o[0].to(self.device, torch.float64)
            kn = kv * torch.exp(o[:, 0])
            tt = tv + kv * o[:, 1]
            Kn[b[:, 0], b[:, 1]] = kn
            Kn[q, b[:, 1] ^ 1] = kn
            Tv[b[:, 0], b[:, 1]] = tt
            Tv[q, b[:, 1] ^ 1] = tt
        return Kn, Tv


# ------------------------------------------------------------------------------------ specimens with internal voids
KMIN = 1e-3                       # minimum stiffness of a kept bond, relative to the mean (timing/t3d/infer3d.py:34)


def voxel_features_v(E, nu, solid, n, s):
    """Voxel features of a specimen with internal voids (timing/t3d/infer3d.py:38-50): void voxels look like the
    padding (log term and Tv term 0, void flag 1); the reference stiffness is the mean shear modulus of the solid."""
    solid = np.asarray(solid, bool)
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    mu = np.where(solid, E / (2 * (1 + nu)), 1.0)
    lam = np.where(solid, E * nu / ((1 + nu) * (1 - 2 * np.where(solid, nu, 0.0))), 1.0)
    kref = mu[solid].mean()
    pad = 2 * s
    X = np.zeros((3,) + (n + 2 * pad,) * 3, np.float32)
    X[2] = 1.0
    core = (slice(pad, pad + n),) * 3
    X[(0,) + core] = np.where(solid, np.log(mu / kref), 0.0).reshape(n, n, n)
    X[(1,) + core] = np.where(solid, (lam - mu) / (18 * mu), 0.0).reshape(n, n, n)
    X[(2,) + core] = (~solid).reshape(n, n, n)
    return X


def mask_cells(Kn, Tv, cell, nc, kmin=KMIN):
    """Coarse Kn / Tv (nc^3, 18) of a specimen with void cells (timing/t3d/infer3d.py:95-103): a slot is kept when it
    stays in the box and both of its cells are solid; other slots are zeroed, and kept Kn are raised to at least
    kmin x (mean Kn of the kept slots).  `cell` (nc^3,) bool; tensors on any device, returned on the same."""
    dev = Kn.device
    cell = torch.as_tensor(np.asarray(cell, bool), device=dev)
    ijk = torch.stack(torch.meshgrid(*(torch.arange(nc, device=dev),) * 3, indexing="ij"), -1).reshape(-1, 3).flip(1)
    off = torch.as_tensor(SC_OFFS, device=dev)
    nb = ijk[:, None, :] + off[None]
    inside = ((nb >= 0) & (nb < nc)).all(-1)
    q = ((nb[..., 2] * nc + nb[..., 1]) * nc + nb[..., 0]).clamp(0, nc ** 3 - 1)
    v = inside & cell[:, None] & cell[q]
    Kn = Kn * v
    Tv = Tv * v
    km = kmin * Kn.sum() / max(int(v.sum()), 1)
    Kn = torch.where(v, torch.maximum(Kn, km), Kn)
    return Kn, Tv
