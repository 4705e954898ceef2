"""Aggregation AMG preconditioner and AMG-PCG of the 3D SC LPM on the GPU, with the interface of timing/t3d/amg3d.py.

AMGGeom(lat, free, agg=3, rule="merge", coarse_max=3000, max_levels=8) builds the geometry-only hierarchy
(amg3d.py:435-481): the free particles (a box) in chunks of `agg` per axis with the remainder merged, the
rigid-body-mode tentative prolongator of every level (gmg.py:70-162) and the node grids.  Level 0 keeps at most 8
shape templates instead of a block per particle: Gram-Schmidt removes the translation part of the rotations about
the lattice centroid, so the block depends only on the aggregate's shape; the coarse near-null space of an aggregate
is the template's R factor shifted by its centroid (R_I = R_shape T(c_I - c_lattice)), the same R the reference's
per-aggregate Gram-Schmidt produces.

AMG3D(geom, Kn, Tv, nu=1, omega=1.32, power_its=10, Minv0=None, nu_c=None, gamma=1, w_from=1, rho0=None) forms
the numeric hierarchy on the GPU (amg3d.py:508-557): particle block Jacobi, rho(D^-1 A) per level by power
iteration from default_rng(0) normal start vectors (amg3d.py:493-505), the Galerkin stencils of every level by
probing the batched K u / stencil products with 27 aggregate colours x 6 modes, the inverses of the diagonal
blocks and the coarsest dense inverse.  vcycle(b) and pcg(b, tol, maxit) run as one C ABI call each (the W-cycle
recursion and the PCG loop are in liblpm.so).  The reference's parameters for its timing runs are AMG_CFG
(timing/t3d/problems.py:19): agg=3, coarse_max=1000, nu=1, gamma=2, power_its=5.
"""
import ctypes
import math

import numpy as np
import torch

import _lpmlib as L
import lpm_ops as ops

MAXLEV = 12
ST_BLOCKS = 14          # stored stencil blocks per node: the diagonal and the 13 forward offsets (d = 13 ... 26)
SCRATCH = 4096
PROLONG_ADD, RESTRICT, PROBE, GATHER = 0, 1, 2, 3
AMG_CFG = dict(agg=3, coarse_max=1000, nu=1, gamma=2, power_its=5)


class AmgStruct(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_int32 * 3), ("L", ctypes.c_int32 * 3), ("g0", ctypes.c_int32 * 3),
                ("agg", ctypes.c_int32), ("tmpl", ctypes.c_void_p), ("toff", ctypes.c_int64 * 8),
                ("minv0", ctypes.c_void_p), ("w0", ctypes.c_double), ("nlev", ctypes.c_int32),
                ("nu", ctypes.c_int32), ("nu_c", ctypes.c_int32), ("gamma", ctypes.c_int32),
                ("w_from", ctypes.c_int32), ("pad_", ctypes.c_int32), ("g", (ctypes.c_int32 * 3) * MAXLEV),
                ("A", ctypes.c_void_p * MAXLEV), ("Q", ctypes.c_void_p * MAXLEV),
                ("Minv", ctypes.c_void_p * MAXLEV), ("w", ctypes.c_double * MAXLEV), ("dinv", ctypes.c_void_p)]


_SIGS = {
    "lpm_residual_workspace": (ctypes.c_size_t, [L.GP, L.I32, L.I32]),
    "lpm_residual": (L.I32, [L.GP, L.I32, L.I32, L.I32, L.I32, L.P, L.Bcast, L.P, L.P, L.Bcast, L.P, L.Bcast, L.P,
                             L.Bcast, L.P, L.P, L.P]),
    "lpm_amg_level0": (L.I32, [L.GP, ctypes.POINTER(AmgStruct), L.I32, L.I32, L.I32, L.P, L.P, L.P]),
    "lpm_amg_levelq": (L.I32, [ctypes.POINTER(AmgStruct), L.I32, L.I32, L.I32, L.I32, L.P, L.P, L.P]),
    "lpm_amg_st_mv": (L.I32, [L.I32, L.I32, L.I32, L.I32, L.I32, L.P, L.P, L.P, L.P, L.P]),
    "lpm_amg_set_stencil_kernel": (L.I32, [L.I32]),
    "lpm_amg_inv_diag": (L.I32, [L.I32, L.I32, L.P, L.P, L.P, L.P]),
    "lpm_amg_bj": (L.I32, [ctypes.c_int64, L.I32, L.I32, L.P, L.P, ctypes.c_double, L.I32, L.P, L.P]),
    "lpm_amg_coarsest": (L.I32, [L.I32, L.I32, L.I32, L.I32, L.P, L.P, L.P, L.P]),
    "lpm_amg_dot": (L.I32, [ctypes.c_int64, L.I32, L.P, L.P, ctypes.POINTER(ctypes.c_double), L.P, L.P]),
    "lpm_amg_workspace": (ctypes.c_size_t, [L.GP, ctypes.POINTER(AmgStruct), L.I32]),
    "lpm_amg_vcycle": (L.I32, [L.GP, ctypes.POINTER(AmgStruct), L.I32, L.I32, L.P, L.P, L.P, L.P, L.P, L.P, L.P]),
    "lpm_amg_pcg": (L.I32, [L.GP, ctypes.POINTER(AmgStruct), L.I32, L.I32, L.P, L.P, L.P, L.P, L.P, ctypes.c_double,
                            L.I32, L.P, ctypes.POINTER(L.I32), ctypes.POINTER(ctypes.c_double), L.P]),
    "lpm_amg_refined_workspace": (ctypes.c_size_t, [L.GP, ctypes.POINTER(AmgStruct)]),
    "lpm_amg_pcg_refined": (L.I32, [L.GP, ctypes.POINTER(AmgStruct), L.I32, L.P, L.P, L.I32, L.P, L.P, L.P, L.P, L.P,
                                    ctypes.c_double, ctypes.c_double, L.I32, L.I32, L.P, ctypes.POINTER(L.I32),
                                    ctypes.POINTER(L.I32), ctypes.POINTER(ctypes.c_double), L.P]),
}
L.SIGS.update(_SIGS)


def _lib():
    lib = L.lib()
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


ST_AUTO, ST_ROWS, ST_NODE, ST_TWO_PASS = 0, 1, 2, 3


def set_stencil_kernel(mode):
    """Kernel of the stencil products y = A x (process-wide): ST_AUTO (by level size), ST_ROWS (one thread per row
    and node), ST_NODE (one thread per node) or ST_TWO_PASS (every stored block read once; inside the cycle)."""
    L.check(_lib().lpm_amg_set_stencil_kernel(int(mode)))


def chunk_index(L_, a, rule="merge"):
    """chunk id of positions 0..L-1 along one axis and the number of chunks (amg3d.py:42-48)."""
    if rule != "merge":
        raise NotImplementedError("only the merged-remainder aggregation (rule='merge') runs on the GPU")
    nch = max(1, L_ // a)
    return np.minimum(np.arange(L_) // a, nch - 1), nch


def _rbm_rows(X):
    """rigid-body modes (3n, 6) of points X (n, 3) about their centroid, rows interleaved by DOF (gmg.py:145-162)."""
    c = X - X.mean(0)
    n = X.shape[0]
    B = np.zeros((3 * n, 6))
    for d in range(3):
        B[d::3, d] = 1.0
    B[0::3, 3], B[1::3, 3] = -c[:, 1], c[:, 0]
    B[1::3, 4], B[2::3, 4] = -c[:, 2], c[:, 1]
    B[0::3, 5], B[2::3, 5] = c[:, 2], -c[:, 0]
    return B


def _mgs(B, tol=1e-8):
    """Batched modified Gram-Schmidt with rank detection (gmg.py:101-130): B (G, rows, m) -> Q, R."""
    G, n, m = B.shape
    Q = np.zeros_like(B)
    R = np.zeros((G, m, m))
    keep = np.zeros((G, m), bool)
    scale = np.sqrt((B ** 2).sum(1)).max(1)
    for c in range(m):
        v = B[:, :, c].copy()
        for d in range(c):
            pr = np.where(keep[:, d], np.einsum("gi,gi->g", Q[:, :, d], v), 0.0)
            R[:, d, c] = pr
            v -= pr[:, None] * Q[:, :, d]
        nrm = np.sqrt((v * v).sum(1))
        ok = (nrm > tol * scale) & (nrm > 0.0)
        keep[:, c] = ok
        R[:, c, c] = np.where(ok, nrm, 0.0)
        Q[:, :, c] = np.where(ok[:, None], v / np.where(ok, nrm, 1.0)[:, None], 0.0)
    if not keep.all():
        raise FloatingPointError("rank-deficient aggregate")
    return Q, R


def _box_points(shape, h):
    sx, sy, sz = shape
    k, j, i = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float64) * h


def _shift(D):
    """T(D): the rotations about the lattice centroid in the basis of rotations about a point D away (6x6)."""
    T = np.tile(np.eye(6), D.shape[:-1] + (1, 1))
    T[..., 0, 3], T[..., 1, 3] = -D[..., 1], D[..., 0]
    T[..., 1, 4], T[..., 2, 4] = -D[..., 2], D[..., 1]
    T[..., 0, 5], T[..., 2, 5] = D[..., 2], -D[..., 0]
    return T


class AMGGeom:
    """Aggregation and tentative prolongators of all levels for one lattice and one free box (amg3d.py:435-481)."""

    def __init__(self, lat, free, agg=3, rule="merge", coarse_max=3000, max_levels=8, device=None):
        if not getattr(lat, "full", True):
            raise NotImplementedError("the GPU AMG takes full box lattices")
        self.lat = lat
        self.dims = (int(lat.nx), int(lat.ny), int(lat.nz))
        self.h = float(getattr(lat, "h", 1.0))
        self.device = torch.device(device) if device is not None else getattr(lat, "device", torch.device("cuda"))
        free = free.detach().cpu().numpy() if torch.is_tensor(free) else np.asarray(free)
        free = free.astype(bool).reshape(-1)
        nx, ny, nz = self.dims
        self.free = free
        self.agg, self.rule, self.coarse_max = int(agg), rule, int(coarse_max)
        fb = free.reshape(nz, ny, nx)
        idx = np.nonzero(fb)
        if len(idx[0]) == 0:
            raise ValueError("no free particles")
        lo = np.array([idx[2].min(), idx[1].min(), idx[0].min()])
        hi = np.array([idx[2].max(), idx[1].max(), idx[0].max()])
        Lx = hi - lo + 1
        if int(np.prod(Lx)) != int(free.sum()):
            raise ValueError("free set must be a box")
        self.lo, self.L = lo, Lx
        ch = [chunk_index(int(Lx[d]), self.agg, rule) for d in range(3)]
        g = tuple(int(c[1]) for c in ch)
        self.g0 = g
        # level-0 shape templates (s = bit a set for the last chunk along axis a) and R factors
        a = self.agg
        lens = [(a if g[d] > 1 else int(Lx[d]), int(Lx[d]) - a * (g[d] - 1)) for d in range(3)]
        tm, self.toff, Rs, shapes = [], [0] * 8, [], []
        off = 0
        for s in range(8):
            shape = tuple(lens[d][(s >> d) & 1] for d in range(3))
            Q, R = _mgs(_rbm_rows(_box_points(shape, self.h))[None])
            tm.append(Q[0].reshape(-1))
            self.toff[s] = off
            off += Q[0].size
            Rs.append(R[0])
            shapes.append(shape)
        self.tmpl64 = np.concatenate(tm)
        # coarse near-null space of the aggregates: R_shape T(centroid offset)
        gi = np.stack(np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij"), -1).reshape(-1, 3)
        gi = gi[:, ::-1]                                   # (x, y, z) chunk coordinates, aggregate id order
        last = (gi == np.array(g) - 1)
        sid = last[:, 0] * 1 + last[:, 1] * 2 + last[:, 2] * 4
        size = np.where(last, np.array([lens[d][1] for d in range(3)]), np.array([lens[d][0] for d in range(3)]))
        cent = (lo + a * gi + 0.5 * (size - 1)) * self.h
        cg = 0.5 * (np.array(self.dims) - 1) * self.h
        Rstack = np.stack(Rs)[sid]
        self.Bc0 = np.einsum("gij,gjk->gik", Rstack, _shift(cent - cg))
        # levels >= 1: node grids, per-node blocks Q and the next near-null space
        self.levels = []
        Bn = self.Bc0
        nagg = int(np.prod(g))
        while 6 * nagg > coarse_max and len(self.levels) < max_levels:
            chn = [chunk_index(g[d], a, rule) for d in range(3)]
            gc = tuple(int(c[1]) for c in chn)
            ni = np.stack(np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij"),
                          -1).reshape(-1, 3)[:, ::-1]
            kc = np.stack([chn[d][0][ni[:, d]] for d in range(3)], 1)
            a_of = (kc[:, 2] * gc[1] + kc[:, 1]) * gc[0] + kc[:, 0]
            nc_ = int(np.prod(gc))
            Qn = np.zeros((nagg, 6, 6))
            Bc2 = np.zeros((nc_, 6, 6))
            order = np.argsort(a_of, kind="stable")
            counts = np.bincount(a_of, minlength=nc_)
            starts = np.concatenate([[0], np.cumsum(counts)])
            for cnt in np.unique(counts):
                ks = np.nonzero(counts == cnt)[0]
                mem = np.stack([order[starts[k]:starts[k] + cnt] for k in ks])      # (G, cnt) node ids, ascending
                Bst = Bn[mem].reshape(len(ks), cnt * 6, 6)
                Qg, Rg = _mgs(Bst)
                Qn[mem.reshape(-1)] = Qg.reshape(len(ks) * cnt, 6, 6)
                Bc2[ks] = Rg
            self.levels.append(dict(g=g, gc=gc, Q64=Qn))
            g, nagg, Bn = gc, nc_, Bc2
        self.g_coarsest = g
        self._dev_cache = {}

    def grids(self):
        return [self.g0] + [Lg["gc"] for Lg in self.levels]

    def describe(self):
        s = ["L0 N=%d free=%d" % (int(np.prod(self.dims)), int(self.free.sum())),
             "L1 %s n=%d" % (self.g0, 6 * int(np.prod(self.g0)))]
        for k, Lg in enumerate(self.levels):
            s.append("L%d %s n=%d" % (k + 2, Lg["gc"], 6 * int(np.prod(Lg["gc"]))))
        return " | ".join(s)

    def dev(self, dtype):
        """device copies (templates, per-node blocks, DOF mask) in dtype"""
        if dtype not in self._dev_cache:
            t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=self.device, dtype=dtype)
            mask = torch.as_tensor(np.repeat(self.free, 3), device=self.device, dtype=torch.uint8)
            self._dev_cache[dtype] = dict(tmpl=t(self.tmpl64), Q=[t(Lg["Q64"]) for Lg in self.levels], mask=mask)
        return self._dev_cache[dtype]


def _power_rho(mv, bj, start, its):
    """rho(M A) by power iteration (amg3d.py:493-505): x = M start; repeat x /= |x|, y = M A x, lam = |y|."""
    x = bj(start, 1.0)
    lam = 1.0
    for _ in range(its):
        nrm = math.sqrt(_dot(x, x))
        y = bj(mv(x), 1.0 / nrm)
        lam = math.sqrt(_dot(y, y))
        x = y
    return lam


_SCR = {}


def _scratch(device):
    key = str(device)
    if key not in _SCR:
        _SCR[key] = torch.empty(SCRATCH // 8, device=device, dtype=torch.float64)
    return _SCR[key]


def _dot(x, y):
    out = ctypes.c_double()
    L.check(_lib().lpm_amg_dot(x.numel(), L.dtype_code(x.dtype), _p(x), _p(y), ctypes.byref(out),
                                _p(_scratch(x.device)), L.stream_of(x)))
    return out.value


class AMG3D:
    """Numeric hierarchy for one specimen (Kn, Tv) on an AMGGeom; vcycle() is the preconditioner, pcg() the solver
    (amg3d.py:508-672).  Kn, Tv: directed (N, 18) FP64 arrays or tensors (symmetric pairs).  dtype: arithmetic of
    the hierarchy, the cycle and the PCG (FP64 as the reference).  starts: power-iteration start vectors per level
    as a dict {level: array} (default: default_rng(0).standard_normal, as the reference), or "device" for normal vectors drawn on the GPU
    (seed 0), which avoids the host RNG on large lattices.  Kn, Tv may also be packed forward planes (9, P)."""

    def __init__(self, geom, Kn, Tv, nu=1, omega=1.32, power_its=10, Minv0=None, nu_c=None, gamma=1,
                 fine_additive=False, w_from=1, rho0=None, dtype=torch.float64, starts=None, rho=None):
        if fine_additive:
            raise NotImplementedError("the additive fine-level correction is not part of the GPU cycle")
        self.geom, self.nu, self.omega = geom, int(nu), float(omega)
        self.nu_c = self.nu if nu_c is None else int(nu_c)
        self.gamma, self.w_from = int(gamma), int(w_from)
        self.dtype = dtype
        dev = geom.device
        dims = geom.dims
        N = int(np.prod(dims))
        self.N = N
        lib = _lib()
        st = L.stream_of(torch.empty(0, device=dev))
        D = geom.dev(dtype)
        self.mask = D["mask"]
        kn = torch.as_tensor(Kn, device=dev, dtype=dtype)
        tv = torch.as_tensor(Tv, device=dev, dtype=dtype)
        if tuple(kn.shape) == (9, ops.packed_len(dims)):
            self.kn, self.tv = kn.contiguous(), tv.contiguous()      # packed forward planes already
        else:
            self.kn, self.tv, _ = ops.pack(dims, kn.reshape(N, 18), tv.reshape(N, 18))
        self.grid = L.grid_of(dims, geom.h)
        dc = L.dtype_code(dtype)
        free = torch.as_tensor(geom.free, device=dev)
        if Minv0 is None:
            self.minv0 = ops.block_jacobi(dims, geom.h, self.kn, self.tv, free=free, layout=L.PACKED)
        else:
            m = torch.as_tensor(Minv0, device=dev, dtype=dtype)
            self.minv0 = ops.sym_from_full(m) if m.shape[-1] == 3 else m
            self.minv0 = (self.minv0 * free[:, None].to(dtype)).contiguous()
        s = AmgStruct()
        for k in range(3):
            s.lo[k], s.L[k], s.g0[k] = int(geom.lo[k]), int(geom.L[k]), int(geom.g0[k])
        s.agg = geom.agg
        s.tmpl = D["tmpl"].data_ptr()
        for k in range(8):
            s.toff[k] = int(geom.toff[k])
        s.minv0 = self.minv0.data_ptr()
        s.nu, s.nu_c, s.gamma, s.w_from = self.nu, self.nu_c, self.gamma, self.w_from
        grids = geom.grids()
        s.nlev = len(grids)
        for k, g in enumerate(grids):
            for d in range(3):
                s.g[k][d] = int(g[d])
        for k, Q in enumerate(D["Q"]):
            s.Q[k] = Q.data_ptr()
        self.s = s
        self.rwork = torch.empty(max(1, lib.lpm_residual_workspace(ctypes.byref(self.grid), 1, dc)) // 8 + 1,
                                 device=dev, dtype=torch.float64)
        device_starts = starts == "device"
        rng_starts = {} if (starts is None or device_starts) else starts
        gen = torch.Generator(device=dev).manual_seed(0) if device_starts else None

        def start_vec(k, shape):
            if k in rng_starts:
                return torch.as_tensor(rng_starts[k], device=dev, dtype=dtype).reshape(shape)
            if device_starts:
                return torch.randn(shape, generator=gen, device=dev, dtype=dtype)
            return torch.as_tensor(np.random.default_rng(0).standard_normal(shape), device=dev, dtype=dtype)

        # rho of the finest level (amg3d.py:528-529)
        if rho0 is not None:
            self.rho0 = float(rho0)
        elif rho is not None:
            self.rho0 = float(rho[0])
        else:
            x0 = start_vec(0, (N, 3))
            self.rho0 = _power_rho(self._kff, lambda v, w: self._bj3(v, w), x0, int(power_its))
        s.w0 = self.omega / self.rho0
        # Galerkin operator of level 0 by probing (replaces _galerkin0, amg3d.py:151-196)
        nagg = int(np.prod(geom.g0))
        A = torch.zeros((ST_BLOCKS * 36, nagg), device=dev, dtype=dtype)
        v = torch.empty((6, N, 3), device=dev, dtype=dtype)
        bc0 = L.Bcast(0, 0)
        for col in range(27):
            L.check(lib.lpm_amg_level0(ctypes.byref(self.grid), ctypes.byref(s), dc, PROBE, col, None, _p(v), st))
            kv = ops.apply(dims, geom.h, v, self.kn, self.tv, layout=L.PACKED)
            L.check(lib.lpm_amg_level0(ctypes.byref(self.grid), ctypes.byref(s), dc, GATHER, col, _p(kv), _p(A), st))
        self.A = [A]
        self.Minv, self.rho = [], []
        for k, Lg in enumerate(geom.levels):
            g = Lg["g"]
            n = int(np.prod(g))
            Mi = torch.empty((36, n), device=dev, dtype=dtype)
            L.check(lib.lpm_amg_inv_diag(n, dc, _p(A), _p(Mi), _p(_scratch(dev)), st))
            self.Minv.append(Mi)
            s.A[k] = A.data_ptr()
            s.Minv[k] = Mi.data_ptr()
            if rho is not None:
                r_k = float(rho[k + 1])
            else:
                xk = start_vec(k + 1, (n, 6))
                r_k = _power_rho(lambda x, A=A, g=g: self._st(A, g, x), lambda x, w, Mi=Mi: self._bj6(Mi, x, w),
                                 xk, int(power_its))
            self.rho.append(r_k)
            s.w[k] = self.omega / r_k
            gc = Lg["gc"]
            nc_ = int(np.prod(gc))
            Ac = torch.zeros((ST_BLOCKS * 36, nc_), device=dev, dtype=dtype)
            vq = torch.empty((6, n, 6), device=dev, dtype=dtype)
            wq = torch.empty_like(vq)
            for col in range(27):
                L.check(lib.lpm_amg_levelq(ctypes.byref(s), k, dc, PROBE, col, None, _p(vq), st))
                L.check(lib.lpm_amg_st_mv(g[0], g[1], g[2], 6, dc, _p(A), _p(vq), None, _p(wq), st))
                L.check(lib.lpm_amg_levelq(ctypes.byref(s), k, dc, GATHER, col, _p(wq), _p(Ac), st))
            self.A.append(Ac)
            A = Ac
        s.A[len(grids) - 1] = A.data_ptr()
        gl = grids[-1]
        m = 6 * int(np.prod(gl))
        self.dinv = torch.empty((m, m), device=dev, dtype=dtype)
        L.check(lib.lpm_amg_coarsest(gl[0], gl[1], gl[2], dc, _p(A), _p(self.dinv), _p(_scratch(dev)), st))
        s.dinv = self.dinv.data_ptr()
        ws = lib.lpm_amg_workspace(ctypes.byref(self.grid), ctypes.byref(s), dc)
        if ws == 0:
            raise RuntimeError("lpm_amg_workspace: %s" % lib.lpm_last_error().decode())
        self.work = torch.empty(ws // 8 + 1, device=dev, dtype=torch.float64)

    # ------------------------------------------------------------------------------------------- operators
    def _kff(self, x):
        """-K_ff x (the residual of a zero right-hand side; the sign drops out of the power iteration)"""
        out = torch.empty_like(x)
        zero = torch.zeros_like(x)
        n3 = self.N * 3
        L.check(_lib().lpm_residual(ctypes.byref(self.grid), 1, 1, L.dtype_code(self.dtype), L.PACKED, _p(x),
                                    L.Bcast(n3, n3), _p(self.kn), _p(self.tv), L.Bcast(0, 0), _p(self.mask),
                                    L.Bcast(n3, n3), _p(zero), L.Bcast(n3, n3), _p(out), _p(self.rwork),
                                    L.stream_of(x)))
        return out

    def apply0(self, x):
        """K_ff x (fixed rows zero)."""
        return -self._kff(x)

    def _bj3(self, b, w):
        out = torch.empty_like(b)
        L.check(_lib().lpm_amg_bj(self.N, 3, L.dtype_code(self.dtype), _p(self.minv0), _p(b), float(w), 0, _p(out),
                                  L.stream_of(b)))
        return out

    def _bj6(self, Mi, b, w):
        out = torch.empty_like(b)
        L.check(_lib().lpm_amg_bj(b.shape[0], 6, L.dtype_code(self.dtype), _p(Mi), _p(b), float(w), 0, _p(out),
                                  L.stream_of(b)))
        return out

    def _st(self, A, g, x):
        y = torch.empty_like(x)
        L.check(_lib().lpm_amg_st_mv(g[0], g[1], g[2], 1, L.dtype_code(self.dtype), _p(A), _p(x), None, _p(y),
                                     L.stream_of(x)))
        return y

    def stencil(self, k):
        """Galerkin stencil of level k + 1 as (n, 27, 6, 6) (the reference's layout), the blocks d < 13 expanded
        from the stored transposes."""
        g = self.geom.grids()[k]
        n = int(np.prod(g))
        S = self.A[k].reshape(ST_BLOCKS, 6, 6, n).permute(3, 0, 1, 2).cpu()
        full = torch.zeros((n, 27, 6, 6), dtype=S.dtype)
        full[:, 13:] = S
        ij = np.stack(np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij"), -1)
        ij = ij.reshape(-1, 3)[:, ::-1]
        for d in range(13):
            off = np.array([d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1])
            j = ij + off
            ok = np.all((j >= 0) & (j < np.array(g)), 1)
            J = (j[:, 2] * g[1] + j[:, 1]) * g[0] + j[:, 0]
            full[torch.as_tensor(np.nonzero(ok)[0]), d] = S[torch.as_tensor(J[ok]), 13 - d].transpose(-1, -2)
        return full

    def vcycle(self, b):
        """z = B b for b (N, 3) zero on fixed particles (amg3d.py:620-646)."""
        b = torch.as_tensor(b, device=self.geom.device, dtype=self.dtype).reshape(self.N, 3).contiguous()
        z = torch.empty_like(b)
        L.check(_lib().lpm_amg_vcycle(ctypes.byref(self.grid), ctypes.byref(self.s), L.dtype_code(self.dtype),
                                      L.PACKED, _p(self.kn), _p(self.tv), _p(self.mask), _p(b), _p(z),
                                      _p(self.work), L.stream_of(b)))
        return z

    def pcg(self, b, tol, maxit=2000):
        """K_ff x = b (b (N, 3) zero on fixed); ||r|| <= tol ||b|| from x = 0 (amg3d.py:649-672).  Returns x, iters,
        relres."""
        b = torch.as_tensor(b, device=self.geom.device, dtype=self.dtype).reshape(self.N, 3).contiguous()
        x = torch.empty_like(b)
        it = L.I32()
        rr = ctypes.c_double()
        L.check(_lib().lpm_amg_pcg(ctypes.byref(self.grid), ctypes.byref(self.s), L.dtype_code(self.dtype), L.PACKED,
                                   _p(self.kn), _p(self.tv), _p(self.mask), _p(b), _p(x), float(tol), int(maxit),
                                   _p(self.work), ctypes.byref(it), ctypes.byref(rr), L.stream_of(b)))
        return x, int(it.value), float(rr.value)

    def pcg_refined(self, b, Kn64, Tv64, tol=1e-8, inner_tol=1e-4, max_outer=20, maxit=5000):
        """FP32 AMG-PCG inner solves (this hierarchy, dtype float32) inside an FP64 iterative refinement: the true
        residual r = mask (b - K x) with the FP64 operator (Kn64, Tv64: packed (9, P) or directed (N, 18)) and the
        update x += d are FP64.  Returns x (N, 3) FP64, total inner iterations, refinement steps, true relres."""
        if self.dtype != torch.float32:
            raise ValueError("pcg_refined runs the FP32 hierarchy inside the FP64 refinement (build with "
                             "dtype=torch.float32)")
        dev, dims, N = self.geom.device, self.geom.dims, self.N
        k64 = torch.as_tensor(Kn64, device=dev, dtype=torch.float64)
        t64 = torch.as_tensor(Tv64, device=dev, dtype=torch.float64)
        if tuple(k64.shape) != (9, ops.packed_len(dims)):
            k64, t64, _ = ops.pack(dims, k64.reshape(N, 18), t64.reshape(N, 18))
        b = torch.as_tensor(b, device=dev, dtype=torch.float64).reshape(N, 3).contiguous()
        lib = _lib()
        ws = lib.lpm_amg_refined_workspace(ctypes.byref(self.grid), ctypes.byref(self.s))
        if ws == 0:
            raise RuntimeError("lpm_amg_refined_workspace: %s" % lib.lpm_last_error().decode())
        if getattr(self, "_rwork", None) is None or self._rwork.numel() * 8 < ws:
            self._rwork = torch.empty(ws // 8 + 1, device=dev, dtype=torch.float64)
        x = torch.empty_like(b)
        it, nout, rr = L.I32(), L.I32(), ctypes.c_double()
        L.check(lib.lpm_amg_pcg_refined(ctypes.byref(self.grid), ctypes.byref(self.s), L.PACKED, _p(k64), _p(t64),
                                        L.PACKED, _p(self.kn), _p(self.tv), _p(self.mask), _p(b), _p(x), float(tol),
                                        float(inner_tol), int(max_outer), int(maxit), _p(self._rwork), ctypes.byref(it),
                                        ctypes.byref(nout), ctypes.byref(rr), L.stream_of(b)))
        return x, int(it.value), int(nout.value), float(rr.value)
