"""Continuous-batching server for independent 2D coarse LPM solves (liblpm.so lpm_serve_*).

A request is one coarse square-lattice solve (Kn, Tv (N, 8) FP32 symmetric, free mask, prescribed displacement).
The server keeps up to `max_slots` requests resident in a device arena; one persistent kernel launch (replayed as a
hipGraph) advances every resident request by a work budget; requests that arrive meanwhile are admitted between
launches as slots and arena space free up (admission by HBM budget).  Each request's arithmetic is independent of
the others, so its result is the same alone and in any mix.

run_stream() drives a timed arrival stream (Poisson or all-at-once) and reports per-request latency (arrival to
result on the host) and throughput.  make_requests() builds the synthetic coarse-request mix: Gaussian-random-field
two-phase fine images coarsened by s x s blocks, Voigt-averaged cell stiffness, two-end bond means times a
lognormal(0, 0.2) factor per undirected bond, tension or shear.
"""
import collections
import ctypes
import time

import numpy as np
import torch

import _lpmlib as L
import geom2d
import synth

DONE, FAILED = 4, 5


class Request:
    def __init__(self, nx, ny, h, kn, tv, mask, up, load, meta=None):
        self.nx, self.ny, self.h = nx, ny, h
        self.kn, self.tv, self.mask, self.up = kn, tv, mask, up
        self.load = load
        self.meta = meta or {}
        self.u = None
        self.iters = self.outer = 0
        self.relres = float("nan")
        self.t_arrive = self.t_admit = self.t_done = None
        self.failed = False

    @property
    def N(self):
        return self.nx * self.ny


class Server:
    def __init__(self, max_slots=512, arena_bytes=None, device="cuda", budget=None, use_graph=True, threads=1024):
        L.require_gpu(torch.empty(0, device=device))
        self.dev = torch.device(device)
        self.lib = L.lib()
        if arena_bytes is None:
            arena_bytes = 4 << 30
        h = ctypes.c_void_p()
        L.check(self.lib.lpm_serve_create(int(max_slots), int(arena_bytes), int(threads), ctypes.byref(h)))
        self.h = h
        self.max_slots = int(max_slots)
        self.budget = int(budget or 64 * 20736)
        self.use_graph = bool(use_graph)
        self.stream = torch.cuda.current_stream(self.dev)
        self._st = np.zeros(self.max_slots, np.int32)
        self._it = np.zeros(self.max_slots, np.int32)
        self._out = np.zeros(self.max_slots, np.int32)
        self._rel = np.zeros(self.max_slots, np.float64)

    def close(self):
        if self.h:
            self.lib.lpm_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sp(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def footprint(self, nx, ny):
        return int(self.lib.lpm_serve_footprint(nx, ny))

    def arena_free(self):
        return int(self.lib.lpm_serve_arena_free(self.h))

    def admit(self, req, tol=1e-8, inner_tol=1e-4, maxit=200000):
        """Slot id, or None when no slot or arena space is free."""
        slot = ctypes.c_int32(-1)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        rc = self.lib.lpm_serve_admit(self.h, req.nx, req.ny, float(req.h), p(req.kn), p(req.tv), p(req.mask),
                                      p(req.up), float(tol), float(inner_tol), int(maxit), self._sp(),
                                      ctypes.byref(slot))
        if rc == L.LPM_E_FULL:
            return None
        L.check(rc)
        return int(slot.value)

    def step(self):
        L.check(self.lib.lpm_serve_step(self.h, self.budget, 1 if self.use_graph else 0, self._sp()))

    def poll(self):
        v = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        L.check(self.lib.lpm_serve_poll(self.h, v(self._st), v(self._it), v(self._out), v(self._rel), self._sp()))
        return self._st, self._it, self._out, self._rel

    def result(self, slot, N):
        u = torch.empty((N, 2), device=self.dev, dtype=torch.float64)
        L.check(self.lib.lpm_serve_result(self.h, int(slot), ctypes.c_void_p(u.data_ptr()), self._sp()))
        return u

    def release(self, slot):
        L.check(self.lib.lpm_serve_release(self.h, int(slot)))


def run_stream(server, reqs, arrivals=None, tol=1e-8, inner_tol=1e-4, maxit=200000):
    """Serve `reqs` arriving at `arrivals` seconds (None: all at t = 0) with continuous batching on a server that
    holds no other requests.  Returns a dict with wall time, solves/s and latency percentiles (ms, arrival to
    result)."""
    order = sorted(range(len(reqs)), key=lambda i: 0.0 if arrivals is None else arrivals[i])
    pending = collections.deque(order)
    waiting = collections.deque()
    running = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = 0
    while pending or waiting or running:
        now = time.perf_counter() - t0
        while pending and (arrivals is None or arrivals[pending[0]] <= now):
            i = pending.popleft()
            reqs[i].t_arrive = 0.0 if arrivals is None else float(arrivals[i])
            waiting.append(i)
        while waiting:
            slot = server.admit(reqs[waiting[0]], tol, inner_tol, maxit)
            if slot is None:
                break
            i = waiting.popleft()
            reqs[i].t_admit = time.perf_counter() - t0
            running[slot] = i
        if not running:
            if waiting:
                raise RuntimeError("request of %d x %d does not fit an empty server (arena %d B free)"
                                   % (reqs[waiting[0]].nx, reqs[waiting[0]].ny, server.arena_free()))
            if pending:
                wait = arrivals[pending[0]] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
            continue
        server.step()
        steps += 1
        st, it, outer, rel = server.poll()
        tnow = time.perf_counter() - t0
        for slot in [s for s in running if st[s] in (DONE, FAILED)]:
            i = running.pop(slot)
            r = reqs[i]
            r.iters, r.outer, r.relres = int(it[slot]), int(outer[slot]), float(rel[slot])
            r.failed = st[slot] == FAILED
            if not r.failed:
                r.u = server.result(slot, r.N)
            server.release(slot)
            r.t_done = tnow
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([(r.t_done - r.t_arrive) * 1e3 for r in reqs])
    return dict(wall_s=wall, solves_per_s=len(reqs) / wall, steps=steps, p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), max_ms=float(lat.max()),
                failed=int(sum(r.failed for r in reqs)), max_relres=float(max(r.relres for r in reqs)))


# ------------------------------------------------------------------------------------------------ synthetic mix
def coarse_request(fine, s, phi, corr, seed, device, load=None, contrast=10.0, h=1.0, sigma=0.2):
    """One coarse request over a fine x fine two-phase Gaussian-random-field image, s x s blocks."""
    rng = np.random.default_rng(seed)
    img = synth.grf_two_phase(fine, fine, corr, phi, seed).reshape(fine, fine)
    E = np.where(img > 0, contrast, 1.0)
    C = geom2d.iso_C(E, np.full(E.shape, 0.2))                                # (fine, fine, 3)
    nc = fine // s
    Cc = C[:nc * s, :nc * s].reshape(nc, s, nc, s, 3).mean(axis=(1, 3)).reshape(-1, 3)  # Voigt mean per cell
    lat = geom2d.SquareGeom(nc, nc, s * h)
    Kn, Tv = geom2d.bond_params_two_end_mean(lat, geom2d.square_params_from_C(Cc))
    fac = np.exp(sigma * rng.standard_normal((lat.N, 8)))
    q = np.where(lat.valid, lat.nbr, 0)
    for k in range(1, 8, 2):                                                  # one factor per undirected bond
        fac[:, k] = np.where(lat.valid[:, k], fac[q[:, k], k - 1], 1.0)
    Kn, Tv = Kn * fac, Tv * fac
    load = load or ("tb_ty" if rng.random() < 0.5 else "tb_sx")
    fixed, up, _ = geom2d.prescribed(load, lat.x, lat.ij, nc, nc, s, h, 1)
    mask = np.repeat(~fixed[:, None], 2, 1).astype(np.uint8)
    t32 = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device, dtype=torch.float32)
    return Request(nc, nc, s * h, t32(Kn), t32(Tv), torch.as_tensor(mask, device=device),
                   torch.as_tensor(up, device=device, dtype=torch.float64), load,
                   dict(fine=fine, s=s, phi=phi, seed=seed))


def make_requests(n, seed, device, fine_range=(96, 288), s_choices=(2, 3, 4), phi_range=(0.15, 0.40), corr=8.0):
    """n requests: fine size uniform in fine_range (a multiple of 12 so that every s divides it), s uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        fine = int(rng.integers(fine_range[0] // 12, fine_range[1] // 12 + 1)) * 12
        s = int(rng.choice(s_choices))
        phi = float(rng.uniform(*phi_range))
        out.append(coarse_request(fine, s, phi, corr, int(rng.integers(1 << 31)), device))
    return out


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


# ------------------------------------------------------------------------------------------------ baseline
def plain_cg(req, tol=1e-8, maxit=200000):
    """The plain per-solve GPU CG the server is compared with: one request at a time, FP64, unpreconditioned
    (as timing/t2d/cg2plain.py:11-49), the HIP K u kernel for every product, alpha and beta formed on the host
    (two device-to-host reads per iteration), no batching, no graph, no persistence.  Returns (u, iterations)."""
    import lpm_ops as ops
    dims, h = (req.nx, req.ny), float(req.h)
    kn, tv = req.kn.double(), req.tv.double()
    m = req.mask.double()
    b = -ops.apply(dims, h, req.up, kn, tv) * m
    bn = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b)
    if bn == 0.0:
        return req.up.clone(), 0
    r = b.clone()
    p = r.clone()
    rr = float((r * r).sum())
    it = 0
    while it < maxit:
        ap = ops.apply(dims, h, p, kn, tv) * m
        alpha = rr / float((p * ap).sum())
        x += alpha * p
        r -= alpha * ap
        rn = float((r * r).sum())
        it += 1
        if rn ** 0.5 <= tol * bn:
            break
        p = r + (rn / rr) * p
        rr = rn
    return req.up + x, it
