"""Continuous-batching server (lpm_serve_*) on the GPU.

Every request's u is compared with the oracle's SuperLU solve of the same FP32 coefficients (promoted to FP64): the
server stops at a true FP64 relative residual of 1e-8, and the bar is a relative displacement error of 1e-6.  The
same request solved alone, in a mix with queueing and slot reuse, with launches split at a small work budget, with
and without the hipGraph replay, gives bit-identical results (each slot's arithmetic is independent of the rest)."""
import numpy as np
import pytest
import torch

import lpm_ref as R
import serve

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small_mix(dev, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        s = int(rng.choice([2, 3, 4]))
        fine = int(rng.integers(2, 8)) * 12
        reqs.append(serve.coarse_request(fine, s, float(rng.uniform(0.15, 0.4)), 4.0, 100 + i, dev))
    return reqs


def oracle_u(r):
    lat = R.Box((r.nx, r.ny), r.h)
    Kn, Tv = r.kn.double().cpu().numpy(), r.tv.double().cpu().numpy()
    free = r.mask.cpu().numpy()[:, 0].astype(bool)
    return R.solve_direct(R.stiffness(lat, Kn, Tv), np.repeat(~free, 2), r.up.cpu().numpy()).reshape(-1, 2)


def solve_all(dev, reqs, **kw):
    srv = serve.Server(device=dev, **{k: v for k, v in kw.items() if k in ("max_slots", "arena_bytes", "budget",
                                                                           "use_graph", "threads")})
    arr = kw.get("arrivals")
    stats = serve.run_stream(srv, reqs, arr)
    srv.close()
    return stats, [r.u.clone() for r in reqs]


def test_mix_matches_oracle_and_alone(dev):
    reqs = small_mix(dev, 12, 1)
    print("sizes", [(r.nx, r.ny) for r in reqs], flush=True)
    stats, mixed = solve_all(dev, reqs, max_slots=5, budget=3000, arrivals=np.linspace(0.0, 0.02, len(reqs)))
    assert stats["failed"] == 0 and stats["max_relres"] <= 1e-8
    for r, u in zip(reqs, mixed):
        assert rel(u, oracle_u(r)) < 1e-6
    for i, r in enumerate(reqs):
        _, alone = solve_all(dev, [r], max_slots=1, budget=10 ** 9, use_graph=False)
        assert torch.equal(alone[0], mixed[i]), "request %d differs alone and in the mix" % i


def test_budget_and_graph_do_not_change_results(dev):
    reqs = small_mix(dev, 6, 2)
    _, a = solve_all(dev, reqs, max_slots=8, budget=500, use_graph=True)
    _, b = solve_all(dev, reqs, max_slots=8, budget=10 ** 8, use_graph=False)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_admission_by_arena_budget(dev):
    reqs = small_mix(dev, 3, 3)
    probe = serve.Server(device=dev, max_slots=1, arena_bytes=1 << 20)
    need = max(probe.footprint(r.nx, r.ny) for r in reqs)
    probe.close()
    srv = serve.Server(device=dev, max_slots=8, arena_bytes=need + 256)   # room for one request at a time
    slot = srv.admit(reqs[0])
    assert slot is not None
    assert srv.admit(reqs[1]) is None                                      # full: nothing changes
    for _ in range(100000):
        srv.step()
        st, _, _, _ = srv.poll()
        if st[slot] in (serve.DONE, serve.FAILED):
            break
    assert st[slot] == serve.DONE
    u0 = srv.result(slot, reqs[0].N)
    srv.release(slot)
    assert srv.arena_free() >= need
    assert rel(u0, oracle_u(reqs[0])) < 1e-6
    stats = serve.run_stream(srv, reqs[1:])                                # the rest queue one at a time
    assert stats["failed"] == 0
    for r in reqs[1:]:
        assert rel(r.u, oracle_u(r)) < 1e-6
    srv.close()


def test_zero_load_request(dev):
    r = small_mix(dev, 1, 4)[0]
    r.up = torch.zeros_like(r.up)
    stats, us = solve_all(dev, [r], max_slots=2)
    assert stats["failed"] == 0 and float(us[0].abs().max()) == 0.0


@pytest.mark.parametrize("threads", [256, 1024])
@pytest.mark.parametrize("fused", [1, 0], ids=["fused", "two_pass"])
def test_banded_and_two_pass_products(dev, threads, fused):
    """The inner iterations' product as one banded pass through LDS (bands of 3 threads / nx - 2 rows: 72-wide
    requests take 9 bands at 256 threads and 2 at 1024, 32-wide ones 2 and 1) and as two passes through device
    memory: each against SuperLU, alone and in a mix."""
    import _lpmlib as L
    print("serve threads", threads, "fused", fused, flush=True)
    L.lib().lpm_serve_set_fused(fused)
    try:
        reqs = [serve.coarse_request(144, 2, 0.3, 4.0, 500 + i, dev) for i in range(2)] + \
               [serve.coarse_request(96, 3, 0.25, 4.0, 600 + i, dev) for i in range(2)]
        stats, mixed = solve_all(dev, reqs, max_slots=3, budget=20000, threads=threads)
        assert stats["failed"] == 0 and stats["max_relres"] <= 1e-8
        for r, u in zip(reqs, mixed):
            assert rel(u, oracle_u(r)) < 1e-6
        _, alone = solve_all(dev, reqs[:1], max_slots=1, budget=10 ** 9, use_graph=False, threads=threads)
        assert torch.equal(alone[0], mixed[0])
    finally:
        L.lib().lpm_serve_set_fused(1)
