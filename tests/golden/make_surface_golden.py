"""Writes tests/golden/{surface3d_n10, coarse3d_nc5, coarse2d_8x6, voids3d_n8}.npz: input / output vectors of the
reference's torch modules (coarse3d/torch_lpm3d.py, coarse2d/torch_lpm.py, cement3d/torch_v1.py), run on the CPU
in FP64.

Run in a container that holds the reference (python tests/golden/make_surface_golden.py [REFERENCE_ROOT]); the tests
only read the .npz files.  What these vectors pin that the mathematical checks of the oracle leave open: the
computeStress surface rule (the opp-flag weights of coarse3d/torch_lpm3d.py:51-53 and cement3d/torch_v1.py:45-50)
at box faces and at removed particles, the block means of the fine reference, the reference's own autograd gradient
of a coarse loss through LinSolve (per directed slot, coarse3d/torch_lpm3d.py:151-175), the 2D dense solve and its
gradient (coarse2d/torch_lpm.py:173-274).  The recipes follow coarse3d/dev/export_check.py:1-35 (with SCT in place of
lin3d) and coarse3d/check_torch3d.py:48-95.
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
sys.path.insert(0, os.path.join(REF, "coarse3d"))
sys.path.insert(0, os.path.join(REF, "cement3d"))
import geom3d as G  # noqa: E402
import torch_lpm3d as T  # noqa: E402
import torch_v1 as TV  # noqa: E402

DEV = "cpu"
f64 = torch.float64


def t(a):
    return torch.as_tensor(a, device=DEV, dtype=f64)


def two_phase_params(lat, rng, frac=0.3, contrast=4.0):
    """Per-particle isotropic C of a random two-phase field -> directed Kn, Tv by the reference's maps."""
    ph = rng.random(lat.N) < frac
    C = G.iso_C(np.where(ph, contrast, 1.0), np.where(ph, 0.2, 0.25))
    return G.bond_params_two_end_mean(lat, G.sc_params_from_C(C, 0.5))


def random_bond_params(lat, rng):
    """Random symmetric per-bond Kn, Tv (the coarse3d/check_torch3d.py:54-61 recipe)."""
    Kn = np.zeros((lat.N, 18))
    Tv = np.zeros((lat.N, 18))
    for i in range(lat.N):
        for k in range(0, 18, 2):
            q = lat.nbr[i, k]
            if q >= 0:
                Kn[i, k] = Kn[q, k ^ 1] = rng.uniform(0.8, 1.2)
                Tv[i, k] = Tv[q, k ^ 1] = rng.uniform(-0.03, 0.03)
    return Kn, Tv


def surface3d():
    """n = 10 two-phase box: operator outputs on a random field and the fine reference at s = 2."""
    n = 10
    rng = np.random.default_rng(5)
    lat = G.LatticeGeom(n, n, n, 1.0)
    Kn, Tv = two_phase_params(lat, rng)
    u = rng.standard_normal((lat.N, 3))
    op = T.SCT(lat, DEV)
    out = dict(n=n, Kn=Kn, Tv=Tv, u=u, Ku=op.apply(t(u), t(Kn), t(Tv)).numpy(),
               stress=op.stress(t(u), t(Kn), t(Tv)).numpy(), energy=op.energy(t(u), t(Kn), t(Tv)).numpy(),
               block_diag=op.block_diag(t(Kn), t(Tv)).numpy(), stw=op.stw.numpy())
    fr = T.FineRefT(n, 2, device=DEV)
    r = fr.run(t(Kn)[None], t(Tv)[None], tol=1e-12)
    for k in ("u", "sig", "w", "F"):
        out["fine_" + k] = r[k][0].numpy()
    out["fine_relres"] = float(r["relres"])
    return out


def coarse3d():
    """Coarse 5^3 model at s = 2 (all 9 loads): run outputs and the gradient of a fixed linear functional of them
    by the reference's autograd (LinSolve backward + autograd of stress / energy / reactions)."""
    nc, s = 5, 2
    rng = np.random.default_rng(1)
    cm = T.CoarseT(nc, s, device=DEV)
    Kn, Tv = random_bond_params(cm.lat, rng)
    L = len(cm.loads)
    Wsig = rng.standard_normal((1, L, cm.N, 6))
    Wu = rng.standard_normal((1, L, cm.N, 3))
    Ww = rng.standard_normal((1, L, cm.N))
    WF = rng.standard_normal((1, L))
    kn = t(Kn)[None].requires_grad_(True)
    tv = t(Tv)[None].requires_grad_(True)
    o = cm.run(kn, tv, tol=1e-13)
    loss = (o["sig"] * t(Wsig)).sum() + 1e3 * (o["u"] * t(Wu)).sum() + (o["w"] * t(Ww)).sum() + (o["F"] * t(WF)).sum()
    loss.backward()
    out = dict(nc=nc, s=s, Kn=Kn, Tv=Tv, Wsig=Wsig, Wu=Wu, Ww=Ww, WF=WF, loss=float(loss.detach()),
               gKn=kn.grad[0].numpy(), gTv=tv.grad[0].numpy())
    for k in ("u", "sig", "w", "F"):
        out[k] = o[k][0].detach().numpy()
    return out


def coarse2d():
    """2D coarse 8 x 6 model at s = 2 (all 6 loads, dense Cholesky) with symmetric per-bond parameters, and the
    gradient of a fixed linear functional of its outputs."""
    sys.path.insert(0, os.path.join(REF, "coarse2d"))
    import torch_lpm as T2
    import lin2d
    ncx, ncy, s = 8, 6, 2
    rng = np.random.default_rng(3)
    m = T2.CoarseT(ncx, ncy, s, device=DEV)
    lat = m.lat
    ph = rng.random(lat.N) < 0.35
    C = lin2d.iso_C(np.where(ph, 5.0, 1.0), np.where(ph, 0.2, 0.25))
    Kn, Tv = lin2d.bond_params_two_end_mean(lat, lin2d.square_params_from_C(C))
    # a per-undirected-bond lognormal factor keeps the set symmetric but not a two-end mean
    for i in range(lat.N):
        for k in range(0, 8, 2):
            q = lat.nbr[i, k]
            if q >= 0:
                f = np.exp(0.2 * rng.standard_normal())
                Kn[i, k] *= f
                Kn[q, k ^ 1] = Kn[i, k]
    B = 2
    Kb = np.stack([Kn, Kn * 1.3])
    Tb = np.stack([Tv, Tv * 0.7])
    kn = t(Kb).requires_grad_(True)
    tv = t(Tb).requires_grad_(True)
    o = m.run(kn, tv)
    out = dict(ncx=ncx, ncy=ncy, s=s, Kn=Kb, Tv=Tb, loads=np.array(m.loads))
    loss = 0.0
    for ld in m.loads:
        W = {k: rng.standard_normal(tuple(v.shape)) for k, v in o[ld].items() if v is not None}
        for k, v in o[ld].items():
            if v is None:
                continue
            out["%s/%s" % (ld, k)] = v.detach().numpy()
            out["%s/W%s" % (ld, k)] = W[k]
            loss = loss + (v * t(W[k])).sum() * (1e3 if k == "u" else 1.0)
    loss.backward()
    out["gKn"] = kn.grad.numpy()
    out["gTv"] = tv.grad.numpy()
    out["loss"] = float(loss.detach())
    assert B == 2
    return out


def voids3d():
    """n = 8 box with removed particles: SCTV operator outputs, FineRefTV block means at s = 2, CoarseTV on 4^3
    with removed cells."""
    n = 8
    rng = np.random.default_rng(11)
    lat = G.LatticeGeom(n, n, n, 1.0)
    Kn, Tv = two_phase_params(lat, rng, frac=0.35, contrast=3.0)
    solid = rng.random(lat.N) >= 0.12
    u = rng.standard_normal((lat.N, 3))
    op = TV.SCTV(lat, DEV)
    op.set_mask(torch.as_tensor(solid[None], device=DEV))
    Km = t(Kn)[None] * op.valid[:, 0]
    Tm = t(Tv)[None] * op.valid[:, 0]
    # the mask's (B, 1, N, 18) validity broadcasts the outputs to (B, 1, N, ...): B = 1, one system
    out = dict(n=n, Kn=Kn, Tv=Tv, solid=solid, u=u, Ku=op.apply(t(u)[None], Km, Tm).reshape(-1, 3).numpy(),
               stress=op.stress(t(u)[None], Km, Tm).reshape(-1, 6).numpy(),
               energy=op.energy(t(u)[None], Km, Tm).reshape(-1).numpy(),
               block_diag=op.block_diag(Km, Tm).reshape(-1, 3, 3).numpy())
    fr = TV.FineRefTV(n, 2, device=DEV)
    r = fr.run(t(Kn)[None], t(Tv)[None], torch.as_tensor(solid[None], device=DEV), tol=1e-12)
    for k in ("u", "sig", "w", "F"):
        out["fine_" + k] = r[k][0].numpy()
    nc, s = 4, 2
    cm = TV.CoarseTV(nc, s, device=DEV)
    Kc, Tc = random_bond_params(cm.lat, np.random.default_rng(12))
    cell = np.ones(cm.N, bool)
    cell[[21, 22, 42]] = False                              # interior cells removed
    r = cm.run(t(Kc)[None], t(Tc)[None], torch.as_tensor(cell[None], device=DEV), tol=1e-12, differentiable=False)
    out.update(nc=nc, s=s, Kc=Kc, Tc=Tc, cell=cell)
    for k in ("u", "sig", "w", "F"):
        out["coarse_" + k] = r[k][0].numpy()
    return out


def main():
    torch.set_num_threads(4)
    for name, fn in (("surface3d_n10", surface3d), ("coarse3d_nc5", coarse3d), ("coarse2d_8x6", coarse2d),
                     ("voids3d_n8", voids3d)):
        out = fn()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, os.path.getsize(os.path.join(HERE, name + ".npz")), "bytes", flush=True)


if __name__ == "__main__":
    main()
