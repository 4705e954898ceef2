"""Writes tests/golden/bondnet_*.npz: input/output vectors of the reference's bond network (coarse3d/net3d.py).

Run in a container that holds the reference (python tests/golden/make_bondnet_golden.py [REFERENCE_ROOT]); the tests
only read the .npz files.  Each fixture records the network's hyper-parameters, its weights (drawn here from a
seeded normal distribution, last head layer included, so that every layer is exercised), the voxel features X of a
random two-phase microstructure (net3d.voxel_features), the trunk output, the bond windows (iv, ic, bonds, perms of
net3d.BondWindows) and the raw head outputs of net3d.BondNet.forward.  What they pin and nothing else is taken from
the reference: the ordering of window slots and stabiliser perms, which the symmetry checks cannot fix.
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
sys.path.insert(0, os.path.join(REF, "coarse3d"))
import net3d  # noqa: E402

CASES = [
    # name, s, nc, batch, ch, D, hid, seed
    ("bondnet_s2_small", 2, 3, 2, 8, 4, 16, 1),
    ("bondnet_s3_small", 3, 3, 1, 6, 3, 8, 2),
    ("bondnet_s2_full", 2, 2, 1, 24, 12, 64, 3),
]


def main():
    torch.set_num_threads(4)
    for name, s, nc, B, ch, D, hid, seed in CASES:
        rng = np.random.default_rng(seed)
        net = net3d.BondNet(s, ch=ch, D=D, hid=hid).double()
        sd = {}
        for k, v in net.state_dict().items():
            if k.endswith("masks"):
                continue
            sd[k] = (rng.standard_normal(tuple(v.shape)) * 0.3).astype(np.float32)
        net.load_state_dict({k: torch.as_tensor(v, dtype=torch.float64) for k, v in sd.items()}, strict=False)
        n = nc * s
        Xs = []
        for b in range(B):
            phase = (rng.random(n ** 3) < 0.4).astype(np.int64)
            Xs.append(net3d.voxel_features(phase, 3.0 + b, 0.3, n, s))
        X = np.stack(Xs)
        with torch.no_grad():
            Xt = torch.as_tensor(X, dtype=torch.float64)
            Fv = net.trunk(Xt).numpy()
            outs = [o.numpy() for o in net(Xt, nc)]
        bw = net3d.BondWindows(nc, s)
        out = dict(s=s, nc=nc, ch=ch, D=D, hid=hid, X=X, Fv=Fv.astype(np.float64))
        out.update({"w/" + k: v for k, v in sd.items()})
        for gi, (g, o) in enumerate(zip(bw.groups, outs)):
            out["g%d/iv" % gi] = g["iv"]
            out["g%d/ic" % gi] = g["ic"]
            out["g%d/bonds" % gi] = g["bonds"]
            out["g%d/perms" % gi] = g["perms"]
            out["g%d/out" % gi] = o
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, {k: v.shape for k, v in out.items() if hasattr(v, "shape") and v.ndim})


if __name__ == "__main__":
    main()
