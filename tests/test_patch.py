"""Patch homogenisation (coarse3d/patch3d.py): the oracle (oracle/patch_ref.py) pinned by the exactness for
homogeneous material (the reference's own check, patch3d.py:246-252), symmetry and positive definiteness; the GPU
patch_C (patch3d.py of the package) against the oracle."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "lpm-solver-trained-coarse-graining_amd"))
import lpm_ref as R  # noqa: E402
import patch_ref as PR  # noqa: E402


def _iso(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = 0.5 * E / (1 + nu)
    return np.array([lam + 2 * mu, lam, mu])


@pytest.mark.parametrize("s", [2, 3])
def test_oracle_exact_for_homogeneous_material(s):
    geom = PR.PatchGeom(s)
    c3 = _iso(1.0, 0.2)
    PP = np.tile(PR.sc_params_from_C(c3, 0.5), (geom.lat.N, 1))
    C = PR.cell_C(geom, PP)
    ref = PR.iso_C66(c3)[0]
    assert np.abs(C - ref).max() <= 1e-10 * ref[0, 0]


def test_oracle_two_phase_symmetric_positive_and_bounded():
    rng = np.random.default_rng(0)
    n, s = 6, 2
    E = np.where(rng.random(n ** 3) < 0.4, 4.0, 1.0)
    Cv = np.stack([_iso(e, 0.2) for e in E])
    C = PR.patch_C(Cv, n, s)
    assert np.abs(C - C.transpose(0, 2, 1)).max() < 1e-12 * np.abs(C).max()
    assert np.linalg.eigvalsh(C).min() > 0
    # between the homogeneous bounds of the softer and the stiffer phase
    lo, hi = PR.iso_C66(_iso(1.0, 0.2))[0], PR.iso_C66(_iso(4.0, 0.2))[0]
    assert np.all(C[:, 0, 0] >= lo[0, 0] * (1 - 1e-9)) and np.all(C[:, 0, 0] <= hi[0, 0] * (1 + 1e-9))


def test_windows_follow_reflect_padding():
    n, s = 4, 2
    Cv = np.arange(n ** 3 * 3, dtype=float).reshape(n ** 3, 3) + 1.0
    PP = PR.windows(Cv, n, s)
    pp = PR.sc_params_from_C(Cv, 0.5).reshape(n, n, n, 3)
    # window of cell (0, 0, 0): window index (z, y, x) = (2, 2, 2) is voxel (0, 0, 0); (1, 2, 2) mirrors voxel z = 1
    W = 3 * s
    w0 = PP[0].reshape(W, W, W, 3)
    assert np.array_equal(w0[2, 2, 2], pp[0, 0, 0]) and np.array_equal(w0[1, 2, 2], pp[1, 0, 0])


@pytest.mark.gpu
def test_gpu_patch_matches_oracle():
    import patch3d
    rng = np.random.default_rng(1)
    n, s = 6, 2
    E = np.where(rng.random(n ** 3) < 0.35, 5.0, 1.0)
    Cv = np.stack([_iso(e, 0.2) for e in E])
    ref = PR.patch_C(Cv, n, s)
    got = patch3d.patch_C(Cv, n, s, chunk=10)             # chunks that do not divide the 27 cells
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("s", [2, 3])
def test_gpu_patch_exact_for_homogeneous_material(s):
    import patch3d
    n = 3 * s
    c3 = _iso(2.0, 0.25)
    C = patch3d.patch_C(np.tile(c3, (n ** 3, 1)), n, s)
    ref = PR.iso_C66(c3)[0]
    assert np.abs(C - ref[None]).max() <= 1e-9 * ref[0, 0]
