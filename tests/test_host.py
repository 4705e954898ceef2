"""Host-side pieces that need no GPU: the C ABI exports every declared symbol, the product's geometry / load
cases / parameter maps agree with the oracle's restatement, and the operators refuse CPU tensors."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import lpm_ref as R
import lpm_ops as ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lpm-solver-trained-coarse-graining_amd")


def _declared():
    src = open(os.path.join(ROOT, "include", "lpm.h")).read()
    return sorted(set(re.findall(r"\b(lpm_[a-z0-9_]+)\s*\(", src)))


def test_abi_exports_every_declared_symbol():
    import _lpmlib
    lib = _lpmlib.lib()
    names = _declared()
    assert "lpm_apply" in names and "lpm_pcg_refined" in names
    for n in names:
        assert hasattr(lib, n), n
    assert lib.lpm_version() >= 1
    for n in _lpmlib.SIGS:
        assert n in names


def test_integration_binding_matches_the_library_structs(monkeypatch):
    """The ctypes binding the integration guide gives maintainers, executed as written: its Grid / Bcast have the
    size and field offsets of the library's own binding (lpm_grid is 48 bytes with the bond-word pointer and its
    strides), and its lpm_apply signature is the one the library is driven through."""
    import _lpmlib as L
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", txt, re.S)
    snippet = next(b for b in blocks if "class Grid" in b)
    monkeypatch.setenv("LPM_LIB", L.LIB_PATH)
    ns = {}
    exec(compile(snippet, "INTEGRATION.md", "exec"), ns)
    G, B = ns["Grid"], ns["Bcast"]
    assert ctypes.sizeof(G) == ctypes.sizeof(L.Grid) == 48
    assert ctypes.sizeof(B) == ctypes.sizeof(L.Bcast) == 16
    for (name, _), (name2, _) in zip(G._fields_, L.Grid._fields_):
        assert name == name2 and getattr(G, name).offset == getattr(L.Grid, name2).offset
    assert len(G._fields_) == len(L.Grid._fields_)
    assert [ctypes.sizeof(t) for t in ns["lib"].lpm_apply.argtypes] == \
        [ctypes.sizeof(t) for t in L.SIGS["lpm_apply"][1]]
    # a bad dimension is refused through the snippet's binding (no GPU call is made)
    g = G(4, 2, 2, 2, 1.0, None, B(0, 0))
    rc = ns["lib"].lpm_apply(ctypes.byref(g), 1, 1, 0, 0, None, B(0, 0), None, None, B(0, 0), None, None)
    assert rc == L.LPM_E_BADARG


@pytest.mark.parametrize("dims", [(256, 256, 256), (70, 11, 40), (7, 6, 5), (1, 1, 1), (48, 48, 48), (20, 14),
                                  (128, 128), (3, 2)])
def test_packed_layout_is_one_definition_and_aligned(dims):
    """The packed plane stride of the C ABI, lpm_ops and synth agree; guard and stride are multiples of 64 values
    (FP32 planes start 256-byte aligned) and the trailing guard is at least the leading one, which covers every
    neighbour offset (one z-plane + one row + one site)."""
    import _lpmlib as L
    import synth
    g = L.grid_of(dims, 1.0)
    P = L.lib().lpm_packed_len(ctypes.byref(g))
    G = ops.packed_guard(dims)
    N = int(np.prod(dims))
    reach = (dims[0] * dims[1] + dims[0] + 1) if len(dims) == 3 else dims[0] + 1
    assert P == ops.packed_len(dims) == synth.plane_len(dims) and G == synth.guard_len(dims)
    assert P % 64 == 0 and G % 64 == 0 and G >= reach and P - N - G >= G


def test_noalloc_solve_refuses_oversize_before_any_gpu_call():
    """The allocation-free PCG entry point checks its workspace against the static scratch before touching the GPU."""
    import _lpmlib as L
    lib = L.lib()
    cap = int(lib.lpm_noalloc_scratch_bytes())
    assert cap == 64 << 20
    g = L.grid_of((200, 200, 200), 1.0)
    assert int(lib.lpm_pcg_workspace(ctypes.byref(g), 1, L.F32)) > cap
    o = ops._opts(False, 1e-8, 10, 10, False)
    z = L.Bcast(0, 0)
    dummy = ctypes.c_void_p(16)
    rc = lib.lpm_pcg_solve_noalloc(ctypes.byref(g), 1, 1, L.F32, L.PACKED, dummy, dummy, z, None, z, None, z, dummy,
                                   dummy, 0, ctypes.byref(o), None, None, None, None)
    assert rc == L.LPM_E_BADARG and b"static scratch" in lib.lpm_last_error()


def test_abi_rejects_bad_arguments_without_a_gpu():
    import _lpmlib as L
    g = L.Grid(4, 2, 2, 2, 1.0)
    rc = L.lib().lpm_apply(ctypes.byref(g), 1, 1, L.F32, L.DIRECTED, None, L.Bcast(0, 0), None, None, L.Bcast(0, 0),
                           None, None)
    assert rc == L.LPM_E_BADARG
    assert b"dim" in L.lib().lpm_last_error()


def test_operators_refuse_cpu_tensors():
    import lpm_ops
    u = torch.zeros(8, 3, dtype=torch.float64)
    k = torch.zeros(8, 18, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="GPU"):
        lpm_ops.apply((2, 2, 2), 1.0, u, k, k)


@pytest.mark.parametrize("dims", [(5, 4, 3), (1, 2, 3)])
def test_lattice_geometry_matches_oracle(dims):
    import geom3d as G
    lat = G.LatticeGeom(*dims, h=2.0)
    o = R.Box(dims, h=2.0)
    assert np.array_equal(lat.ijk, o.ijk) and np.array_equal(lat.nbr, o.nbr) and np.array_equal(lat.valid, o.valid)
    assert np.allclose(lat.x, o.x) and np.allclose(lat.ndir, o.ndir) and np.allclose(lat.l0, o.l0)
    assert lat.V == o.V and np.array_equal(lat.opp, o.opp)


@pytest.mark.parametrize("load", ["k_xx", "k_yy", "k_zz", "k_yz", "k_xz", "k_xy", "tb_tz", "tb_sx", "lr_tx"])
def test_load_cases_match_oracle(load):
    import geom3d as G
    lat = G.LatticeGeom(8, 8, 8, 1.0)
    IJK = lat.ijk // 2
    a = G.prescribed(load, lat.x, IJK, (4, 4, 4), 2, 1.0)
    b = R.prescribed3(load, lat.x, IJK, (4, 4, 4), 2, 1.0)
    assert np.array_equal(a[0], b[0]) and np.allclose(a[1], b[1], rtol=0, atol=1e-18)
    assert (a[2] is None) == (b[2] is None)
    if a[2] is not None:
        assert np.array_equal(a[2], b[2])


def test_parameter_maps_match_oracle():
    import geom3d as G
    rng = np.random.default_rng(0)
    lat = G.LatticeGeom(5, 4, 6)
    E = rng.uniform(1, 10, lat.N)
    nu = rng.uniform(0.1, 0.3, lat.N)
    C = G.iso_C(E, nu)
    assert np.allclose(C, R.iso_C(E, nu))
    pp = G.sc_params_from_C(C, 0.5)
    assert np.allclose(pp, R.sc_params_from_C(C, 0.5))
    Kn, Tv = G.bond_params_two_end_mean(lat, pp)
    Kr, Tr = R.two_end_mean(R.Box((5, 4, 6)), pp)
    assert np.allclose(Kn, Kr, rtol=1e-15) and np.allclose(Tv, Tr, rtol=1e-15)
    # generalised calcKnTv of an isotropic cell reduces to the isotropic map (both shells, one Tv)
    cp = G.cell_params_from_C66(G.iso_C66(C), 0.5)
    assert np.allclose(cp[:, 0], pp[:, 0]) and np.allclose(cp[:, 3], pp[:, 1]) and np.allclose(cp[:, 9], pp[:, 2])


@pytest.mark.parametrize("dims", [(5, 4, 3), (6, 5), (2, 1, 3), (1, 2)])
def test_csr_pattern_is_the_structure_of_BtHB(dims):
    """The block pattern of the probed assembly equals the structural nonzeros of the oracle's B^T H B for generic
    coefficients (the group rule of coarse3d/lin3d.py:310-376)."""
    lat = R.Box(dims)
    Kn, Tv = R.random_symmetric_params(lat, np.random.default_rng(0))
    K = R.stiffness(lat, Kn, Tv).tocsr()
    K.eliminate_zeros()
    d = lat.dim
    blocks = set(zip(*(np.array((K != 0).nonzero()) // d)))
    offsets, present, rowptr = ops.csr_pattern(dims)
    strides = np.cumprod([1] + list(dims[:-1]))
    mine = {(int(p), int(p + offsets[o] @ strides)) for p, o in zip(*np.nonzero(present))}
    assert mine == {(int(a), int(b)) for a, b in blocks}
    assert rowptr[-1] == present.sum()


def test_geom2d_matches_oracle():
    import geom2d
    lat = geom2d.SquareGeom(7, 5, 0.5)
    ref = R.Box((7, 5), 0.5)
    assert np.array_equal(lat.nbr, ref.nbr) and np.allclose(lat.ndir, ref.ndir) and np.allclose(lat.x, ref.x)
    assert np.allclose(lat.l0, ref.l0)
    for ld in geom2d.LOADS:
        a = geom2d.prescribed(ld, lat.x, lat.ij, 7, 5, 1, 0.5, 1)
        b = R.prescribed2(ld, ref.x, ref.ijk, 7, 5, 1, 0.5, 1)
        assert np.array_equal(a[0], b[0]) and np.allclose(a[1], b[1])
        assert (a[2] is None) == (b[2] is None) and (a[2] is None or np.array_equal(a[2], b[2]))
    C = geom2d.iso_C(np.array([1.0, 10.0]), np.array([0.2, 0.3]))
    assert np.allclose(C, R.iso_C(np.array([1.0, 10.0]), np.array([0.2, 0.3])))
    pp = geom2d.square_params_from_C(C[[0, 1, 0, 1] * 9][:35])
    assert np.allclose(pp, R.square_params_from_C(C[[0, 1, 0, 1] * 9][:35]))
    Kn, Tv = geom2d.bond_params_two_end_mean(lat, pp)
    Kr, Tr = R.two_end_mean(ref, pp)
    assert np.allclose(Kn, Kr) and np.allclose(Tv, Tr)


def test_interface_cells_match_a_voxel_loop():
    """geom3d.interface_cells against a plain loop over voxels and their face neighbours inside the box."""
    import geom3d as G3
    rng = np.random.default_rng(8)
    for n, s in ((6, 2), (9, 3), (8, 4)):
        phase = (rng.random(n ** 3) < 0.3).astype(np.int64)
        p = phase.reshape(n, n, n)
        nc = n // s
        want = np.zeros(nc ** 3, bool)
        for z in range(n):
            for y in range(n):
                for x in range(n):
                    for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                        zz, yy, xx = z + dz, y + dy, x + dx
                        if 0 <= zz < n and 0 <= yy < n and 0 <= xx < n and p[zz, yy, xx] != p[z, y, x]:
                            want[((z // s) * nc + y // s) * nc + x // s] = True
        assert np.array_equal(G3.interface_cells(phase, n, s), want)
