"""CPU-side checks of the product library: it loads, exports every symbol
the C header declares, and its host layer (parameter files, filament
layout and generation) behaves as the reference's."""
import math
import os
import re

import numpy as np
import pytest

import simcore_amd as sc
from oracle import oracle_py as O
from tests.helpers import C0_YAML, ROOT

HEADER = os.path.join(ROOT, "include", "simcore_amd.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(sc_\w+)\s*\(", text, re.M))
    assert len(declared) >= 20
    assert declared == set(sc.EXPORTS)
    for name in declared:
        assert hasattr(sc.lib, name), name
    assert sc.lib.sc_abi_version() == 1


def test_params_from_yaml_c0():
    p, spec = sc.params_from_yaml(C0_YAML)
    assert (p.n_dim, p.n_periodic) == (3, 3)
    assert p.system_radius == 20 and p.delta == 1e-4 and p.seed == 4242
    assert p.wca_eps == 1 and p.wca_sig == 1 and p.f_cutoff == 2000 and p.stoch_flag == 1
    # SetMinCellLength(max(1.5*min_bond_length, diameter))
    assert p.min_cell_length == max(1.5 * 0.65, 1.0)
    assert list(p.unit_cell) == [40, 0, 0, 0, 40, 0, 0, 0, 40]
    # invert_sym_3d_matrix of diag(L): (1/L^3)*L^2, not 1/L
    inv = (1.0 / 40.0 ** 3) * 40.0 ** 2
    assert p.unit_cell_inv[0] == inv and p.unit_cell_inv[4] == inv
    assert spec["num"] == 64 and spec["length"] == 19 and spec["persistence_length"] == 100
    assert spec["friction_ratio"] == 2 and spec["n_steps"] == 1000


def test_yaml_sequence_of_species_and_defaults():
    p, spec = sc.params_from_yaml(os.path.join(GOLDEN, "params_sequence.yaml"))
    assert p.system_radius == 30 and p.delta == 0.001   # delta from defaults
    assert spec["num"] == 10 and spec["length"] == 12    # first named block
    assert spec["persistence_length"] == 400              # filament default
    assert spec["diameter"] == 0.8                        # shared species block


def test_yaml_rejections():
    with pytest.raises(sc.SimcoreError):
        sc.params_from_yaml(os.path.join(GOLDEN, "params_2d.yaml"))
    with pytest.raises(sc.SimcoreError):
        sc.params_from_yaml(os.path.join(GOLDEN, "params_malformed.yaml"))
    with pytest.raises(sc.SimcoreError):
        sc.params_from_yaml(os.path.join(GOLDEN, "does_not_exist.yaml"))


@pytest.mark.parametrize("length,mbl,expect", [
    (19, 0.65, (19, 1.0)),          # C0: floor(19/0.975) = 19 bonds
    (50, 1.0, (33, 50 / 33)),       # floor(50/1.5) = 33
    (1.0, 0.5, (2, 0.5)),           # floor(1/0.75) = 1 -> at least 2 bonds
    (600, 1.0, (333, 500 / 333)),   # clamped to max_length 500
])
def test_filament_layout(length, mbl, expect):
    nb, bl = sc.filament_layout(length, mbl, 5.0 if length > 2 else 0.0, 500.0)
    assert nb == expect[0]
    assert bl == pytest.approx(expect[1], rel=1e-15)


def test_generate_filaments_geometry_and_no_overlap():
    p = sc.params_init(8.0, min_cell_length=1.0)
    n = [10, 7, 12, 10]
    fil = sc.generate_filaments(p, n, [9.0, 6.0, 11.0, 9.0], [20.0, 5.0, 100.0, 0.0],
                                1.0, 2.0, seed=5)
    assert list(fil.site_off) == [0, 10, 17, 29, 39]
    r = fil.r.reshape(3, -1).T
    for f in range(4):
        s = slice(fil.site_off[f], fil.site_off[f + 1])
        d = np.linalg.norm(np.diff(r[s], axis=0), axis=1)
        np.testing.assert_allclose(d, 1.0, rtol=1e-12)
        # SetDiffusion: gamma_perp = 4L / (3 n ln(L/d)), gamma_par = gamma_perp/2
        L = n[f] - 1.0
        assert fil.fil[f].gamma_perp == pytest.approx(4 * L / (3 * n[f] * math.log(L)))
        assert fil.fil[f].gamma_par == pytest.approx(fil.fil[f].gamma_perp / 2)
    # no two bonds of different filaments closer than a diameter
    h = np.array(p.unit_cell)
    bonds = []
    for f in range(4):
        for i in range(fil.site_off[f], fil.site_off[f + 1] - 1):
            c = 0.5 * (r[i] + r[i + 1])
            u = r[i + 1] - r[i]
            bonds.append((f, c / 16.0, u / np.linalg.norm(u), np.linalg.norm(u)))
    for (f1, s1, u1, l1), (f2, s2, u2, l2) in [(a, b) for i, a in enumerate(bonds)
                                              for b in bonds[i + 1:]]:
        if f1 == f2:
            continue
        s1w, s2w = s1 - np.round(s1), s2 - np.round(s2)
        _, m2, _, _ = O.min_distance(h, s1w, u1, l1, s2w, u2, l2)
        assert m2 >= 1.0 - 1e-9
    # same seed, same system
    fil2 = sc.generate_filaments(p, n, [9.0, 6.0, 11.0, 9.0], [20.0, 5.0, 100.0, 0.0],
                                 1.0, 2.0, seed=5)
    np.testing.assert_array_equal(fil.r, fil2.r)


def test_create_without_device_fails_cleanly():
    import ctypes
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    p = sc.params_init(10.0)
    h = ctypes.c_void_p()
    rc = sc.lib.sc_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -2 and not h.value


def test_ensemble_create_without_device_fails_cleanly():
    import ctypes
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    from simcore_amd._lib import ScEnsembleParams
    prm = ScEnsembleParams(1 << 20, 2, 8)
    h = ctypes.c_void_p()
    rc = sc.lib.sc_ensemble_create(ctypes.byref(prm), ctypes.byref(h))
    assert rc == -2 and not h.value


def test_mixed_precision_limits_cells_per_side():
    """The mixed path's compact bond records pack 10 bits per cell
    coordinate: more than 1024 cells per side is refused at creation
    (before any device is touched); FP64 systems keep the 2^31-cell limit."""
    from ctypes import byref, c_void_p
    h = c_void_p()
    # side 1200 sigma, cells >= 2^(1/6) sigma: 1069 per side
    p = sc.params_init(600.0, delta=1e-4, min_cell_length=1.0, seed=1,
                       precision=sc.SC_PREC_MIXED)
    assert sc.lib.sc_create(byref(p), byref(h)) == -1
    # side 2600 sigma: 2316 per side, more than 2^31 cells
    p = sc.params_init(1300.0, delta=1e-4, min_cell_length=1.0, seed=1,
                       precision=sc.SC_PREC_FP64)
    assert sc.lib.sc_create(byref(p), byref(h)) == -1
