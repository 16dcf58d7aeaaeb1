"""Edge inputs of the load path and the smallest systems, against the CPU
oracle on the same inputs and noise stream (C ABI through simcore_amd).

sc_load_filaments accepts filaments of 3-256 sites (two bonds at least,
as Filament::InitFilamentLength gives, src/filament.cpp:352-377; 256 the
register integrators' limit) and refuses anything else with SC_ERR_ARG
before it touches the device."""
import numpy as np
import pytest

import simcore_amd as sc
from oracle.oracle_py import OracleSystem
from simcore_amd._lib import SC_ERR_ARG
from tests import scenarios as S
from tests.helpers import culled

pytestmark = pytest.mark.gpu


def _straight(n, y, z=0.0):
    s = np.zeros((n, 3))
    s[:, 0] = np.arange(n) - 0.5 * (n - 1)
    s[:, 1] = y
    s[:, 2] = z
    return s


@pytest.mark.parametrize("sites,fragment", [(2, "at least 2 bonds"), (257, "at most 256")])
def test_load_refuses_filaments_outside_3_to_256_sites(sites, fragment):
    print(f"case: load a {sites}-site filament", flush=True)
    p = sc.params_init(20.0, delta=1e-4, min_cell_length=1.0, seed=1)
    fil = S.make_filaments([_straight(5, -3.0), _straight(sites, 3.0)], diameter=0.5)
    with pytest.raises(sc.SimcoreError) as e:
        sc.FilamentSystem(p, fil, precision="fp64")
    assert e.value.code == SC_ERR_ARG
    assert fragment in str(e.value)


def test_load_refuses_a_site_offset_not_starting_at_zero():
    print("case: site_off[0] = 1", flush=True)
    p = sc.params_init(10.0, delta=1e-4, min_cell_length=1.0, seed=1)
    fil = S.make_filaments([_straight(5, -3.0), _straight(5, 3.0)])
    fil.site_off = fil.site_off + 1
    with pytest.raises(sc.SimcoreError) as e:
        sc.FilamentSystem(p, fil, precision="fp64")
    assert e.value.code == SC_ERR_ARG


def _bit_exact_run(p, fil, chunks):
    dev = sc.FilamentSystem(p, fil, precision="fp64", use_graph=True)
    ora = OracleSystem(p, fil)
    d1, d2 = dev.candidates()
    o1, o2 = culled(*ora.candidates(), fil.site_off)
    np.testing.assert_array_equal(d1, o1)
    np.testing.assert_array_equal(d2, o2)
    for n in chunks:
        dev.step(n)
        assert ora.step(n) == 0
        cd, co = dev.counters(), ora.counters()
        assert cd["i_step"] == co["i_step"]
        assert cd["n_rebuilds"] == co["n_rebuilds"]
        for a, b in zip(dev.sites(), ora.sites()):
            np.testing.assert_array_equal(a, b)
    fd, td = dev.bond_forces()
    fo, to = ora.bond_forces()
    np.testing.assert_array_equal(fd, fo)
    np.testing.assert_array_equal(td, to)
    np.testing.assert_array_equal(dev.tensions(), ora.tensions())
    out = dev.counters()
    dev.close()
    return out


def test_single_filament_fp64_bit_exact():
    """One 64-site filament in a small box: its only candidates are its own
    non-adjacent bonds, and it wraps the periodic box."""
    print("case: one 64-site filament, fp64, 60 steps", flush=True)
    p = sc.params_init(6.0, delta=1e-4, min_cell_length=1.0, seed=11)
    fil = sc.generate_filaments(p, [64], 63.0, 5.0, 1.0, 2.0, seed=5)
    c = _bit_exact_run(p, fil, (1, 2, 17, 40))
    assert c["n_candidates"] > 0


def test_three_site_filaments_fp64_bit_exact():
    """The shortest filaments accepted (two bonds each), densely packed."""
    print("case: 200 three-site filaments, fp64, 60 steps", flush=True)
    p = sc.params_init(7.0, delta=1e-4, min_cell_length=1.0, seed=12)
    fil = sc.generate_filaments(p, [3] * 200, 2.0, 20.0, 1.0, 2.0, seed=6)
    c = _bit_exact_run(p, fil, (1, 2, 17, 40))
    assert c["n_rebuilds"] > 1
    assert c["n_candidates"] > 200


def test_single_filament_mixed_within_tolerance():
    print("case: one 200-site filament, mixed, 20 steps", flush=True)
    p = sc.params_init(8.0, delta=1e-5, min_cell_length=1.0, seed=13)
    fil = sc.generate_filaments(p, [200], 199.0, 50.0, 1.0, 2.0, seed=7)
    dev = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    ora = OracleSystem(p, fil)
    dev.step(20)
    assert ora.step(20) == 0
    # same positions bound as the C0 mixed trajectory test: FP32 pair
    # forces round to ~6e-8 relative; 20 steps at dt 1e-5 move a site by
    # far less than 1e-6 sigma against the FP64 path
    d = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 20 steps", d)
    assert d <= 1e-6
    dev.close()
