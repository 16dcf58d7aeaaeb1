"""Crosslinked network on the device against the CPU path: spring forces,
their application order, the rebinding events and the per-bond lists the
device rebuilds after them."""
import numpy as np
import pytest

import simcore_amd as sc
from oracle.oracle_py import OracleSystem
from tests.helpers import box_for_fraction

pytestmark = pytest.mark.gpu

# Mixed-path position tolerance over the first steps: FP32 pair forces
# round to ~6e-8 relative, and WCA forces reach the 2000 kT/sigma cap at
# contact, so one step (dt/2 = 5e-5, gamma_perp ~ 0.4) can move a site by a
# few 1e-8 sigma against the FP64 path; 1e-7 sigma keeps a 3x margin.
POS_TOL_SHORT = 1e-7

XL = dict(k_spring=10.0, rest_length=1.0, k_compress=-1.0, k_off=500.0, r_bind=2.0, seed=99)


def _network(nx=400, seed=3):
    n = np.full(150, 16, dtype=np.int32)
    R = box_for_fraction(n, 1.0, 1.0, 0.08)
    p = sc.params_init(R, delta=1e-4, min_cell_length=1.0, seed=17)
    fil = sc.generate_filaments(p, n, 15.0, 50.0, 1.0, 2.0, seed=seed)
    b, lam = sc.generate_crosslinks(p, fil, nx, r_bind=2.0, seed=seed + 1)
    return p, fil, b, lam


def _pair(precision="fp64", use_graph=False, spatial=False):
    p, fil, b, lam = _network()
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=use_graph,
                            spatial_order=spatial)
    dev.load_crosslinks(b, lam, **XL)
    ora = OracleSystem(p, fil)
    ora.load_crosslinks(b, lam, **XL)
    return p, fil, dev, ora


def test_crosslinks_fp64_bit_exact_with_rebinding():
    print("case: 150x16 network, 400 crosslinks, k_off 500, fp64 60 steps", flush=True)
    p, fil, dev, ora = _pair()
    dev.step(2)
    assert ora.step(2) == 0
    db, dl, df, dev_ev = dev.crosslinks()
    ob, ol, of = ora.crosslinks()
    np.testing.assert_array_equal(df, of)
    for a, b in zip(dev.bond_forces(), ora.bond_forces()):
        np.testing.assert_array_equal(a, b)
    dev.step(58)
    assert ora.step(58) == 0
    db, dl, df, dev_ev = dev.crosslinks()
    ob, ol, of = ora.crosslinks()
    print("events", dev_ev, ora.xlink_events())
    assert dev_ev == ora.xlink_events() > 0
    np.testing.assert_array_equal(db, ob)
    np.testing.assert_array_equal(dl, ol)
    np.testing.assert_array_equal(df, of)
    for a, b in zip(dev.sites(), ora.sites()):
        np.testing.assert_array_equal(a, b)


def test_crosslinks_mixed_spatial_order_within_tolerance():
    print("case: network mixed spatial order 1 and 100 steps", flush=True)
    p, fil, dev, ora = _pair(precision="mixed", use_graph=True, spatial=True)
    dev.step(2)
    assert ora.step(2) == 0
    d2 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 2 steps", d2)
    assert d2 <= POS_TOL_SHORT
    dev.step(98)
    assert ora.step(98) == 0
    d100 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 100 steps", d100)
    assert d100 <= 1e-5
    db, dl, df, ev = dev.crosslinks()
    ob, ol, of = ora.crosslinks()
    assert ev == ora.xlink_events()
    np.testing.assert_array_equal(db, ob)
    np.testing.assert_allclose(dl, ol, rtol=0, atol=0)
    np.testing.assert_allclose(df, of, rtol=1e-5, atol=1e-5 * np.abs(of).max())


def test_crosslinks_graph_matches_eager():
    print("case: network fp64 graph vs eager 41 steps", flush=True)
    p, fil, b, lam = _network()
    e = sc.FilamentSystem(p, fil, precision="fp64")
    g = sc.FilamentSystem(p, fil, precision="fp64", use_graph=True)
    e.load_crosslinks(b, lam, **XL)
    g.load_crosslinks(b, lam, **XL)
    e.step(41)
    g.step(41)
    for x, y in zip(e.sites(), g.sites()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(e.crosslinks(), g.crosslinks()):
        np.testing.assert_array_equal(x, y)


def test_crosslinks_checkpoint_resume():
    print("case: network mixed checkpoint at 21, +30 steps", flush=True)
    p, fil, b, lam = _network()
    a = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    a.load_crosslinks(b, lam, **XL)
    a.step(21)
    blob = a.checkpoint()
    a.step(30)
    c = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    c.load_crosslinks(b, lam, **XL)
    c.restore(blob)
    c.step(30)
    for x, y in zip(a.sites(), c.sites()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a.crosslinks(), c.crosslinks()):
        np.testing.assert_array_equal(x, y)


def test_crosslinks_restore_into_system_loaded_elsewhere():
    """The rebinding search reads the cell membership of the last rebuild:
    a restore must bring that back too.  The restoring system was loaded
    with other positions (its own cell list unrelated to the saved state);
    after the restore both runs must agree bit for bit, rebinding included."""
    print("case: network mixed, restore into a system loaded from other positions", flush=True)
    p, fil, b, lam = _network()
    _, fil_other, _, _ = _network(seed=8)
    a = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    a.load_crosslinks(b, lam, **XL)
    a.step(21)
    blob = a.checkpoint()
    ev0 = a.crosslinks()[3]
    a.step(30)
    assert a.crosslinks()[3] > ev0  # rebinding happened after the checkpoint
    c = sc.FilamentSystem(p, fil_other, precision="mixed", use_graph=True)
    c.load_crosslinks(b, lam, **XL)
    c.restore(blob)
    c.step(30)
    for x, y in zip(a.sites(), c.sites()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a.crosslinks(), c.crosslinks()):
        np.testing.assert_array_equal(x, y)
