"""Pins of the oracle's crosslink restatement: the spring law in closed
form, Newton's third law on the bonds, the anchor lever torque, the
binding-model statistics and ranges, and the host generator's geometry."""
import math

import numpy as np
import pytest

import simcore_amd as sc
from oracle import oracle_py as O
from tests.helpers import box_for_fraction


def _box(L):
    h = np.diag([L, L, L]).reshape(-1).astype(float)
    hi = np.diag([1 / L, 1 / L, 1 / L]).reshape(-1).astype(float)
    return h, hi


def test_spring_force_closed_forms():
    h, hi = _box(10.0)
    # stretched along x: force on anchor 0 points to anchor 1, |F| = k (d - r0)
    F, L = O.xlink_force(h, hi, [0, 0, 0], [1.7, 0, 0], 10.0, 1.0)
    assert L == pytest.approx(1.7)
    np.testing.assert_allclose(F, [7.0, 0, 0], rtol=1e-14)
    # compressed, with and without a compression constant
    F, _ = O.xlink_force(h, hi, [0, 0, 0], [0, 0.4, 0], 10.0, 1.0)
    np.testing.assert_allclose(F, [0, -6.0, 0], rtol=1e-14)
    F, _ = O.xlink_force(h, hi, [0, 0, 0], [0, 0.4, 0], 10.0, 1.0, kc=2.0)
    np.testing.assert_allclose(F, [0, -1.2, 0], rtol=1e-14)
    # minimum image across the periodic boundary
    F, L = O.xlink_force(h, hi, [4.8, 0, 0], [-4.7, 0, 0], 10.0, 0.2)
    assert L == pytest.approx(0.5)
    np.testing.assert_allclose(F, [3.0, 0, 0], rtol=1e-12)
    # arbitrary direction: F parallel to the separation, magnitude law
    rng = np.random.default_rng(3)
    for _ in range(20):
        p0, p1 = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        F, L = O.xlink_force(h, hi, p0, p1, 3.0, 0.7)
        d = p1 - p0
        assert L == pytest.approx(np.linalg.norm(d), rel=1e-13)
        np.testing.assert_allclose(F, 3.0 * (L - 0.7) * d / L, rtol=1e-12, atol=1e-14)


def _two_filaments():
    p = sc.params_init(8.0, delta=1e-5, seed=5)
    # two parallel straight filaments 1.5 sigma apart
    n = 6
    r = np.zeros((3, 2 * n))
    r[0, :n] = np.arange(n) - 2.5
    r[0, n:] = np.arange(n) - 2.5
    r[1, n:] = 1.5
    fil = sc.Filaments(np.array([0, n, 2 * n], dtype=np.int32), r.reshape(-1).copy(),
                       (sc.ScFilamentParams * 2)())
    for f in fil.fil:
        f.bond_length, f.persistence_length, f.diameter = 1.0, 10.0, 1.0
        f.gamma_par, f.gamma_perp = 0.2, 0.4
    return p, fil


def test_bond_forces_and_torques_of_one_crosslink():
    p, fil = _two_filaments()
    p.stoch_flag = 0
    o = O.OracleSystem(p, fil)
    # bond 1 of filament 0 (lambda 0.3) to bond 7 = bond 2 of filament 1 (0.8)
    o.load_crosslinks([1, 7], [0.3, 0.8], k_spring=10.0, rest_length=1.0, k_off=0.0)
    assert o.step(1) == 0
    pos, _, u, ln = (x.reshape(3, -1) if x.ndim == 1 and x.size % 3 == 0 and x.size != 10 else x
                     for x in o.bonds())
    pos, u = pos.reshape(3, -1), u.reshape(3, -1)
    assert o.step(1) == 0        # step 2 (full step) computes the forces
    b, lam, xf = o.crosslinks()
    P0 = pos[:, 1] - (0.5 * ln[1] - 0.3) * u[:, 1]
    P1 = pos[:, 7] - (0.5 * ln[7] - 0.8) * u[:, 7]
    d = P1 - P0
    L = np.linalg.norm(d)
    F = 10.0 * (L - 1.0) * d / L
    np.testing.assert_allclose(xf[0, 0:3], F, rtol=1e-12)
    np.testing.assert_allclose(xf[0, 3:6], np.cross((0.3 - 0.5 * ln[1]) * u[:, 1], F),
                               rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(xf[0, 6:9], np.cross((0.8 - 0.5 * ln[7]) * u[:, 7], -F),
                               rtol=1e-10, atol=1e-14)
    # against the same run without the crosslink: bonds 1 and 7 carry +F /
    # -F and the lever torques on top of the pair forces (straight filaments:
    # bonds two apart touch at r = sigma), every other bond is unchanged
    ref = O.OracleSystem(p, fil)
    assert ref.step(2) == 0
    bf, bt = (x.reshape(3, -1) for x in o.bond_forces())
    rf, rt = (x.reshape(3, -1) for x in ref.bond_forces())
    np.testing.assert_allclose(bf[:, 1] - rf[:, 1], xf[0, 0:3], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(bf[:, 7] - rf[:, 7], -xf[0, 0:3], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(bt[:, 1] - rt[:, 1], xf[0, 3:6], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(bt[:, 7] - rt[:, 7], xf[0, 6:9], rtol=1e-12, atol=1e-12)
    others = [k for k in range(bf.shape[1]) if k not in (1, 7)]
    np.testing.assert_array_equal(bf[:, others], rf[:, others])
    # Newton's third law for the pair of anchors
    np.testing.assert_array_equal(xf[0, 0:3] + (-xf[0, 0:3]), np.zeros(3))
    # an odd step re-applies the stored force unchanged, the next even step
    # recomputes it from the moved bonds
    assert o.step(1) == 0
    np.testing.assert_array_equal(o.crosslinks()[2], xf)
    assert o.step(1) == 0
    assert not np.array_equal(o.crosslinks()[2], xf)


def _network(nx=400, k_off=0.0, seed=3):
    n = np.full(150, 16, dtype=np.int32)
    R = box_for_fraction(n, 1.0, 1.0, 0.08)
    p = sc.params_init(R, delta=1e-4, min_cell_length=1.0, seed=17)
    fil = sc.generate_filaments(p, n, 15.0, 50.0, 1.0, 2.0, seed=seed)
    b, lam = sc.generate_crosslinks(p, fil, nx, r_bind=2.0, seed=seed + 1)
    return p, fil, b, lam


def test_generator_geometry():
    p, fil, b, lam = _network()
    assert np.all(b[:, 0] != b[:, 1])
    o = O.OracleSystem(p, fil)
    pos, sp, u, ln = o.bonds()
    sp, u = sp.reshape(3, -1), u.reshape(3, -1)
    L = 2 * p.system_radius
    assert np.all(lam >= 0) and np.all(lam <= 1.0 + 1e-12)
    for x in range(len(b)):
        b0, b1 = b[x]
        P = sp[:, b0] * L - (0.5 * ln[b0] - lam[x, 0]) * u[:, b0]
        d = sp[:, b1] * L - P
        d -= L * np.round(d / L)
        t = np.clip(d @ u[:, b1], -0.5 * ln[b1], 0.5 * ln[b1])
        assert np.linalg.norm(d - t * u[:, b1]) <= 2.0 + 1e-9


def test_binding_model_rate_and_range():
    """Events per crosslink per full step follow p_off = 1 - exp(-k_off dt);
    every rebound spring starts within the binding range."""
    p, fil, b, lam = _network(nx=400)
    o = O.OracleSystem(p, fil)
    k_off = 500.0
    o.load_crosslinks(b, lam, k_off=k_off, r_bind=2.0, seed=99)
    steps = 40
    assert o.step(steps) == 0
    p_off = 1 - math.exp(-k_off * p.delta)
    expect = 400 * (steps // 2) * p_off
    got = o.xlink_events()
    sd = math.sqrt(400 * (steps // 2) * p_off * (1 - p_off))
    print("events", got, "expected", expect)
    assert abs(got - expect) < 5 * sd
    # springs after the run: none longer than the range plus a bond length
    # (the new anchor sits anywhere on the new bond) plus the drift
    bb, ll, _ = o.crosslinks()
    assert np.all(bb[:, 0] != bb[:, 1])
    pos, sp, u, ln = o.bonds()
    sp, u = sp.reshape(3, -1), u.reshape(3, -1)
    L = 2 * p.system_radius
    P0 = sp[:, bb[:, 0]] * L - (0.5 * ln[bb[:, 0]] - ll[:, 0]) * u[:, bb[:, 0]]
    P1 = sp[:, bb[:, 1]] * L - (0.5 * ln[bb[:, 1]] - ll[:, 1]) * u[:, bb[:, 1]]
    d = P1 - P0
    d -= L * np.round(d / L)
    assert np.linalg.norm(d, axis=0).max() < 2.0 + 1.0 + 0.3
