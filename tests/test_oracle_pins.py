"""Pins of the CPU oracle against what mathematics and published algorithms
fix: known-answer vectors, brute force, closed forms and invariants.  Each
check is chosen so that a plausible slip (dropped term, wrong sign or index,
swapped operand) fails it."""
import itertools
import math

import numpy as np
import pytest

import simcore_amd as sc
from oracle import oracle_py as O
from tests.helpers import bond_filament, c0_system


# ---------------------------------------------------------------- noise ----
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                    [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                  0x24126EA1]


def test_noise_is_uniform_minus_half():
    xs = np.array([O.noise(123, s, m) for s in range(2000) for m in range(10)]).ravel()
    assert xs.min() > -0.5 and xs.max() < 0.5
    # uniform(-1/2, 1/2): mean 0, variance 1/12 (the reference's
    # RandomUniform() - 0.5, src/filament.cpp:529-530)
    assert abs(xs.mean()) < 4 * math.sqrt(1 / 12 / xs.size)
    assert abs(xs.var() - 1 / 12) < 5e-3
    # distinct counters give distinct draws
    assert not np.array_equal(O.noise(1, 0, 0), O.noise(1, 0, 1))
    assert not np.array_equal(O.noise(1, 0, 0), O.noise(1, 1, 0))
    assert not np.array_equal(O.noise(1, 0, 0), O.noise(2, 0, 0))


# ------------------------------------------------------------ cell list ----
def _adjacent(a, b, n):
    return all(((x - y) % n) in (0, 1, n - 1) for x, y in zip(a, b))


@pytest.mark.parametrize("n", [3, 4, 5, 7])
def test_half_shell_covers_each_adjacent_cell_pair_once(n):
    seen = {}
    for x, y, z in itertools.product(range(n), repeat=3):
        nb = O.cell_neighbors(n, x, y, z)
        assert len(nb) == 13
        home = (x * n + y) * n + z
        for k in nb:
            key = frozenset((home, int(k)))
            assert len(key) == 2, "a cell never neighbours itself"
            seen[key] = seen.get(key, 0) + 1
    dec = lambda k: (k // (n * n), (k // n) % n, k % n)
    for key, cnt in seen.items():
        a, b = (dec(k) for k in key)
        assert _adjacent(a, b, n)
        assert cnt == 1, f"pair {a},{b} emitted {cnt} times"
    n_adjacent_pairs = sum(
        1 for a, b in itertools.combinations(itertools.product(range(n), repeat=3), 2)
        if _adjacent(a, b, n))
    assert len(seen) == n_adjacent_pairs


def test_half_shell_first_entries_order():
    # AssignCellNeighbors(false) order: (x+1,y,z) first, then the y+1 row
    # from x-1, then the z+1 plane from (x-1, y-1).
    n = 6
    nb = O.cell_neighbors(n, 2, 2, 2)
    key = lambda x, y, z: (x * n + y) * n + z
    assert list(nb[:4]) == [key(3, 2, 2), key(1, 3, 2), key(2, 3, 2), key(3, 3, 2)]
    assert nb[4] == key(1, 1, 3) and nb[12] == key(3, 3, 3)


def test_cell_key_faces_and_clamp():
    n = 35
    k = lambda s: O.cell_key(n, np.array([s, 0.0, 0.0])) // (n * n)
    assert k(-0.5) == 0
    assert k(0.5) == n - 1                 # floor(n) == n clamps to n-1
    assert k(np.nextafter(0.5, 0)) == n - 1
    assert k(-0.5 - 2.0 ** -54) == 0       # out-of-range guard
    # monotone and onto over a fine sweep
    s = np.linspace(-0.5, 0.5, 20001)
    ks = np.array([k(v) for v in s])
    assert np.all(np.diff(ks) >= 0)
    assert set(ks.tolist()) == set(range(n))
    # each cell holds the centres whose exact position lies inside it
    for v in s[::97]:
        c = k(v)
        assert c / n - 1e-12 <= v + 0.5 <= (c + 1) / n + 1e-12


def test_candidate_set_matches_brute_force():
    p, fil = c0_system(num=24, seed=5)
    o = O.OracleSystem(p, fil, n_threads=2)
    n = o.counters()["n_cells_1d"]
    keys = o.cell_keys()
    b1, b2 = o.candidates()           # pre-cull list (rebuild just happened)
    cells = [(k // (n * n), (k // n) % n, k % n) for k in keys]
    nb = o.nb
    brute = set()
    for i in range(nb):
        for j in range(i + 1, nb):
            if _adjacent(cells[i], cells[j], n):
                brute.add(frozenset((i, j)))
    got = [frozenset((int(a), int(b))) for a, b in zip(b1, b2)]
    assert len(got) == len(set(got)), "no pair twice"
    assert set(got) == brute
    # ordering: home cells ascending, same-cell pairs in bond order
    assert np.all(np.diff(keys[b1]) >= 0)
    same = keys[b1] == keys[b2]
    assert np.all(b1[same] < b2[same])
    # after one step the site-sharing pairs are gone and nothing else is
    o.step(1)
    c1, c2 = o.candidates()
    bf = bond_filament(fil.site_off)
    shares = lambda a, b: abs(int(a) - int(b)) == 1 and bf[a] == bf[b]
    expect = [(a, b) for a, b in zip(b1, b2) if not shares(a, b)]
    assert list(zip(c1.tolist(), c2.tolist())) == [(int(a), int(b)) for a, b in expect]


# --------------------------------------------------------- min distance ----
def _seg_dist_brute(s1, u1, l1, s2, u2, l2, L):
    d = (np.asarray(s2) - np.asarray(s1))
    d = (d - np.round(d)) * L
    t = np.linspace(-0.5, 0.5, 801)
    p1 = t[:, None, None] * l1 * u1[None, None, :]
    p2 = d[None, None, :] + t[None, :, None] * l2 * u2[None, None, :]
    dist = np.linalg.norm(p2 - p1, axis=2)
    i, j = np.unravel_index(np.argmin(dist), dist.shape)

    def f(a, b):
        return np.linalg.norm(d + b * u2 - a * u1)
    # golden-section refinement in both parameters (coordinate descent)
    a, b = t[i] * l1, t[j] * l2
    for _ in range(60):
        for which in (0, 1):
            lo, hi = (-l1 / 2, l1 / 2) if which == 0 else (-l2 / 2, l2 / 2)
            g = (math.sqrt(5) - 1) / 2
            x1, x2 = hi - g * (hi - lo), lo + g * (hi - lo)
            for _ in range(80):
                f1 = f(x1, b) if which == 0 else f(a, x1)
                f2 = f(x2, b) if which == 0 else f(a, x2)
                if f1 < f2:
                    hi, x2 = x2, x1
                    x1 = hi - g * (hi - lo)
                else:
                    lo, x1 = x1, x2
                    x2 = lo + g * (hi - lo)
            if which == 0:
                a = 0.5 * (lo + hi)
            else:
                b = 0.5 * (lo + hi)
    return f(a, b)


def _unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v)


def test_min_distance_against_brute_force():
    rng = np.random.default_rng(3)
    L = 10.0
    h = np.diag([L, L, L]).ravel()
    for _ in range(60):
        s1 = rng.uniform(-0.5, 0.5, 3)
        s2 = s1 + rng.normal(0, 0.12, 3)
        s2 = s2 - np.round(s2)
        u1, u2 = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        l1, l2 = rng.uniform(0.3, 2.0, 2)
        rmin, m2, c1, c2 = O.min_distance(h, s1, u1, l1, s2, u2, l2)
        ref = _seg_dist_brute(s1, u1, l1, s2, u2, l2, L)
        assert math.sqrt(m2) == pytest.approx(ref, abs=1e-7)
        # r_min = dr - contact1 + contact2 with contacts on the segments
        ds = (s2 - s1) - np.round(s2 - s1)
        np.testing.assert_allclose(rmin, ds * L - c1 + c2, atol=1e-12)
        assert np.linalg.norm(c1) <= l1 / 2 + 1e-12 and np.linalg.norm(c2) <= l2 / 2 + 1e-12


def test_min_distance_closed_forms():
    L = 20.0
    h = np.diag([L, L, L]).ravel()
    ex, ey, ez = np.eye(3)
    # perpendicular, crossing at their centres, offset 0.7 along z
    _, m2, c1, c2 = O.min_distance(h, [0, 0, 0], ex, 1, [0, 0, 0.7 / L], ey, 1)
    assert math.sqrt(m2) == pytest.approx(0.7, abs=1e-14)
    np.testing.assert_allclose(c1, 0, atol=1e-15)
    # parallel (near-parallel branch), side by side at distance 0.9
    _, m2, _, _ = O.min_distance(h, [0, 0, 0], ex, 1, [0.2 / L, 0.9 / L, 0], ex, 1)
    assert math.sqrt(m2) == pytest.approx(0.9, abs=1e-14)
    # collinear, end to end with a gap of 0.25
    _, m2, c1, c2 = O.min_distance(h, [0, 0, 0], ex, 1, [1.25 / L, 0, 0], ex, 1)
    assert math.sqrt(m2) == pytest.approx(0.25, abs=1e-14)
    # across the periodic boundary
    _, m2, _, _ = O.min_distance(h, [0.49, 0, 0], ey, 1, [-0.49, 0, 0], ey, 1)
    assert math.sqrt(m2) == pytest.approx(0.02 * L, abs=1e-12)


# ------------------------------------------------------------ potential ----
def test_wca_force_closed_forms():
    eps, sig, fcut = 1.0, 1.0, 2000.0
    rc = 2 ** (1 / 6) * sig
    c12, c6 = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    ex = np.array([1.0, 0, 0])
    z = np.zeros(3)
    ok, f = O.pair_force(c12, c6, rc * rc, fcut, sig * ex, sig ** 2, z, z)
    assert ok == 1
    # force on obj1 points away from obj2, |F(sigma)| = 24 eps / sigma
    np.testing.assert_allclose(f, [-24 * eps / sig, 0, 0], rtol=1e-14)
    ok, f = O.pair_force(c12, c6, rc * rc, fcut, rc * ex, rc * rc, z, z)
    assert abs(f[0]) < 1e-12
    ok, f = O.pair_force(c12, c6, rc * rc, fcut, 1.2 * ex, 1.44, z, z)
    assert ok == 0                        # beyond the cutoff: untouched
    ok, f = O.pair_force(c12, c6, rc * rc, fcut, 0.5 * ex, 0.25, z, z)
    np.testing.assert_allclose(f, [-fcut, 0, 0])   # clamped at f_cutoff

    def U(r):
        return 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6) + eps
    for r in (0.95, 1.0, 1.05, 1.1):
        ok, f = O.pair_force(c12, c6, rc * rc, fcut, r * ex, r * r, z, z)
        h = 1e-6
        dudr = (U(r + h) - U(r - h)) / (2 * h)
        assert f[0] == pytest.approx(dudr, rel=1e-7)   # F1 = +dU/dr along r12


def test_max_force_fallback():
    fcut = 2000.0
    # overlapping centre lines: direction from the contacts
    c1 = np.array([0.1, 0, 0])
    c2 = np.array([0, 0.2, 0])
    ok, f = O.pair_force(1, 1, 1.3, fcut, np.zeros(3), 0.0, c1, c2)
    assert ok == 1
    d = c1 - c2
    np.testing.assert_allclose(f, fcut * d / np.linalg.norm(d))
    ok, f = O.pair_force(1, 1, 1.3, fcut, np.zeros(3), 0.0, np.zeros(3), np.zeros(3))
    assert ok == 0                        # no direction: force left as it was


# ------------------------------------------------------------ tridiagonal --
@pytest.mark.parametrize("n", [2, 3, 5, 31, 100, 255])
def test_thomas_against_dense_solve(n):
    rng = np.random.default_rng(n)
    off = rng.uniform(-1, 1, n - 1)
    diag = 2.0 + np.abs(np.concatenate([[0], off])) + np.abs(np.concatenate([off, [0]]))
    diag *= rng.uniform(1.0, 1.5, n)
    rhs = rng.normal(size=n)
    A = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    x = O.tridiag(off, diag, off.copy(), rhs)
    np.testing.assert_allclose(x, np.linalg.solve(A, rhs), rtol=1e-11, atol=1e-12)


# --------------------------------------------------------------- bending --
def _bonds_soa(u):
    return np.ascontiguousarray(np.asarray(u).T.ravel())


def test_bending_straight_filament():
    n, kappa, bl = 12, 50.0, 1.0
    u = np.tile(_unit([1, 2, 3]), (n - 1, 1))
    f, ke = O.bending_forces(_bonds_soa(u), kappa, bl)
    assert np.abs(f).max() < 1e-12 * kappa
    # det_t[i] = i+1, det_b[j] = n-j+1  =>  g_inv[i] = (i+1)(n-i-2)/n
    i = np.arange(n - 2)
    g = (i + 1) * (n - i - 2) / n
    np.testing.assert_allclose(ke, (kappa + bl * g) / bl ** 2, rtol=1e-14)


def test_bending_metric_is_inverse_of_G():
    rng = np.random.default_rng(1)
    for n in (3, 4, 6, 8):
        u = np.array([_unit(rng.normal(size=3)) for _ in range(n - 1)])
        kappa, bl = 7.0, 0.8
        _, ke = O.bending_forces(_bonds_soa(u), kappa, bl)
        g_inv = (ke * bl ** 2 - kappa) / bl
        cos = np.sum(u[:-1] * u[1:], axis=1)
        G = 2 * np.eye(n - 1) - np.diag(cos, 1) - np.diag(cos, -1)
        Ginv = np.linalg.inv(G)
        np.testing.assert_allclose(g_inv, [Ginv[i, i + 1] for i in range(n - 2)], rtol=1e-10)


def test_bending_kappa_part_is_minus_gradient():
    rng = np.random.default_rng(2)
    n, bl, kappa = 7, 1.0, 3.0
    u = np.array([_unit(np.array([1, 0, 0]) + 0.4 * rng.normal(size=3)) for _ in range(n - 1)])
    r = np.vstack([np.zeros(3), np.cumsum(bl * u, axis=0)])

    def energy(rr):
        uu = np.diff(rr, axis=0)
        uu /= np.linalg.norm(uu, axis=1)[:, None]
        return kappa / bl * np.sum(1 - np.sum(uu[:-1] * uu[1:], axis=1))
    f_k, _ = O.bending_forces(_bonds_soa(u), kappa, bl)
    f_0, _ = O.bending_forces(_bonds_soa(u), 0.0, bl)
    fk = (f_k - f_0).reshape(3, n).T
    h = 1e-6
    for k in range(n):
        for d in range(3):
            rp, rm = r.copy(), r.copy()
            rp[k, d] += h
            rm[k, d] -= h
            grad = (energy(rp) - energy(rm)) / (2 * h)
            assert fk[k, d] == pytest.approx(-grad, abs=1e-6)


# ----------------------------------------------------------- integrator --
def test_random_force_projection_removes_stretching():
    p, fil = c0_system(num=8, seed=9)
    o = O.OracleSystem(p, fil, n_threads=1)
    _, _, u, _ = o.bonds()
    o.step(1)  # midstep: unprojected noise, projection
    _, _, fr = o.sites()
    fr = fr.reshape(3, -1).T
    u = u.reshape(3, -1).T
    so = fil.site_off
    worst = 0.0
    scale = np.abs(fr).max()
    for f in range(fil.nf):
        s0, s1 = so[f], so[f + 1]
        b0 = s0 - f
        for i in range(s1 - s0 - 1):
            worst = max(worst, abs(np.dot(fr[s0 + i + 1] - fr[s0 + i], u[b0 + i])))
    assert worst < 1e-10 * scale


def test_tension_keeps_bond_lengths_to_second_order():
    """Deterministic midstep from a bent, force-free-of-noise state: the
    bond-length change must scale as dt^2 (a sign or index slip in the
    tension solve leaves an O(dt) stretch)."""
    p, fil = c0_system(num=6, seed=4)
    p.stoch_flag = 0
    devs = []
    for dt in (2e-4, 1e-4):
        p.delta = dt
        o = O.OracleSystem(p, fil, n_threads=1)
        _, _, _, l0 = o.bonds()
        o.step(1)
        _, _, _, l1 = o.bonds()
        devs.append(np.abs(l1 - l0).max())
    assert devs[0] > 0
    assert devs[0] / devs[1] == pytest.approx(4.0, rel=0.15)


def test_unphysical_displacement_is_an_error():
    p, fil = c0_system(num=8, seed=9)
    o = O.OracleSystem(p, fil, n_threads=1)
    o.step(2)                            # ends with a full step
    r, _, _ = o.sites()
    r[: fil.site_off[1]] += 2.0e4        # translate filament 0 by 2e4 in x
    o.set_sites(r)
    # step 3 checks the displacement (the last update was a full step):
    # 2 * (2e4)^2 > 10000 -> the reference throws
    assert o.step(1) == -1


def test_bond_forces_to_sites_keep_net_force_and_perpendicular_torque():
    """A14 (src/filament.cpp:992-1048): each bond's force F and torque T
    become two site forces whose sum is F and whose moment about the bond
    centre is the torque's part perpendicular to the bond, T - u(u.T);
    contributions of different bonds add."""
    from oracle.oracle_py import bond_site_forces
    rng = np.random.default_rng(5)
    n, bl = 7, 0.83
    nb = n - 1
    u = rng.normal(size=(3, nb))
    u /= np.linalg.norm(u, axis=0)
    F = rng.normal(size=(3, nb))
    T = rng.normal(size=(3, nb))
    total = bond_site_forces(u.ravel(), F.ravel(), T.ravel(), bl).reshape(3, n)
    acc = np.zeros((3, n))
    for i in range(nb):
        Fi = np.zeros((3, nb))
        Ti = np.zeros((3, nb))
        Fi[:, i] = F[:, i]
        Ti[:, i] = T[:, i]
        fs = bond_site_forces(u.ravel(), Fi.ravel(), Ti.ravel(), bl).reshape(3, n)
        acc += fs
        others = [j for j in range(n) if j not in (i, i + 1)]
        assert np.all(fs[:, others] == 0.0)
        fa, fb = fs[:, i], fs[:, i + 1]
        np.testing.assert_allclose(fa + fb, F[:, i], rtol=0, atol=1e-13)
        half = 0.5 * bl * u[:, i]  # site i at centre - half, site i+1 at centre + half
        moment = np.cross(-half, fa) + np.cross(half, fb)
        t_perp = T[:, i] - u[:, i] * np.dot(u[:, i], T[:, i])
        np.testing.assert_allclose(moment, t_perp, rtol=0, atol=1e-12)
    np.testing.assert_allclose(total, acc, rtol=0, atol=1e-12)


def test_bond_lengths_stay_within_the_renormalisation_threshold():
    """A21 (src/filament.cpp:840-951): after every step each bond length
    stays within 1e-3 relative of the filament's bond length (the update
    re-places the sites whenever a bond leaves that band)."""
    from tests.helpers import uneven_system
    p, fil = uneven_system()
    ora = O.OracleSystem(p, fil)
    so = np.asarray(fil.site_off)
    bl = fil.fil_array("bond_length")
    worst = 0.0
    for _ in range(20):
        assert ora.step(5) == 0
        r = ora.sites()[0]
        ns = len(r) // 3
        x = r.reshape(3, ns)
        for f in range(len(so) - 1):
            seg = x[:, so[f]:so[f + 1]]
            ln = np.linalg.norm(np.diff(seg, axis=1), axis=0)
            worst = max(worst, float(np.max(np.abs(ln - bl[f]) / bl[f])))
    print("worst relative bond-length deviation", worst)
    assert worst <= 1e-3


# ------------------------------------------- trajectory-level rules ----
# Closed-form scenarios (tests/scenarios.py) for the rules the reference's
# source fixes but no published vector does: the sticky pair force, the
# double-counted rebuild trigger, the alternating renormalisation and the
# MakePairs emission order.
from tests import scenarios as S  # noqa: E402


def test_sticky_force_persists_until_rebuild():
    p, fil = S.sticky_setup()
    S.run_sticky(O.OracleSystem(p, fil, n_threads=1))


def test_sticky_scenario_rejects_a_fresh_evaluation():
    """The scenario's expectation is not met by a pass that re-evaluates an
    out-of-range pair to zero: at h1 the closed form is 0, not F(h0)."""
    assert S.wca_force_magnitude(S.STICKY_H0) > 1.0
    assert S.STICKY_H1 > S.RCUT > S.STICKY_H0
    assert math.sqrt(0.25 + S.STICKY_H0 ** 2) > S.RCUT


def test_thermo_stress_energy_and_pressure_closed_form():
    p, fil = S.sticky_setup()
    S.run_thermo(O.OracleSystem(p, fil, n_threads=1))


def test_thermo_scenario_rejects_a_dropped_or_fresh_term():
    """The expectations differ from what a stress without the sticky steps
    (h0 f0 / 4 on zz), a fresh evaluation (0 at h1) or a missing density
    term would give."""
    f0 = S.wca_force_magnitude(S.STICKY_H0)
    assert S.STICKY_H0 * f0 > 1.0 and S.wca_energy(S.STICKY_H0) > 0.1
    assert S.wca_energy(S.RCUT) == pytest.approx(0.0, abs=1e-12)


def test_rebuild_trigger_counts_each_displacement_twice():
    p, fil = S.trigger_setup()
    S.run_trigger(O.OracleSystem(p, fil, n_threads=1), cycles=2)


def test_renormalisation_alternates_head_and_tail():
    p, fil, r0, u = S.renorm_setup()
    S.run_renorm(O.OracleSystem(p, fil, n_threads=1), r0, u)


@pytest.mark.parametrize("radius,n1d", [(1.8, 3), (2.4, 4), (3.0, 5), (3.5, 6)])
def test_candidate_order_is_the_reference_walk(radius, n1d):
    p = sc.params_init(radius, delta=1e-4, min_cell_length=1.0, seed=3)
    fil = S.make_filaments(S.random_sticks(60, 4, radius, seed=n1d))
    o = O.OracleSystem(p, fil, n_threads=2)
    assert o.counters()["n_cells_1d"] == n1d
    keys = o.cell_keys()
    b1, b2 = o.candidates()               # as emitted (the cull comes later)
    walk = S.reference_pair_walk(keys, n1d)
    assert len(walk) > 1000
    assert list(zip(b1.tolist(), b2.tolist())) == walk


def test_noise_buffer_of_zeros_is_the_noiseless_step():
    """A caller buffer of zero draws gives zero random forces: the same
    trajectory, bit for bit, as stoch_flag off (src/filament.cpp:518-548 with
    xi = 0), and a buffer equal to the counter draws the default one."""
    p, fil = c0_system(num=12, seed=4)
    quiet = sc.params_init(p.system_radius, delta=p.delta, min_cell_length=p.min_cell_length,
                           stoch_flag=0, seed=p.seed)
    a = O.OracleSystem(quiet, fil, n_threads=1)
    b = O.OracleSystem(p, fil, n_threads=1)
    b.set_noise_buffer(np.zeros(15 * 3 * fil.ns), 15)
    assert a.step(30) == 0 and b.step(30) == 0
    np.testing.assert_array_equal(a.sites()[0], b.sites()[0])
    c = O.OracleSystem(p, fil, n_threads=1)
    d = O.OracleSystem(p, fil, n_threads=1)
    xi = np.array([O.noise(p.seed, s, m) for m in range(15) for s in range(fil.ns)]).ravel()
    d.set_noise_buffer(xi, 15)
    assert c.step(30) == 0 and d.step(30) == 0
    np.testing.assert_array_equal(c.sites()[0], d.sites()[0])


def _bend_disorder(o, so):
    """Mean 1 - cos of consecutive bond directions over all filaments."""
    r = o.sites()[0].reshape(3, -1)
    c = []
    for f in range(len(so) - 1):
        u = np.diff(r[:, so[f]:so[f + 1]], axis=1)
        u /= np.linalg.norm(u, axis=0)
        c.append(1.0 - np.mean(np.sum(u[:, 1:] * u[:, :-1], axis=0)))
    return float(np.mean(c))


def test_stiff_chains_need_the_small_time_step():
    """Why the stiff (lp up to 1000 sigma) workloads run at dt 1e-5: eight
    64-site filaments with lp = 1000 sigma, bond length 1.  At dt 1e-5 the
    bending disorder <1 - cos theta> stays at the worm-like-chain value
    1 - (coth(lp/b) - b/lp) ~ b/lp = 1e-3; at dt 1e-4 the explicit bending
    update is unstable and the chains lose their orientation within 400
    steps (<1 - cos theta> of order 1, a random walk)."""
    wlc = 1.0 - (1.0 / np.tanh(1000.0) - 1.0 / 1000.0)
    out = {}
    for dt in (1e-5, 1e-4):
        p = sc.params_init(40.0, delta=dt, min_cell_length=1.0, seed=3)
        fil = sc.generate_filaments(p, [64] * 8, 63.0, 1000.0, 1.0, 2.0, seed=4)
        o = O.OracleSystem(p, fil, n_threads=1)
        assert o.step(400) == 0
        out[dt] = _bend_disorder(o, np.asarray(fil.site_off))
        o.close()
    print(out, wlc)
    assert 0.5 * wlc < out[1e-5] < 2.0 * wlc
    assert out[1e-4] > 0.3
