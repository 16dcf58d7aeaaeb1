"""Closed-form scenarios for the trajectory-level rules of the reference's
step: sticky pair forces, the double-counted rebuild trigger, the
alternating bond renormalisation and the MakePairs emission order.

Each scenario drives a system object with the interface shared by
oracle.oracle_py.OracleSystem (CPU) and simcore_amd.FilamentSystem (HIP):
step(n), sites(), set_sites(r), bond_forces(), counters(), dr_state(),
candidates(), cell_keys().  The expected values come from the physics of
the set-up (closed forms) and from the reference's rules as its source
states them, never from either implementation."""
import itertools
import math

import numpy as np

import simcore_amd as sc
from simcore_amd._lib import ScFilamentParams

RCUT = 2.0 ** (1.0 / 6.0)


def make_filaments(sites, bond_length=1.0, lp=50.0, diameter=1.0, friction_ratio=2.0):
    """Filaments from explicit site lists (each an (n, 3) array), with the
    frictions of Filament::SetDiffusion (src/filament.cpp:320-348):
    gamma_perp = 4L / (3 n ln(L/d)), gamma_par = gamma_perp / friction_ratio."""
    nf = len(sites)
    n = np.array([len(s) for s in sites], dtype=np.int32)
    site_off = np.zeros(nf + 1, dtype=np.int32)
    site_off[1:] = np.cumsum(n)
    ns = int(site_off[-1])
    r = np.empty(3 * ns)
    fil = (ScFilamentParams * nf)()
    for f, s in enumerate(sites):
        s = np.asarray(s, dtype=np.float64)
        for k in range(3):
            r[k * ns + site_off[f]: k * ns + site_off[f + 1]] = s[:, k]
        L = (len(s) - 1) * bond_length
        gp = 4.0 * L / (3.0 * len(s) * math.log(L / diameter))
        fil[f].bond_length = bond_length
        fil[f].persistence_length = lp
        fil[f].gamma_perp = gp
        fil[f].gamma_par = gp / friction_ratio
        fil[f].diameter = diameter
    return sc.Filaments(site_off, r, fil)


def soa(sites):
    """(n, 3) site list(s) -> the ABI's SoA layout."""
    s = np.concatenate([np.asarray(x, dtype=np.float64) for x in sites])
    return np.ascontiguousarray(s.T).ravel()


def wca_force_magnitude(r, eps=1.0, sig=1.0):
    """|F| of the WCA pair at distance r (repulsive), -dU/dr of
    U = 4 eps ((sig/r)^12 - (sig/r)^6) + eps."""
    x6 = (sig / r) ** 6
    return 24.0 * eps / r * (2.0 * x6 * x6 - x6)


# ------------------------------------------------------------- sticky ----
STICKY_H0, STICKY_H1 = 1.05, 1.25


def sticky_setup():
    """Filament A along x, filament B along y crossing over A's first bond
    at height h: only the pair (A0, B0) is within the WCA cutoff (the other
    cross pairs meet at sqrt(1/4 + h^2) > r_c), with the minimum distance
    between the two bond centres, so the closed-form force is exact."""
    p = sc.params_init(10.0, delta=1e-4, min_cell_length=1.0, stoch_flag=0, seed=5)
    fil = make_filaments(sticky_sites(STICKY_H0), lp=20.0)
    return p, fil


def sticky_sites(h, shift_b=(0.0, 0.0, 0.0)):
    a = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (2.0, 0.0, 0.0)]
    b = [(0.5 + shift_b[0], y + shift_b[1], h + shift_b[2]) for y in (-0.5, 0.5, 1.5)]
    return [np.array(a), np.array(b)]


def run_sticky(sys_, rtol=1e-12, atol=1e-12):
    """src/interaction_manager.cpp:394 returns before touching the pair's
    force when dr^2 > r_c^2, and :469-470 / :500-517 apply the pair's stored
    force every step: a pair that leaves the cutoff keeps pushing with the
    force of its last in-range evaluation until a rebuild creates fresh
    (zero-force) pair records.  Bonds: A0 = 0, A1 = 1, B0 = 2, B1 = 3."""
    f0 = wca_force_magnitude(STICKY_H0)
    expect = np.zeros((3, 4))
    expect[2, 0], expect[2, 2] = -f0, +f0       # A0 pushed down, B0 up
    sys_.step(1)                                 # midstep: pair at h0 in range
    F, T = (x.reshape(3, -1) for x in sys_.bond_forces())
    np.testing.assert_allclose(F, expect, rtol=rtol, atol=atol)
    np.testing.assert_allclose(T, 0.0, atol=atol)
    rebuilds = sys_.counters()["n_rebuilds"]
    # separate the bonds beyond the cutoff: every later step keeps the
    # h0 force on both bonds (a fresh evaluation would give exactly 0)
    moved = soa(sticky_sites(STICKY_H1))
    for k in (2, 3, 4):
        sys_.set_sites(moved)
        sys_.step(1)
        F, _ = (x.reshape(3, -1) for x in sys_.bond_forces())
        np.testing.assert_allclose(F, expect, rtol=rtol, atol=atol, err_msg=f"step {k}")
        assert sys_.counters()["n_rebuilds"] == rebuilds
    # a displacement past the rebuild threshold: the step rebuilds the pair
    # list, the new records start at zero force and stay there out of range
    sys_.set_sites(soa(sticky_sites(STICKY_H1, shift_b=(0.5, 0.0, 0.0))))
    sys_.step(1)
    assert sys_.counters()["n_rebuilds"] == rebuilds + 1
    F, T = (x.reshape(3, -1) for x in sys_.bond_forces())
    assert np.all(F == 0.0) and np.all(T == 0.0)


def wca_energy(r, eps=1.0, sig=1.0):
    """U(r) = 4 eps ((sig/r)^12 - (sig/r)^6) + eps, the shifted WCA energy."""
    x6 = (sig / r) ** 6
    return 4.0 * eps * (x6 * x6 - x6) + eps


def run_thermo(sys_, rtol=1e-12, atol=1e-12):
    """The thermo output on the sticky set-up.  Each step adds every pair's
    stored stress -dr_i F_j to InteractionManager::stress_ and its stored
    energy to both bonds (src/interaction_manager.cpp:500-531; the stress
    and pote are set where the force is, include/simcore/
    wca_potential.hpp:28-31, so they are sticky too).  The one pair, A0
    below B0 at distance h0 along z with the force -f0 z on A0, has stress
    diag(0, 0, h0 f0) and energy U(h0) on A0 and B0; after n_thermo = 4
    steps CalculatePressure (:534-559) gives diag(n_b/V) + diag(0, 0,
    4 h0 f0) / (4 V) with n_b = 4 bonds and V = 20^3; a rebuild clears both."""
    f0 = wca_force_magnitude(STICKY_H0)
    u0 = wca_energy(STICKY_H0)
    nb, vol = 4, 20.0 ** 3
    sys_.thermo_enable()
    e_expect = np.array([u0, 0.0, u0, 0.0])
    sys_.step(1)
    np.testing.assert_allclose(sys_.bond_energies(), e_expect, rtol=rtol, atol=atol)
    moved = soa(sticky_sites(STICKY_H1))
    for k in (2, 3, 4):
        sys_.set_sites(moved)
        sys_.step(1)
        np.testing.assert_allclose(sys_.bond_energies(), e_expect, rtol=rtol, atol=atol,
                                   err_msg=f"step {k}")
    out = sys_.calculate_pressure(4)
    assert out["volume"] == vol
    pt = np.diag([nb / vol, nb / vol, (nb + STICKY_H0 * f0) / vol])
    np.testing.assert_allclose(out["pressure_tensor"], pt, rtol=rtol, atol=atol)
    np.testing.assert_allclose(out["pressure"], np.trace(pt) / 3, rtol=rtol)
    rebuilds = sys_.counters()["n_rebuilds"]
    sys_.set_sites(soa(sticky_sites(STICKY_H1, shift_b=(0.5, 0.0, 0.0))))
    sys_.step(1)
    assert sys_.counters()["n_rebuilds"] == rebuilds + 1
    assert np.all(sys_.bond_energies() == 0.0)
    out = sys_.calculate_pressure(1)
    np.testing.assert_allclose(out["pressure_tensor"], np.eye(3) * nb / vol, rtol=rtol,
                               atol=atol)


# ------------------------------------------------------------ trigger ----
TRIGGER_R = 10.0


def trigger_setup():
    """One straight 3-site filament, no noise: nothing moves it but the rigid
    shifts the scenario applies (its two bonds share a site, so it has no
    pair; a straight filament has no bending force and zero tension)."""
    p = sc.params_init(TRIGGER_R, delta=1e-4, min_cell_length=1.0, stoch_flag=0, seed=5)
    u = np.array([2.0, 3.0, 6.0]) / 7.0
    fil = make_filaments([np.array([i * u for i in range(3)])], lp=20.0)
    return p, fil


def trigger_threshold(n1d):
    cell_len = 2.0 * TRIGGER_R / n1d
    return 0.25 * cell_len * cell_len


def run_trigger(sys_, cycles=2, check_acc_every_step=True):
    """Rebuild rule of src/interaction_manager.cpp:284-321 with
    Mesh::UpdateDrTot / GetDrTot (src/mesh.cpp:533-556) and
    Object::UpdateDrTot / GetDrTot (src/object.cpp:274-287): each check adds
    |r - r0|^2 to a site's accumulator TWICE (UpdateDrTot, then again inside
    GetDrTot), the accumulator is cumulative until a rebuild zeroes it, the
    check runs only at steps that follow a full step, and a rebuild happens
    when the maximum exceeds 0.25 cell_len^2.  The filament is shifted
    rigidly by delta after every full step, so check m sees |r - r0| =
    m delta and the accumulator holds 2 delta^2 sum_{j<=m} j^2:
      double count   2 delta^2 (1, 5, 14, ...)  -> rebuild at check 3
      single count     delta^2 (1, 5, 14, 30)   -> check 4
      not cumulative 2 delta^2 (1, 4, 9, 16)    -> check 4
    with delta^2 = thr / 22.  Checks after midsteps as well would double
    the accumulator at even steps (caught by the accumulator values)."""
    n1d = sys_.counters()["n_cells_1d"]
    thr = trigger_threshold(n1d)
    d2 = thr / 22.0
    assert 2 * d2 * 5 < thr < 2 * d2 * 14          # double count: check 3
    assert d2 * 14 < thr < d2 * 30                 # single count: check 4
    assert 2 * d2 * 9 < thr < 2 * d2 * 16          # not cumulative: check 4
    step_dir = np.array([1.0, -1.0, 1.0]) / math.sqrt(3.0)
    delta = math.sqrt(d2) * step_dir
    r = sys_.sites()[0].reshape(3, -1).copy()
    rebuilds = sys_.counters()["n_rebuilds"]
    sys_.step(2)                                   # midstep + full step, no motion
    step = 2
    for cycle in range(cycles):
        shifts = 0
        while True:
            r += delta[:, None]
            shifts += 1
            sys_.set_sites(r.ravel())
            sys_.step(1)                           # a midstep: the check runs first
            step += 1
            c = sys_.counters()
            acc = sys_.dr_state()
            acc = acc[0] if isinstance(acc, tuple) else acc
            expect = 2.0 * d2 * sum(j * j for j in range(1, shifts + 1))
            if shifts < 3:
                assert c["n_rebuilds"] == rebuilds, f"cycle {cycle} check {shifts}: early rebuild"
                np.testing.assert_allclose(acc, expect, rtol=1e-12,
                                           err_msg=f"accumulator at check {shifts}")
            else:
                assert c["n_rebuilds"] == rebuilds + 1, f"cycle {cycle}: no rebuild at check 3"
                assert np.all(acc == 0.0), "a rebuild restarts the accumulator"
                rebuilds += 1
            sys_.step(1)                           # full step: no check before it
            step += 1
            if check_acc_every_step and shifts < 3:
                acc2 = sys_.dr_state()
                acc2 = acc2[0] if isinstance(acc2, tuple) else acc2
                np.testing.assert_allclose(acc2, expect, rtol=1e-12,
                                           err_msg="no check after a midstep")
            if shifts == 3:
                break
    return step


# ------------------------------------------------------ renormalisation ----
def renorm_setup():
    """One straight 3-site filament along (2,3,6)/7, bond length 1, no noise
    (no pair, no bending force, zero tension: only the renormalisation moves
    it)."""
    p = sc.params_init(10.0, delta=1e-4, min_cell_length=1.0, stoch_flag=0, seed=5)
    u = np.array([2.0, 3.0, 6.0]) / 7.0
    r0 = np.array([[0.3, -0.2, 0.1] + i * u for i in range(3)])
    return p, make_filaments([r0], lp=20.0), r0, u


def run_renorm(sys_, r0, u, tol=1e-14):
    """src/filament.cpp:897-926: when a bond leaves 1e-3 of its length the
    switch normalize_switch_ (initially true, include/simcore/filament.hpp)
    is flipped FIRST; true re-places the sites from tail to head (site 0
    kept, site i = site i-1 + l u_{i-1}), false from head to tail (last
    site kept, site i-1 = site i - l u_{i-1}).  Bond 1 stretched by 2e-3:
      head -> tail keeps the stretched head: all sites move by 2e-3 l u
      tail -> head keeps the tail: the original sites
    The full step restarts from the midstep-start positions (still
    stretched), so a midstep/full-step pair renormalises twice."""
    stretched = r0.copy()
    stretched[2] += 2e-3 * u
    sys_.step(2)
    sys_.set_sites(soa([stretched]))
    sys_.step(1)    # first renormalisation: switch true -> false: head to tail
    got = sys_.sites()[0].reshape(3, -1).T
    np.testing.assert_allclose(got, r0 + 2e-3 * u, rtol=0, atol=tol,
                               err_msg="first renormalisation keeps the head")
    sys_.step(1)    # second: false -> true: tail to head
    got = sys_.sites()[0].reshape(3, -1).T
    np.testing.assert_allclose(got, r0, rtol=0, atol=tol,
                               err_msg="second renormalisation keeps the tail")
    sys_.step(2)    # bond lengths exact: nothing moves
    got2 = sys_.sites()[0].reshape(3, -1).T
    np.testing.assert_allclose(got2, r0, rtol=0, atol=tol)
    # and the third flips back: head to tail
    sys_.set_sites(soa([stretched]))
    sys_.step(1)
    got = sys_.sites()[0].reshape(3, -1).T
    np.testing.assert_allclose(got, r0 + 2e-3 * u, rtol=0, atol=tol,
                               err_msg="third renormalisation keeps the head")


# --------------------------------------------------- MakePairs order ----
def reference_pair_walk(keys, n1d):
    """Candidate pairs in the order CellList::MakePairs emits them
    (src/cell_list.cpp:127-137, 206-262; src/cell.cpp:14-70), written from
    the reference's loop structure: cells in x, y, z order (x outermost);
    each occupied cell pairs its own objects i < j in insertion (bond)
    order, then every object of the home cell with every object of each
    neighbour cell, neighbours in the order AssignCellNeighbors(false)
    inserts them -- for zp in (z, z+1), yp from (y if zp == z else y-1) to
    y+1, xp from (x+1 if yp == y and zp == z else x-1) to x+1, periodic
    wrap, the cell itself skipped."""
    n = n1d
    cells = {}
    for b, k in enumerate(keys.tolist()):
        cells.setdefault(k, []).append(b)

    def neighbours(x, y, z):
        out = []
        for zp in range(z, z + 2):
            nz = zp % n
            y_begin = y - 1 if zp > z else y
            for yp in range(y_begin, y + 2):
                ny = yp % n
                x_begin = x - 1 if (yp > y or zp > z) else x + 1
                for xp in range(x_begin, x + 2):
                    nx = xp % n
                    if (nx, ny, nz) == (x, y, z):
                        continue
                    out.append((nx * n + ny) * n + nz)
        return out

    pairs = []
    for x, y, z in itertools.product(range(n), repeat=3):
        home = cells.get((x * n + y) * n + z)
        if not home:
            continue
        for i in range(len(home) - 1):
            for j in range(i + 1, len(home)):
                pairs.append((home[i], home[j]))
        for k in neighbours(x, y, z):
            other = cells.get(k)
            if not other:
                continue
            for a in home:
                for b in other:
                    pairs.append((a, b))
    return pairs


def random_sticks(n_filaments, n_sites, radius, seed):
    """Filaments of n_sites sites (bond length 1) at uniform random positions
    and orientations in a cube of half side `radius`, without overlap
    rejection (the pair walk does not depend on the physics)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_filaments):
        c = rng.uniform(-radius, radius, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        out.append(np.array([c + i * u for i in range(n_sites)]))
    return out
