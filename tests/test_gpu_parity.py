"""Parity of the HIP path against the CPU oracle on the same inputs and the
same injected noise stream (C ABI calls through simcore_amd)."""
import numpy as np
import pytest

import simcore_amd as sc
from oracle.oracle_py import OracleSystem
from tests.helpers import c0_system, culled

pytestmark = pytest.mark.gpu

# Mixed-path position tolerance over the first steps: FP32 pair forces
# round to ~6e-8 relative, and WCA forces reach the 2000 kT/sigma cap at
# contact, so one step (dt/2 = 5e-5, gamma_perp ~ 0.4) can move a site by a
# few 1e-8 sigma against the FP64 path; 1e-7 sigma keeps a 3x margin.
POS_TOL_SHORT = 1e-7


def _pair(num=None, seed=7, precision="fp64", use_graph=False, capacity_factor=0.0):
    p, fil = c0_system(num=num, seed=seed)
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=use_graph,
                            capacity_factor=capacity_factor)
    ora = OracleSystem(p, fil)
    return p, fil, dev, ora


def test_initial_cell_keys_and_candidates_bit_exact():
    p, fil, dev, ora = _pair()
    print("case: C0 64x20 seed 7, load-time rebuild", flush=True)
    np.testing.assert_array_equal(dev.cell_keys(), ora.cell_keys())
    ob1, ob2 = culled(*ora.candidates(), fil.site_off)
    db1, db2 = dev.candidates()
    np.testing.assert_array_equal(db1, ob1)
    np.testing.assert_array_equal(db2, ob2)
    # bond state from the same site positions: bit-identical FP64
    for a, b in zip(dev.bonds(), ora.bonds()):
        np.testing.assert_array_equal(a, b)


def test_mixed_compact_bond_records_at_load():
    """Mixed path: 32-byte bond records (32-bit fixed-point offset inside the
    bond's cell, packed cell coordinates, octahedral orientation, two-float
    length).  Against the FP64 records of the CPU path: scaled positions
    within 1e-9 of a cell (1/35 of the box per cell here), orientations
    within 5e-9, lengths to 1e-12 relative, identical cell keys and
    candidates, centres from the FP64 sites bit-identical."""
    print("case: C0 64x20 seed 7 mixed, load-time records", flush=True)
    p, fil, dev, ora = _pair(precision="mixed")
    dpos, dsp, du, dln = dev.bonds()
    opos, osp, ou, oln = ora.bonds()
    np.testing.assert_array_equal(dpos, opos)
    d = np.abs(dsp - osp)
    d = np.minimum(d, 1.0 - d)  # periodic in scaled units
    n1d = dev.counters()["n_cells_1d"]
    print("max scaled-position error in cells", d.max() * n1d,
          "orientation", np.abs(du - ou).max(), "length", np.abs(dln / oln - 1).max())
    assert d.max() * n1d <= 1e-9
    assert np.abs(du - ou).max() <= 5e-9
    np.testing.assert_allclose(dln, oln, rtol=1e-12)
    np.testing.assert_array_equal(dev.cell_keys(), ora.cell_keys())
    ob1, ob2 = culled(*ora.candidates(), fil.site_off)
    db1, db2 = dev.candidates()
    np.testing.assert_array_equal(db1, ob1)
    np.testing.assert_array_equal(db2, ob2)


def test_fp64_one_step_bit_exact():
    p, fil, dev, ora = _pair()
    print("case: C0 fp64 1 step", flush=True)
    dev.step(1)
    assert ora.step(1) == 0
    fd, td = dev.bond_forces()
    fo, to = ora.bond_forces()
    np.testing.assert_array_equal(fd, fo)
    np.testing.assert_array_equal(td, to)
    np.testing.assert_array_equal(dev.tensions(), ora.tensions())
    for a, b in zip(dev.sites(), ora.sites()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(dev.slot_forces(), ora.slot_forces())


def test_fp64_100_steps_bit_exact():
    p, fil, dev, ora = _pair()
    print("case: C0 fp64 100 steps", flush=True)
    # cumulative 1, 3, 7, 21, 51, 99, 100.  After odd steps both sides hold
    # the same displacement accumulators (the device adds a full step's
    # displacement right after it, the CPU path at the next step's start).
    for chunk in (1, 2, 4, 14, 30, 48, 1):
        dev.step(chunk)
        assert ora.step(chunk) == 0
        cd, co = dev.counters(), ora.counters()
        assert cd["i_step"] == co["i_step"]
        assert cd["n_rebuilds"] == co["n_rebuilds"]
        np.testing.assert_array_equal(dev.cell_keys(), ora.cell_keys())
        db1, db2 = dev.candidates()
        ob1, ob2 = ora.candidates()
        np.testing.assert_array_equal(db1, ob1)
        np.testing.assert_array_equal(db2, ob2)
        for a, b in zip(dev.sites(), ora.sites()):
            np.testing.assert_array_equal(a, b)
        if cd["i_step"] % 2 == 1:
            np.testing.assert_array_equal(dev.dr_state(), ora.dr_state()[0])
    assert dev.counters()["n_rebuilds"] > 1


def test_graph_replay_matches_eager():
    print("case: C0 fp64 graph vs eager 60 steps", flush=True)
    _, _, eager, _ = _pair()
    _, _, graph, _ = _pair(use_graph=True)
    for _ in range(3):
        eager.step(20)
        graph.step(20)
        for a, b in zip(eager.sites(), graph.sites()):
            np.testing.assert_array_equal(a, b)
    assert eager.counters() == graph.counters()


@pytest.mark.parametrize("precision,spatial", [("fp64", False), ("mixed", True)])
def test_capacity_growth_mid_run_is_exact(precision, spatial):
    """A candidate capacity that holds the load-time list exactly overflows
    at the first rebuild after the box's contents are compressed by 0.75
    (about 2.4x the candidates): the step aborts on the device, the host
    grows the buffers, recaptures the graphs and re-runs the step.  Sites,
    tensions, counters and candidates stay bit-identical to a run that
    never overflowed, in graph mode, and the growth happened mid-run."""
    print(f"case: C0 {precision} capacity = load-time list, compressed after step 4",
          flush=True)
    p, fil = c0_system()
    ref = sc.FilamentSystem(p, fil, precision=precision, use_graph=True, spatial_order=spatial)
    n0 = ref.counters()["n_candidates"]
    small = sc.FilamentSystem(p, fil, precision=precision, use_graph=True,
                              spatial_order=spatial, capacity_factor=-float(n0))
    b0 = small.device_bytes()
    assert b0 < ref.device_bytes()
    ref.step(4)
    small.step(4)
    assert small.device_bytes() == b0
    r = ref.sites()[0] * 0.75
    ref.set_sites(r)
    small.set_sites(r)
    ref.step(1)
    small.step(1)
    c = small.counters()
    assert c["n_rebuilds"] == 2 and c["n_candidates"] > 2 * n0
    assert small.device_bytes() > b0, "the in-step growth path did not run"
    ref.step(35)
    small.step(35)
    for a, b in zip(ref.sites(), small.sites()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ref.tensions(), small.tensions())
    assert ref.counters() == small.counters()
    for a, b in zip(ref.candidates(), small.candidates()):
        np.testing.assert_array_equal(a, b)


def test_mixed_forces_within_tolerance():
    """FP32 pair math: per-bond force error relative to the FP64 force
    scale <= 1e-5 after one step from identical state."""
    print("case: C0 mixed 1 step", flush=True)
    p, fil, dev, ora = _pair(precision="mixed")
    dev.step(1)
    assert ora.step(1) == 0
    fd, td = dev.bond_forces()
    fo, to = ora.bond_forces()
    scale = np.linalg.norm(fo)
    assert scale > 0
    err = np.linalg.norm(fd - fo) / scale
    print("mixed relative force error", err)
    assert err <= 1e-5
    terr = np.linalg.norm(td - to) / np.linalg.norm(to)
    assert terr <= 1e-5


def test_unphysical_displacement_raises_at_the_same_step():
    print("case: C0 8 filaments, filament 0 moved by 2e4 after step 2", flush=True)
    p, fil, dev, ora = _pair(num=8, seed=9)
    dev.step(2)
    assert ora.step(2) == 0
    r = dev.sites()[0]
    r[: fil.site_off[1]] += 2.0e4
    dev.set_sites(r)
    ora.set_sites(r)
    assert ora.step(1) == -1
    with pytest.raises(sc.UnphysicalResult):
        dev.step(1)
    assert dev.counters()["i_step"] == 2


def test_set_sites_keeps_trajectories_identical():
    print("case: C0 small perturbation after a full step", flush=True)
    p, fil, dev, ora = _pair(num=16, seed=13)
    dev.step(6)
    assert ora.step(6) == 0
    r = dev.sites()[0]
    r[5] += 0.3
    dev.set_sites(r)
    ora.set_sites(r)
    dev.step(40)
    assert ora.step(40) == 0
    for a, b in zip(dev.sites(), ora.sites()):
        np.testing.assert_array_equal(a, b)
    assert dev.counters()["n_rebuilds"] == ora.counters()["n_rebuilds"]


def test_mixed_trajectory_within_tolerance():
    """Mixed path (FP32 pair math, FP64 integrator with the algebraically
    reduced forms and parallel cyclic reduction) against the FP64 CPU path
    from the same state and noise: positions within POS_TOL_SHORT after one
    step and 1e-5 sigma after 100 steps; candidate sets identical."""
    print("case: C0 mixed trajectory 1 and 100 steps", flush=True)
    p, fil, dev, ora = _pair(precision="mixed")
    dev.step(1)
    assert ora.step(1) == 0
    d1 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 1 step", d1)
    assert d1 <= POS_TOL_SHORT
    np.testing.assert_allclose(dev.tensions(), ora.tensions(), rtol=1e-5,
                               atol=1e-5 * np.abs(ora.tensions()).max())
    dev.step(99)
    assert ora.step(99) == 0
    d100 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 100 steps", d100)
    assert d100 <= 1e-5
    assert dev.counters()["n_rebuilds"] == ora.counters()["n_rebuilds"]
    db1, db2 = dev.candidates()
    ob1, ob2 = ora.candidates()
    np.testing.assert_array_equal(db1, ob1)
    np.testing.assert_array_equal(db2, ob2)


def _uneven(kind, precision, use_graph=False):
    from tests.helpers import long_system, uneven_system
    p, fil = uneven_system() if kind == "short" else long_system()
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=use_graph)
    return p, fil, dev


@pytest.mark.parametrize("kind", ["short", "long"])
def test_uneven_lengths_fp64_bit_exact(kind):
    """Every integrator length bin (3..256 sites) in one system: FP64 path
    bit-identical to the CPU path over 60 steps incl. rebuilds."""
    print(f"case: uneven {kind} fp64 60 steps", flush=True)
    p, fil, dev = _uneven(kind, "fp64")
    ora = OracleSystem(p, fil)
    dev.step(60)
    assert ora.step(60) == 0
    for a, b in zip(dev.sites(), ora.sites()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(dev.tensions(), ora.tensions())
    assert dev.counters()["n_rebuilds"] == ora.counters()["n_rebuilds"]
    db1, db2 = dev.candidates()
    ob1, ob2 = ora.candidates()
    np.testing.assert_array_equal(db1, ob1)
    np.testing.assert_array_equal(db2, ob2)


@pytest.mark.parametrize("kind", ["short", "long"])
def test_uneven_lengths_mixed_trajectory(kind):
    """Mixed path on every length bin against the FP64 CPU path: positions
    within POS_TOL_SHORT after one step and 1e-5 sigma after 100 steps,
    tensions within 1e-5 relative after one step."""
    print(f"case: uneven {kind} mixed 1 and 100 steps", flush=True)
    p, fil, dev = _uneven(kind, "mixed")
    ora = OracleSystem(p, fil)
    dev.step(1)
    assert ora.step(1) == 0
    d1 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 1 step", d1)
    assert d1 <= POS_TOL_SHORT
    np.testing.assert_allclose(dev.tensions(), ora.tensions(), rtol=1e-5,
                               atol=1e-5 * np.abs(ora.tensions()).max())
    dev.step(1)
    assert ora.step(1) == 0
    d2 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 2 steps", d2)
    assert d2 <= POS_TOL_SHORT
    dev.step(98)
    assert ora.step(98) == 0
    d100 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 100 steps", d100)
    assert d100 <= 1e-5
    assert dev.counters()["n_rebuilds"] == ora.counters()["n_rebuilds"]


def test_uneven_lengths_mixed_graph_matches_eager():
    print("case: uneven short mixed graph vs eager 31 steps", flush=True)
    p, fil, eager = _uneven("short", "mixed")
    _, _, graph = _uneven("short", "mixed", use_graph=True)
    eager.step(31)
    graph.step(31)
    for a, b in zip(eager.sites(), graph.sites()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(eager.tensions(), graph.tensions())


def test_spatial_order_getters_use_caller_order():
    """Filaments stored in space-filling-curve order: every getter still
    reports the caller's order (bit-identical FP64 bond state and cell keys
    at load, the same candidate set)."""
    print("case: C0 fp64 spatial order, load-time state", flush=True)
    p, fil = c0_system()
    dev = sc.FilamentSystem(p, fil, precision="fp64", spatial_order=True)
    ora = OracleSystem(p, fil)
    np.testing.assert_array_equal(dev.sites()[0], fil.r)
    np.testing.assert_array_equal(dev.cell_keys(), ora.cell_keys())
    for a, b in zip(dev.bonds(), ora.bonds()):
        np.testing.assert_array_equal(a, b)
    ob1, ob2 = culled(*ora.candidates(), fil.site_off)
    db1, db2 = dev.candidates()
    unordered = lambda a, b: set(map(frozenset, zip(a.tolist(), b.tolist())))
    assert unordered(db1, db2) == unordered(ob1, ob2)


@pytest.mark.parametrize("kind", ["c0", "short", "long"])
def test_spatial_order_mixed_trajectory(kind):
    """Mixed path with spatial storage order against the CPU path: same
    tolerances as in caller order (the noise is keyed by caller sites)."""
    print(f"case: {kind} mixed spatial order 1 and 100 steps", flush=True)
    from tests.helpers import long_system, uneven_system
    p, fil = {"c0": c0_system, "short": uneven_system, "long": long_system}[kind]()
    dev = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True, spatial_order=True)
    ora = OracleSystem(p, fil)
    dev.step(1)
    assert ora.step(1) == 0
    d1 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 1 step", d1)
    assert d1 <= POS_TOL_SHORT
    fd, _ = dev.bond_forces()
    fo, _ = ora.bond_forces()
    assert np.linalg.norm(fd - fo) / np.linalg.norm(fo) <= 1e-5
    dev.step(99)
    assert ora.step(99) == 0
    d100 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 100 steps", d100)
    assert d100 <= 1e-5
    assert dev.counters()["n_rebuilds"] == ora.counters()["n_rebuilds"]
    # same pairs; within a cell the storage order decides which bond is obj1
    db1, db2 = dev.candidates()
    ob1, ob2 = ora.candidates()
    unordered = lambda a, b: set(map(frozenset, zip(a.tolist(), b.tolist())))
    assert unordered(db1, db2) == unordered(ob1, ob2)


@pytest.mark.parametrize("precision,spatial", [("fp64", False), ("mixed", True)])
def test_checkpoint_resume_is_bit_identical(precision, spatial):
    """Save after 31 steps (odd: next is a full step), continue 40 steps;
    a fresh handle restored from the checkpoint reaches the same state."""
    print(f"case: C0 {precision} spatial={spatial} checkpoint at 31, +40 steps", flush=True)
    p, fil = c0_system()
    a = sc.FilamentSystem(p, fil, precision=precision, use_graph=True, spatial_order=spatial)
    a.step(31)
    blob = a.checkpoint()
    a.step(40)
    b = sc.FilamentSystem(p, fil, precision=precision, use_graph=True, spatial_order=spatial)
    b.restore(blob)
    assert b.counters()["i_step"] == 31
    b.step(40)
    for x, y in zip(a.sites(), b.sites()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.tensions(), b.tensions())
    assert a.counters() == b.counters()
    if precision == "fp64":
        ora = OracleSystem(p, fil)
        assert ora.step(71) == 0
        np.testing.assert_array_equal(b.sites()[0], ora.sites()[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_leaves_trajectory_unchanged(precision):
    print(f"case: C0 {precision} autotune after 31 steps, then 40 steps", flush=True)
    from tests.helpers import uneven_system
    p, fil = uneven_system()
    a = sc.FilamentSystem(p, fil, precision=precision, use_graph=True)
    b = sc.FilamentSystem(p, fil, precision=precision, use_graph=True)
    a.step(31)
    b.step(31)
    rep = b.autotune(steps=6)
    print(rep)
    # 8 pair shapes, then 3 waves per block x side streams, then 3 counts of
    # overlapped bond ranges; the mixed path's packed integrator is tried as
    # one launch (no side stream: no filament above 128 sites) and as one
    # launch per lane-group width (two here, 16- and 32-lane groups: one
    # side stream); then the four fastest shapes are timed twice more
    expect = (23 if precision == "fp64" else 8 + 3 * 1 + 3 * 2 + 3) + 8
    assert rep["n_trials"] == expect and rep["step_ms_best"] <= rep["step_ms_start"]
    assert b.counters()["i_step"] == 31
    a.step(40)
    b.step(40)
    for x, y in zip(a.sites(), b.sites()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("layout", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["short", "long"])
def test_integrator_layouts_within_tolerance(kind, layout):
    """Both integrator layouts for filaments of up to 64 sites (one site per
    lane, two sites per lane -- the default), with non-default block shapes,
    against the CPU path at the mixed tolerances."""
    print(f"case: {kind} mixed, layout {layout}, 1 and 100 steps", flush=True)
    from tests.helpers import long_system, uneven_system
    p, fil = uneven_system() if kind == "short" else long_system()
    dev = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    dev.set_launch_shape(pair_block=128, pair_xcd=1, integrate_waves=2,
                         integrate_layout=layout)
    ora = OracleSystem(p, fil)
    dev.step(1)
    assert ora.step(1) == 0
    d1 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 1 step", d1)
    assert d1 <= POS_TOL_SHORT
    np.testing.assert_allclose(dev.tensions(), ora.tensions(), rtol=1e-5,
                               atol=1e-5 * np.abs(ora.tensions()).max())
    dev.step(99)
    assert ora.step(99) == 0
    d100 = np.abs(dev.sites()[0] - ora.sites()[0]).max()
    print("max |dr| after 100 steps", d100)
    assert d100 <= 1e-5


@pytest.mark.parametrize("spatial", [False, True])
@pytest.mark.parametrize("kind", ["short", "long", "c0"])
def test_packed_integrator_matches_length_bins(kind, spatial):
    """The packed layout (every filament of up to 128 sites in a lane segment
    of its own width, several lengths per wavefront, lane moves chosen per
    wavefront; one launch, or one per lane-group width) runs the
    length-binned layout's arithmetic: sites, tensions and counters
    bit-identical over 41 steps, with and without side streams."""
    print(f"case: {kind} mixed spatial={spatial}, packed vs binned integrator, 41 steps",
          flush=True)
    from tests.helpers import long_system, uneven_system
    p, fil = {"c0": c0_system, "short": uneven_system, "long": long_system}[kind]()
    runs = []
    for layout, streams in ((1, 0), (2, 0), (2, 1), (3, 0), (3, 1)):
        s = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True, spatial_order=spatial)
        s.set_launch_shape(pair_block=0, integrate_waves=2, integrate_layout=layout,
                           integrate_streams=streams)
        s.step(41)
        runs.append((s.sites(), s.tensions(), s.counters()))
        s.close()
    for j in range(1, len(runs)):
        for x, y in zip(runs[0][0], runs[j][0]):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(runs[0][1], runs[j][1])
        assert runs[0][2] == runs[j][2]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_balanced_pair_kernel_is_bit_identical(precision):
    """The load-balanced pair kernel sums the same contributions in the same
    order as the per-bond kernel: identical results; FP64 also against the
    CPU path."""
    print(f"case: uneven short {precision} balanced pair kernel, 41 steps", flush=True)
    from tests.helpers import uneven_system
    p, fil = uneven_system()
    a = sc.FilamentSystem(p, fil, precision=precision, use_graph=True)
    b = sc.FilamentSystem(p, fil, precision=precision, use_graph=True)
    b.set_launch_shape(pair_block=0, pair_xcd=1)
    a.step(41)
    b.step(41)
    for x, y in zip(a.sites(), b.sites()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a.bond_forces(), b.bond_forces()):
        np.testing.assert_array_equal(x, y)
    if precision == "fp64":
        ora = OracleSystem(p, fil)
        assert ora.step(41) == 0
        np.testing.assert_array_equal(b.sites()[0], ora.sites()[0])



@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_length_bins_on_side_streams_are_bit_identical(precision):
    """The integrator's length bins forked onto side streams (graph and
    eager) give the same results as one stream."""
    print(f"case: uneven short {precision}, bins on side streams, 41 steps", flush=True)
    from tests.helpers import uneven_system
    p, fil = uneven_system()
    runs = []
    for streams, graph in ((0, True), (1, True), (3, True), (3, False)):
        s = sc.FilamentSystem(p, fil, precision=precision, use_graph=graph)
        s.set_launch_shape(pair_block=0, integrate_waves=2, integrate_streams=streams)
        s.step(41)
        runs.append((s.sites(), s.tensions(), s.counters()))
        s.close()
    for j in (1, 2, 3):
        for x, y in zip(runs[0][0], runs[j][0]):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(runs[0][1], runs[j][1])
        assert runs[0][2] == runs[j][2]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["uneven2000", "long"])
def test_overlapped_ranges_are_bit_identical(precision, kind):
    """The pair pass cut into bond ranges with the integrator of finished
    ranges on a side stream (graph and eager; 2, 4 and 8 ranges), and cut
    after the long filaments with their integrator beside the rest of the
    pass, gives the results of the sequential order: sites, tensions,
    counters and bond forces identical over 41 steps with rebuilds; FP64
    equal to the CPU path."""
    from simcore_amd import workloads as W
    from tests.helpers import long_system
    if kind == "uneven2000":
        p, fil = W.uneven(nf=2000)[:2]
    else:
        p, fil = long_system()
    print(f"case: {kind} {precision}, overlapped pair/integrator ranges, 41 steps", flush=True)
    runs = []
    for parts, graph, early in ((1, True, 0), (2, True, 0), (4, True, 0), (8, True, 0),
                                (4, False, 0), (1, True, 1), (1, False, 1)):
        s = sc.FilamentSystem(p, fil, precision=precision, use_graph=graph,
                              spatial_order=precision == "mixed")
        s.set_launch_shape(pair_block=0, pair_xcd=1, integrate_waves=2,
                           integrate_layout=2 if precision == "mixed" else 1, pair_parts=parts,
                           early_long=early)
        assert s.launch_shape()["pair_parts"] == parts
        assert s.launch_shape()["early_long"] == early
        s.step(41)
        runs.append((s.sites(), s.tensions(), s.counters(), s.bond_forces()))
        s.close()
    assert runs[0][2]["n_rebuilds"] >= 2 or kind == "long"
    for j in range(1, len(runs)):
        for x, y in zip(runs[0][0], runs[j][0]):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(runs[0][1], runs[j][1])
        assert runs[0][2] == runs[j][2]
        for x, y in zip(runs[0][3], runs[j][3]):
            np.testing.assert_array_equal(x, y)
    if precision == "fp64":
        ora = OracleSystem(p, fil)
        assert ora.step(41) == 0
        np.testing.assert_array_equal(runs[2][0][0], ora.sites()[0])


def test_pair_timing_probe_leaves_state_unchanged():
    """sc_time_pair repeats the pair pass on an unchanged state: the run
    continues exactly as one that was never timed."""
    print("case: uneven short mixed, 21 steps, pair timing probe, 20 steps", flush=True)
    from tests.helpers import uneven_system
    p, fil = uneven_system()
    a = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    b = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    a.step(21)
    b.step(21)
    assert b.time_pair(5) > 0.0
    a.step(20)
    b.step(20)
    for x, y in zip(a.sites(), b.sites()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a.bond_forces(), b.bond_forces()):
        np.testing.assert_array_equal(x, y)
