"""Thermodynamic output on the HIP path: the pair stress summed into the
pressure tensor (InteractionManager::CalculatePressure,
src/interaction_manager.cpp:534-559) and each bond's pair energy
(Object::AddPotential, :500-531), against the closed-form sticky scenario
and against the CPU oracle on a dense system with rebuilds.

FP64: bond energies are summed per bond in the reference's candidate order,
so they equal the oracle's bit for bit; the stress is summed per block and
then over blocks, a different order from the oracle's serial sum, so the
pressure tensor agrees to 1e-12 relative.  Mixed precision: energies and
stress from FP32 slot values, to 1e-5 of the largest bond energy and 1e-6
relative on the pressure."""
import numpy as np
import pytest

import simcore_amd as sc
from oracle.oracle_py import OracleSystem
from simcore_amd import workloads as W
from simcore_amd._lib import SimcoreError
from tests import scenarios as S

pytestmark = pytest.mark.gpu

STEPS = 60


@pytest.mark.parametrize("precision,graph", [("fp64", False), ("fp64", True),
                                             ("mixed", True)])
def test_thermo_closed_form(precision, graph):
    print(f"case: thermo sticky scenario, {precision}, graph={graph}", flush=True)
    p, fil = S.sticky_setup()
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=graph)
    if precision == "fp64":
        S.run_thermo(dev)
    else:
        S.run_thermo(dev, rtol=1e-5, atol=1e-6)
    dev.close()


def _dense():
    p, fil = W.dense(nf=400)[:2]
    return p, fil


@pytest.mark.parametrize("shape", ["per-bond-256", "per-bond-64", "balanced"])
@pytest.mark.parametrize("graph", [False, True])
def test_thermo_matches_oracle_fp64(shape, graph):
    """400 x 32 sites at phi 0.05, 60 steps with rebuilds: pressure tensor
    within 1e-12, bond energies bit-identical, for each pair-kernel shape
    (several wavefronts per block, one, and the balanced kernel)."""
    p, fil = _dense()
    print(f"case: thermo vs oracle fp64, {shape}, graph={graph}, {fil.nb} bonds", flush=True)
    ora = OracleSystem(p, fil)
    dev = sc.FilamentSystem(p, fil, precision="fp64", use_graph=graph)
    bs = {"per-bond-256": 256, "per-bond-64": 64, "balanced": 0}[shape]
    dev.set_launch_shape(pair_block=bs)
    dev.thermo_enable()
    ora.thermo_enable()
    dev.step(STEPS)
    assert ora.step(STEPS) == 0
    assert dev.counters()["n_rebuilds"] == ora.counters()["n_rebuilds"] >= 2
    ed, eo = dev.bond_energies(), ora.bond_energies()
    assert np.count_nonzero(eo) > fil.nb // 4
    np.testing.assert_array_equal(ed, eo)
    pd, po = dev.calculate_pressure(STEPS), ora.calculate_pressure(STEPS)
    print(pd["pressure_tensor"], po["pressure_tensor"])
    assert pd["volume"] == po["volume"]
    np.testing.assert_allclose(pd["pressure_tensor"], po["pressure_tensor"], rtol=1e-12,
                               atol=1e-12 * np.abs(po["pressure_tensor"]).max())
    np.testing.assert_allclose(pd["pressure"], po["pressure"], rtol=1e-12)
    # the virial is not negligible beside the density term
    rho = fil.nb / po["volume"]
    assert po["pressure"] - rho > 0.1 * rho
    # the sum restarts after each read
    dev.step(2)
    assert ora.step(2) == 0
    np.testing.assert_allclose(dev.calculate_pressure(2)["pressure_tensor"],
                               ora.calculate_pressure(2)["pressure_tensor"], rtol=1e-12,
                               atol=1e-12 * np.abs(po["pressure_tensor"]).max())
    dev.close()
    ora.close()


def test_thermo_mixed_within_tolerance():
    """The benchmarked mixed path (graph, spatial storage order) over 2
    steps from the same state: energies within 1e-5 of the largest, the
    pressure tensor within 1e-6 relative."""
    p, fil = _dense()
    print(f"case: thermo mixed vs oracle, {fil.nb} bonds", flush=True)
    ora = OracleSystem(p, fil)
    dev = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True, spatial_order=True)
    dev.thermo_enable()
    ora.thermo_enable()
    dev.step(2)
    assert ora.step(2) == 0
    eo = ora.bond_energies()
    np.testing.assert_allclose(dev.bond_energies(), eo, rtol=0, atol=1e-5 * eo.max())
    pd, po = dev.calculate_pressure(2), ora.calculate_pressure(2)
    np.testing.assert_allclose(pd["pressure_tensor"], po["pressure_tensor"], rtol=1e-6,
                               atol=1e-6 * np.abs(po["pressure_tensor"]).max())
    dev.close()
    ora.close()


def test_thermo_state_errors_and_autotune():
    """Reading before enabling is a state error; tuning trials neither add
    to the stress sum nor change the energies, and a launch shape changed
    while the sum accumulates does not lose blocks of it."""
    p, fil = _dense()
    print(f"case: thermo state and autotune, {fil.nb} bonds", flush=True)
    dev = sc.FilamentSystem(p, fil, precision="fp64", use_graph=True)
    with pytest.raises(SimcoreError):
        dev.calculate_pressure(1)
    with pytest.raises(SimcoreError):
        dev.bond_energies()
    ora = OracleSystem(p, fil)
    dev.thermo_enable()
    ora.thermo_enable()
    dev.step(10)
    assert ora.step(10) == 0
    e10 = dev.bond_energies()
    dev.autotune(steps=4)
    np.testing.assert_array_equal(dev.bond_energies(), e10)
    dev.set_launch_shape(pair_block=0)
    dev.step(6)
    dev.set_launch_shape(pair_block=512)
    dev.step(4)
    assert ora.step(10) == 0
    np.testing.assert_array_equal(dev.bond_energies(), ora.bond_energies())
    pd, po = dev.calculate_pressure(20), ora.calculate_pressure(20)
    np.testing.assert_allclose(pd["pressure_tensor"], po["pressure_tensor"], rtol=1e-12,
                               atol=1e-12 * np.abs(po["pressure_tensor"]).max())
    dev.thermo_enable(False)
    with pytest.raises(SimcoreError):
        dev.calculate_pressure(1)
    dev.close()
    ora.close()
