"""The closed-form scenarios of tests/scenarios.py on the HIP path: sticky
pair forces, the double-counted rebuild trigger, the alternating
renormalisation and the MakePairs emission order, against the same closed
forms the CPU oracle is pinned by."""
import numpy as np
import pytest

import simcore_amd as sc
from tests import scenarios as S
from tests.helpers import culled

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,graph", [("fp64", False), ("fp64", True),
                                             ("mixed", True)])
def test_sticky_force_persists_until_rebuild(precision, graph):
    print(f"case: sticky pair scenario, {precision}, graph={graph}", flush=True)
    p, fil = S.sticky_setup()
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=graph)
    if precision == "fp64":
        S.run_sticky(dev)
    else:
        # FP32 slot and sums: the closed form to FP32 rounding
        S.run_sticky(dev, rtol=1e-6, atol=1e-5)
    dev.close()


@pytest.mark.parametrize("precision,graph", [("fp64", False), ("fp64", True),
                                             ("mixed", True)])
def test_rebuild_trigger_counts_each_displacement_twice(precision, graph):
    """The device folds a full step's displacement right after that step,
    the reference at the next step's start: the accumulators agree after
    odd steps, which is where the scenario reads them."""
    print(f"case: rebuild trigger scenario, {precision}, graph={graph}", flush=True)
    p, fil = S.trigger_setup()
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=graph)
    S.run_trigger(dev, cycles=2, check_acc_every_step=False)
    dev.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_renormalisation_alternates_head_and_tail(precision):
    print(f"case: renormalisation scenario, {precision}", flush=True)
    p, fil, r0, u = S.renorm_setup()
    dev = sc.FilamentSystem(p, fil, precision=precision, use_graph=True)
    S.run_renorm(dev, r0, u, tol=1e-14 if precision == "fp64" else 1e-12)
    dev.close()


@pytest.mark.parametrize("radius,n1d", [(1.8, 3), (2.4, 4), (3.0, 5), (3.5, 6)])
def test_candidate_order_is_the_reference_walk(radius, n1d):
    """Caller storage order: the device's list is the reference walk with
    the site-sharing pairs culled (the device culls at emission)."""
    print(f"case: MakePairs walk on {n1d}^3 cells", flush=True)
    p = sc.params_init(radius, delta=1e-4, min_cell_length=1.0, seed=3)
    fil = S.make_filaments(S.random_sticks(60, 4, radius, seed=n1d))
    dev = sc.FilamentSystem(p, fil, precision="fp64")
    assert dev.counters()["n_cells_1d"] == n1d
    walk = S.reference_pair_walk(dev.cell_keys(), n1d)
    w1 = np.array([a for a, _ in walk], dtype=np.int32)
    w2 = np.array([b for _, b in walk], dtype=np.int32)
    e1, e2 = culled(w1, w2, fil.site_off)
    d1, d2 = dev.candidates()
    np.testing.assert_array_equal(d1, e1)
    np.testing.assert_array_equal(d2, e2)
    dev.close()
