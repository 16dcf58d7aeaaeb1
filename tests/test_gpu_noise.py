"""Caller-supplied thermal noise (sc_set_noise, buffer mode): a buffer equal
to the counter stream's draws reproduces the default trajectory bit for bit;
an arbitrary buffer drives the HIP path and the CPU oracle identically; the
buffer's end is enforced."""
import numpy as np
import pytest

import simcore_amd as sc
from oracle import oracle_py as O
from tests.helpers import c0_system, uneven_system

pytestmark = pytest.mark.gpu


def counter_draws(seed, n_sites, m0, n_mid):
    """The counter stream's xi for midsteps m0 .. m0+n_mid-1, midstep-major,
    3 per site (the oracle's restatement of the stream)."""
    out = np.empty((n_mid, n_sites, 3))
    for j in range(n_mid):
        for s in range(n_sites):
            out[j, s] = O.noise(seed, s, m0 + j)
    return out.ravel()


@pytest.mark.parametrize("precision,spatial", [("fp64", False), ("mixed", True)])
def test_buffer_of_the_counter_draws_is_bit_identical(precision, spatial):
    print(f"case: C0 {precision} spatial={spatial}, noise buffer = counter draws, 40 steps",
          flush=True)
    p, fil = c0_system(num=24, seed=5)
    a = sc.FilamentSystem(p, fil, precision=precision, use_graph=True, spatial_order=spatial)
    b = sc.FilamentSystem(p, fil, precision=precision, use_graph=True, spatial_order=spatial)
    a.step(3)
    b.step(3)
    # after 3 steps the next midstep is step 5, midstep index 2
    b.set_noise(counter_draws(p.seed, fil.ns, 2, 19))
    a.step(37)
    b.step(37)
    for x, y in zip(a.sites(), b.sites()):
        np.testing.assert_array_equal(x, y)
    assert a.counters() == b.counters()


def test_arbitrary_buffer_matches_the_oracle():
    print("case: uneven short fp64, numpy noise buffer, 40 steps vs oracle", flush=True)
    p, fil = uneven_system()
    rng = np.random.default_rng(9)
    xi = rng.uniform(-0.5, 0.5, size=20 * 3 * fil.ns)
    dev = sc.FilamentSystem(p, fil, precision="fp64")
    ora = O.OracleSystem(p, fil)
    dev.set_noise(xi)
    ora.set_noise_buffer(xi, 20)
    dev.step(40)
    assert ora.step(40) == 0
    for x, y in zip(dev.sites(), ora.sites()):
        np.testing.assert_array_equal(x, y)
    # the draws differ from the counter stream's, so does the trajectory
    ref = sc.FilamentSystem(p, fil, precision="fp64")
    ref.step(40)
    assert not np.array_equal(ref.sites()[0], dev.sites()[0])


def test_stepping_past_the_buffer_is_refused():
    print("case: C0 fp64, 2-midstep buffer", flush=True)
    p, fil = c0_system(num=8, seed=3)
    dev = sc.FilamentSystem(p, fil, precision="fp64")
    dev.set_noise(np.zeros(2 * 3 * fil.ns))
    dev.step(4)          # midsteps 1 and 3 (indices 0, 1)
    dev.step(0)
    with pytest.raises(sc.SimcoreError):
        dev.step(1)      # step 5 needs midstep index 2
    assert dev.counters()["i_step"] == 4
    dev.set_noise()      # back to the counter stream
    dev.step(2)
    assert dev.counters()["i_step"] == 6
