"""Why the uneven-length configuration (persistence length up to 1000 sigma)
runs at dt = 1e-5: the midstep scheme (src/filament.cpp:394-428) is an
explicit integrator, and its stiffest bending mode relaxes at about
16 kappa / (l^3 gamma_perp) per unit time, so a step beyond 2 / that rate
amplifies the mode instead of damping it.  Recorded on the CPU oracle."""
import math

import numpy as np

import simcore_amd as sc
from oracle.oracle_py import OracleSystem

LP, N_SITES, N_FIL = 1000.0, 64, 8


def _mean_one_minus_cos(ora, site_off):
    r = ora.sites()[0].reshape(3, -1)
    vals = []
    for f in range(len(site_off) - 1):
        s = r[:, site_off[f]:site_off[f + 1]]
        u = np.diff(s, axis=1)
        u /= np.linalg.norm(u, axis=0)
        vals.append(1.0 - np.mean(np.sum(u[:, 1:] * u[:, :-1], axis=0)))
    return float(np.mean(vals))


def _run(dt, chunks):
    p = sc.params_init(40.0, delta=dt, min_cell_length=1.0, seed=3)
    fil = sc.generate_filaments(p, [N_SITES] * N_FIL, float(N_SITES - 1), LP, 1.0, 2.0, seed=4)
    ora = OracleSystem(p, fil, n_threads=1)
    out = []
    for _ in range(chunks):
        assert ora.step(200) == 0
        out.append(_mean_one_minus_cos(ora, fil.site_off))
    gperp = fil.fil_array("gamma_perp")[0]
    return out, gperp


def test_lp1000_is_unstable_at_dt_1e4_and_stable_at_1e5():
    # equilibrium of a discrete worm-like chain, P(cos) ~ exp((lp/l) cos):
    # <cos> = coth(lp/l) - l/lp
    kappa = LP / 1.0
    eq = 1.0 - (1.0 / math.tanh(kappa) - 1.0 / kappa)
    fine, gperp = _run(1e-5, 3)
    rate = 16.0 * LP / gperp          # l = 1
    assert 1e-5 < 2.0 / rate < 1e-4
    print("dt 1e-5: mean 1 - cos", fine, "equilibrium", eq, "limit 2/rate", 2.0 / rate)
    for v in fine:
        assert 0.6 * eq < v < 1.5 * eq
    coarse, _ = _run(1e-4, 2)
    print("dt 1e-4: mean 1 - cos", coarse)
    assert min(coarse) > 100 * eq      # the bending modes have blown up
