"""Long-run statistics of the device path against the CPU path and against
the chain's equilibrium: the two runs use independent noise streams, so
they are compared as distributions (two-sample KS tests), as the
acceptance of a stochastic integrator requires.

System: 400 filaments x 16 sites (bond length 1, lp 10 sigma) in a dilute
periodic box, inserted from the equilibrium bond-angle distribution
P(cos) ~ exp((lp/l) cos) (src/filament.cpp:350-383); 4000 steps (0.2 tau of
dt 1e-4)."""
import numpy as np
import pytest
from scipy import stats

import simcore_amd as sc
from oracle.oracle_py import OracleSystem
from tests.helpers import box_for_fraction

pytestmark = pytest.mark.gpu

NF, NS, LP = 400, 16, 10.0
STEPS, EVERY = 4000, 500


def _system(noise_seed):
    R = box_for_fraction([NS] * NF, 1.0, 1.0, 0.005)
    p = sc.params_init(R, delta=1e-4, min_cell_length=1.0, seed=noise_seed)
    fil = sc.generate_filaments(p, [NS] * NF, float(NS - 1), LP, 1.0, 2.0, seed=77)
    return p, fil


def _host_stats(r, so):
    r = r.reshape(3, -1)
    first, last = so[:-1], so[1:] - 1
    e2e = r[:, last] - r[:, first]
    com = np.stack([np.add.reduceat(r[q], so[:-1]) / np.diff(so) for q in range(3)])
    u = np.diff(r, axis=1)
    u /= np.linalg.norm(u, axis=0)
    keep = np.ones(u.shape[1], bool)
    keep[so[1:-1] - 1] = False          # differences across filaments
    u = u[:, keep]
    nb = np.diff(so) - 1
    ends = np.cumsum(nb)
    c = np.sum(u[:, :-1] * u[:, 1:], axis=0)
    ok = np.ones(c.shape[0], bool)
    ok[ends[:-1] - 1] = False           # pairs across filaments
    return e2e, com, c[ok]


def _cos_from_hist(h, n_bins):
    centres = (np.arange(n_bins) + 0.5) * (2.0 / n_bins) - 1.0
    tot = h.sum(axis=0)
    return centres, tot


def test_long_run_distributions_match_cpu_path_and_equilibrium():
    print("case: 400 x 16 sites, lp 10, 4000 steps, device mixed vs CPU fp64", flush=True)
    p, fil = _system(noise_seed=9001)
    dev = sc.FilamentSystem(p, fil, precision="mixed", use_graph=True)
    p2, fil2 = _system(noise_seed=9002)
    ora = OracleSystem(p2, fil2)
    so = fil.site_off
    e0, c0 = dev.filament_stats()
    n_bins = 200
    hist = np.zeros((NS - 2, n_bins), dtype=np.int64)
    d_e2e, o_e2e, d_cos, o_cos = [], [], [], []
    for k in range(STEPS // EVERY):
        dev.step(EVERY)
        assert ora.step(EVERY) == 0
        e, c = dev.filament_stats()
        d_e2e.append(np.sum(e * e, axis=0))
        dev.theta_histogram(n_bins, NS - 2, hist)
        oe, oc, ocos = _host_stats(ora.sites()[0], so)
        o_e2e.append(np.sum(oe * oe, axis=0))
        o_cos.append(ocos)
        # device angles on the host too, for the KS test
        _, _, dcos = _host_stats(dev.sites()[0], so)
        d_cos.append(dcos)
    # MSD of the centres of mass over the run
    _, c_end = dev.filament_stats()
    d_msd = np.sum((c_end - c0) ** 2, axis=0)
    _, oc_end, _ = _host_stats(ora.sites()[0], so)
    _, oc0, _ = _host_stats(fil2.r, so)
    o_msd = np.sum((oc_end - oc0) ** 2, axis=0)

    d_e2e, o_e2e = np.concatenate(d_e2e), np.concatenate(o_e2e)
    d_cos, o_cos = np.concatenate(d_cos), np.concatenate(o_cos)
    ks_e2e = stats.ks_2samp(d_e2e, o_e2e)
    ks_cos = stats.ks_2samp(d_cos, o_cos)
    ks_msd = stats.ks_2samp(d_msd, o_msd)
    print("KS e2e", ks_e2e, "KS cos", ks_cos, "KS msd", ks_msd, flush=True)
    assert ks_e2e.pvalue > 1e-3
    assert ks_cos.pvalue > 1e-3
    assert ks_msd.pvalue > 1e-3

    # equilibrium of the discrete chain: <cos> = coth(k) - 1/k, k = lp/l;
    # <R^2> = N l^2 (1+c)/(1-c) - 2 l^2 c (1-c^N)/(1-c)^2 (N bonds)
    kap = LP / 1.0
    cm = 1.0 / np.tanh(kap) - 1.0 / kap
    nbond = NS - 1
    r2 = nbond * (1 + cm) / (1 - cm) - 2 * cm * (1 - cm ** nbond) / (1 - cm) ** 2
    print("<cos> device", d_cos.mean(), "oracle", o_cos.mean(), "theory", cm)
    print("<R2> device", d_e2e.mean(), "oracle", o_e2e.mean(), "theory", r2)
    assert abs(d_cos.mean() - cm) < 0.01
    assert abs(d_e2e.mean() - r2) / r2 < 0.05
    # the device histogram is the binned device angles
    centres, tot = _cos_from_hist(hist, n_bins)
    assert tot.sum() == d_cos.size
    assert abs(np.sum(centres * tot) / tot.sum() - d_cos.mean()) < 2.0 / n_bins
    # diffusion: MSD of the centre of mass, device vs CPU means agree
    assert abs(d_msd.mean() - o_msd.mean()) / o_msd.mean() < 0.15
    dev.close()
    ora.close()
