"""Error of the mixed-precision path against the FP64 path on identical
state: both are loaded from the same site positions (so both build the same
candidate list at load), stepped the same number of steps, and compared per
bond and per filament.

The force error of a bond is |F_mixed - F_fp64| relative to the force scale
of the system, the root mean square of the FP64 net pair force over the
bonds that have one.  A torque error dT acts on the bond's two sites as the
force pair +-(u x dT)/l (Filament::ApplyInteractionForces,
src/filament.cpp:992-1048), so the torque error of a bond is |dT|/l
relative to the same force scale.  A bond's own
net force is not a usable denominator: a pair near the WCA cutoff carries a
force that tends to zero there, so a bond whose only pair sits at r_c - 1e-5
has a net force of ~1e-4 kT/sigma, and the same absolute error that is 1e-7
of a typical force is 1e-3 of that one.  Beside the scaled maxima the
result reports the per-bond error relative to the bond's gross pair force
sum_c |F_c| (the FP64 slot forces of its candidates), its 99th percentile
and maximum, which show those near-cutoff bonds."""
import numpy as np

from .system import FilamentSystem


def _per_bond_gross(ref):
    b1, b2 = ref.candidates()
    slots = ref.slot_forces().reshape(3, -1)
    mag = np.sqrt((slots * slots).sum(axis=0))
    g = np.zeros(ref.nb)
    np.add.at(g, b1, mag)
    np.add.at(g, b2, mag)
    return g


def compare(ref, test):
    """Errors of `test` against `ref` (FP64) after both took the same steps
    from the same state."""
    fr, tr = (x.reshape(3, -1) for x in ref.bond_forces())
    ft, tt = (x.reshape(3, -1) for x in test.bond_forces())
    g = _per_bond_gross(ref)
    df = np.sqrt(((ft - fr) ** 2).sum(axis=0))
    dt = np.sqrt(((tt - tr) ** 2).sum(axis=0))
    live = g > 0
    net_f = np.sqrt((fr * fr).sum(axis=0))
    net_t = np.sqrt((tr * tr).sum(axis=0))
    fscale = float(np.sqrt(np.mean(net_f[live] ** 2))) if live.any() else 1.0
    tscale = float(np.sqrt(np.mean(net_t[live] ** 2))) if live.any() else 1.0
    length = ref.bonds()[3]
    egross = df[live] / g[live]
    tens_r, tens_t = ref.tensions(), test.tensions()
    so = np.asarray(ref.filaments.site_off, dtype=np.int64)
    bf = np.repeat(np.arange(len(so) - 1), np.diff(so) - 1)
    tmax = np.zeros(len(so) - 1)
    np.maximum.at(tmax, bf, np.abs(tens_r))
    tl = tmax[bf] > 0
    et_ten = np.abs(tens_t - tens_r)[tl] / tmax[bf][tl]
    return {
        "bonds": int(ref.nb),
        "bonds_with_pair_forces": int(live.sum()),
        "force_scale": fscale,
        "torque_scale": tscale,
        "force_err_max": float(df.max() / fscale) if df.size else 0.0,
        "torque_err_max": float((dt / length).max() / fscale) if dt.size else 0.0,
        "torque_err_max_vs_torque_rms": float(dt.max() / tscale) if dt.size else 0.0,
        "force_err_global": float(np.linalg.norm(ft - fr) / max(np.linalg.norm(fr), 1e-300)),
        "force_rel_gross_p99": float(np.percentile(egross, 99)) if egross.size else 0.0,
        "force_rel_gross_max": float(egross.max()) if egross.size else 0.0,
        "force_mismatch_without_pairs": int(np.count_nonzero(df[~live])),
        "tension_rel_max": float(et_ten.max()) if et_ten.size else 0.0,
        "definition": "force_err_max = max over bonds |F_mixed - F_fp64| / force_scale, "
                      "force_scale = RMS FP64 net pair force per bond over bonds with pairs; "
                      "torque_err_max = max over bonds |dT| / l / force_scale (the site-force "
                      "error the torque error causes); force_rel_gross = |dF| / sum_c |F_c| "
                      "per bond; tension per filament relative to its largest |tension|",
    }


def mixed_vs_fp64(params, filaments, steps=1, spatial_order=True, drift_steps=0):
    """Loads an FP64 and a mixed system (graph-replayed, the mixed one in
    spatial storage order when asked) from the same positions, steps both,
    and returns compare()'s errors; with drift_steps > 0 also the site
    position differences after that many further steps (their maximum, 99.9th
    percentile and the number of sites beyond 1e-6 sigma: the dynamics
    amplify differences, and a bond length that crosses the renormalisation
    threshold of src/filament.cpp:934-951 in one precision and not the other
    re-places its filament by up to 1e-3 of a bond, so a large system shows a
    few such sites over 100 steps)."""
    ref = FilamentSystem(params, filaments, precision="fp64", use_graph=True)
    test = FilamentSystem(params, filaments, precision="mixed", use_graph=True,
                          spatial_order=spatial_order)
    try:
        ref.step(steps)
        test.step(steps)
        out = compare(ref, test)
        out["steps"] = steps
        if drift_steps:
            ref.step(drift_steps)
            test.step(drift_steps)
            out["position_drift_steps"] = steps + drift_steps
            d = np.abs(ref.sites()[0] - test.sites()[0]).reshape(3, -1).max(axis=0)
            out["position_drift_max"] = float(d.max())
            out["position_drift_p999"] = float(np.percentile(d, 99.9))
            out["sites_drift_over_1e-6"] = int(np.count_nonzero(d > 1e-6))
            out["same_rebuilds"] = ref.counters()["n_rebuilds"] == test.counters()["n_rebuilds"]
        return out
    finally:
        ref.close()
        test.close()
