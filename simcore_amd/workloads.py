"""Synthetic filament systems of the benchmark configurations, generated with
the reference-style random insertion (sc_generate_filaments): the dense
single system, the uneven-length system and its crosslinked variant, at
full size or reduced to a given filament count with the same densities,
length and stiffness distributions."""
import numpy as np

from .system import generate_crosslinks, generate_filaments, params_init


def box_for_fraction(n_sites, bond_length, diameter, phi):
    """system_radius for volume fraction phi, with the reference's filament
    volume pi/4 d^2 L + pi/6 d^3 (src/filament.cpp:385-392)."""
    n_sites = np.asarray(n_sites, dtype=float)
    L = (n_sites - 1) * bond_length
    v = np.sum(0.25 * np.pi * diameter ** 2 * L + np.pi / 6.0 * diameter ** 3)
    return 0.5 * (v / phi) ** (1.0 / 3.0)


def dense(nf=20000, n_sites=32, lp=50.0, phi=0.05, delta=1e-4, seed=11,
          noise_seed=20240001):
    """Dense single system: nf filaments of n_sites sites (bond length 1,
    diameter 1), persistence length lp, periodic cube at volume fraction
    phi, WCA.  Returns (params, filaments, description)."""
    R = box_for_fraction([n_sites] * nf, 1.0, 1.0, phi)
    p = params_init(R, delta=delta, min_cell_length=1.0, seed=noise_seed)
    fil = generate_filaments(p, [n_sites] * nf, float(n_sites - 1), lp, 1.0, 2.0, seed=seed)
    desc = {"workload": f"dense single system: {nf} filaments x {n_sites} sites, lp {lp:g} "
                        f"sigma, volume fraction {phi:g}, periodic cube, WCA, dt {delta:g}",
            "n_filaments": nf, "sites_per_filament": n_sites, "persistence_length": lp,
            "volume_fraction": phi, "box_side": 2 * R, "bonds": int(fil.nb)}
    return p, fil, desc


def uneven_lengths(nf, rng_seed=20240002):
    """Site counts round(LogNormal(ln 24, 0.8)) clipped to [8, 256] and
    persistence lengths U[10, 1000] sigma, one pair per filament."""
    rng = np.random.default_rng(rng_seed)
    n = np.clip(np.rint(rng.lognormal(np.log(24.0), 0.8, nf)).astype(np.int64), 8, 256)
    lp = rng.uniform(10.0, 1000.0, nf)
    return n, lp


def uneven(nf=50000, phi=0.08, delta=1e-5, rng_seed=20240002, seed=12,
           noise_seed=20240003):
    """Uneven-length system: nf filaments with log-normal site counts
    (median 24, sigma_ln 0.8, clipped to [8, 256]), persistence length
    U[10, 1000] sigma, volume fraction phi.  dt 1e-5: the midstep scheme is
    explicit, and lp = 1000 sigma puts the stiffest bending mode at about
    16 kappa / (l^3 gamma_perp) ~ 7e4 per unit time, beyond the 2/dt
    stability limit at dt 1e-4."""
    n, lp = uneven_lengths(nf, rng_seed)
    R = box_for_fraction(n, 1.0, 1.0, phi)
    p = params_init(R, delta=delta, min_cell_length=1.0, seed=noise_seed)
    fil = generate_filaments(p, n, (n - 1).astype(float), lp, 1.0, 2.0, seed=seed)
    desc = {"workload": f"uneven lengths: {nf} filaments, sites lognormal(median 24, "
                        f"sigma_ln 0.8) clipped to [8,256], lp U[10,1000] sigma, "
                        f"volume fraction {phi:g}, periodic cube, WCA, dt {delta:g}",
            "n_filaments": nf, "volume_fraction": phi, "box_side": 2 * R,
            "bonds": int(fil.nb)}
    return p, fil, desc


def crosslinked(nf=50000, n_crosslinks=200000, seed=13, **kw):
    """The uneven-length system plus n_crosslinks harmonic crosslinks between
    random bond pairs within 2 sigma.  Returns (params, filaments,
    description, (bonds, lambdas), crosslink parameters)."""
    p, fil, desc = uneven(nf=nf, **kw)
    b, lam = generate_crosslinks(p, fil, n_crosslinks, r_bind=2.0, seed=seed)
    desc = dict(desc)
    desc["workload"] = (f"crosslinked network: the uneven-length system plus {n_crosslinks} "
                        "harmonic crosslinks (k 10 kT/sigma^2, rest length 1 sigma) between "
                        "random bond pairs within 2 sigma at uniform positions, unbinding at "
                        "rate 0.1/tau with rebinding to a random in-range bond, "
                        f"dt {p.delta:g}")
    desc["crosslinks"] = n_crosslinks
    xl_params = dict(k_spring=10.0, rest_length=1.0, k_off=0.1, r_bind=2.0, seed=14)
    return p, fil, desc, (b, lam), xl_params
