"""Shared system builders for the tests (C0-like parity systems)."""
import os

import numpy as np

import simcore_amd as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C0_YAML = os.path.join(ROOT, "simcore_amd", "configs", "c0_plain_fp64.yaml")


def c0_system(num=None, seed=7, noise_seed=None):
    """Params + filaments of the 64 x 20-site parity configuration."""
    p, spec = sc.params_from_yaml(C0_YAML)
    if noise_seed is not None:
        p.seed = noise_seed
    nb, bl = sc.filament_layout(spec["length"], spec["min_bond_length"], spec["min_length"],
                                spec["max_length"])
    nf = int(spec["num"] if num is None else num)
    fil = sc.generate_filaments(p, [nb + 1] * nf, spec["length"], spec["persistence_length"],
                                spec["diameter"], spec["friction_ratio"], seed=seed)
    return p, fil


def box_for_fraction(n_sites, bond_length, diameter, phi):
    """system_radius giving volume fraction phi with the reference's
    filament volume pi/4 d^2 L + pi/6 d^3 (src/filament.cpp:385-392)."""
    n_sites = np.asarray(n_sites, dtype=float)
    L = (n_sites - 1) * bond_length
    v = np.sum(0.25 * np.pi * diameter ** 2 * L + np.pi / 6.0 * diameter ** 3)
    return 0.5 * (v / phi) ** (1.0 / 3.0)


def candidate_set_pairs(b1, b2):
    return set(zip(b1.tolist(), b2.tolist()))


def bond_filament(site_off):
    """Filament index of every bond (bonds of f: [so[f]-f, so[f+1]-f-1))."""
    so = np.asarray(site_off, dtype=np.int64)
    nbf = np.diff(so) - 1
    return np.repeat(np.arange(len(nbf)), nbf)


def culled(b1, b2, site_off):
    """Drop the pairs that share a site (Bond::HasNeighbor), as the first
    pair pass after a rebuild does (interaction_manager.cpp:485-497)."""
    bf = bond_filament(site_off)
    ex = (np.abs(b1.astype(np.int64) - b2) == 1) & (bf[b1] == bf[b2])
    return b1[~ex], b2[~ex]


# site counts straddling the integrator's length bins (<=8, 16, 32, 64,
# 128, 256 sites) and their edges
SHORT_SITES = [3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33, 47, 63, 64]
LONG_SITES = [65, 100, 127, 128, 129, 200, 255, 256]


def uneven_system(sites=SHORT_SITES, copies=3, seed=21, phi=0.04, radius=None,
                  delta=1e-4, noise_seed=777):
    """Filaments of the given site counts (each `copies` times) in one
    periodic box, lp in [10, 100] sigma, bond length 1, d = 1; the box is
    sized for volume fraction phi unless `radius` is given."""
    n = np.array(list(sites) * copies, dtype=np.int32)
    rng = np.random.default_rng(seed)
    lp = rng.uniform(10.0, 100.0, len(n))
    R = box_for_fraction(n, 1.0, 1.0, phi) if radius is None else radius
    p = sc.params_init(R, delta=delta, min_cell_length=1.0, seed=noise_seed)
    fil = sc.generate_filaments(p, n, (n - 1).astype(float), lp, 1.0, 2.0, seed=seed)
    return p, fil


def long_system(seed=23, noise_seed=778):
    """Every length class up to 256 sites, two of each, in a box wider than
    the longest filament, plus short ones."""
    return uneven_system(sites=SHORT_SITES + LONG_SITES, copies=2, seed=seed,
                         radius=140.0, noise_seed=noise_seed)
