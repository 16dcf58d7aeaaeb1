"""Host-side mirror of the reference's filament species / simulation
interface for the BD hot path, driving the HIP path through the C ABI.

`FilamentSystem` corresponds to one reference `Simulation` holding one
`FilamentSpecies` (src/simulation.cpp, src/filament_species.cpp): `step(n)`
is n iterations of RunSimulation's body, and the getters are the object
accessors the reference exposes (positions, bond forces, tensions, the
interaction manager's pair list and cell assignments).
"""
import ctypes
from ctypes import byref, c_double, c_int32, c_int64, c_size_t, c_void_p
from dataclasses import dataclass

import numpy as np

from ._lib import (SC_ERR_UNPHYSICAL, SC_OK, SC_PREC_FP64, SC_PREC_MIXED,
                   ScCrosslinkParams, ScFilamentParams, ScLaunchShape, ScParams, ScTuneReport,
                   SimcoreError, UnphysicalResult, lib)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(c_int32))


def _check(rc, h=None, what=""):
    if rc == SC_OK:
        return
    msg = what
    if h is not None and h.value:
        msg = f"{what}: {lib.sc_last_error(h).decode()}"
    if rc == SC_ERR_UNPHYSICAL:
        raise UnphysicalResult(rc, msg)
    raise SimcoreError(rc, msg)


def params_init(system_radius, **kw):
    """sc_params with the reference defaults for a periodic cube of side
    2*system_radius; keyword arguments override fields."""
    p = ScParams()
    _check(lib.sc_params_init(byref(p), float(system_radius)), what="sc_params_init")
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def params_from_yaml(path):
    """(sc_params, species dict) from a reference-style YAML file."""
    p = ScParams()
    spec = (c_double * 9)()
    _check(lib.sc_params_from_yaml(str(path).encode(), byref(p), spec),
           what=f"sc_params_from_yaml({path})")
    keys = ["num", "length", "persistence_length", "diameter", "min_bond_length",
            "friction_ratio", "n_steps", "min_length", "max_length"]
    return p, dict(zip(keys, list(spec)))


def filament_layout(length, min_bond_length, min_length=5.0, max_length=500.0):
    """(n_bonds, bond_length) as Filament::InitFilamentLength computes them."""
    nb = c_int32()
    bl = c_double()
    _check(lib.sc_filament_layout(length, min_bond_length, min_length, max_length,
                                  byref(nb), byref(bl)), what="sc_filament_layout")
    return nb.value, bl.value


@dataclass
class Filaments:
    """Filament set in the C ABI's layout."""
    site_off: np.ndarray      # int32, nf+1
    r: np.ndarray             # float64, 3*Ns (SoA)
    fil: object               # ctypes array of ScFilamentParams

    @property
    def nf(self):
        return len(self.site_off) - 1

    @property
    def ns(self):
        return int(self.site_off[-1])

    @property
    def nb(self):
        return self.ns - self.nf

    def fil_array(self, field):
        return np.array([getattr(f, field) for f in self.fil])


def generate_filaments(params, n_sites, length, persistence_length, diameter=None,
                       friction_ratio=2.0, seed=1):
    """Random reference-style insertion with overlap rejection."""
    n_sites = np.ascontiguousarray(n_sites, dtype=np.int32)
    nf = len(n_sites)
    length = np.ascontiguousarray(np.broadcast_to(length, (nf,)), dtype=np.float64)
    lp = np.ascontiguousarray(np.broadcast_to(persistence_length, (nf,)), dtype=np.float64)
    d = np.ascontiguousarray(np.broadcast_to(1.0 if diameter is None else diameter, (nf,)),
                             dtype=np.float64)
    tot = c_int64()
    _check(lib.sc_generate_filaments(byref(params), nf, _ip(n_sites), _dp(length), _dp(lp),
                                     _dp(d), friction_ratio, seed, None, None, None,
                                     byref(tot)), what="sc_generate_filaments")
    site_off = np.zeros(nf + 1, dtype=np.int32)
    r = np.zeros(3 * tot.value, dtype=np.float64)
    fil = (ScFilamentParams * nf)()
    _check(lib.sc_generate_filaments(byref(params), nf, _ip(n_sites), _dp(length), _dp(lp),
                                     _dp(d), friction_ratio, seed, _ip(site_off), _dp(r), fil,
                                     byref(tot)), what="sc_generate_filaments")
    return Filaments(site_off, r, fil)


def generate_crosslinks(params, filaments, n, r_bind=2.0, seed=1):
    """(bonds (n,2), lambda (n,2)): n crosslinks between a random bond and a
    random other bond within r_bind of the first anchor."""
    so = np.ascontiguousarray(filaments.site_off, dtype=np.int32)
    r = np.ascontiguousarray(filaments.r, dtype=np.float64)
    b = np.empty(2 * n, dtype=np.int32)
    lam = np.empty(2 * n)
    _check(lib.sc_generate_crosslinks(byref(params), filaments.nf, _ip(so), _dp(r), filaments.fil,
                                      int(n), float(r_bind), int(seed), _ip(b), _dp(lam)),
           what="sc_generate_crosslinks")
    return b.reshape(-1, 2), lam.reshape(-1, 2)


def crosslink_params(k_spring=10.0, rest_length=1.0, k_compress=-1.0, k_off=0.1, r_bind=2.0,
                     seed=1):
    return ScCrosslinkParams(k_spring, rest_length, k_compress, k_off, r_bind, seed)


class FilamentSystem:
    """One system's device state (sc_handle)."""

    def __init__(self, params, filaments, precision="fp64", use_graph=False,
                 capacity_factor=0.0, spatial_order=False):
        p = ScParams()
        ctypes.pointer(p)[0] = params
        p.precision = {"fp64": SC_PREC_FP64, "mixed": SC_PREC_MIXED}[precision]
        p.use_graph = int(bool(use_graph))
        p.spatial_order = int(bool(spatial_order))
        p.capacity_factor = capacity_factor
        self.params = p
        self.filaments = filaments
        self.h = c_void_p()
        _check(lib.sc_create(byref(p), byref(self.h)), what="sc_create")
        so = np.ascontiguousarray(filaments.site_off, dtype=np.int32)
        r = np.ascontiguousarray(filaments.r, dtype=np.float64)
        _check(lib.sc_load_filaments(self.h, filaments.nf, _ip(so), _dp(r), filaments.fil),
               self.h, "sc_load_filaments")
        self.ns, self.nb, self.nf = filaments.ns, filaments.nb, filaments.nf

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        if self.h is not None and self.h.value:
            lib.sc_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping ------------------------------------------------------------
    def step(self, n=1):
        _check(lib.sc_step(self.h, int(n)), self.h, "sc_step")

    def step_async(self, n=1):
        """Enqueue n steps and return at once."""
        _check(lib.sc_step_async(self.h, int(n)), self.h, "sc_step_async")

    def poll(self):
        """True when every enqueued step has completed (never blocks)."""
        done = ctypes.c_int()
        _check(lib.sc_step_poll(self.h, byref(done)), self.h, "sc_step_poll")
        return bool(done.value)

    # -- taps ----------------------------------------------------------------
    def sites(self):
        r = np.empty(3 * self.ns)
        rp = np.empty(3 * self.ns)
        fr = np.empty(3 * self.ns)
        _check(lib.sc_get_sites(self.h, _dp(r), _dp(rp), _dp(fr)), self.h, "sc_get_sites")
        return r, rp, fr

    def bonds(self):
        pos, sp, u = (np.empty(3 * self.nb) for _ in range(3))
        ln = np.empty(self.nb)
        _check(lib.sc_get_bonds(self.h, _dp(pos), _dp(sp), _dp(u), _dp(ln)), self.h,
               "sc_get_bonds")
        return pos, sp, u, ln

    def bond_forces(self):
        f = np.empty(3 * self.nb)
        t = np.empty(3 * self.nb)
        _check(lib.sc_get_bond_forces(self.h, _dp(f), _dp(t)), self.h, "sc_get_bond_forces")
        return f, t

    def tensions(self):
        t = np.empty(self.nb)
        _check(lib.sc_get_tensions(self.h, _dp(t)), self.h, "sc_get_tensions")
        return t

    def cell_keys(self):
        k = np.empty(self.nb, dtype=np.int32)
        _check(lib.sc_get_cell_keys(self.h, _ip(k)), self.h, "sc_get_cell_keys")
        return k

    def candidates(self):
        n = c_size_t()
        _check(lib.sc_get_candidates(self.h, None, None, 0, byref(n)), self.h,
               "sc_get_candidates")
        b1 = np.empty(n.value, dtype=np.int32)
        b2 = np.empty(n.value, dtype=np.int32)
        _check(lib.sc_get_candidates(self.h, _ip(b1), _ip(b2), n.value, byref(n)), self.h,
               "sc_get_candidates")
        return b1, b2

    def slot_forces(self):
        n = self.counters()["n_candidates"]
        f = np.empty(3 * n)
        _check(lib.sc_get_slot_forces(self.h, _dp(f), n), self.h, "sc_get_slot_forces")
        return f

    def dr_state(self):
        a = np.empty(self.ns)
        _check(lib.sc_get_dr_state(self.h, _dp(a)), self.h, "sc_get_dr_state")
        return a

    def counters(self):
        i_step, nreb, nc = c_int64(), c_int64(), c_int64()
        n1d = c_int32()
        _check(lib.sc_get_counters(self.h, byref(i_step), byref(nreb), byref(nc), byref(n1d)),
               self.h, "sc_get_counters")
        return {"i_step": i_step.value, "n_rebuilds": nreb.value,
                "n_candidates": nc.value, "n_cells_1d": n1d.value}

    def load_crosslinks(self, bonds, lam, **kw):
        """Attach crosslinks (caller bond numbering); keyword arguments as
        crosslink_params."""
        b = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1)
        la = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        xp = crosslink_params(**kw)
        _check(lib.sc_load_crosslinks(self.h, len(b) // 2, _ip(b), _dp(la), byref(xp)), self.h,
               "sc_load_crosslinks")
        self.nx = len(b) // 2

    def crosslinks(self):
        """(bonds (nx,2), lambda (nx,2), stored forces (nx,9), events)."""
        nx = getattr(self, "nx", 0)
        b = np.empty(2 * nx, dtype=np.int32)
        la = np.empty(2 * nx)
        f = np.empty(9 * nx)
        ev = c_int64()
        _check(lib.sc_get_crosslinks(self.h, _ip(b), _dp(la), _dp(f), byref(ev)), self.h,
               "sc_get_crosslinks")
        return b.reshape(-1, 2), la.reshape(-1, 2), f.reshape(-1, 9), ev.value

    def set_launch_shape(self, pair_block=256, pair_xcd=0, integrate_waves=1,
                         integrate_layout=1, integrate_streams=0, pair_parts=1, early_long=0):
        """Launch shapes (sc_set_launch_shape); only integrate_layout changes
        results (the rounding of the mixed integrator)."""
        v = ScLaunchShape(pair_block, pair_xcd, integrate_waves, integrate_layout,
                          integrate_streams, pair_parts, early_long)
        _check(lib.sc_set_launch_shape(self.h, byref(v)), self.h, "sc_set_launch_shape")

    def time_pair(self, launches=50):
        """Average duration (ms) of one pair-pass launch, timed back to back
        with HIP events on the system's stream (sc_time_pair); the state is
        left unchanged."""
        ms = c_double()
        _check(lib.sc_time_pair(self.h, int(launches), byref(ms)), self.h, "sc_time_pair")
        return ms.value

    def launch_shape(self):
        v = ScLaunchShape()
        _check(lib.sc_get_launch_shape(self.h, byref(v)), self.h, "sc_get_launch_shape")
        return {k: getattr(v, k) for k, _ in ScLaunchShape._fields_}

    def autotune(self, steps=20):
        """Pick the fastest result-invariant launch shapes (sc_autotune); the
        kept shape's fields plus n_trials and the step times before/after."""
        rep = ScTuneReport()
        _check(lib.sc_autotune(self.h, int(steps), byref(rep)), self.h, "sc_autotune")
        out = {k: getattr(rep.shape, k) for k, _ in ScLaunchShape._fields_}
        out.update(n_trials=rep.n_trials, step_ms_start=rep.step_ms_start,
                   step_ms_best=rep.step_ms_best)
        return out

    def checkpoint(self):
        """Complete dynamic state as bytes (sc_checkpoint_save)."""
        n = c_size_t()
        _check(lib.sc_checkpoint_save(self.h, None, 0, byref(n)), self.h, "sc_checkpoint_save")
        buf = ctypes.create_string_buffer(n.value)
        _check(lib.sc_checkpoint_save(self.h, buf, n.value, byref(n)), self.h,
               "sc_checkpoint_save")
        return buf.raw

    def restore(self, blob):
        """Continue from a checkpoint of a system with the same layout."""
        buf = ctypes.create_string_buffer(blob, len(blob))
        _check(lib.sc_checkpoint_load(self.h, buf, len(blob)), self.h, "sc_checkpoint_load")

    def filament_stats(self):
        """(end-to-end vectors, centres of mass), each (3, nf)."""
        e2e = np.empty(3 * self.nf)
        com = np.empty(3 * self.nf)
        _check(lib.sc_filament_stats(self.h, _dp(e2e), _dp(com)), self.h, "sc_filament_stats")
        return e2e.reshape(3, -1), com.reshape(3, -1)

    def theta_histogram(self, n_bins, n_rows, hist=None):
        """cos(theta) histogram (n_rows, n_bins), accumulated into `hist`."""
        if hist is None:
            hist = np.zeros((n_rows, n_bins), dtype=np.int64)
        assert hist.shape == (n_rows, n_bins) and hist.dtype == np.int64
        _check(lib.sc_theta_histogram(self.h, int(n_bins), int(n_rows),
                                      hist.ctypes.data_as(ctypes.POINTER(c_int64))),
               self.h, "sc_theta_histogram")
        return hist

    def thermo_enable(self, on=True):
        """Sum the pair stress and the bond energies on every step
        (sc_thermo_enable); enabling starts a fresh stress sum."""
        _check(lib.sc_thermo_enable(self.h, 1 if on else 0), self.h, "sc_thermo_enable")

    def calculate_pressure(self, n_thermo):
        """InteractionManager::CalculatePressure over the stress summed since
        the last call (sc_calculate_pressure): {"pressure_tensor" (3x3),
        "pressure", "volume"}; resets the sum."""
        pt = np.empty(9)
        pr, vol = c_double(), c_double()
        _check(lib.sc_calculate_pressure(self.h, int(n_thermo), _dp(pt), byref(pr), byref(vol)),
               self.h, "sc_calculate_pressure")
        return {"pressure_tensor": pt.reshape(3, 3), "pressure": pr.value, "volume": vol.value}

    def bond_energies(self):
        """Each bond's pair potential energy of the last step, caller bond
        order (sc_get_bond_energies)."""
        e = np.empty(self.nb)
        _check(lib.sc_get_bond_energies(self.h, _dp(e)), self.h, "sc_get_bond_energies")
        return e

    def set_noise(self, xi=None, seed=None):
        """Caller-supplied thermal noise (sc_set_noise): `xi` holds the draws
        of the next len(xi) / (3 n_sites) midsteps, 3 per site, sites in
        caller order, midstep-major; without `xi` the counter stream is used,
        keyed by `seed` (default: the key the system was created with)."""
        if xi is None:
            key = self.params.seed if seed is None else int(seed)
            _check(lib.sc_set_noise(self.h, 0, key, None, 0), self.h, "sc_set_noise")
            return
        xi = np.ascontiguousarray(xi, dtype=np.float64).ravel()
        _check(lib.sc_set_noise(self.h, 1, 0, _dp(xi), xi.size), self.h, "sc_set_noise")

    def set_sites(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        _check(lib.sc_set_sites(self.h, _dp(r)), self.h, "sc_set_sites")

    def device_bytes(self):
        b = c_size_t()
        _check(lib.sc_device_bytes(self.h, byref(b)), self.h, "sc_device_bytes")
        return b.value

    def profile(self, on=True):
        _check(lib.sc_profile_enable(self.h, int(on)), self.h, "sc_profile_enable")

    def profile_read(self):
        ms = (c_double * 4)()
        n = (c_int64 * 4)()
        _check(lib.sc_profile_read(self.h, ms, n, 4), self.h, "sc_profile_read")
        names = ["pair", "commit", "integrate", "rebuild"]
        return {k: (ms[i], n[i]) for i, k in enumerate(names)}
