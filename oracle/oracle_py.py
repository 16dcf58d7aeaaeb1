"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("h", c_double * 9), ("h_inv", c_double * 9), ("system_radius", c_double),
        ("min_cell_length", c_double), ("delta", c_double), ("wca_eps", c_double),
        ("wca_sig", c_double), ("f_cutoff", c_double), ("stoch_flag", c_int),
        ("n_threads", c_int), ("seed", c_uint64),
    ]


class OrcXlinkParams(ctypes.Structure):
    _fields_ = [("k_spring", c_double), ("rest_length", c_double), ("k_compress", c_double),
                ("k_off", c_double), ("r_bind", c_double), ("seed", c_uint64)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    dp, ip = POINTER(c_double), POINTER(c_int32)
    V = ctypes.c_void_p
    sig = {
        "orc_create": ([POINTER(OrcParams), c_int, ip, dp, dp, dp, dp, dp, dp, dp, dp], V),
        "orc_destroy": ([V], None),
        "orc_step": ([V, c_int], c_int),
        "orc_n_sites": ([V], c_int64), "orc_n_bonds": ([V], c_int64),
        "orc_n_candidates": ([V], c_int64), "orc_rebuild_count": ([V], c_int64),
        "orc_step_index": ([V], c_int64), "orc_n_cells_1d": ([V], c_int),
        "orc_cell_length": ([V], c_double),
        "orc_get_candidates": ([V, ip, ip], None),
        "orc_get_slot_forces": ([V, dp], None),
        "orc_get_cell_keys": ([V, ip], None),
        "orc_get_sites": ([V, dp, dp, dp], None),
        "orc_get_bonds": ([V, dp, dp, dp, dp], None),
        "orc_get_bond_forces": ([V, dp, dp], None),
        "orc_get_tensions": ([V, dp], None),
        "orc_get_dr": ([V, dp, dp], None),
        "orc_set_sites": ([V, dp], None),
        "orc_set_noise_buffer": ([V, dp, c_int64], None),
        "orc_take_stress": ([V, dp], None),
        "orc_get_bond_energies": ([V, dp], None),
        "orc_min_distance": ([dp, dp, dp, c_double, dp, dp, c_double, dp], None),
        "orc_pair_force": ([c_double, c_double, c_double, c_double, dp, c_double, dp, dp, dp],
                           c_int),
        "orc_tridiag": ([c_int, dp, dp, dp, dp], None),
        "orc_cell_neighbors": ([c_int, c_int, c_int, c_int, ip], c_int),
        "orc_cell_key": ([c_int, dp], c_int32),
        "orc_philox4x32": ([POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)], None),
        "orc_noise": ([c_uint64, c_int64, c_int64, dp], None),
        "orc_bending_forces": ([c_int, dp, c_double, c_double, dp, dp], None),
        "orc_bond_site_forces": ([c_int, c_double, dp, dp, dp, dp], None),
        "orc_load_crosslinks": ([V, c_int64, ip, dp, POINTER(OrcXlinkParams)], c_int),
        "orc_get_crosslinks": ([V, ip, dp, dp], None),
        "orc_xlink_events": ([V], c_int64),
        "orc_xlink_force": ([dp, dp, dp, dp, c_double, c_double, c_double, dp], c_double),
    }
    for n, (a, r) in sig.items():
        f = getattr(L, n)
        f.argtypes = a
        f.restype = r
    return L


lib = _load()


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _ip(a):
    return a.ctypes.data_as(POINTER(c_int32))


class OracleSystem:
    """CPU restatement of one simcore filament system."""

    def __init__(self, params, filaments, n_threads=0):
        """params: simcore_amd ScParams (only the physical fields are read);
        filaments: simcore_amd Filaments."""
        p = OrcParams()
        for i in range(9):
            p.h[i] = params.unit_cell[i]
            p.h_inv[i] = params.unit_cell_inv[i]
        p.system_radius = params.system_radius
        p.min_cell_length = params.min_cell_length
        p.delta = params.delta
        p.wca_eps = params.wca_eps
        p.wca_sig = params.wca_sig
        p.f_cutoff = params.f_cutoff
        p.stoch_flag = params.stoch_flag
        p.n_threads = n_threads
        p.seed = params.seed
        so = np.ascontiguousarray(filaments.site_off, dtype=np.int32)
        r = np.ascontiguousarray(filaments.r, dtype=np.float64)
        ns = int(so[-1])
        nf = len(so) - 1
        arr = {k: np.ascontiguousarray(filaments.fil_array(k), dtype=np.float64)
               for k in ("bond_length", "persistence_length", "gamma_par", "gamma_perp",
                         "diameter")}
        self._keep = (so, r, arr)
        self.s = lib.orc_create(byref(p), nf, _ip(so), _dp(r[:ns]), _dp(r[ns:2 * ns]),
                                _dp(r[2 * ns:]), _dp(arr["bond_length"]),
                                _dp(arr["persistence_length"]), _dp(arr["gamma_par"]),
                                _dp(arr["gamma_perp"]), _dp(arr["diameter"]))
        if not self.s:
            raise ValueError("orc_create rejected the input")
        self.ns = ns
        self.nf = nf
        self.nb = ns - nf
        self._h = np.array(params.unit_cell[:9], dtype=np.float64).reshape(3, 3)

    def close(self):
        if self.s:
            lib.orc_destroy(self.s)
            self.s = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        return lib.orc_step(self.s, int(n))

    def sites(self):
        r, rp, fr = (np.empty(3 * self.ns) for _ in range(3))
        lib.orc_get_sites(self.s, _dp(r), _dp(rp), _dp(fr))
        return r, rp, fr

    def bonds(self):
        pos, sp, u = (np.empty(3 * self.nb) for _ in range(3))
        ln = np.empty(self.nb)
        lib.orc_get_bonds(self.s, _dp(pos), _dp(sp), _dp(u), _dp(ln))
        return pos, sp, u, ln

    def bond_forces(self):
        f, t = np.empty(3 * self.nb), np.empty(3 * self.nb)
        lib.orc_get_bond_forces(self.s, _dp(f), _dp(t))
        return f, t

    def tensions(self):
        t = np.empty(self.nb)
        lib.orc_get_tensions(self.s, _dp(t))
        return t

    def cell_keys(self):
        k = np.empty(self.nb, dtype=np.int32)
        lib.orc_get_cell_keys(self.s, _ip(k))
        return k

    def candidates(self):
        n = lib.orc_n_candidates(self.s)
        b1 = np.empty(n, dtype=np.int32)
        b2 = np.empty(n, dtype=np.int32)
        lib.orc_get_candidates(self.s, _ip(b1), _ip(b2))
        return b1, b2

    def slot_forces(self):
        n = lib.orc_n_candidates(self.s)
        f = np.empty(3 * n)
        lib.orc_get_slot_forces(self.s, _dp(f))
        return f

    def dr_state(self):
        a = np.empty(self.ns)
        d = np.empty(self.nf)
        lib.orc_get_dr(self.s, _dp(a), _dp(d))
        return a, d

    def set_sites(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        lib.orc_set_sites(self.s, _dp(r))

    def take_stress(self):
        """The pair stress summed over pairs and steps since the last call
        (3x3, row-major), reset."""
        out = np.empty(9)
        lib.orc_take_stress(self.s, _dp(out))
        return out.reshape(3, 3)

    def thermo_enable(self, on=True):
        """The restatement always sums the stress; enabling starts a fresh
        sum, as on the device."""
        self.take_stress()

    def calculate_pressure(self, n_thermo):
        """InteractionManager::CalculatePressure (src/interaction_manager.cpp:
        534-559) over the stress summed since the last call: pressure tensor
        delta_ij n_objs / V + stress_ij / (V n_thermo), n_objs the bond
        count (Mesh::GetCount, src/mesh.cpp:399), V the periodic box's unit
        cell volume (Space::UpdateVolume, src/space.cpp:233-237), the
        cofactor expansion of Space::CalculateUnitCellQuantities
        (src/space.cpp:104-118); pressure the trace over n_dim; resets the
        sum."""
        stress = self.take_stress()
        u = [float(x) for x in self._h.ravel()]
        vol = (u[0] * ((u[4] * u[8]) - (u[7] * u[5])) - u[1] * (u[3] * u[8] - u[6] * u[5])
               + u[2] * (u[3] * u[7] - u[6] * u[4]))
        inv_v = 1.0 / vol
        pt = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                if i == j:
                    pt[i, j] += self.nb * inv_v
                pt[i, j] += stress[i, j] * inv_v / n_thermo
        return {"pressure_tensor": pt, "pressure": float(np.trace(pt) / 3), "volume": vol}

    def bond_energies(self):
        """Each bond's pair potential energy of the last step."""
        out = np.empty(self.nb)
        lib.orc_get_bond_energies(self.s, _dp(out))
        return out

    def set_noise_buffer(self, xi, n_mid):
        """Pre-drawn xi for the next n_mid midsteps (3 per site, site-major);
        n_mid = 0: back to the counter stream."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        self._xi = xi
        lib.orc_set_noise_buffer(self.s, _dp(xi), int(n_mid))

    def load_crosslinks(self, bonds, lam, k_spring=10.0, rest_length=1.0, k_compress=-1.0,
                        k_off=0.1, r_bind=2.0, seed=1):
        bonds = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1)
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        xp = OrcXlinkParams(k_spring, rest_length, k_compress, k_off, r_bind, seed)
        rc = lib.orc_load_crosslinks(self.s, len(bonds) // 2, _ip(bonds), _dp(lam), byref(xp))
        if rc != 0:
            raise ValueError("orc_load_crosslinks rejected the input")
        self.nx = len(bonds) // 2

    def crosslinks(self):
        """(bonds (nx,2), lambda (nx,2), stored forces (nx,9))."""
        b = np.empty(2 * self.nx, dtype=np.int32)
        lam = np.empty(2 * self.nx)
        f = np.empty(9 * self.nx)
        lib.orc_get_crosslinks(self.s, _ip(b), _dp(lam), _dp(f))
        return b.reshape(-1, 2), lam.reshape(-1, 2), f.reshape(-1, 9)

    def xlink_events(self):
        return lib.orc_xlink_events(self.s)

    def counters(self):
        return {"i_step": lib.orc_step_index(self.s),
                "n_rebuilds": lib.orc_rebuild_count(self.s),
                "n_candidates": lib.orc_n_candidates(self.s),
                "n_cells_1d": lib.orc_n_cells_1d(self.s)}


def min_distance(h, s1, u1, l1, s2, u2, l2):
    out = np.empty(10)
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (h, s1, u1, s2, u2)]
    lib.orc_min_distance(_dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), l1, _dp(arrs[3]),
                         _dp(arrs[4]), l2, _dp(out))
    return out[:3], out[3], out[4:7], out[7:10]


def pair_force(c12, c6, rcut2, fcut, dr, dr2, c1, c2):
    f = np.zeros(3)
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (dr, c1, c2)]
    ok = lib.orc_pair_force(c12, c6, rcut2, fcut, _dp(a[0]), dr2, _dp(a[1]), _dp(a[2]), _dp(f))
    return ok, f


def tridiag(a, b, c, d):
    a, b, c, d = (np.array(x, dtype=np.float64) for x in (a, b, c, d))
    lib.orc_tridiag(len(b), _dp(a), _dp(b), _dp(c), _dp(d))
    return d


def cell_neighbors(n, x, y, z):
    out = np.empty(26, dtype=np.int32)
    k = lib.orc_cell_neighbors(n, x, y, z, _ip(out))
    return out[:k]


def cell_key(n, s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib.orc_cell_key(n, _dp(s))


def philox(ctr, key):
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    lib.orc_philox4x32(c, k, o)
    return list(o)


def noise(seed, site, m):
    xi = np.empty(3)
    lib.orc_noise(seed, site, m, _dp(xi))
    return xi


def bending_forces(u_soa, kappa, bl):
    nb = len(u_soa) // 3
    n = nb + 1
    f = np.empty(3 * n)
    ke = np.empty(max(n - 2, 1))
    u = np.ascontiguousarray(u_soa, dtype=np.float64)
    lib.orc_bending_forces(n, _dp(u), kappa, bl, _dp(f), _dp(ke))
    return f, ke[:n - 2]


def bond_site_forces(u_soa, f_soa, t_soa, bl):
    """Site forces (3 x n SoA) of one filament from per-bond orientation,
    force and torque (3 x nb SoA each)."""
    nb = len(u_soa) // 3
    n = nb + 1
    out = np.empty(3 * n)
    u = np.ascontiguousarray(u_soa, dtype=np.float64)
    f = np.ascontiguousarray(f_soa, dtype=np.float64)
    t = np.ascontiguousarray(t_soa, dtype=np.float64)
    lib.orc_bond_site_forces(n, bl, _dp(u), _dp(f), _dp(t), _dp(out))
    return out


def xlink_force(h, h_inv, p0, p1, k, r0, kc=-1.0):
    """(force on anchor 0, spring length) of one crosslink spring."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    hi = np.ascontiguousarray(h_inv, dtype=np.float64)
    p0 = np.ascontiguousarray(p0, dtype=np.float64)
    p1 = np.ascontiguousarray(p1, dtype=np.float64)
    out = np.empty(3)
    length = lib.orc_xlink_force(_dp(h), _dp(hi), _dp(p0), _dp(p1), k, r0, kc, _dp(out))
    return out, length
