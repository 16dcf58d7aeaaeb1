"""Ensemble runner (continuous batching of independent filament systems),
a thin mirror of the sc_ensemble_* C ABI.

The reference runs the n_runs of a parameter file one after another
(src/simulation_manager.cpp:461-545); here requests are queued, admitted by
device-memory footprint while slots are free, advanced in chunks -- up to
`batch_size` mixed-precision systems per set of grouped launches, the rest
on their own streams -- and retired as they finish.
"""
import ctypes
from ctypes import byref, c_int32, c_int64, c_void_p

import numpy as np

from ._lib import (SC_OK, SC_REQ_DONE, ScEnsembleParams, ScEnsembleStats,
                   ScParams, ScRequestInfo, SimcoreError, lib)
from .system import _dp, _ip

STATUS = {0: "queued", 1: "running", 2: "done"}


class Ensemble:
    def __init__(self, hbm_budget=0, max_active=16, steps_per_launch=64, batch_size=0):
        prm = ScEnsembleParams(int(hbm_budget), int(max_active), int(steps_per_launch),
                               int(batch_size))
        self.h = c_void_p()
        rc = lib.sc_ensemble_create(byref(prm), byref(self.h))
        if rc != SC_OK:
            raise SimcoreError(rc, "sc_ensemble_create")
        self._ns = {}

    def _check(self, rc, what):
        if rc != SC_OK:
            raise SimcoreError(rc, f"{what}: {lib.sc_ensemble_last_error(self.h).decode()}")

    def close(self):
        if self.h is not None and self.h.value:
            lib.sc_ensemble_destroy(self.h)
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

    def clock(self):
        """Seconds since the runner was created."""
        return lib.sc_ensemble_clock(self.h)

    def submit(self, params, filaments, n_steps, precision=None, use_graph=True,
               t_submit=-1.0, spatial_order=None):
        """Queue one system; returns its request id."""
        p = ScParams()
        ctypes.pointer(p)[0] = params
        if precision is not None:
            p.precision = {"fp64": 0, "mixed": 1}[precision]
        if spatial_order is not None:
            p.spatial_order = int(bool(spatial_order))
        p.use_graph = int(bool(use_graph))
        so = np.ascontiguousarray(filaments.site_off, dtype=np.int32)
        r = np.ascontiguousarray(filaments.r, dtype=np.float64)
        rid = c_int64()
        self._check(lib.sc_ensemble_submit(self.h, byref(p), filaments.nf, _ip(so), _dp(r),
                                           filaments.fil, int(n_steps), float(t_submit),
                                           byref(rid)), "sc_ensemble_submit")
        self._ns[rid.value] = filaments.ns
        return rid.value

    def progress(self):
        """One scheduling round; returns (running, queued)."""
        nr, nq = c_int32(), c_int32()
        self._check(lib.sc_ensemble_progress(self.h, byref(nr), byref(nq)),
                    "sc_ensemble_progress")
        return nr.value, nq.value

    def drain(self, timeout_s=0.0):
        self._check(lib.sc_ensemble_drain(self.h, float(timeout_s)), "sc_ensemble_drain")

    def request(self, rid):
        info = ScRequestInfo()
        self._check(lib.sc_ensemble_request(self.h, int(rid), byref(info)),
                    "sc_ensemble_request")
        d = {k: getattr(info, k) for k, _ in ScRequestInfo._fields_}
        d["state"] = STATUS.get(info.status, "failed")
        return d

    def result(self, rid):
        """Final site positions (3*Ns, SoA) of a finished request."""
        r = np.empty(3 * self._ns[rid])
        self._check(lib.sc_ensemble_result(self.h, int(rid), _dp(r)), "sc_ensemble_result")
        return r

    def done(self, rid):
        return self.request(rid)["status"] == SC_REQ_DONE

    def stats(self):
        st = ScEnsembleStats()
        self._check(lib.sc_ensemble_stats_get(self.h, byref(st)), "sc_ensemble_stats_get")
        return {k: getattr(st, k) for k, _ in ScEnsembleStats._fields_}
