"""The reference's binary .spec / .posit trajectory files for a filament
species and its .thermo file: writers over the C ABI (sc_output_header,
sc_spec_frame, sc_posit_frame, sc_thermo_header, sc_thermo_frame) and plain
readers of the same layouts (the reference's analysis readers,
analysis/simcore_tools.py, parse the trajectory files the same way)."""
import ctypes
import struct
from ctypes import byref

import numpy as np

from ._lib import SC_OK, ScParams, SimcoreError, lib
from .system import _dp, _ip


def file_name(run_name, species_name, kind):
    """<run>_filament_<species>.<kind> (SpeciesBase::InitPositFile,
    src/species_base.cpp:13-18)."""
    return f"{run_name}_filament_{species_name}.{kind}"


def write_header(path, n_steps, n_every, delta):
    rc = lib.sc_output_header(str(path).encode(), int(n_steps), int(n_every), float(delta))
    if rc != SC_OK:
        raise SimcoreError(rc, f"sc_output_header({path})")


def write_spec_frame(path, filaments, r):
    so = np.ascontiguousarray(filaments.site_off, dtype=np.int32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    rc = lib.sc_spec_frame(str(path).encode(), filaments.nf, _ip(so), _dp(r), filaments.fil)
    if rc != SC_OK:
        raise SimcoreError(rc, f"sc_spec_frame({path})")


def write_posit_frame(path, params, filaments, r):
    so = np.ascontiguousarray(filaments.site_off, dtype=np.int32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    p = ScParams()
    ctypes.pointer(p)[0] = params
    rc = lib.sc_posit_frame(str(path).encode(), byref(p), filaments.nf, _ip(so), _dp(r),
                            filaments.fil)
    if rc != SC_OK:
        raise SimcoreError(rc, f"sc_posit_frame({path})")


def _header(buf):
    n_steps, n_every, delta = struct.unpack_from("=iid", buf, 0)
    return {"n_steps": n_steps, "n_every": n_every, "delta": delta}, 16


def read_spec(path):
    """(header, frames); a frame is a list of dicts with diameter, length,
    bond_length, positions (n_sites x 3), persistence_length, poly."""
    buf = open(path, "rb").read()
    head, pos = _header(buf)
    frames = []
    while pos < len(buf):
        (num,) = struct.unpack_from("=i", buf, pos)
        pos += 4
        fr = []
        for _ in range(num):
            d, length, bl, n = struct.unpack_from("=dddi", buf, pos)
            pos += 28
            x = np.frombuffer(buf, dtype=np.float64, count=3 * n, offset=pos).reshape(n, 3)
            pos += 24 * n
            lp, poly = struct.unpack_from("=dB", buf, pos)
            pos += 9
            fr.append({"diameter": d, "length": length, "bond_length": bl,
                       "positions": x.copy(), "persistence_length": lp, "poly": poly})
        frames.append(fr)
    return head, frames


def read_posit(path):
    """(header, frames); a frame is an (n, 11) array: position[3],
    scaled_position[3], orientation[3], diameter, length."""
    buf = open(path, "rb").read()
    head, pos = _header(buf)
    frames = []
    while pos < len(buf):
        (num,) = struct.unpack_from("=i", buf, pos)
        pos += 4
        a = np.frombuffer(buf, dtype=np.float64, count=11 * num, offset=pos).reshape(num, 11)
        pos += 88 * num
        frames.append(a.copy())
    return head, frames


def write_thermo_header(path, n_steps, n_thermo, delta, n_dim=3):
    """<run>.thermo header (OutputManagerBase::InitThermo,
    include/simcore/output_manager.hpp:156-168)."""
    rc = lib.sc_thermo_header(str(path).encode(), int(n_steps), int(n_thermo), float(delta),
                              int(n_dim))
    if rc != SC_OK:
        raise SimcoreError(rc, f"sc_thermo_header({path})")


def write_thermo_frame(path, unit_cell, pressure_tensor, pressure, volume):
    """One frame of OutputManagerBase::WriteThermo
    (include/simcore/output_manager.hpp:170-183)."""
    uc = np.ascontiguousarray(unit_cell, dtype=np.float64).ravel()
    pt = np.ascontiguousarray(pressure_tensor, dtype=np.float64).ravel()
    if uc.size != 9 or pt.size != 9:
        raise ValueError("unit cell and pressure tensor are 3 x 3")
    rc = lib.sc_thermo_frame(str(path).encode(), _dp(uc), _dp(pt), float(pressure),
                             float(volume))
    if rc != SC_OK:
        raise SimcoreError(rc, f"sc_thermo_frame({path})")


def read_thermo(path):
    """(header, frames); header n_steps, n_thermo, delta, n_dim; a frame is
    a dict of unit_cell (3x3), pressure_tensor (3x3), pressure, volume."""
    buf = open(path, "rb").read()
    n_steps, n_thermo, delta, n_dim = struct.unpack_from("=iidi", buf, 0)
    head = {"n_steps": n_steps, "n_thermo": n_thermo, "delta": delta, "n_dim": n_dim}
    pos, frames = 20, []
    while pos < len(buf):
        a = np.frombuffer(buf, dtype=np.float64, count=20, offset=pos)
        pos += 160
        frames.append({"unit_cell": a[:9].reshape(3, 3).copy(),
                       "pressure_tensor": a[9:18].reshape(3, 3).copy(),
                       "pressure": float(a[18]), "volume": float(a[19])})
    return head, frames
