"""The reference's .spec / .posit layouts, written from host arrays and read
back; byte counts follow the field lists of Mesh::WriteSpec,
Filament::WriteSpec and Filament::WritePosit."""
import numpy as np

import simcore_amd as sc
from simcore_amd import output


def _fil():
    p = sc.params_init(15.0, delta=2e-4)
    n = np.array([5, 9, 3, 12], dtype=np.int32)
    fil = sc.generate_filaments(p, n, (n - 1).astype(float), [20.0, 5.0, 100.0, 50.0],
                                1.0, 2.0, seed=3)
    return p, fil


def test_spec_round_trip_and_size(tmp_path):
    p, fil = _fil()
    path = tmp_path / output.file_name("run", "fil", "spec")
    output.write_header(path, 1000, 10, p.delta)
    r2 = fil.r + 0.25
    output.write_spec_frame(path, fil, fil.r)
    output.write_spec_frame(path, fil, r2)
    n = np.diff(fil.site_off)
    per_frame = 4 + sum(8 * 3 + 4 + 24 * k + 8 + 1 for k in n)
    assert path.stat().st_size == 16 + 2 * per_frame
    head, frames = output.read_spec(path)
    assert head == {"n_steps": 1000, "n_every": 10, "delta": p.delta}
    assert len(frames) == 2
    for fr, r in zip(frames, (fil.r, r2)):
        rr = r.reshape(3, -1)
        for f, rec in enumerate(fr):
            s0, s1 = fil.site_off[f], fil.site_off[f + 1]
            np.testing.assert_array_equal(rec["positions"], rr[:, s0:s1].T)
            assert rec["bond_length"] == fil.fil[f].bond_length
            assert rec["length"] == fil.fil[f].bond_length * (s1 - s0 - 1)
            assert rec["persistence_length"] == fil.fil[f].persistence_length
            assert rec["diameter"] == fil.fil[f].diameter and rec["poly"] == 0


def test_posit_fields(tmp_path):
    p, fil = _fil()
    path = tmp_path / "x.posit"
    output.write_header(path, 50, 5, p.delta)
    output.write_posit_frame(path, p, fil, fil.r)
    head, frames = output.read_posit(path)
    assert path.stat().st_size == 16 + 4 + 88 * fil.nf
    a = frames[0]
    rr = fil.r.reshape(3, -1)
    L = 2 * p.system_radius
    for f in range(fil.nf):
        s0, s1 = fil.site_off[f], fil.site_off[f + 1]
        c = rr[:, s0:s1].mean(axis=1)
        np.testing.assert_allclose(a[f, 0:3], c, rtol=0, atol=1e-12)
        s = c / L
        s -= np.where(s < 0, np.trunc(s - 0.5), np.trunc(s + 0.5))
        np.testing.assert_allclose(a[f, 3:6], s, rtol=0, atol=1e-12)
        assert np.all(np.abs(a[f, 3:6]) <= 0.5)
        u = np.diff(rr[:, s0:s1], axis=1)
        u /= np.linalg.norm(u, axis=0)
        us = u.sum(axis=1) + u[:, -1]
        np.testing.assert_allclose(a[f, 6:9], us / np.linalg.norm(us), atol=1e-12)
        assert a[f, 9] == fil.fil[f].diameter
        assert a[f, 10] == fil.fil[f].bond_length * (s1 - s0 - 1)


def test_thermo_file_layout(tmp_path):
    """OutputManagerBase::InitThermo / WriteThermo
    (include/simcore/output_manager.hpp:156-183): a 20-byte header (int,
    int, double, int) and 160 bytes per frame (9 + 9 + 1 + 1 doubles)."""
    import struct
    path = tmp_path / "run.thermo"
    output.write_thermo_header(path, 5000, 100, 2e-4)
    rng = np.random.default_rng(1)
    frames = []
    for k in range(3):
        uc = np.diag([20.0, 20.0, 20.0]) + k
        pt = rng.normal(size=(3, 3))
        frames.append((uc, pt, float(np.trace(pt) / 3), 8000.0 + k))
        output.write_thermo_frame(path, *frames[-1])
    raw = path.read_bytes()
    assert len(raw) == 20 + 3 * 160
    assert struct.unpack_from("=iidi", raw, 0) == (5000, 100, 2e-4, 3)
    head, got = output.read_thermo(path)
    assert head == {"n_steps": 5000, "n_thermo": 100, "delta": 2e-4, "n_dim": 3}
    for (uc, pt, pr, vol), g in zip(frames, got):
        np.testing.assert_array_equal(g["unit_cell"], uc)
        np.testing.assert_array_equal(g["pressure_tensor"], pt)
        assert g["pressure"] == pr and g["volume"] == vol
    # the second frame's pressure tensor starts 20 + 160 + 72 bytes in
    assert struct.unpack_from("=d", raw, 20 + 160 + 72)[0] == frames[1][1][0, 0]
