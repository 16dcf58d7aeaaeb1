"""The .spec / .posit files this library writes, against what the reference's
own readers (SpecReader / PositReader, analysis/simcore_tools.py) returned
for the same files: tests/golden/reference_reader_fixture.npz, made by
tests/golden/make_reference_reader_fixture.py.  The files are rewritten here
from the fixture's system; the reference readers' values must equal the
written positions, parameters and per-filament records, and this library's
reader must return the same numbers.  The reference readers go through a
text file parsed by pandas, whose float parser can miss the last bit, so
their values are compared to 1e-15; this library's binary reader is exact."""
import os

import numpy as np

import simcore_amd as sc
from simcore_amd import output
from simcore_amd._lib import ScFilamentParams

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "reference_reader_fixture.npz")


def _system(fx):
    p = sc.params_init(float(fx["system_radius"]), delta=float(fx["delta"]), seed=11)
    rows = fx["fil"]
    fil = (ScFilamentParams * len(rows))()
    for f, (bl, lp, gpar, gperp, d) in enumerate(rows):
        fil[f].bond_length, fil[f].persistence_length = bl, lp
        fil[f].gamma_par, fil[f].gamma_perp, fil[f].diameter = gpar, gperp, d
    return p, sc.Filaments(fx["site_off"], fx["r1"], fil), fx["r2"]


def test_reference_readers_read_our_files(tmp_path):
    fx = np.load(FIXTURE)
    p, fil, r2 = _system(fx)
    spec = tmp_path / output.file_name("fixture", "fil", "spec")
    posit = tmp_path / output.file_name("fixture", "fil", "posit")
    for path in (spec, posit):
        output.write_header(path, 2, 1, p.delta)
    for r in (fil.r, r2):
        output.write_spec_frame(spec, fil, r)
        output.write_posit_frame(posit, p, fil, r)

    # the reference's SpecReader: header, first filament's record, and every
    # site position (filament by filament, x y z per site) of both frames
    names = list(fx["spec_params_names"])
    sp = dict(zip(names, fx["spec_params"]))
    assert (sp["n_steps"], sp["n_spec"], sp["delta"]) == (2, 1, p.delta)
    assert sp["n_filaments"] == fil.nf and sp["n_sites"] == 7
    assert sp["diameter"] == fil.fil[0].diameter
    assert sp["bond_length"] == fil.fil[0].bond_length
    assert sp["length"] == fil.fil[0].bond_length * 6
    assert sp["persistence_length"] == fil.fil[0].persistence_length
    ns = fil.ns
    head, frames = output.read_spec(spec)
    assert head["n_steps"] == 2 and head["n_every"] == 1
    for k, (fr, r) in enumerate(zip(frames, (fil.r, r2))):
        written = r.reshape(3, ns).T.ravel()
        np.testing.assert_allclose(fx["spec_sites"][k], written, rtol=1e-15, atol=1e-15)
        ours = np.concatenate([rec["positions"].ravel() for rec in fr])
        np.testing.assert_array_equal(ours, written)

    # the reference's PositReader: header and, per frame, n then 11 values per
    # filament (position, scaled position, orientation, diameter, length)
    pp = dict(zip(list(fx["posit_params_names"]), fx["posit_params"]))
    assert (pp["n_steps"], pp["n_posit"], pp["delta"]) == (2, 1, p.delta)
    head, frames = output.read_posit(posit)
    for k, fr in enumerate(frames):
        row = fx["posit_rows"][k]
        assert row[0] == fil.nf
        ref = row[1:].reshape(fil.nf, 11)
        np.testing.assert_allclose(ref, fr, rtol=1e-15, atol=1e-15)
        np.testing.assert_array_equal(ref[:, 9], fil.fil_array("diameter"))
        np.testing.assert_allclose(ref[:, 10], fil.fil_array("bond_length") * 6, rtol=1e-15)
        # orientation: a unit vector; scaled position inside the box
        np.testing.assert_allclose(np.linalg.norm(ref[:, 6:9], axis=1), 1.0, rtol=1e-12)
        assert np.all(np.abs(ref[:, 3:6]) <= 0.5)
