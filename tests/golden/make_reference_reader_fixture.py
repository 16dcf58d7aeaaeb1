"""Makes reference_reader_fixture.npz: what the reference's own readers return
for trajectory files written by this library.

A small filament system (six filaments of seven sites, host-generated) is
written as a .spec and a .posit file with two frames each by
sc_output_header / sc_spec_frame / sc_posit_frame; the reference's
SpecReader and PositReader (analysis/simcore_tools.py) convert them to their
human-readable form and parse that, and the values they return are saved
together with the system the files were written from.  Run in a container
with the reference checkout at /root/reference:

    python tests/golden/make_reference_reader_fixture.py

tests/test_reference_reader_fixture.py rewrites the files from the saved
system and checks the library's writer and reader against these values.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import simcore_amd as sc  # noqa: E402
from simcore_amd import output  # noqa: E402

REFERENCE_ANALYSIS = "/root/reference/analysis"


def system():
    """The fixture's system: six 7-site filaments, lp 20..70, box side 20."""
    p = sc.params_init(10.0, delta=2e-4, seed=11)
    n = np.full(6, 7, dtype=np.int32)
    lp = 20.0 + 10.0 * np.arange(6)
    fil = sc.generate_filaments(p, n, (n - 1).astype(float), lp, 1.0, 2.0, seed=5)
    r2 = fil.r + 0.125 * np.arange(fil.r.size) / fil.r.size
    return p, fil, r2


def write_files(directory, p, fil, r2):
    spec = os.path.join(directory, output.file_name("fixture", "fil", "spec"))
    posit = os.path.join(directory, output.file_name("fixture", "fil", "posit"))
    for path in (spec, posit):
        output.write_header(path, 2, 1, p.delta)
    for r in (fil.r, r2):
        output.write_spec_frame(spec, fil, r)
        output.write_posit_frame(posit, p, fil, r)
    return spec, posit


def main():
    sys.path.insert(0, REFERENCE_ANALYSIS)
    from simcore_tools import PositReader, SpecReader  # the reference's readers

    p, fil, r2 = system()
    with tempfile.TemporaryDirectory() as d:
        spec, posit = write_files(d, p, fil, r2)
        sr = SpecReader(spec)
        sr.make_human_readable()
        sparams, specs = sr.get_params_specs()
        pr = PositReader(posit)
        pr.make_human_readable()
        pparams, posits = pr.get_params_posit()
    out = {
        "site_off": np.asarray(fil.site_off, dtype=np.int32),
        "r1": np.asarray(fil.r), "r2": np.asarray(r2),
        "fil": np.array([[f.bond_length, f.persistence_length, f.gamma_par, f.gamma_perp,
                          f.diameter] for f in fil.fil]),
        "system_radius": np.array(p.system_radius), "delta": np.array(p.delta),
        "spec_params_names": np.array(list(sparams.columns)),
        "spec_params": sparams.to_numpy(dtype=float)[0],
        "spec_sites": specs.to_numpy(dtype=float),
        "posit_params_names": np.array(list(pparams.columns)),
        "posit_params": pparams.to_numpy(dtype=float)[0],
        "posit_rows": posits.to_numpy(dtype=float),
    }
    path = os.path.join(HERE, "reference_reader_fixture.npz")
    np.savez(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
