"""Output formats against the reference's own writers.

tests/golden/formats_cav24x20.npz holds the split binary VTK files that
STEP_CUF's out2d.phi/rho/temp.f90 write and the dumpFile.dat of
STEP_7/dump_out.f90, after 50 steps of a 24x20 cavity, produced by the
reference sources compiled into oracle/_ref/vtkbin (make_golden.py).  The
oracle reproduces the reference state bit for bit, so the Python writer
fed with the oracle's fields must produce the same bytes, and the dump
reader must return the oracle's fields exactly.  (STEP_CUF's out2d.vel.f90
needs a constant of STEP_CUF's own module and is not in that build; the
velocity file's layout is checked by reading it back.)
"""
import json

import numpy as np

from bgk_pcm_amd import vtk
from bgk_pcm_amd.params import Params
from golden_io import GOLDEN
from oracle_py import Oracle

FIX = np.load(GOLDEN / "formats_cav24x20.npz", allow_pickle=False)
P = Params(**json.loads(str(FIX["params"])))
N = int(FIX["nsteps"])


def oracle_after():
    o = Oracle(P)
    o.step(N)
    return o


def test_binary_vtk_is_byte_identical_to_the_reference(tmp_path):
    o = oracle_after()
    for kind, field in (("phi", "phi"), ("rho", "rho"), ("T", "T")):
        path = vtk.write_binary_vtk(tmp_path, kind, N, o.field(field))
        want = FIX["bin_" + vtk.binary_vtk_name(kind, N)].tobytes()
        assert open(path, "rb").read() == want, kind


def test_binary_vtk_reads_back():
    o = oracle_after()
    hdr, data = vtk.read_binary_vtk_bytes(FIX["bin_" + vtk.binary_vtk_name("T", N)].tobytes())
    assert (hdr["nx"], hdr["ny"]) == (P.nx, P.ny)
    assert np.array_equal(data, o.field("T")[1:-1, 1:-1])


def test_velocity_file_layout(tmp_path):
    o = oracle_after()
    path = vtk.write_binary_vtk(tmp_path, "vel", N, o.field("u"), o.field("v"))
    hdr, data = vtk.read_binary_vtk(path)
    assert hdr["vector"] and hdr["lines"][-1] == "VECTORS velocity double"
    assert np.array_equal(data[..., 0], o.field("u")[1:-1, 1:-1])
    assert np.array_equal(data[..., 1], o.field("v")[1:-1, 1:-1])
    assert np.all(data[..., 2] == 0.0)


def test_reference_dump_reads_exactly(tmp_path):
    o = oracle_after()
    path = tmp_path / "dumpFile.dat"
    path.write_text(str(FIX["text_dumpFile"]))
    d = vtk.read_dump(path, P.nx, P.ny)
    assert d["it"] == N and d["phi"] is None
    for f in ["u", "v", "rho", "T"] + [f"f{k}" for k in range(9)] + [f"g{k}" for k in range(9)]:
        assert np.array_equal(d[f], o.field(f)), f


def test_extended_dump_round_trip(tmp_path):
    o = oracle_after()
    fields = {f: o.field(f) for f in ["u", "v", "rho", "T", "phi"] +
              [f"f{k}" for k in range(9)] + [f"g{k}" for k in range(9)]}
    vtk.write_dump(tmp_path / "d.dat", N, fields)
    d = vtk.read_dump(tmp_path / "d.dat", P.nx, P.ny)
    for f, a in fields.items():
        assert np.array_equal(d[f], a), f
