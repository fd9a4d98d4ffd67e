"""Device reductions and gathers against the oracle's serial sums.

The reference sums serially (field_check.f90:11-21, mass_check.f90:12-29,
diag.f90:63-104); the device sums are compensated and deterministic, so the
two agree to the serial sum's own rounding error: relative 1e-13 is written
here for sizes of a few thousand nodes.  Counts are exact.  Gathers
(probe.f90, prof_x.f90, prof_y.f90) are exact copies.
"""
import numpy as np
import pytest

from bgk_pcm_amd.lattice import BgkError, Lattice
from bgk_pcm_amd.params import Params
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

REL = 1e-13


@pytest.fixture(scope="module")
def pair():
    p = Params(nx=64, ny=48, re=1.0, length=48.0)
    o = Oracle(p)
    o.step(600)
    lat = Lattice(p)
    lat.step(600)
    yield p, o, lat
    lat.close()


def test_field_and_mass_check(pair):
    p, o, lat = pair
    np.testing.assert_allclose(lat.field_check(), o.field_check(), rtol=REL, atol=0)
    got, want = lat.mass_check(), o.mass_check()
    np.testing.assert_allclose(got[1:], want[1:], rtol=REL, atol=0)
    # mass percent defect: a difference of nearly equal sums
    assert abs(got[0] - want[0]) <= 1e-10 * max(1.0, abs(want[0]))


def test_diag_sums_and_values(pair):
    p, o, lat = pair
    got, want = lat.diag_sums(), o.diag_sums()
    np.testing.assert_allclose(got[:2], want[:2], rtol=REL, atol=1e-12)
    assert got[2] == want[2] and got[3] == want[3]
    v, ok = lat.diag(600)
    ov, ook = o.diag(600)
    assert ok == ook
    np.testing.assert_allclose(v, ov, rtol=1e-12, atol=1e-15)


def test_probe_and_profiles(pair):
    p, o, lat = pair
    fields = ("u", "v", "rho", "T", "phi")
    want = [o.field(f)[p.ny // 2, p.nx // 2] for f in fields]
    assert lat.probe() == want
    px = lat.prof_x()
    for c, f in enumerate(fields):
        assert np.array_equal(px[:, c], o.field(f)[p.ny // 2, 1:p.nx + 1])
    py = lat.prof_y()
    for c, f in enumerate(fields):
        assert np.array_equal(py[:, c], o.field(f)[1:p.ny + 1, p.nx // 2])


def test_gather_rejects_nodes_outside_the_lattice(pair):
    p, o, lat = pair
    with pytest.raises(BgkError) as e:
        lat.gather("rho", [p.nx + 2], [0])
    assert e.value.code == -1


def test_diag_needs_three_columns():
    with Lattice(Params(nx=1, ny=4)) as lat:
        with pytest.raises(BgkError) as e:
            lat.diag_sums()
        assert e.value.code == -4
