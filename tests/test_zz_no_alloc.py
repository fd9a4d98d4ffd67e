"""bgk_field_check_planes, the allocation-free variant of bgk_field_check,
through the C ABI: on device planes the caller owns (torch tensors here), it
gives bgk_field_check's means bit for bit, on two calls in a row, the second
on a smaller lattice (the shared static scratch keeps the first call's
partials, which the second must not pick up)."""
import ctypes

import numpy as np
import pytest
import torch

from bgk_pcm_amd.lattice import BgkError, Lattice, load_library
from bgk_pcm_amd.params import Params

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)


def planes_on_device(lat, pitch):
    """rho, T, phi of a lattice as (Ny+2) x pitch device tensors (the
    columns past Nx+1 hold garbage the sums must not read)."""
    out = []
    for f in ("rho", "T", "phi"):
        a = np.full((lat.ny + 2, pitch), 1.0e300)
        a[:, :lat.nx + 2] = lat.download(f)
        out.append(torch.from_numpy(a).to("cuda"))
    return out


def field_check_planes(planes, nx, ny, pitch, stream):
    lib = load_library()
    out = (ctypes.c_double * 3)()
    rc = lib.bgk_field_check_planes(*(ctypes.c_void_p(t.data_ptr()) for t in planes), nx, ny,
                                    pitch, ctypes.c_void_p(stream), ctypes.cast(out, _D))
    if rc != 0:
        raise BgkError(rc, lib.bgk_last_error().decode())
    return list(out)


def test_planes_variant_equals_field_check_on_two_calls():
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    for nx, ny, steps, extra in ((300, 1100, 120, 37), (40, 30, 60, 5)):
        print(f"field_check_planes vs field_check: {nx}x{ny}, {steps} steps", flush=True)
        p = Params(nx=nx, ny=ny, length=float(ny), grav=1e-3)
        with Lattice(p) as lat:
            lat.step(steps)
            want = lat.field_check()
            planes = planes_on_device(lat, nx + 2 + extra)
        torch.cuda.synchronize()
        got = field_check_planes(planes, nx, ny, nx + 2 + extra, stream.cuda_stream)
        assert np.array_equal(np.array(got), np.array(want)), (got, want)


def test_planes_variant_rejects_bad_arguments_and_flags_nan():
    torch.cuda.set_device(0)
    p = Params(nx=20, ny=10)
    with Lattice(p) as lat:
        planes = planes_on_device(lat, 22)
    with pytest.raises(BgkError) as e:
        field_check_planes(planes, 20, 10, 21, None)  # pitch < Nx + 2
    assert e.value.code == -1
    planes[1][5, 7] = float("nan")
    torch.cuda.synchronize()
    with pytest.raises(BgkError) as e:
        field_check_planes(planes, 20, 10, 22, None)
    assert e.value.code == -6
