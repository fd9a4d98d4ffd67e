"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, its parameter struct matches the header byte for byte, and
it refuses to run (rather than falling back to the CPU) when no GPU is
visible."""
import ctypes
import os
import pathlib
import re
import subprocess

import pytest

from bgk_pcm_amd import lattice
from bgk_pcm_amd.params import BgkParamsC, Params

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "bgk_pcm.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(bgk_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = lattice.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(lattice.SIGNATURES), set(names) ^ set(lattice.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", str(lattice.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(bgk_\w+)$", out, re.M))
    assert set(names) <= exported


def test_abi_version():
    assert lattice.load_library().bgk_abi_version() == 4


@pytest.mark.parametrize("cname,ctype", [("bgk_params", BgkParamsC),
                                         ("bgk_ext", lattice.BgkExtC)])
def test_struct_layout_matches_header(tmp_path, cname, ctype):
    fields = [f for f, _ in ctype._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"',
           "int main(void){", f'printf("%zu\\n", sizeof({cname}));']
    src += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    src += ["return 0;}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(ctype)
    for f, off in zip(fields, vals[1:]):
        assert getattr(ctype, f).offset == off, f


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_gpu_means_error_not_fallback():
    with pytest.raises(lattice.BgkError) as e:
        lattice.Lattice(Params(nx=8, ny=8))
    assert e.value.code == -5
    assert "device" in str(e.value)


def test_invalid_parameters_are_rejected_before_touching_the_device():
    lib = lattice.load_library()
    h = ctypes.c_void_p()
    for bad in (Params(nx=0, ny=8), Params(nx=8, ny=8, west_bc=7), Params(nx=8, ny=8, east_bc=1),
                Params(nx=8, ny=8, tau_f=0.0), Params(nx=8, ny=8, precision=4)):
        c = bad.to_c()
        assert lib.bgk_create(ctypes.byref(c), ctypes.byref(h)) == -1
        assert lib.bgk_last_error()
    assert lib.bgk_step(None, 1) == -1


def test_bench_refuses_more_than_one_gpu():
    """The time step runs on one GPU: bench.py --gpus 2 exits with status 2
    and says so, before touching a device."""
    root = pathlib.Path(__file__).resolve().parents[1]
    r = subprocess.run(["python3", str(root / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
