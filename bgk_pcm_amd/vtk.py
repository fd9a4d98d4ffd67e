"""On-disk formats of the reference driver, for the Python interface.

- Split binary VTK files of STEP_CUF (out2d.phi.f90, out2d.rho.f90,
  out2d.temp.f90, out2d.vel.f90): an ASCII header written with the
  reference's Fortran edit descriptors, then the interior nodes (j outer, i
  inner) as big-endian float64.  `write_binary_vtk` reproduces the files
  byte for byte; `read_binary_vtk` reads them (and the Fortran driver's).
- The restart file dumpFile.dat of STEP_7/dump_out.f90 (it, then per node a
  record u, v, rho, f0..f8 and a record T, g0..g8, list-directed), followed
  by a block with phi of every node, which the reference's restart lacks.
  `read_dump` reads both the reference's files (phi None) and the extended
  ones; `write_dump` writes the extended format with 17 significant digits
  (any Fortran list-directed read accepts it).

Arrays are (Ny+2, Nx+2) indexed [j, i] like the rest of the package.
"""
from __future__ import annotations

import os
import re

import numpy as np

TITLES = {"phi": ("Outphi_", "SCALARS phi double"),
          "rho": ("Outrho_", "SCALARS rho double"),
          "T": ("OutTem_", "SCALARS Temperature double"),
          "vel": ("Outvel_", "VECTORS velocity double")}


def _i(x: int, w: int) -> str:
    s = str(int(x))
    return s.rjust(w) if len(s) <= w else "*" * w  # Fortran Iw


def binary_vtk_header(nx: int, ny: int, title: str) -> bytes:
    lines = ["# vtk DataFile Version 2.0", "Campo", "BINARY", "DATASET STRUCTURED_POINTS",
             "DIMENSIONS" + _i(nx, 8) + _i(ny, 8) + _i(1, 8),          # (A,3I8)
             "ORIGIN " + _i(1, 10) + " " + _i(1, 10) + " " + _i(1, 10),  # (A7,I10,A1,...)
             "SPACING " + _i(1, 10) + " " + _i(1, 10) + " " + _i(1, 10),
             "POINT_DATA " [:10] + _i(nx * ny, 10),                     # (A10,I10)
             title]
    if title.startswith("SCALARS"):
        lines.append("LOOKUP_TABLE default")
    return ("\n".join(lines) + "\n").encode()


def binary_vtk_name(kind: str, it: int) -> str:
    return f"{TITLES[kind][0]}{it:08d}.vtk"


def write_binary_vtk(directory: str | os.PathLike, kind: str, it: int, field: np.ndarray,
                     v: np.ndarray | None = None) -> str:
    """Write one split file; kind in phi, rho, T (field (Ny+2, Nx+2)) or vel
    (field = u, v = v).  Returns the path."""
    ny, nx = field.shape[0] - 2, field.shape[1] - 2
    path = os.path.join(directory, binary_vtk_name(kind, it))
    inner = field[1:-1, 1:-1]
    if kind == "vel":
        data = np.stack([inner, v[1:-1, 1:-1], np.zeros_like(inner)], axis=-1)
    else:
        data = inner
    with open(path, "wb") as f:
        f.write(binary_vtk_header(nx, ny, TITLES[kind][1]))
        f.write(np.ascontiguousarray(data, dtype=">f8").tobytes())
    return path


def read_binary_vtk(path: str | os.PathLike) -> tuple[dict, np.ndarray]:
    """(header, data): data (Ny, Nx) for scalars, (Ny, Nx, 3) for vectors."""
    return read_binary_vtk_bytes(open(path, "rb").read())


def read_binary_vtk_bytes(raw: bytes) -> tuple[dict, np.ndarray]:
    lines, pos = [], 0
    while True:
        end = raw.index(b"\n", pos)
        lines.append(raw[pos:end].decode())
        pos = end + 1
        if lines[-1].startswith("LOOKUP_TABLE") or lines[-1].startswith("VECTORS"):
            break
    nx, ny, _ = (int(x) for x in lines[4][len("DIMENSIONS"):].split())
    vec = lines[-1].startswith("VECTORS")
    data = np.frombuffer(raw[pos:], dtype=">f8").astype(np.float64)
    data = data.reshape(ny, nx, 3) if vec else data.reshape(ny, nx)
    return {"lines": lines, "nx": nx, "ny": ny, "vector": vec}, data


_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][-+]?\d+)?")


def read_dump(path: str | os.PathLike, nx: int, ny: int) -> dict:
    """Fields of dumpFile.dat: it, u, v, rho, T, f0..f8, g0..g8 and phi (None
    for the reference's own files, which end after the node records)."""
    nums = _NUM.findall(open(path).read())
    vals = np.array([float(x.replace("d", "e").replace("D", "e")) for x in nums])
    n = (nx + 2) * (ny + 2)
    it = int(vals[0])
    rec = vals[1:1 + 22 * n].reshape(ny + 2, nx + 2, 22)
    out = {"it": it, "u": rec[..., 0], "v": rec[..., 1], "rho": rec[..., 2],
           "T": rec[..., 12]}
    for k in range(9):
        out[f"f{k}"] = rec[..., 3 + k]
        out[f"g{k}"] = rec[..., 13 + k]
    rest = vals[1 + 22 * n:]
    out["phi"] = rest[:n].reshape(ny + 2, nx + 2) if rest.size >= n else None
    return out


def write_dump(path: str | os.PathLike, it: int, fields: dict) -> None:
    """Write the extended restart file from fields as returned by read_dump
    (phi required)."""
    ny, nx = fields["u"].shape[0] - 2, fields["u"].shape[1] - 2
    g = lambda x: repr(float(x))
    with open(path, "w") as f:
        f.write(f" {it}\n")
        for j in range(ny + 2):
            for i in range(nx + 2):
                a = [fields["u"][j, i], fields["v"][j, i], fields["rho"][j, i]] + \
                    [fields[f"f{k}"][j, i] for k in range(9)]
                b = [fields["T"][j, i]] + [fields[f"g{k}"][j, i] for k in range(9)]
                f.write(" " + " ".join(g(x) for x in a) + "\n")
                f.write(" " + " ".join(g(x) for x in b) + "\n")
        for j in range(ny + 2):
            for i in range(nx + 2):
                f.write(" " + g(fields["phi"][j, i]) + "\n")
