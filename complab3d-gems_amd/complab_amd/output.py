"""The reference's output contract: one VTK ImageData file per field per
snapshot and the diagnostics CSV.

Files (complab_functions.hh:82-166; written at complab.cpp:2268-2286 every
save_VTK_interval steps and once more after the last step, 3270-3310):

* ``<species>_<iteration, 7 digits>.vti`` -- array "Density" (Float64): the
  concentration over the reaction box x = 1..nx-2, every y and z;
* ``<microbe>_<iteration>.vti`` -- the microbe's biomass, the same way;
* ``<mask_filename>_<iteration>.vti`` -- the mask codes as "Density";
* ``nsLattice_<iteration>.vti`` (flow runs) -- "velocityNorm" and the
  3-component "velocity" (Float32, lattice units).

Each file is an ImageData piece with WholeExtent "0 nx-3 0 ny-1 0 nz-1",
origin 0 and spacing 1 (the writer's dx argument), point data in x-fastest
order, each array inline binary in base64: the UInt32 byte count encoded on
its own, then the values encoded on their own (the inline convention of the
VTK XML format; the reference's collector reads both conventions,
ABC_sweep_package/tools/collect_to_h5.py:108-135).

A closed voxel has no populations here; the reference's lattice answers the
fictitious density of the dynamics attached there (a Palabos BounceBack cell
its stored rho, a NoDynamics cell its own, complab3d_processors_part3.hh:429-
440).  Those values live in Palabos, which is not available here: the writer
takes them as `closed_density` (NoDynamics, BounceBack), by default (1, 0),
the Palabos constructors' defaults as this build states them -- parity
unpinned.  Velocities are zero at closed voxels, as computeVelocity gives there.

``summary.csv`` (complab3d_diagnostics.hh:255-290): the reference's comment
header, then ``iteration,porosity,open_voxels`` and, per field,
``_total,_mean,_min,_max,_held``; numbers as printf %.17g, the file reopened
per row so that a run cut short leaves a readable file.  Rows are taken after
whole steps (the reference takes them after the step's reaction, before its
CA and streaming, complab.cpp:2661-2790).
"""
import base64
import os
import re
import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import params as P

_VTK_TYPES = {"Float64": np.float64, "Float32": np.float32, "Int32": np.int32, "UInt8": np.uint8}


def snapshot_name(prefix: str, iteration: int) -> str:
    """createFileName(prefix, iteration, 7) + '.vti'."""
    return "%s%07d.vti" % (prefix, iteration)


def _b64_block(raw: bytes) -> str:
    return base64.b64encode(struct.pack("<I", len(raw))).decode("ascii") + base64.b64encode(raw).decode("ascii")


def write_vti(path: str, arrays: Sequence[Tuple[str, np.ndarray, str]], spacing: float = 1.0):
    """One ImageData file.  `arrays` holds (name, values, VTK type); values are
    (nz, ny, nx) for a scalar or (3, nz, ny, nx) for a vector, all over the same
    grid; they are written x fastest."""
    if not arrays:
        raise ValueError("no array to write")
    shape = arrays[0][1].shape[-3:]
    nz, ny, nx = shape
    ext = "0 %d 0 %d 0 %d" % (nx - 1, ny - 1, nz - 1)
    sp = "%.17g" % spacing
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n',
           '<ImageData WholeExtent="%s" Origin="0 0 0" Spacing="%s %s %s">\n' % (ext, sp, sp, sp),
           '<Piece Extent="%s">\n' % ext, '<PointData>\n']
    for name, vals, vtype in arrays:
        if vtype not in _VTK_TYPES:
            raise ValueError("VTK type %s is not written here" % vtype)
        a = np.asarray(vals)
        if a.shape[-3:] != shape:
            raise ValueError("array %s has shape %s, the grid is %s" % (name, a.shape, shape))
        if a.ndim == 4:
            ncomp = a.shape[0]
            flat = np.ascontiguousarray(np.moveaxis(a, 0, -1), dtype=_VTK_TYPES[vtype])   # (nz, ny, nx, c)
        elif a.ndim == 3:
            ncomp = 1
            flat = np.ascontiguousarray(a, dtype=_VTK_TYPES[vtype])
        else:
            raise ValueError("array %s: a scalar (nz, ny, nx) or a vector (c, nz, ny, nx)" % name)
        out.append('<DataArray type="%s" Name="%s" format="binary" encoding="base64" NumberOfComponents="%d">\n'
                   % (vtype, name, ncomp))
        out.append(_b64_block(flat.astype(flat.dtype.newbyteorder("<"), copy=False).tobytes()))
        out.append('\n</DataArray>\n')
    out += ['</PointData>\n', '</Piece>\n', '</ImageData>\n', '</VTKFile>\n']
    tmp = path + ".part"
    with open(tmp, "w") as f:
        f.write("".join(out))
    os.replace(tmp, path)


def read_vti(path: str) -> Tuple[Dict[str, np.ndarray], Tuple[int, int, int]]:
    """The arrays of an inline-binary ImageData file as written by write_vti:
    scalars (nz, ny, nx), vectors (c, nz, ny, nx); and (nx, ny, nz)."""
    with open(path) as f:
        text = f.read()
    m = re.search(r'WholeExtent="([-\d\s]+)"', text)
    e = [int(v) for v in m.group(1).split()]
    nx, ny, nz = e[1] - e[0] + 1, e[3] - e[2] + 1, e[5] - e[4] + 1
    out = {}
    for mt in re.finditer(r'<DataArray\b([^>]*)>(.*?)</DataArray>', text, re.S):
        attrs, body = mt.group(1), re.sub(r"\s", "", mt.group(2))
        name = re.search(r'Name="([^"]+)"', attrs).group(1)
        dt = np.dtype(_VTK_TYPES[re.search(r'type="(\w+)"', attrs).group(1)]).newbyteorder("<")
        nc = int(re.search(r'NumberOfComponents="(\d+)"', attrs).group(1))
        nbytes = struct.unpack("<I", base64.b64decode(body[:8]))[0]
        raw = base64.b64decode(body[8:])
        if len(raw) != nbytes:
            raise ValueError("%s: array %s holds %d bytes, its header says %d" % (path, name, len(raw), nbytes))
        a = np.frombuffer(raw, dtype=dt)
        out[name] = a.reshape(nz, ny, nx) if nc == 1 else np.moveaxis(a.reshape(nz, ny, nx, nc), -1, 0)
    return out, (nx, ny, nz)


def box_density(case, field: int, index: int, mask: np.ndarray, closed_density=(1.0, 0.0)) -> np.ndarray:
    """A density field over the reaction box as the reference's lattice
    reports it: the value at open voxels, the fictitious densities at closed
    ones (mask code SOLID: NoDynamics, any other closed code: BounceBack)."""
    d = case.read(field, index)[:, :, 1:-1].copy()
    m = mask[:, :, 1:-1]
    d[m == P.SOLID] = closed_density[0]
    d[(m != P.SOLID) & (m < P.PORE)] = closed_density[1]
    return d


class Snapshots:
    """Writes one snapshot of every field of a case into `out_dir`."""

    def __init__(self, out_dir: str, species: Sequence[str], microbes: Sequence[str] = (),
                 mask_filename: str = "maskLattice", flow: bool = False, closed_density=(1.0, 0.0),
                 energy: Sequence = ()):
        self.out_dir = out_dir
        self.energy = list(energy)                  # (microbe index, name) of the gated organisms
        self.species = list(species)
        self.microbes = list(microbes)
        self.mask_filename = mask_filename
        self.flow = flow
        self.closed_density = closed_density
        self.written: List[str] = []
        os.makedirs(out_dir, exist_ok=True)

    def _put(self, prefix, iteration, arrays):
        path = os.path.join(self.out_dir, snapshot_name(prefix, iteration))
        write_vti(path, arrays)
        self.written.append(path)

    def write(self, case, iteration: int):
        mask = case.read(P.FIELD_MASK)
        for s, name in enumerate(self.species):
            self._put(name + "_", iteration,
                      [("Density", box_density(case, P.FIELD_CONC, s, mask, self.closed_density), "Float64")])
        for m, name in enumerate(self.microbes):
            self._put(name + "_", iteration,
                      [("Density", box_density(case, P.FIELD_BIOMASS, m, mask, self.closed_density), "Float64")])
        self._put(self.mask_filename + "_", iteration,
                  [("Density", mask[:, :, 1:-1].astype(np.float64), "Float64")])
        if self.flow:
            u = np.stack([case.read(P.FIELD_VEL, c)[:, :, 1:-1] for c in range(3)])
            closed = mask[:, :, 1:-1] < P.PORE
            u[:, closed] = 0.0
            norm = np.sqrt((u * u).sum(axis=0))
            self._put("nsLattice_", iteration, [("velocityNorm", norm, "Float32"), ("velocity", u, "Float32")])
        # the energy snapshot (writeEnergyFields, complab.cpp:2152-2163): dG and
        # F_T of each gated organism at the standing concentrations
        for m, name in self.energy:
            self._put("dG_" + name + "_", iteration, [("DeltaG", case.read(P.FIELD_DG, m)[:, :, 1:-1], "Float64")])
            self._put("FT_" + name + "_", iteration, [("F_T", case.read(P.FIELD_FT, m)[:, :, 1:-1], "Float64")])

    def write_rates(self, case, iteration: int):
        """rate_<species>_<iteration>.vti: the captured step's rate of every
        species ("Rate", Float64, x = 1..nx-2; writeRateVTI,
        complab_functions.hh:163-166)."""
        for s, name in enumerate(self.species):
            self._put("rate_" + name + "_", iteration, [("Rate", case.read(P.FIELD_RATE, s)[:, :, 1:-1], "Float64")])


SUMMARY_HEADER = (
    "# CompLB3D run summary, one row per diagnostic interval\n"
    "# totals are over the physical domain x=1..nx-2; porosity and the means\n"
    "# use the OPEN voxel count, taken from the live mask so it tracks a\n"
    "# pore space that precipitation or dissolution is changing.\n"
    "# <name>_held is the mass sitting in wall and grain voxels, which\n"
    "# computeDensity cannot report because BounceBack and NoDynamics answer\n"
    "# from a stored number. The conserved quantity is _total + _held, and\n"
    "# that is what <conserve> is checked against.\n")


class SummaryCsv:
    """summary.csv of the diagnostics rows (complab3d_diagnostics.hh:255-290)."""

    def __init__(self, path: str):
        self.path = path
        self.rows = 0

    def append(self, row):
        if self.rows == 0:
            d = os.path.dirname(self.path)
            if d:
                os.makedirs(d, exist_ok=True)
            cols = "iteration,porosity,open_voxels"
            for f in row.fields:
                n = f.name
                cols += ",%s_total,%s_mean,%s_min,%s_max,%s_held" % (n, n, n, n, n)
            with open(self.path, "w") as fh:
                fh.write(SUMMARY_HEADER + cols + "\n")
        line = "%d,%.17g,%d" % (row.iteration, row.porosity, row.open_voxels)
        for f in row.fields:
            line += ",%.17g,%.17g,%.17g,%.17g,%.17g" % (f.total, f.mean, f.minv, f.maxv, f.held)
        with open(self.path, "a") as fh:
            fh.write(line + "\n")
        self.rows += 1


def read_summary(path: str):
    """(column names, rows as float arrays) of a summary.csv."""
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh if not ln.startswith("#")]
    cols = lines[0].split(",")
    rows = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]]) if len(lines) > 1 else np.zeros((0, len(cols)))
    return cols, rows
