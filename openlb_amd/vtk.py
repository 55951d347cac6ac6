"""VTK output of a lattice in OpenLB's SuperVTMwriter3D layout.

SuperVTMwriter3D (io/superVtmWriter3D.hh) writes, per output step iT and
cuboid iC:
  - DIR/data/NAME_iT%07d.vtm, a vtkMultiBlockDataSet naming the pieces
    (preambleVTM/dataVTM/closeVTM, lines 384-443);
  - DIR/data/NAME_iT%07diC%05d.vti, a VTK ImageData piece whose WholeExtent
    runs from -overlap to extent+overlap-1 with the cuboid origin and spacing
    (preambleVTI, lines 313-339), one Float32 DataArray per functor with the
    values of every point, x fastest, components interleaved (dataArray,
    lines 445-536): base64 of a zlib stream behind the prefix
    {1, raw bytes, raw bytes, compressed bytes} (compress, the default),
    base64 of the raw floats behind their byte count (binary), or text;
  - DIR/NAME.pvd, the master collection, one DataSet line per write
    (createMasterFile, dataPVDmaster: the closing 25 bytes are rewritten).
The functors are those of the lattice functor library that a case adds most:
velocity and density (SuperLatticeVelocity3D / Density3D, lattice units),
physVelocity and physPressure (SuperLatticePhysVelocity3D / PhysPressure3D,
through the converter's velocity and pressure factors).  Values come from
Lattice.compute_rho_u, which evaluates every cell through the momenta of its
dynamics as Cell::computeRhoU does; overlap points on a periodic axis take the
value of their periodic image, as the reference's communicate() before
writing makes them.
"""
from __future__ import annotations

import base64
import os
import struct
import zlib

import numpy as np

FUNCTORS = ("velocity", "density", "physVelocity", "physPressure")
NCOMP = {"velocity": 3, "density": 1, "physVelocity": 3, "physPressure": 1}


def _g(v: float) -> str:
    """a number as std::ostream prints it by default (precision 6, %g)"""
    return "%g" % v


def file_name(name: str, iT: int | None = None, iC: int | None = None) -> str:
    """createFileName (io/fileName.hh:34-74)"""
    s = name
    if iT is not None:
        s += "_iT%07d" % iT
    if iC is not None:
        s += "iC%05d" % iC
    return s


def data_array(name: str, values: np.ndarray, ncomp: int, mode: str = "compress") -> str:
    """one DataArray element (dataArray, superVtmWriter3D.hh:445-536);
    `values` are the point values in file order (x fastest, components
    interleaved), written as Float32"""
    raw = np.ascontiguousarray(values, dtype="<f4").reshape(-1).tobytes()
    head = (f'<DataArray type="Float32" Name="{name}" NumberOfComponents="{ncomp}" ')
    if mode == "compress":
        packed = zlib.compress(raw, -1)
        prefix = struct.pack("<4I", 1, len(raw), len(raw), len(packed))
        body = base64.b64encode(prefix).decode() + base64.b64encode(packed).decode()
        head += 'format="binary" encoding="base64">\n'
    elif mode == "binary":
        body = base64.b64encode(struct.pack("<I", len(raw))).decode() + base64.b64encode(raw).decode()
        head += 'format="binary" encoding="base64">\n'
    elif mode == "ascii":
        f = np.frombuffer(raw, dtype="<f4")
        body = "".join(_g(float(v)) + " " for v in f)
        head += 'format="ascii" >\n'
    else:
        raise ValueError("mode is compress, binary or ascii")
    return head + body + "\n</DataArray>\n"


def vti_text(arrays, extent, overlap: int, origin, delta: float, mode: str = "compress") -> str:
    """a whole .vti piece: `arrays` is a list of (name, ncomp, values)"""
    lo = -overlap
    hi = [e + overlap - 1 for e in extent]
    ext = f"{lo} {hi[0]} {lo} {hi[1]} {lo} {hi[2]}"
    t = '<?xml version="1.0"?>\n<VTKFile type="ImageData" version="0.1" '
    t += ('byte_order="LittleEndian" compressor="vtkZLibDataCompressor">\n' if mode == "compress"
          else 'byte_order="LittleEndian">\n')
    t += (f'<ImageData WholeExtent="{ext}" Origin="{_g(origin[0])} {_g(origin[1])} {_g(origin[2])}" '
          f'Spacing="{_g(delta)} {_g(delta)} {_g(delta)}">\n')
    t += f'<Piece Extent="{ext}">\n<PointData>\n'
    for name, ncomp, values in arrays:
        t += data_array(name, values, ncomp, mode)
    t += "</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n"
    return t


def vtm_text(name: str, iT: int, ncuboids: int = 1) -> str:
    t = ('<?xml version="1.0"?>\n<VTKFile type="vtkMultiBlockDataSet" version="1.0" '
         'byte_order="LittleEndian">\n<vtkMultiBlockDataSet>\n')
    for iC in range(ncuboids):
        t += (f'<Block index="{iC}" >\n<DataSet index= "0" file="{file_name(name, iT, iC)}.vti">\n'
              "</DataSet>\n</Block>\n")
    return t + "</vtkMultiBlockDataSet>\n</VTKFile>\n"


_PVD_CLOSE = "</Collection>\n</VTKFile>\n"


def point_values(lattice, overlap: int = 1, velocity_factor: float = 1.0, pressure_factor: float = 1.0,
                 char_pressure: float = 0.0, functors=FUNCTORS):
    """(name, ncomp, float32 values in file order) of each functor over the
    points -overlap .. extent+overlap-1 of the lattice"""
    p = lattice.padding
    if overlap > p:
        raise ValueError("the overlap written cannot exceed the lattice padding")
    rho, u = lattice.compute_rho_u()
    shape = lattice.shape
    value = np.float64 if lattice.storage == "f64" else np.float32
    rho = np.asarray(rho, dtype=value).reshape(shape)
    u = np.asarray(u, dtype=value).reshape(shape + (3,))
    idx = []
    for a, n in enumerate(lattice.extent):
        r = np.arange(-overlap, n + overlap)
        if lattice.periodic[a]:
            r = np.mod(r, n)
        idx.append(r + p)
    # file order: z slowest, then y, x fastest
    sel = np.ix_(idx[0], idx[1], idx[2])
    rho_p = rho[sel].transpose(2, 1, 0)
    u_p = u[sel[0], sel[1], sel[2], :].transpose(2, 1, 0, 3)
    out = []
    for name in functors:
        if name == "velocity":
            v = u_p
        elif name == "density":
            v = rho_p
        elif name == "physVelocity":
            v = u_p * value(velocity_factor)
        elif name == "physPressure":
            # (computeRho() - 1.0) / invCs2 in double, held in the value type
            # (latticePhysPressure3D.hh:73-76), then getPhysPressure
            # (core/unitConverter.h:359-362)
            lp = ((rho_p.astype(np.float64) - 1.0) / 3.0).astype(value)
            v = value(pressure_factor) * lp + value(char_pressure)
        else:
            raise ValueError(f"unknown functor {name}")
        out.append((name, NCOMP[name], np.asarray(v, dtype=np.float32).reshape(-1)))
    return out


class SuperVTMwriter3D:
    """SuperVTMwriter3D(name, overlap, binary, compress) for one lattice
    (one cuboid) placed at `origin` with spacing `delta`."""

    def __init__(self, name: str, out_dir: str = "./tmp/vtkData", overlap: int = 1, binary: bool = True,
                 compress: bool = True, origin=(0.0, 0.0, 0.0), delta: float = 1.0,
                 velocity_factor: float = 1.0, pressure_factor: float = 1.0, char_pressure: float = 0.0):
        self.name = name
        self.dir = out_dir
        self.overlap = overlap
        self.mode = "compress" if compress else ("binary" if binary else "ascii")
        self.origin = tuple(origin)
        self.delta = delta
        self.factors = (velocity_factor, pressure_factor, char_pressure)
        self.functors: list[str] = []

    def add_functor(self, name: str) -> None:
        if name not in NCOMP:
            raise ValueError(f"unknown functor {name}")
        self.functors.append(name)

    addFunctor = add_functor

    def create_master_file(self) -> None:
        os.makedirs(os.path.join(self.dir, "data"), exist_ok=True)
        with open(os.path.join(self.dir, file_name(self.name) + ".pvd"), "w") as fh:
            fh.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" '
                     'byte_order="LittleEndian">\n<Collection>\n' + _PVD_CLOSE)

    createMasterFile = create_master_file

    def write(self, lattice, iT: int) -> None:
        if not self.functors:
            raise RuntimeError("No functor to write")
        os.makedirs(os.path.join(self.dir, "data"), exist_ok=True)
        pvd = os.path.join(self.dir, file_name(self.name) + ".pvd")
        if os.path.exists(pvd):
            with open(pvd, "r+b") as fh:
                fh.seek(-len(_PVD_CLOSE), os.SEEK_END)
                fh.write((f'<DataSet timestep="{iT}" group="" part="" '
                          f'file="data/{file_name(self.name, iT)}.vtm"/>\n' + _PVD_CLOSE).encode())
        with open(os.path.join(self.dir, "data", file_name(self.name, iT) + ".vtm"), "w") as fh:
            fh.write(vtm_text(self.name, iT))
        vf, pf, cp = self.factors
        arrays = point_values(lattice, self.overlap, vf, pf, cp, self.functors)
        with open(os.path.join(self.dir, "data", file_name(self.name, iT, 0) + ".vti"), "w") as fh:
            fh.write(vti_text(arrays, lattice.extent, self.overlap, self.origin, self.delta, self.mode))


def read_vti(path: str) -> dict:
    """the DataArrays of a .vti piece written in any of the three modes, as
    float32 arrays (a reader for tests and post-processing)"""
    with open(path) as fh:
        return read_vti_text(fh.read())


def read_vti_text(text: str) -> dict:
    import re

    out = {}
    for m in re.finditer(r'<DataArray type="Float32" Name="([^"]+)" NumberOfComponents="(\d+)" '
                         r'format="(\w+)"[^>]*>\n(.*?)\n</DataArray>', text, re.S):
        name, ncomp, fmt, body = m.group(1), int(m.group(2)), m.group(3), m.group(4)
        if fmt == "ascii":
            out[name] = np.array(body.split(), dtype=np.float32)
            continue
        compressed = 'compressor="vtkZLibDataCompressor"' in text
        if compressed:
            prefix = base64.b64decode(body[:24])
            _, raw, _, packed = struct.unpack("<4I", prefix)
            data = zlib.decompress(base64.b64decode(body[24:]))
            assert len(data) == raw
        else:
            raw = struct.unpack("<I", base64.b64decode(body[:8]))[0]
            data = base64.b64decode(body[8:])
            assert len(data) == raw
        out[name] = np.frombuffer(data, dtype="<f4").copy()
    return out
