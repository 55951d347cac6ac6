"""OpenLB checkpoints of the populations, read and written from the logical
view of a lattice.

SuperLattice::save (core/serializer.h:181, io/serializerIO.hh:40-64) writes
the lattice's serializable blocks as text: base64 of the 8-byte binary size,
then one base64 stream of all blocks.  For a single-block D3Q19 lattice
without per-cell fields those blocks are the 19 population columns in order,
each a CyclicColumn block [shift: int64][count: uint64][data]
(core/platform/cpu/sisd/column.hh:67-89, cpu/simd/column.hh:65-87):
  - CPU_SISD: count = padded cells, data = count values, population i of
    cell n at data[(n + shift) mod count];
  - CPU_SIMD: count = padded cells rounded up to whole 4 KiB pages, data =
    the ring written twice (2 * count values; the platform maps it twice),
    population i of cell n at data[(n + shift) mod count].
Values are the shifted populations f_i - t_i in the value type T.
"""
from __future__ import annotations

import base64

import numpy as np

Q = 19
PAGE = 4096


def page_aligned_count(ncells: int, itemsize: int) -> int:
    """getPageAlignedCount (core/platform/cpu/simd/column.h:134-145)"""
    size = ((ncells * itemsize - 1) // PAGE + 1) * PAGE
    return size // itemsize


def decode(text: bytes) -> bytes:
    """the binary blocks of a checkpoint file's text"""
    head = base64.b64decode(text[:12])
    size = int.from_bytes(head[:8], "little")
    body = base64.b64decode(text[12:])
    if len(body) != size:
        raise ValueError(f"checkpoint body holds {len(body)} bytes, its header says {size}")
    return body


def encode(body: bytes) -> bytes:
    return base64.b64encode(len(body).to_bytes(8, "little")) + base64.b64encode(body)


def read(path: str, ncells: int, value_dtype=np.float32) -> np.ndarray:
    """populations (19 x ncells, logical order) of a checkpoint of a lattice
    with `ncells` padded cells and value type `value_dtype`"""
    with open(path, "rb") as fh:
        body = decode(fh.read())
    item = np.dtype(value_dtype).itemsize
    if len(body) % Q:
        raise ValueError("a population checkpoint holds 19 equal column blocks")
    block = len(body) // Q
    out = np.empty((Q, ncells), dtype=value_dtype)
    for i in range(Q):
        b = body[i * block:(i + 1) * block]
        shift = int.from_bytes(b[0:8], "little", signed=True)
        count = int.from_bytes(b[8:16], "little")
        copies = (block - 16) // (count * item) if count else 0
        if count < ncells or copies not in (1, 2) or 16 + copies * count * item != block:
            raise ValueError("column block does not fit this lattice and value type")
        data = np.frombuffer(b[16:16 + count * item], dtype=value_dtype)
        idx = (np.arange(ncells, dtype=np.int64) + shift) % count
        out[i] = data[idx]
    return out


def write(path: str, populations: np.ndarray, layout: str = "cpu_simd", value_dtype=np.float32) -> None:
    """a checkpoint that SuperLattice::load of the given platform layout
    ("cpu_simd" or "cpu_sisd") reads back as these populations"""
    f = np.ascontiguousarray(populations, dtype=value_dtype)
    if f.shape[0] != Q:
        raise ValueError("populations must be 19 x ncells")
    ncells = f.shape[1]
    item = np.dtype(value_dtype).itemsize
    if layout == "cpu_simd":
        count, copies = page_aligned_count(ncells, item), 2
    elif layout == "cpu_sisd":
        count, copies = ncells, 1
    else:
        raise ValueError("layout is cpu_simd or cpu_sisd")
    parts = []
    for i in range(Q):
        ring = np.zeros(count, dtype=value_dtype)
        ring[:ncells] = f[i]
        parts.append((0).to_bytes(8, "little", signed=True))
        parts.append(count.to_bytes(8, "little"))
        parts.append(np.tile(ring, copies).tobytes())
    with open(path, "wb") as fh:
        fh.write(encode(b"".join(parts)))


def save_lattice(lattice, path: str, layout: str = "cpu_simd") -> None:
    """write a lattice's populations (evaluation view) as an OpenLB checkpoint"""
    f = lattice.get_populations()
    write(path, f, layout, np.float64 if lattice.storage == "f64" else np.float32)


def load_lattice(lattice, path: str) -> None:
    """set a lattice's populations from an OpenLB checkpoint"""
    dtype = np.float64 if lattice.storage == "f64" else np.float32
    lattice.set_populations(read(path, lattice.ncells, dtype))
