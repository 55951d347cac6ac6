"""Reader of the binary dumps written by oracle/ref_driver.cpp."""
from __future__ import annotations

import numpy as np

_DT = {"f4": np.float32, "f8": np.float64, "i4": np.int32, "i8": np.int64}


def read_dump(path: str):
    with open(path, "rb") as fh:
        data = fh.read()
    nl = data.index(b"\n")
    head = data[:nl].decode().split()
    assert head[0] == "OLBDUMP" and head[1] == "v1", head[:2]
    meta = {}
    for kv in head[2:]:
        k, v = kv.split("=", 1)
        try:
            meta[k] = int(v)
        except ValueError:
            try:
                meta[k] = float(v)
            except ValueError:
                meta[k] = v
    pos = nl + 1
    arrays = {}
    while pos < len(data):
        nl = data.index(b"\n", pos)
        _, name, dt, count = data[pos:nl].decode().split()
        count = int(count)
        dtype = np.dtype(_DT[dt])
        start = nl + 1
        arrays[name] = np.frombuffer(data[start:start + count * dtype.itemsize], dtype=dtype).copy()
        pos = start + count * dtype.itemsize
    return meta, arrays
