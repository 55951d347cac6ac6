"""Reader of the binary dumps written by oracle/ref_driver.cpp.

Arrays are memory-mapped (read-only), so dumps of 512^3 blocks (10 GB per
population array) can be reduced in chunks without loading them whole.
"""
from __future__ import annotations

import numpy as np

_DT = {"f4": np.float32, "f8": np.float64, "i4": np.int32, "i8": np.int64}


def _value(v: str):
    for conv in (int, float):
        try:
            return conv(v)
        except ValueError:
            pass
    return v


def read_dump(path: str):
    arrays = {}
    with open(path, "rb") as fh:
        head = fh.readline().decode().split()
        assert head[0] == "OLBDUMP" and head[1] == "v1", head[:2]
        meta = {k: _value(v) for k, v in (kv.split("=", 1) for kv in head[2:])}
        while True:
            line = fh.readline()
            if not line:
                break
            _, name, dt, count = line.decode().split()
            count = int(count)
            dtype = np.dtype(_DT[dt])
            start = fh.tell()
            arrays[name] = np.memmap(path, dtype=dtype, mode="r", offset=start, shape=(count,))
            fh.seek(start + count * dtype.itemsize)
    return meta, arrays
