"""Array bundle format shared with oracle/ref_harness.cpp.

A bundle is a sequence of records ``{char name[32]; int32 dtype; int64 count;
payload}`` with dtype 0=f32, 1=i32, 2=u64, 3=u8, 4=u32.
"""
import struct

import numpy as np

_DT = {0: np.float32, 1: np.int32, 2: np.uint64, 3: np.uint8, 4: np.uint32}
_CODE = {np.dtype(v): k for k, v in _DT.items()}


def write_bundle(path, arrays):
    with open(path, "wb") as f:
        for name, arr in arrays.items():
            a = np.ascontiguousarray(arr)
            code = _CODE[a.dtype]
            nm = name.encode()[:31].ljust(32, b"\0")
            f.write(nm)
            f.write(struct.pack("<iq", code, a.size))
            f.write(a.tobytes())


def read_bundle(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off + 44 <= len(data):
        name = data[off:off + 32].split(b"\0")[0].decode()
        code, cnt = struct.unpack_from("<iq", data, off + 32)
        off += 44
        dt = np.dtype(_DT[code])
        nbytes = cnt * dt.itemsize
        out[name] = np.frombuffer(data[off:off + nbytes], dtype=dt).copy()
        off += nbytes
    return out
