"""Binary checkpoints of a Hermite-4 run (the format of
gravidy_amd/host/hip/H4Checkpoint.hpp, shared with the gravidy-hip driver).

Layout, little endian: b"H4HCKPT" + version byte 1, then int64/float64
header fields in the order of HEADER, then r[n][4], v[n][4], f[n][6], t[n],
dt[n], a2[n][3], a3[n][3] as float64, then the CRC-32 of all preceding
bytes as a uint64.
"""
import os
import zlib

import numpy as np

MAGIC = b"H4HCKPT\x01"

# (name, kind): "u" uint64, "i" int64, "d" float64 -- after the magic
HEADER = (("n", "u"), ("e2", "d"), ("eta", "d"), ("dt_min", "d"), ("dt_max", "d"),
          ("precision", "i"), ("itime", "d"), ("atime", "d"), ("interval", "d"),
          ("output_factor", "i"), ("snapshot_number", "i"), ("energy_ini", "d"),
          ("iterations", "i"), ("nsteps", "i"), ("interactions", "i"))
ARRAYS = (("r", 4), ("v", 4), ("f", 6), ("t", 1), ("dt", 1), ("a2", 3), ("a3", 3))
_DT = {"u": "<u8", "i": "<i8", "d": "<f8"}


class CheckpointError(ValueError):
    pass


def checksum(data):
    """CRC-32 (ISO-HDLC, as zlib) of the bytes, zero-extended to 64 bits."""
    return zlib.crc32(data) & 0xFFFFFFFF


def write_checkpoint(path, state):
    """state: dict with the HEADER fields and ARRAYS (any float arrays of
    shape (n, k) or (n,)).  Written via a temporary file renamed in place."""
    n = int(state["n"])
    parts = [MAGIC]
    for name, kind in HEADER:
        parts.append(np.array([state[name]], dtype=_DT[kind]).tobytes())
    for name, k in ARRAYS:
        a = np.ascontiguousarray(state[name], dtype="<f8").reshape(-1)
        if a.size != n * k:
            raise CheckpointError(f"{name}: {a.size} values, expected {n * k}")
        parts.append(a.tobytes())
    body = b"".join(parts)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(body)
        fh.write(np.array([checksum(body)], dtype="<u8").tobytes())
    os.replace(tmp, path)


def read_checkpoint(path):
    with open(path, "rb") as fh:
        data = fh.read()
    if data[:7] != MAGIC[:7]:
        raise CheckpointError(f"{path} is not a Hermite4HIP checkpoint")
    if data[7:8] != MAGIC[7:8]:
        raise CheckpointError(f"{path}: unsupported checkpoint version")
    if len(data) < 16:
        raise CheckpointError(f"{path}: truncated")
    n = int(np.frombuffer(data, "<u8", 1, 8)[0])
    want = 8 + 8 * len(HEADER) + 8 * n * sum(k for _, k in ARRAYS) + 8
    if n == 0 or len(data) != want:
        raise CheckpointError(f"{path}: size does not match its particle count (truncated?)")
    body = data[:-8]
    if int(np.frombuffer(data, "<u8", 1, len(data) - 8)[0]) != checksum(body):
        raise CheckpointError(f"{path}: checksum mismatch (corrupted checkpoint)")
    out = {}
    off = 8
    for name, kind in HEADER:
        v = np.frombuffer(data, _DT[kind], 1, off)[0]
        out[name] = float(v) if kind == "d" else int(v)
        off += 8
    for name, k in ARRAYS:
        a = np.frombuffer(data, "<f8", n * k, off).copy()
        out[name] = a.reshape(n, k) if k > 1 else a
        off += 8 * n * k
    return out
