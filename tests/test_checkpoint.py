"""Binary checkpoint format (CPU): round trip, corruption and truncation
detection, and agreement between the Python writer and the C++ reader
(gravidy_amd/host/h4ckpt-info, built by the host Makefile)."""
import os
import subprocess
import zlib

import numpy as np
import pytest

from gravidy_amd.checkpoint import (ARRAYS, HEADER, CheckpointError, checksum, read_checkpoint,
                                    write_checkpoint)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "gravidy_amd", "host", "h4ckpt-info")


def sample_state(n=37, seed=3):
    rng = np.random.default_rng(seed)
    st = {"n": n, "e2": 9.99999905104687e-09, "eta": 0.009999999776482582,
          "dt_min": 2.0 ** -23, "dt_max": 0.125, "precision": 1, "itime": 0.25,
          "atime": 0.25 + 2.0 ** -12, "interval": 0.125, "output_factor": 3,
          "snapshot_number": 3, "energy_ini": -0.2500123456789, "iterations": 1234,
          "nsteps": 98765, "interactions": 98765 * n}
    for name, k in ARRAYS:
        st[name] = rng.normal(size=(n, k)) if k > 1 else rng.random(n)
    return st


def test_checksum_is_zlib_crc32():
    assert checksum(b"123456789") == 0xCBF43926  # the CRC-32 check value
    assert checksum(b"") == zlib.crc32(b"") == 0


def test_round_trip(tmp_path):
    st = sample_state()
    path = str(tmp_path / "a.ckpt")
    write_checkpoint(path, st)
    n = st["n"]
    assert os.path.getsize(path) == 8 + 8 * len(HEADER) + 8 * n * 22 + 8
    back = read_checkpoint(path)
    for name, _ in HEADER:
        assert back[name] == st[name], name
    for name, k in ARRAYS:
        assert np.array_equal(back[name].reshape(n, k), st[name].reshape(n, k)), name
    assert not os.path.exists(path + ".tmp")


def test_detects_corruption_and_truncation(tmp_path):
    path = str(tmp_path / "a.ckpt")
    write_checkpoint(path, sample_state())
    data = bytearray(open(path, "rb").read())
    bad = bytearray(data)
    bad[200] ^= 0x10
    open(path, "wb").write(bad)
    with pytest.raises(CheckpointError, match="checksum"):
        read_checkpoint(path)
    open(path, "wb").write(data[:-9])
    with pytest.raises(CheckpointError, match="size"):
        read_checkpoint(path)
    bad = bytearray(data)
    bad[0:1] = b"X"
    open(path, "wb").write(bad)
    with pytest.raises(CheckpointError, match="not a"):
        read_checkpoint(path)
    bad = bytearray(data)
    bad[7] = 2
    open(path, "wb").write(bad)
    with pytest.raises(CheckpointError, match="version"):
        read_checkpoint(path)


def test_rejects_arrays_of_the_wrong_size(tmp_path):
    st = sample_state()
    st["t"] = st["t"][:-1]
    with pytest.raises(CheckpointError, match="t:"):
        write_checkpoint(str(tmp_path / "a.ckpt"), st)


@pytest.mark.skipif(not os.path.exists(TOOL), reason="h4ckpt-info not built")
def test_cpp_reader_agrees_with_python_writer(tmp_path):
    st = sample_state(n=101, seed=9)
    path = str(tmp_path / "b.ckpt")
    write_checkpoint(path, st)
    out = subprocess.run([TOOL, path], capture_output=True, text=True, check=True).stdout
    got = dict(line.split(" ", 1) for line in out.strip().splitlines())
    for name, kind in HEADER:
        want = st[name]
        if kind == "d":
            assert float(got[name]) == want, name
        else:
            assert int(got[name]) == want, name
    s = 0.0
    for x in st["r"].reshape(-1):  # the tool sums sequentially
        s += float(x)
    assert float(got["sum_r"]) == s
    data = bytearray(open(path, "rb").read())
    data[-20] ^= 1
    open(path, "wb").write(data)
    bad = subprocess.run([TOOL, path], capture_output=True, text=True)
    assert bad.returncode == 1 and "checksum" in bad.stderr
