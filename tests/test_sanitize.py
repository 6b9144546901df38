"""CPU sanitizer build (AddressSanitizer + UndefinedBehaviorSanitizer) of
the code that runs on the host and reads untrusted files: the oracle
(oracle/sanitize_check.c drives every function on small and degenerate
systems), and the gravidy-hip host side without the GPU (io-check: options
and the -r .info resume, the initial-condition / snapshot reader, the
output writers, the binary checkpoint writer and reader; h4ckpt-info: the
checkpoint reader) on good, corrupt and truncated inputs.  Each program
must end with its own exit status and no sanitizer report."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
HOST_DIR = os.path.join(ROOT, "gravidy_amd", "host")
H4O_CHECK = os.path.join(ORACLE_DIR, "_build", "sanitize", "h4o-check")
IO_CHECK = os.path.join(HOST_DIR, "_build", "sanitize", "io-check")
CKPT_INFO = os.path.join(HOST_DIR, "_build", "sanitize", "h4ckpt-info")

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")

pytestmark = pytest.mark.skipif(shutil.which("make") is None, reason="make not available")


@pytest.fixture(scope="module")
def built():
    for d in (ORACLE_DIR, HOST_DIR):
        p = subprocess.run(["make", "-C", d, "sanitize"], capture_output=True, text=True,
                           timeout=600)
        if p.returncode != 0 and "sanitize" in p.stderr and "fsanitize" in p.stderr:
            pytest.skip(f"sanitizer runtime not available: {p.stderr[-300:]}")
        assert p.returncode == 0, p.stderr
    return True


def run(*args, cwd=None):
    p = subprocess.run(list(args), capture_output=True, text=True, timeout=300, env=ENV, cwd=cwd)
    for marker in ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:"):
        assert marker not in p.stderr, p.stderr[-3000:]
    return p


def test_oracle_under_sanitizers(built):
    p = run(H4O_CHECK)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "h4o-check ok" in p.stdout


def _ic(tmp_path, name="ic.txt", n=32, seed=1):
    from gravidy_amd import ic
    r, v = ic.plummer(n, seed)
    path = tmp_path / name
    ic.write_ic(path, r, v)
    return path


def test_host_io_good_input_and_text_resume(built, tmp_path):
    path = _ic(tmp_path)
    p = run(IO_CHECK, str(tmp_path / "a.ckpt"), "-i", str(path), "-t", "0.25")
    assert p.returncode == 0, p.stderr
    assert p.stdout.startswith("n 32 time 0")
    # the .info it wrote names snapshot 1 (SnapshotNumber appended); a
    # snapshot of that number makes it a resumable run
    os.rename(str(path) + ".out.snapshot_00", str(path) + ".out.snapshot_01")
    with open(str(path) + ".out.snapshot_01") as fh:
        body = fh.read().split("\n", 1)[1]
    with open(str(path) + ".out.snapshot_01", "w") as fh:
        fh.write("# Time:0.2500\n" + body)
    p = run(IO_CHECK, str(tmp_path / "b.ckpt"), "-r", str(path) + ".out.info", "-t", "0.5")
    assert p.returncode == 0, p.stderr
    assert p.stdout.startswith("n 32 time 0.25 snapshot ")
    assert os.path.exists(str(path) + ".out.info.old")


@pytest.mark.parametrize("content", [
    "0 0.5 1 0 0 0 0\n",                        # 7 columns
    "\n",                                       # empty line first
    "# only a comment\n",                       # no particle
    "x" * 100000 + "\n",                        # one huge token
    "\x00\x01\x02\xff\xfe binary\n",            # bytes that are no text
    "0 nan inf -inf 1e999 -1e999 0x1p3 7\n",     # accepted by strtod, no particle lost
])
def test_host_io_bad_input(built, tmp_path, content):
    path = tmp_path / "bad.txt"
    path.write_bytes(content.encode("latin-1"))
    p = run(IO_CHECK, str(tmp_path / "c.ckpt"), "-i", str(path))
    assert p.returncode in (0, 1)


@pytest.mark.parametrize("info", [
    "",                                                     # no keys
    "# Bogus:   1\n",                                       # unknown key
    "# IntegrationTime: 0.25\n# SnapshotNumber: 99999999999999999999\n",
    "no colon at all" * 1000,
    "# InputFilename:" + "a" * 5000 + "\n# OutputFilename: /nonexistent\n",
])
def test_host_io_bad_info_file(built, tmp_path, info):
    path = tmp_path / "run.info"
    path.write_text(info)
    p = run(IO_CHECK, str(tmp_path / "d.ckpt"), "-r", str(path), "-t", "1")
    assert p.returncode == 1


def _checkpoint(tmp_path, n=5):
    from gravidy_amd.checkpoint import HEADER, write_checkpoint
    rng = np.random.default_rng(3)
    st = {name: 0 for name, _ in HEADER}
    st.update({"n": n, "e2": 1e-8, "eta": 0.01, "dt_min": 2.0 ** -23, "dt_max": 0.125,
               "itime": 0.5, "atime": 0.5 + 2.0 ** -10, "interval": 0.125})
    for name, k in (("r", 4), ("v", 4), ("f", 6), ("t", 1), ("dt", 1), ("a2", 3), ("a3", 3)):
        st[name] = rng.normal(size=(n, k)) if k > 1 else rng.normal(size=n)
    path = str(tmp_path / "good.ckpt")
    write_checkpoint(path, st)
    return path


def test_checkpoint_reader_good_corrupt_truncated(built, tmp_path):
    good = _checkpoint(tmp_path)
    p = run(CKPT_INFO, good)
    assert p.returncode == 0 and "n 5" in p.stdout, p.stderr
    data = open(good, "rb").read()
    cases = {
        "truncated_header": data[:12],
        "truncated_body": data[:-9],
        "extended": data + b"\0" * 8,
        "bit_flip": data[:200] + bytes([data[200] ^ 1]) + data[201:],
        "bad_magic": b"X" + data[1:],
        "bad_version": data[:7] + b"\x09" + data[8:],
        "huge_n": data[:8] + (2 ** 40).to_bytes(8, "little") + data[16:],
        "zero_n": data[:8] + (0).to_bytes(8, "little") + data[16:],
        "empty": b"",
    }
    for name, blob in cases.items():
        path = tmp_path / f"{name}.ckpt"
        path.write_bytes(blob)
        p = run(CKPT_INFO, str(path))
        assert p.returncode == 1, (name, p.stdout, p.stderr)
        assert "h4ckpt-info:" in p.stderr, name
