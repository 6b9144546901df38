"""Command-line handling of the gravidy-hip driver (CPU: every case here is
rejected or answered before the first GPU call)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "gravidy_amd", "host", "gravidy-hip")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="gravidy-hip not built")


def run(*args):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=60)


def test_help_lists_the_reference_flags_and_the_backend_options():
    p = run("--help")
    assert p.returncode != 0
    for flag in ("--input", "--time", "--output", "--softening", "--eta", "--screen",
                 "--interval", "--all", "--gpu", "--precision", "--dt-min", "--dt-max",
                 "--checkpoint", "--restart"):
        assert flag in p.stderr, flag


@pytest.mark.parametrize("args,message", [
    ((), "requires an argument -- 'input'"),
    (("-i", "/nonexistent/ic.txt"), "No such file"),
    (("-i", "{ic}", "-r", "x.info"), "not part of this build"),
    (("-i", "{ic}", "-l"), "Lagrange radii are not part of this build"),
    (("-i", "{ic}", "-g", "2"), "one GPU"),
    (("-i", "{ic}", "--precision", "half"), "unknown precision"),
    (("-i", "{ic}", "--dt-min", "0"), "dt-min"),
    (("-i", "{ic}", "--dt-min", "0.5", "--dt-max", "0.25"), "dt-min"),
    (("-i", "{ic}", "-z", "0"), "interval must be positive"),
    (("-i", "{ic}", "--restart", "/nonexistent/run.ckpt"), "No such file"),
])
def test_bad_options_are_rejected(tmp_path, args, message):
    ic = tmp_path / "ic.txt"
    ic.write_text("0 0.5 1 0 0 0 0 0\n1 0.5 -1 0 0 0 0 0\n")
    p = run(*[a.format(ic=ic) for a in args])
    assert p.returncode != 0
    assert message in p.stderr, p.stderr
