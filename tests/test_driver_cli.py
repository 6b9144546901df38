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
    for flag in ("--input", "--time", "--resume", "--output", "--softening", "--eta", "--screen",
                 "--interval", "--all", "--gpu", "--precision", "--dt-min", "--dt-max",
                 "--checkpoint", "--restart"):
        assert flag in p.stderr, flag


@pytest.mark.parametrize("args,message", [
    ((), "requires an argument -- 'input'"),
    (("-i", "/nonexistent/ic.txt"), "No such file"),
    (("-r", "/nonexistent/run.info"), "No such file"),
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


INFO_KEYS = ("NumberParticles", "Softening", "EtaTimestep", "IntegrationTime", "PrintScreen",
             "InputFilename", "OutputFilename", "SnapshotNumber")


def write_info(path, ic, out, snap, extra=None):
    """An .info file as the Logger writes it (Logger.cpp:136-217, SnapshotNumber
    appended at the end of a run)."""
    vals = {"NumberParticles": "2", "Softening": "0.0001", "EtaTimestep": "0.01",
            "IntegrationTime": "0.25", "PrintScreen": "0", "InputFilename": str(ic),
            "OutputFilename": str(out), "SnapshotNumber": str(snap)}
    lines = ["#".ljust(2) + (k + ":").ljust(20) + vals[k].rjust(20) for k in INFO_KEYS]
    if extra:
        lines.append("#".ljust(2) + extra.ljust(20) + "1".rjust(20))
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("case,args,message", [
    ("bad key", ("-t", "0.5"), "Invalid option -- Bogus in the info file"),
    ("time not later", ("-t", "0.25"), "greater than the one specifid in the info file"),
    ("no snapshot", ("-t", "0.5"), "Check the OutputFilename and SnapshotNumber"),
    ("with --restart", ("-t", "0.5", "--restart", "{info}"), "exclude each other"),
])
def test_text_resume_errors(tmp_path, case, args, message):
    """-r <run>.info (OptionsParser.cpp:122-324): unknown keys, an end time not
    after the file's IntegrationTime and a missing snapshot are refused before
    any GPU work, with the reference's messages."""
    ic = tmp_path / "ic.txt"
    ic.write_text("0 0.5 1 0 0 0 0 0\n1 0.5 -1 0 0 0 0 0\n")
    out = tmp_path / "ic.txt.out"
    info = tmp_path / "run.info"
    write_info(info, ic, out, 2, extra="Bogus:" if case == "bad key" else None)
    if case != "no snapshot":
        (tmp_path / "ic.txt.out.snapshot_02").write_text("# Time:0.2500\n")
    p = run("-r", str(info), *[a.format(info=info) for a in args])
    assert p.returncode != 0
    assert message in p.stderr, p.stderr
