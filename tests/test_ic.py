"""Initial-condition tools: seeded Plummer sampling and the reference's
8-column file format (NbodySystem.cpp:79-207)."""
import math

import numpy as np
import pytest

from gravidy_amd import ic


def test_plummer_is_seeded_and_float_massed():
    r1, v1 = ic.plummer(512, 7)
    r2, v2 = ic.plummer(512, 7)
    r3, _ = ic.plummer(512, 8)
    assert np.array_equal(r1, r2) and np.array_equal(v1, v2)
    assert not np.array_equal(r1, r3)
    assert np.all(r1[:, 3] == np.float32(1 / 512))
    assert np.all(v1[:, 3] == 0.0)


def test_plummer_centre_of_mass_frame_and_truncation():
    r, v = ic.plummer(4096, 11)
    m = r[:, 3:]
    assert np.abs((m * r[:, :3]).sum(axis=0)).max() < 1e-12
    assert np.abs((m * v[:, :3]).sum(axis=0)).max() < 1e-12
    # r < 10 Plummer radii before the shift to Henon units (3 pi / 16)
    assert np.linalg.norm(r[:, :3], axis=1).max() < 10 * ic.PLUMMER_SCALE + 0.05


def test_plummer_henon_units(oracle):
    r, v = ic.plummer(4096, 12)
    ek, ep = oracle.get_energy(r, v)
    # virial ratio 1/2 and E = -1/4 up to finite-N noise
    assert ek / -ep == pytest.approx(0.5, abs=0.03)
    assert ek + ep == pytest.approx(-0.25, abs=0.015)


def test_file_round_trip_is_exact(tmp_path):
    r, v = ic.plummer(300, 5)
    p = tmp_path / "ic.txt"
    ic.write_ic(p, r, v)
    ids, r2, v2, time = ic.read_ic(p)
    assert time is None
    assert np.array_equal(ids, np.arange(300))
    assert np.array_equal(r, r2) and np.array_equal(v, v2)


def test_snapshot_header_time(tmp_path):
    r, v = ic.plummer(4, 5)
    p = tmp_path / "snap.txt"
    ic.write_ic(p, r, v, time=0.125)
    _, _, _, time = ic.read_ic(p)
    assert time == 0.125


def test_mass_column_is_parsed_as_float(tmp_path):
    p = tmp_path / "m.txt"
    p.write_text("0 0.1 1 0 0 0 0 0\n1 0.1 -1 0 0 0 0 0\n")
    _, r, _, _ = ic.read_ic(p)
    assert r[0, 3] == float(np.float32(0.1)) != 0.1


def test_wrong_column_count_is_an_error(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("0 0.5 1 0 0 0 0\n")
    with pytest.raises(ic.ICFormatError):
        ic.read_ic(p)


def test_salpeter_masses_and_central_particle():
    m = ic.salpeter_masses(10000, 3)
    # the mass range [0.1, 100] keeps its ratio after rescaling
    assert m.max() / m.min() <= 1000.0 * (1 + 1e-6)
    assert m.sum() == pytest.approx(1.0, rel=1e-6)
    assert np.all(m == m.astype(np.float32))
    # dN/dm ~ m^-2.35: mean of m/m_lo is (1.35/0.35) (1 - x^-0.35)/(1 - x^-1.35), x = 1000
    scale = m.min() / 0.1
    x = 1000.0
    want = (1.35 / 0.35) * (1 - x ** -0.35) / (1 - x ** -1.35)
    assert (m / scale / 0.1).mean() == pytest.approx(want, rel=0.05)
    r, v = ic.plummer_salpeter_bh(1000, 4)
    assert r.shape == (1001, 4)
    assert list(r[0]) == [0.0, 0.0, 0.0, float(np.float32(0.01))]
    assert list(v[0]) == [0.0, 0.0, 0.0, 0.0]


def test_log_uniform_sizes():
    n = ic.log_uniform_sizes(64, 1)
    assert n.shape == (64,)
    assert n.min() >= 1024 and n.max() <= 32768
    assert abs(np.log(n).mean() - 0.5 * (math.log(1024) + math.log(32768))) < 0.5
