"""The one-workgroup block step of small systems (N <= 1024, FP64 and
mixed precision) against the oracle and against the multi-workgroup
kernels.

A step it makes must meet the same bars as the multi-workgroup step:
predictor, old forces and times bit-identical to the oracle, forces within
REL_TOL (FP64) or the mixed-precision statistics, and the same block-step
schedule over whole runs.  The device loop, the loop launched without
graphs and one h4h_step call per block step must agree bit for bit (each
applies the same rule: a step with 1..max_active active particles is made
in one workgroup).
"""
import numpy as np
import pytest

from gravidy_amd import Hermite4HIP, ic

from test_gpu_parity import (E2, MIXED_ACC_MAX, MIXED_JERK_MAX, bits_equal, block_state,
                             check_forces, check_mixed)

pytestmark = pytest.mark.gpu

# Block time at which only the particles with dt = 2^-12 are active
# (about a tenth of a block_state).
T_SMALL = 1.0 + 2.0 ** -12


def small_state(n, seed):
    r, v, f, t, dt = block_state(n, seed, T_SMALL)
    dt[0] = 2.0 ** -12  # at least one active particle
    t[0] = T_SMALL - dt[0]
    return r, v, f, t, dt


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("n,seed,limit", [(1, 1, 256), (2, 2, 256), (3, 3, 256), (255, 4, 256),
                                          (700, 5, 256), (1024, 6, 256), (1024, 6, "nact"),
                                          (1024, 6, "nact-1"), (1024, 6, 0)])
def test_small_step_matches_oracle(oracle, prec, n, seed, limit):
    """One h4h_step from an injected state: the step is made in one
    workgroup exactly when 1 <= nact <= max_active, and either way the
    predictor, old forces and new times are bit-identical to the oracle's
    and the forces are within the precision's bound."""
    r, v, f, t, dt = small_state(n, seed)
    mo = oracle.find_particles_to_move(T_SMALL, t, dt)
    nact = int(mo.size)
    assert 1 <= nact <= 256
    max_active = {"nact": nact, "nact-1": nact - 1}.get(limit, limit)
    with Hermite4HIP(r, v, E2, precision=prec) as h:
        h.set_small_path(max_active)
        h.upload(r, v, f, t, dt)
        h.stats(reset=True)
        k, atime = h.step(T_SMALL)
        st = h.stats()
        p, pm = h.download_predictor()
        s = h.download()
    assert k == nact
    assert st["block_steps"] == 1
    assert st["small_steps"] == (1 if nact <= max_active else 0)
    po, pmo = oracle.predicted_pos_vel(T_SMALL, t, r, v, f)
    assert bits_equal(p, po) and bits_equal(pm, pmo)
    fo = oracle.update_acc_jrk(mo, po, pmo, E2, f)
    label = f"small step n={n} nact={nact} max_active={max_active} {prec}"
    if prec == "fp64":
        check_forces(s["f"], fo, mo, po, pmo, E2, label)
    elif not fo[mo].any():
        assert not s["f"][mo].any()  # one particle: no pair
    elif nact >= 64:
        check_mixed(s["f"], fo, mo, label)
    else:
        # too few particles for the statistics: per-particle bounds only
        ra = np.linalg.norm(s["f"][mo, :3] - fo[mo, :3], axis=1) / np.linalg.norm(fo[mo, :3], axis=1)
        rj = np.linalg.norm(s["f"][mo, 3:] - fo[mo, 3:], axis=1) / np.linalg.norm(fo[mo, 3:], axis=1)
        print(f"{label}: acc rel err max {ra.max():.3e}, jerk {rj.max():.3e}")
        assert ra.max() <= MIXED_ACC_MAX and rj.max() <= MIXED_JERK_MAX
    assert bits_equal(s["old"][mo], f[mo])
    rest = np.setdiff1d(np.arange(n), mo)
    assert bits_equal(s["f"][rest], f[rest])
    z3 = np.zeros((n, 3))
    # The corrector is bit-identical to the oracle's on the forces the step made ...
    a2c, a3c, rc, vc, tc, dtc = oracle.correction_pos_vel(T_SMALL, mo, po, s["f"], f, z3, z3,
                                                          r, v, t, dt)
    assert bits_equal(s["r"][:, :3], rc[:, :3]) and bits_equal(s["v"][:, :3], vc[:, :3])
    assert bits_equal(s["a2"][mo], a2c[mo]) and bits_equal(s["a3"][mo], a3c[mo])
    assert bits_equal(s["t"], tc) and bits_equal(s["dt"], dtc)
    # ... and the new block steps are those of the oracle's own forces.
    _, _, _, _, to, dto = oracle.correction_pos_vel(T_SMALL, mo, po, fo, f, z3, z3, r, v, t, dt)
    assert bits_equal(s["t"], to) and bits_equal(s["dt"], dto)
    assert atime == oracle.next_integration_time(to, dto)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_small_path_loop_modes_agree(prec):
    """N = 1000 to t = 0.25: the graph-replayed device loop, the loop
    launched without graphs and one h4h_step call per block step give the
    same schedule, bit-identical state and the same number of one-workgroup
    steps; the loop honours max_steps and the output threshold."""
    r, v = ic.plummer(1000, 321)
    out = {}
    for mode, graphs in (("run", True), ("step", True), ("run", False)):
        with Hermite4HIP(r, v, E2, precision=prec) as h:
            h.set_graphs(graphs)
            res = h.integration(0.25, mode=mode)
            out[mode, graphs] = (res, h.download(), h.stats())
    (ra, sa, sta) = out["run", True]
    n_steps = len(ra["step_time"])
    print(f"N=1000 {prec}: {n_steps} block steps, {sta['small_steps']} in one workgroup")
    assert 0 < sta["small_steps"] < n_steps
    for key in (("step", True), ("run", False)):
        rb, sb, stb = out[key]
        assert np.array_equal(ra["step_time"], rb["step_time"]), key
        assert np.array_equal(ra["step_nact"], rb["step_nact"]), key
        assert bits_equal(sa["r"], sb["r"]) and bits_equal(sa["dt"], sb["dt"]), key
        assert bits_equal(sa["f"], sb["f"]), key
        np.testing.assert_array_equal(ra["energy"], rb["energy"])
        assert stb["small_steps"] == sta["small_steps"], key
    with Hermite4HIP(r, v, E2, precision=prec) as h:
        h.init_acc_jrk()
        atime = h.init_dt()
        info, st, sn = h.run(0.0, atime, 0.25, 1.0, 5)
        assert info["steps"] == 5 and info["stopped"] == 3
        assert np.array_equal(st, ra["step_time"][:5]) and np.array_equal(sn, ra["step_nact"][:5])
        info, st, _ = h.run(info["itime"], info["atime"], 0.25, 0.0625, 1 << 16)
        assert info["stopped"] == 1 and st[-1] >= 0.0625 and np.all(st[:-1] < 0.0625)
        assert np.array_equal(st, ra["step_time"][5:5 + len(st)])


@pytest.mark.parametrize("n", [1000, 1024])
def test_small_path_matches_multi_workgroup_path(n):
    """FP64 to t = 0.25 with the one-workgroup step on and off: the same
    block-step schedule, and states and energies that differ only by the
    order of the pair sums."""
    r, v = ic.plummer(n, 4321 + n)
    out = {}
    for max_active in (256, 0):
        with Hermite4HIP(r, v, E2, precision="fp64") as h:
            h.set_small_path(max_active)
            res = h.integration(0.25)
            out[max_active] = (res, h.download(), h.stats())
    (ra, sa, sta), (rb, sb, stb) = out[256], out[0]
    print(f"N={n}: {len(ra['step_time'])} block steps, {sta['small_steps']} in one workgroup")
    assert sta["small_steps"] > 0 and stb["small_steps"] == 0
    assert np.array_equal(ra["step_time"], rb["step_time"])
    assert np.array_equal(ra["step_nact"], rb["step_nact"])
    assert bits_equal(sa["t"], sb["t"]) and bits_equal(sa["dt"], sb["dt"])
    np.testing.assert_allclose(sa["r"][:, :3], sb["r"][:, :3], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(sa["v"][:, :3], sb["v"][:, :3], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(ra["energy"], rb["energy"], rtol=1e-12)


def test_exact_precision_keeps_the_multi_workgroup_path(oracle):
    """FP64_EXACT is bit-identical to the oracle only with the exact update
    kernel: its steps never go through the one-workgroup path."""
    r, v = ic.plummer(512, 99)
    with Hermite4HIP(r, v, E2, precision="exact") as h:
        h.set_small_path(256)
        run = h.integration(0.125)
        st = h.stats()
        s = h.download()
    ref = oracle.integrate(r, v, E2, 0.125)
    assert st["small_steps"] == 0
    assert np.array_equal(run["step_time"], ref["step_time"])
    assert bits_equal(s["r"], ref["r"])


@pytest.mark.parametrize("graphs", [True, False])
def test_ensemble_of_small_clusters_matches_oracle(oracle, graphs):
    """An ensemble of clusters of at most 1024 particles, each stepping in a
    workgroup of its own: each cluster's step count, interactions and
    cumulative energy error agree with the oracle's, and graph replay and
    direct launches give the same state bit for bit."""
    from gravidy_amd import EnsembleHIP
    sizes = [64, 300, 700, 1000, 1024]
    systems = [ic.plummer(n, 1300 + n) for n in sizes]
    out = {}
    for g in (graphs, not graphs):
        with EnsembleHIP(systems, E2, precision="fp64") as ens:
            ens.set_graphs(g)
            res = ens.integration(0.25)
            out[g] = (res, ens.download(), ens.stats())
    res, state, st = out[graphs]
    assert st["small_steps"] > 0
    for c, (sa, sb) in enumerate(zip(state, out[not graphs][1])):
        for key in ("r", "v", "f", "t", "dt"):
            assert bits_equal(sa[key], sb[key]), (c, key)
    for c, (r, v) in enumerate(systems):
        ref = oracle.integrate(r, v, E2, 0.25)
        e, eo = res["energy"][:, c], ref["energy"]
        cum = abs((e[-1] - e[0]) / e[0])
        cum_ref = abs((eo[-1] - eo[0]) / eo[0])
        print(f"cluster {c} n={sizes[c]}: steps {res['steps'][c]}/{len(ref['step_time'])} "
              f"CumE {cum:.4e} oracle {cum_ref:.4e}")
        assert res["steps"][c] == len(ref["step_time"])
        assert res["interactions"][c] == ref["interactions"]
        assert abs(cum - cum_ref) <= 0.05 * cum_ref + 1e-15


def test_ensemble_with_a_large_cluster_keeps_the_multi_workgroup_path():
    """A cluster of more than 1024 particles in the ensemble: no cluster
    steps in a workgroup of its own (the large one would wait for them)."""
    from gravidy_amd import EnsembleHIP
    systems = [ic.plummer(n, 1400 + n) for n in (300, 1500)]
    with EnsembleHIP(systems, E2, precision="fp64") as ens:
        ens.integration(0.0625)
        assert ens.stats()["small_steps"] == 0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_small_step_without_softening(oracle, prec):
    """e2 = 0: the pair of a particle with itself (r = 0) is skipped, so the
    one-workgroup step stays finite and matches the oracle."""
    n = 700
    r, v, f, t, dt = small_state(n, 8)
    mo = oracle.find_particles_to_move(T_SMALL, t, dt)
    with Hermite4HIP(r, v, 0.0, precision=prec) as h:
        h.set_small_path(256)
        h.upload(r, v, f, t, dt)
        h.stats(reset=True)
        k, _ = h.step(T_SMALL)
        assert h.stats()["small_steps"] == 1
        s = h.download()
    assert k == mo.size
    po, pmo = oracle.predicted_pos_vel(T_SMALL, t, r, v, f)
    fo = oracle.update_acc_jrk(mo, po, pmo, 0.0, f)
    assert np.all(np.isfinite(s["f"]))
    if prec == "fp64":
        check_forces(s["f"], fo, mo, po, pmo, 0.0, f"small step e2=0 n={n} fp64")
    else:
        check_mixed(s["f"], fo, mo, f"small step e2=0 n={n} mixed")


@pytest.mark.parametrize("prec,e2,mass", [("fp64", 2.0 ** -598, 2.0 ** 127),
                                          ("mixed", 2.0 ** -58, 2.0 ** 60),
                                          ("fp64", 2.0 ** -598, 1.0),
                                          ("mixed", 2.0 ** -58, 1.0)])
def test_small_step_self_pair_with_large_masses(oracle, prec, e2, mass):
    """Softening just above the range where the one-workgroup step drops its
    self-pair test, with masses whose weight m / e2^1.5 at r = 0 overflows
    (2^127 in FP64, 2^60 in the FP32 pair terms; mass 1 is the control,
    where the test may be dropped): the step keeps the test for the large
    masses, so the forces stay finite and match the oracle's."""
    n = 256
    r, v, f, t, dt = small_state(n, 9)
    r = r.copy()
    r[:, 3] = mass
    mo = oracle.find_particles_to_move(T_SMALL, t, dt)
    assert 1 <= mo.size <= 64
    with Hermite4HIP(r, v, e2, precision=prec) as h:
        h.set_small_path(256)
        h.upload(r, v, f, t, dt)
        h.stats(reset=True)
        k, _ = h.step(T_SMALL)
        assert h.stats()["small_steps"] == 1
        s = h.download()
    assert k == mo.size
    po, pmo = oracle.predicted_pos_vel(T_SMALL, t, r, v, f)
    fo = oracle.update_acc_jrk(mo, po, pmo, e2, f)
    assert np.all(np.isfinite(fo[mo]))
    assert np.all(np.isfinite(s["f"][mo])), "self pair gave a non-finite force"
    label = f"small step e2=2^{np.log2(e2):.0f} m={mass:g} n={n} nact={mo.size} {prec}"
    if prec == "fp64":
        check_forces(s["f"], fo, mo, po, pmo, e2, label)
    else:
        ra = np.linalg.norm(s["f"][mo, :3] - fo[mo, :3], axis=1) / np.linalg.norm(fo[mo, :3], axis=1)
        rj = np.linalg.norm(s["f"][mo, 3:] - fo[mo, 3:], axis=1) / np.linalg.norm(fo[mo, 3:], axis=1)
        print(f"{label}: acc rel err max {ra.max():.3e}, jerk {rj.max():.3e}")
        assert ra.max() <= MIXED_ACC_MAX and rj.max() <= MIXED_JERK_MAX
