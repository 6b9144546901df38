"""Parity at the benchmark configurations' own recipes (through the C ABI).

Plummer N=65536, eps=1e-4: the FP64 path checked step by step against
the oracle on injected states (each block step starts both sides from the
oracle's state), from the initial conditions and again from a state the
device loop reached on its own; energies of the same states to 1e-13.

Plummer N=262144: active samples of the update against the oracle (FP64
within 1e-13, mixed within its contract), and the mixed path's energy drift
over the whole run to t=0.25.

Salpeter masses + central particle at reduced N: Salpeter masses on [0.1, 100] (total 1) in a
Plummer sphere plus a central particle of mass 0.01 at rest, eps=1e-5,
block steps down to 2^-30: the exact path bit-identical to the oracle over
a whole run (schedule, state, energies), the FP64 path with the same
schedule and the same cumulative energy error to 2 significant digits.

The ensemble recipe: 64 seeded clusters, N log-uniform in [1024, 32768],
with the launch picked by the autotuner, each cluster checked by
properties and the smallest ones against the oracle.

Restarts from a binary checkpoint against the oracle's uninterrupted run.
"""
import numpy as np
import pytest

from gravidy_amd import Hermite4HIP, ic, softening_e2
from test_gpu_parity import bits_equal, check_forces, check_mixed, p6, pm_of

pytestmark = pytest.mark.gpu

E2 = softening_e2(1e-4)


def oracle_step(oracle, T, r, v, f, t, dt, e2):
    """Hermite4CPU_integration.cpp:81-124, one block step on the oracle."""
    n = r.shape[0]
    move = oracle.find_particles_to_move(T, t, dt)
    p, pm = oracle.predicted_pos_vel(T, t, r, v, f)
    fn = oracle.update_acc_jrk(move, p, pm, e2, f)
    _, _, rn, vn, tn, dtn = oracle.correction_pos_vel(
        T, move, p, fn, f, np.zeros((n, 3)), np.zeros((n, 3)), r, v, t, dt)
    return move, p, pm, fn, rn, vn, tn, dtn


def injected_steps(oracle, h, state, T, steps, e2, label):
    """Runs `steps` block steps, each from the oracle's state uploaded to the
    context; checks the active set, forces (1e-13, two-level), positions,
    velocities, times, steps and the next block time.  Returns the oracle
    state and next block time after the last step."""
    r, v, f, t, dt = state
    for k in range(steps):
        h.upload(r, v, f, t, dt)
        nact, atime = h.step(T)
        s = h.download()
        move, p, pm, fn, rn, vn, tn, dtn = oracle_step(oracle, T, r, v, f, t, dt, e2)
        assert nact == move.size, (label, k, nact, move.size)
        check_forces(s["f"], fn, move, p, pm, e2, f"{label} step {k} T={T!r} nact={nact}")
        assert np.array_equal(s["t"], tn), (label, k)
        assert np.array_equal(s["dt"], dtn), (label, k)
        ar = np.abs(s["r"][:, :3] - rn[:, :3]).max() / np.abs(rn[:, :3]).max()
        av = np.abs(s["v"][:, :3] - vn[:, :3]).max() / np.abs(vn[:, :3]).max()
        assert ar <= 1e-15 and av <= 1e-13, (label, k, ar, av)
        T_next = oracle.next_integration_time(tn, dtn)
        assert atime == T_next, (label, k, atime, T_next)
        r, v, f, t, dt, T = rn, vn, fn, tn, dtn, T_next
    return (r, v, f, t, dt), T


@pytest.mark.slow
def test_plummer65536_fp64_per_step_against_oracle(oracle):
    n = 65536
    r, v = ic.plummer(n, 65536)
    fo = oracle.init_acc_jrk(p6(r, v), pm_of(r), E2)
    to, dto, T = oracle.init_dt(fo)
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.init_acc_jrk()
        f = h.download()["f"]
        check_forces(f, fo, np.arange(n), p6(r, v), pm_of(r), E2, "N=65536 init_acc_jrk fp64")
        ek, ep = h.get_energy()
        eko, epo = oracle.get_energy(r, v)
        assert ek == pytest.approx(eko, rel=1e-13) and ep == pytest.approx(epo, rel=1e-13)

        state, T = injected_steps(oracle, h, (r, v, fo, to, dto), T, 24, E2, "N=65536 from IC")

        # A state the device loop reached by itself (t = 1/16), then more
        # injected steps from there.
        h.upload(r, v)
        h.init_acc_jrk()
        atime = h.init_dt()
        info, _, _ = h.run(0.0, atime, 0.0625, 0.0625, 1 << 20)
        assert info["stopped"] in (1, 2)
        s = h.download()
        ek, ep = h.get_energy()
        eko, epo = oracle.get_energy(s["r"], s["v"])
        assert ek == pytest.approx(eko, rel=1e-13) and ep == pytest.approx(epo, rel=1e-13)
        print(f"N=65536 device loop: {info['steps']} block steps to t={info['itime']}")
        state = (s["r"], s["v"], s["f"], s["t"], s["dt"])
        injected_steps(oracle, h, state, info["atime"], 16, E2, "N=65536 at t=1/16")


SB_T_END = 2.0 ** -6
SB_INTERVAL = 2.0 ** -8


@pytest.fixture(scope="module")
def salpeter_bh_case(oracle):
    """Salpeter + central-mass recipe at 16384 stars + the central particle, and the
    oracle's run of it."""
    r, v = ic.plummer_salpeter_bh(16384, 1048576)
    e2, dt_min = softening_e2(1e-5), 2.0 ** -30
    ref = oracle.integrate(r, v, e2, SB_T_END, interval=SB_INTERVAL, dt_min=dt_min)
    print(f"Salpeter+BH n={r.shape[0]}: {len(ref['step_time'])} block steps, "
          f"smallest dt {ref['dt'].min()!r}")
    return r, v, e2, dt_min, ref


@pytest.mark.slow
def test_salpeter_bh_reduced_exact_is_bitwise(salpeter_bh_case):
    r, v, e2, dt_min, ref = salpeter_bh_case
    with Hermite4HIP(r, v, e2, precision="exact", dt_min=dt_min) as h:
        run = h.integration(SB_T_END, interval=SB_INTERVAL)
        s = h.download()
    assert np.array_equal(run["step_time"], ref["step_time"])
    assert np.array_equal(run["step_nact"], ref["step_nact"])
    assert bits_equal(s["r"], ref["r"]) and bits_equal(s["v"][:, :3], ref["v"][:, :3])
    assert bits_equal(s["t"], ref["t"]) and bits_equal(s["dt"], ref["dt"])
    np.testing.assert_allclose(run["energy"], ref["energy"], rtol=1e-13)


@pytest.mark.slow
def test_salpeter_bh_reduced_fp64_energy_error_matches(salpeter_bh_case):
    r, v, e2, dt_min, ref = salpeter_bh_case
    with Hermite4HIP(r, v, e2, precision="fp64", dt_min=dt_min) as h:
        run = h.integration(SB_T_END, interval=SB_INTERVAL)
    same = (run["step_time"].shape == ref["step_time"].shape
            and np.array_equal(run["step_time"], ref["step_time"])
            and np.array_equal(run["step_nact"], ref["step_nact"]))
    cum = abs((run["energy"][-1] - run["energy"][0]) / run["energy"][0])
    cum_ref = abs((ref["energy"][-1] - ref["energy"][0]) / ref["energy"][0])
    print(f"Salpeter+BH fp64: schedule identical={same}, CumE {cum:.6e} oracle {cum_ref:.6e}")
    assert same
    assert float(f"{cum:.1e}") == float(f"{cum_ref:.1e}") or abs(cum - cum_ref) <= 0.05 * cum_ref


@pytest.mark.slow
def test_plummer262144_mixed_energy_drift_within_2x_fp64():
    """Plummer N=262144, eps=1e-4, to t=0.25 from identical initial
    conditions: the mixed path's cumulative energy error at t_end is within
    2x the FP64 path's (the per-particle force error of the mixed kernel at
    this N is checked in test_gpu_parity)."""
    import time
    n = 262144
    r, v = ic.plummer(n, 262144)
    out = {}
    for precision in ("fp64", "mixed"):
        t0 = time.perf_counter()
        with Hermite4HIP(r, v, E2, precision=precision) as h:
            run = h.integration(0.25, max_steps=400000)
        el = time.perf_counter() - t0
        e = run["energy"]
        assert run["energy_time"][-1] == 0.25, run["energy_time"]
        out[precision] = abs((e[-1] - e[0]) / e[0])
        print(f"N=262144 {precision}: {len(run['step_time'])} block steps in {el:.1f} s, "
              f"{run['interactions'] / el:.3e} pairs/s, CumE {out[precision]:.3e}")
    assert out["mixed"] <= 2.0 * out["fp64"], out


@pytest.fixture(scope="module")
def salpeter_bh_full():
    """The Salpeter + central-mass initial conditions at full size: 1048576 Salpeter-mass stars
    in a Plummer sphere plus the central particle (1048577 bodies)."""
    r, v = ic.plummer_salpeter_bh(1048576, 1048576)
    return r, v, softening_e2(1e-5)


@pytest.mark.slow
@pytest.mark.parametrize("precision", ["exact", "fp64", "mixed"])
def test_salpeter_bh_full_size_update_on_a_sample(oracle, salpeter_bh_full, precision):
    """nact = 256 active particles (the central one and 255 seeded others)
    against all 1048577 j: exact bit-identical, FP64 within
    1e-13 (two-level), mixed within the RMS contract."""
    r, v, e2 = salpeter_bh_full
    n = r.shape[0]
    rng = np.random.default_rng(5)
    move = np.sort(np.concatenate([[0], rng.choice(np.arange(1, n), 255, replace=False)]))
    move = move.astype(np.uint32)
    with Hermite4HIP(r, v, e2, precision=precision, dt_min=2.0 ** -30) as h:
        h.set_active(move)
        h.update_acc_jrk(move.size)
        f = h.download()["f"]
    fo = oracle.update_acc_jrk(move, p6(r, v), pm_of(r), e2, np.zeros((n, 6)))
    label = f"Salpeter+BH n={n} nact=256 {precision}"
    if precision == "exact":
        assert bits_equal(f[move], fo[move]), label
    elif precision == "fp64":
        check_forces(f, fo, move, p6(r, v), pm_of(r), e2, label)
    else:
        check_mixed(f, fo, move, label)


@pytest.mark.slow
def test_salpeter_bh_full_size_initial_forces_conserve_momentum(salpeter_bh_full):
    """All 1048577 particles active (1.1e12 pairs): sum_i m_i a_i and
    sum_i m_i j_i vanish up to rounding (pair terms are antisymmetric), a
    size-independent check of the whole N^2 pass."""
    r, v, e2 = salpeter_bh_full
    m = r[:, 3].astype(np.float32).astype(np.float64)
    with Hermite4HIP(r, v, e2, precision="fp64", dt_min=2.0 ** -30) as h:
        h.init_acc_jrk()
        f = h.download()["f"]
    for c, name in ((0, "acc"), (3, "jerk")):
        net = np.abs((m[:, None] * f[:, c:c + 3]).sum(axis=0)).max()
        scale = (m * np.linalg.norm(f[:, c:c + 3], axis=1)).sum()
        print(f"Salpeter+BH n={r.shape[0]} {name}: |sum m f| / sum m|f| = {net / scale:.3e}")
        assert net <= 1e-12 * scale, (name, net, scale)


@pytest.mark.parametrize("precision", ["exact", "fp64", "mixed"])
def test_python_restart_is_bit_exact(tmp_path, precision):
    """Hermite4HIP.integration(checkpoint=..., restart=...): t = 0 -> 0.25,
    then a new context continues from the checkpoint to 0.5; schedule and
    final state equal the run 0 -> 0.5 in one go."""
    r, v = ic.plummer(2048, 515)
    with Hermite4HIP(r, v, E2, precision=precision) as h:
        whole = h.integration(0.5)
        sw = h.download()
    ck = str(tmp_path / "half.ckpt")
    with Hermite4HIP(r, v, E2, precision=precision) as h:
        first = h.integration(0.25, checkpoint=ck)
    with Hermite4HIP(r, v, E2, precision=precision) as h:
        second = h.integration(0.5, restart=ck)
        sr = h.download()
    steps = np.concatenate([first["step_time"], second["step_time"]])
    nact = np.concatenate([first["step_nact"], second["step_nact"]])
    assert np.array_equal(steps, whole["step_time"]) and np.array_equal(nact, whole["step_nact"])
    for key in ("r", "v", "f", "t", "dt"):
        assert bits_equal(sr[key], sw[key]), key
    assert second["energy_ini"] == whole["energy_ini"]
    assert np.array_equal(second["energy"][1:], whole["energy"][3:])


@pytest.mark.slow
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plummer262144_update_sample_against_oracle(oracle, precision):
    """Plummer N=262144 (the benchmark's IC): 256 seeded active particles
    against all 262144 j, FP64 within 1e-13 (two-level), mixed within its
    RMS contract and per-particle caps."""
    n = 262144
    r, v = ic.plummer(n, 262144)
    rng = np.random.default_rng(2621)
    move = np.sort(rng.choice(n, size=256, replace=False)).astype(np.uint32)
    with Hermite4HIP(r, v, E2, precision=precision) as h:
        h.set_active(move)
        h.update_acc_jrk(move.size)
        f = h.download()["f"]
    fo = oracle.update_acc_jrk(move, p6(r, v), pm_of(r), E2, np.zeros((n, 6)))
    label = f"Plummer n={n} nact=256 {precision}"
    if precision == "fp64":
        check_forces(f, fo, move, p6(r, v), pm_of(r), E2, label)
    else:
        check_mixed(f, fo, move, label)


C4_COUNT, C4_SEED = 64, 64


@pytest.mark.slow
def test_ensemble_recipe_64_clusters(oracle):
    """The ensemble configuration's own recipe: 64 Plummer clusters with N
    log-uniform in [1024, 32768] from seed 64 (cluster k from seed
    64000 + k), FP64, the launch picked by h4h_tune_launch, integrated to
    t = 0.125.  Every cluster reaches t_end with finite energies; graph
    replay and direct launches of the same launch give bit-identical
    states; the 4 smallest clusters take the oracle's number of block
    steps and interactions and reach its cumulative energy error to 2
    significant digits (each cluster is one run of
    Hermite4CPU_integration.cpp:38-130)."""
    from gravidy_amd import EnsembleHIP
    sizes = ic.log_uniform_sizes(C4_COUNT, C4_SEED)
    assert sizes.min() >= 1024 and sizes.max() <= 32768
    systems = [ic.plummer(int(nk), C4_SEED * 1000 + k) for k, nk in enumerate(sizes)]
    t_end = 0.125
    out = {}
    choice = None
    for graphs in (True, False):
        with EnsembleHIP(systems, E2, precision="fp64") as ens:
            ens.set_graphs(graphs)
            if choice is None:
                ens.init()
                choice = ens.tune_launch(16)
                print(f"ensemble launch: {choice}")
            else:
                ens.set_launch_choice(choice["workgroups"], choice["max_jsplit"],
                                      choice["iw_fill16"])
            res = ens.integration(t_end)
            out[graphs] = (res, ens.download())
    res, states = out[True]
    assert list(res["energy_time"]) == [0.0, t_end]
    assert np.all(np.isfinite(res["energy"]))
    assert np.all(res["itime"] == t_end), res["itime"]
    for c, st in enumerate(states):
        assert np.all(st["t"] <= t_end) and np.all(st["t"] + st["dt"] > t_end), c
    res2, states2 = out[False]
    assert np.array_equal(res["steps"], res2["steps"])
    assert np.array_equal(res["energy"], res2["energy"])
    for c, (a, b) in enumerate(zip(states, states2)):
        for key in ("r", "v", "f", "t", "dt"):
            assert bits_equal(a[key], b[key]), (c, key)
    cum = np.abs((res["energy"][-1] - res["energy"][0]) / res["energy"][0])
    print(f"ensemble recipe: {int(sizes.sum())} particles, block steps "
          f"{int(res['steps'].min())}..{int(res['steps'].max())}, CumE median "
          f"{np.median(cum):.3e} max {cum.max():.3e}")
    for c in np.argsort(sizes)[:4]:
        r, v = systems[c]
        ref = oracle.integrate(r, v, E2, t_end)
        e = ref["energy"]
        cum_ref = abs((e[-1] - e[0]) / e[0])
        print(f"cluster {c} n={sizes[c]}: steps {res['steps'][c]}/{len(ref['step_time'])} "
              f"interactions {res['interactions'][c]}/{ref['interactions']} "
              f"CumE {cum[c]:.4e} oracle {cum_ref:.4e}")
        assert res["steps"][c] == len(ref["step_time"]), c
        assert res["interactions"][c] == ref["interactions"], c
        assert float(f"{cum[c]:.1e}") == float(f"{cum_ref:.1e}") or \
            abs(cum[c] - cum_ref) <= 0.05 * cum_ref, c


def test_python_restart_exact_matches_oracle(oracle, tmp_path):
    """Exact precision, N = 1024 to t = 0.25 with a checkpoint, then a new
    context restarted from it to t = 0.5: the two runs' schedules together,
    the final state (bit for bit) and the energies at every output time
    (1e-13) are those of the oracle's uninterrupted run to t = 0.5
    (Hermite4CPU_integration.cpp:38-130)."""
    r, v = ic.plummer(1024, 6161)
    ref = oracle.integrate(r, v, E2, 0.5)
    ck = str(tmp_path / "quarter.ckpt")
    with Hermite4HIP(r, v, E2, precision="exact") as h:
        first = h.integration(0.25, checkpoint=ck)
    with Hermite4HIP(r, v, E2, precision="exact") as h:
        second = h.integration(0.5, restart=ck)
        s = h.download()
    steps = np.concatenate([first["step_time"], second["step_time"]])
    nact = np.concatenate([first["step_nact"], second["step_nact"]])
    assert np.array_equal(steps, ref["step_time"])
    assert np.array_equal(nact, ref["step_nact"])
    # a restarted run carries the checkpoint's counters on
    assert second["interactions"] == ref["interactions"]
    assert bits_equal(s["r"], ref["r"]) and bits_equal(s["v"][:, :3], ref["v"][:, :3])
    assert bits_equal(s["t"], ref["t"]) and bits_equal(s["dt"], ref["dt"])
    assert bits_equal(s["f"], ref["f"])
    times = np.concatenate([first["energy_time"], second["energy_time"][1:]])
    energies = np.concatenate([first["energy"], second["energy"][1:]])
    assert np.array_equal(times, ref["energy_time"])
    np.testing.assert_allclose(energies, ref["energy"], rtol=1e-13)
    assert second["energy_ini"] == first["energy_ini"]
    np.testing.assert_allclose(second["energy_ini"], ref["energy"][0], rtol=1e-13)
