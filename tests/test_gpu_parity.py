"""Parity of the HIP backend (through the C ABI) with the CPU oracle.

FP64_EXACT must be bit-identical to the oracle (Hermite4CPU's rounding):
forces, predictor, corrector, timesteps and whole-run schedules.

FP64 (the fast kernel: refined v_rsq_f64, FMA, j split) must stay within
REL_TOL = 1e-13 per particle on state-injected steps:
|a_gpu - a_ref| <= 1e-13 |a_ref|, and the same for the jerk.  A particle
whose pair terms cancel (|a_i| or |j_i| much smaller than sum_j of the
term magnitudes) cannot be reproduced to 1e-13 of its own magnitude by
any summation order other than the oracle's, because the oracle's
sequential sum carries an error of that size itself; for those particles
the bound is 1e-13 * sum_j |a_ij| (resp. |j_ij|), the condition-scaled
bound of a floating-point sum.

The condition-scaled bound is for the few particles whose pair terms
cancel: at most FALLBACK_MAX(k) of k checked particles (8, or 0.05 % of
large samples) may need it; more than that is a wrong force, not
cancellation.

MIXED (FP32 pair terms): every pair term carries a few FP32 roundings
(~1e-7 of itself), so the error is stated as a statistic over particles
plus a per-particle cap: the root-mean-square of |a_gpu - a_ref| / |a_ref|
must be <= 1e-7 (the target relative force error) and that of the jerk
<= 5e-7, and no particle may exceed MIXED_ACC_MAX = 2e-6 (acceleration)
or MIXED_JERK_MAX = 2e-5 (jerk), about twice the largest errors measured
over every particle of the N = 262144 Plummer sphere (1.0e-6 and 1.0e-5).
Its cumulative energy error over a whole run must stay within 2x the FP64
path's.  Energies are compared to 1e-13 relative (only the summation order
differs).
"""
import numpy as np
import pytest

from gravidy_amd import Hermite4HIP, ic, softening_e2

pytestmark = pytest.mark.gpu

E2 = softening_e2(1e-4)
REL_TOL = 1e-13
MIXED_ACC_RMS = 1e-7
MIXED_JERK_RMS = 5e-7
MIXED_ACC_MAX = 2e-6
MIXED_JERK_MAX = 2e-5


def fallback_max(k):
    """Largest number of the k checked particles allowed to fall back to the
    condition-scaled FP64 bound."""
    return max(8, k // 2000)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def rel_err(a, b):
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    den = np.where(den == 0, 1.0, den)
    return float((num / den).max())


def term_scales(i, p, pm, e2):
    """(sum_j |a_ij|, sum_j |j_ij|) for particle i (numpy, FP64)."""
    d = p[:, :3] - p[i, :3]
    w = p[:, 3:] - p[i, 3:]
    r2 = (d * d).sum(axis=1) + e2
    rinv = 1.0 / np.sqrt(r2)
    rv = (d * w).sum(axis=1)
    m = pm.astype(np.float64)
    a = m[:, None] * d * rinv[:, None] ** 3
    t = m[:, None] * (w * rinv[:, None] ** 3 - 3 * rv[:, None] * d * rinv[:, None] ** 5)
    a[i] = 0.0
    t[i] = 0.0
    return np.linalg.norm(a, axis=1).sum(), np.linalg.norm(t, axis=1).sum()


def check_forces(f, fo, idx, p, pm, e2, label, tol=REL_TOL):
    """Per particle: |da| <= tol |a| and |dj| <= tol |j|; where the pair
    terms cancel so that this fails, the condition-scaled bound
    |da| <= tol sum_j |a_ij| (and the same for the jerk) must hold."""
    da = np.linalg.norm(f[idx, :3] - fo[idx, :3], axis=1)
    dj = np.linalg.norm(f[idx, 3:] - fo[idx, 3:], axis=1)
    na = np.linalg.norm(fo[idx, :3], axis=1)
    nj = np.linalg.norm(fo[idx, 3:], axis=1)
    ra = da / np.where(na == 0, 1.0, na)
    rj = dj / np.where(nj == 0, 1.0, nj)
    over = np.nonzero((da > tol * na) | (dj > tol * nj))[0]
    # The condition-scaled bound is for the few particles whose terms cancel;
    # many particles over the bound are wrong forces, not cancellation.
    assert over.size <= fallback_max(idx.size), (
        f"{label}: {over.size} of {idx.size} particles above {tol:g} norm-wise, "
        f"max rel err acc {ra.max():.3e} jerk {rj.max():.3e}")
    worst = 0.0
    for k in over:
        sa, sj = term_scales(int(idx[k]), p, pm, e2)
        worst = max(worst, da[k] / sa, dj[k] / sj)
    print(f"{label}: rel err acc max {ra.max():.3e} p99 {np.percentile(ra, 99):.3e} "
          f"median {np.median(ra):.3e}; jerk max {rj.max():.3e} p99 {np.percentile(rj, 99):.3e} "
          f"median {np.median(rj):.3e}; {over.size} above {tol:g} norm-wise, "
          f"condition-scaled max {worst:.3e}")
    assert worst <= tol


def p6(r, v):
    return np.concatenate([r[:, :3], v[:, :3]], axis=1)


def pm_of(r):
    return r[:, 3].astype(np.float32)


def block_state(n, seed, T):
    """Plummer state with random forces and a block-time layout in which
    t_i + dt_i >= T for every i and equality holds for dt_i <= 2^-9."""
    rng = np.random.default_rng(seed)
    r, v = ic.plummer(n, seed)
    f = rng.normal(size=(n, 6)) * 0.3
    k = rng.integers(3, 13, size=n)
    dt = 2.0 ** -k.astype(np.float64)
    t = np.floor((T - 1e-12) / dt) * dt
    return r, v, f, t, dt


# --- initial forces -------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 255, 256, 1000, 1024, 3000])
def test_init_acc_jrk_exact_is_bitwise(oracle, n):
    r, v = ic.plummer(n, 100 + n)
    with Hermite4HIP(r, v, E2, precision="exact") as h:
        h.init_acc_jrk()
        f = h.download()["f"]
    fo = oracle.init_acc_jrk(p6(r, v), pm_of(r), E2)
    assert bits_equal(f, fo)


@pytest.mark.parametrize("n", [2, 255, 1024, 3000, 8192])
def test_init_acc_jrk_fp64_within_tolerance(oracle, n):
    r, v = ic.plummer(n, 200 + n)
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.init_acc_jrk()
        f = h.download()["f"]
    fo = oracle.init_acc_jrk(p6(r, v), pm_of(r), E2)
    check_forces(f, fo, np.arange(n), p6(r, v), pm_of(r), E2, f"init n={n} fp64")


@pytest.mark.parametrize("nact", [1, 10, 37, 100])
def test_update_fp64_small_active_sets_without_softening(oracle, nact):
    n = 4096
    rng = np.random.default_rng(nact)
    r, v = ic.plummer(n, 55)
    move = np.sort(rng.choice(n, size=nact, replace=False)).astype(np.uint32)
    with Hermite4HIP(r, v, 0.0, precision="fp64") as h:
        h.set_active(move)
        h.update_acc_jrk(nact)
        f = h.download()["f"]
    fo = oracle.update_acc_jrk(move, p6(r, v), pm_of(r), 0.0, np.zeros((n, 6)))
    assert np.all(np.isfinite(f))
    check_forces(f, fo, move, p6(r, v), pm_of(r), 0.0, f"update e2=0 nact={nact} fp64")


def test_init_acc_jrk_without_softening(oracle):
    r, v = ic.plummer(512, 77)
    fo = oracle.init_acc_jrk(p6(r, v), pm_of(r), 0.0)
    for prec in ("exact", "fp64"):
        with Hermite4HIP(r, v, 0.0, precision=prec) as h:
            h.init_acc_jrk()
            f = h.download()["f"]
        assert np.all(np.isfinite(f))
        if prec == "exact":
            assert bits_equal(f, fo)
        else:
            assert rel_err(f[:, :3], fo[:, :3]) <= REL_TOL


# --- one state-injected block step ---------------------------------------

@pytest.mark.parametrize("n,seed", [(1024, 1), (1000, 2), (4096, 3)])
def test_block_step_pieces_exact_are_bitwise(oracle, n, seed):
    T = 0.375 + 2.0 ** -9
    r, v, f, t, dt = block_state(n, seed, T)
    with Hermite4HIP(r, v, E2, precision="exact") as h:
        h.upload(r, v, f, t, dt)
        nact = h.find_particles_to_move(T)
        move = h.get_active()
        mo = oracle.find_particles_to_move(T, t, dt)
        assert nact == mo.size > 0
        assert np.array_equal(move, mo)

        h.save_old_acc_jrk()
        h.predicted_pos_vel(T)
        p, pm = h.download_predictor()
        po, pmo = oracle.predicted_pos_vel(T, t, r, v, f)
        assert bits_equal(p, po) and bits_equal(pm, pmo)

        h.update_acc_jrk(nact)
        fo = oracle.update_acc_jrk(mo, po, pmo, E2, f)
        h.correction_pos_vel(T, nact)
        atime = h.next_integration_time()
        s = h.download()
        a2o, a3o, ro, vo, to, dto = oracle.correction_pos_vel(
            T, mo, po, fo, f, np.zeros((n, 3)), np.zeros((n, 3)), r, v, t, dt)
        assert bits_equal(s["f"], fo)
        assert bits_equal(s["old"][mo], f[mo])
        assert bits_equal(s["a2"][mo], a2o[mo]) and bits_equal(s["a3"][mo], a3o[mo])
        assert bits_equal(s["r"], ro) and bits_equal(s["v"][:, :3], vo[:, :3])
        assert bits_equal(s["t"], to) and bits_equal(s["dt"], dto)
        assert atime == oracle.next_integration_time(to, dto)


# nact picks the lane-group width: 1, 3, 10 -> 4; 37 -> 16; 100, 300 -> 64;
# 700, 5000 -> 256 i slots per workgroup.
@pytest.mark.parametrize("n,nact", [(1024, 1), (1024, 37), (4096, 700), (65536, 1),
                                    (65536, 3), (65536, 10), (65536, 37), (65536, 100),
                                    (65536, 300), (65536, 5000), (300000, 64)])
def test_update_fp64_jsplit_within_tolerance(oracle, n, nact):
    rng = np.random.default_rng(n + nact)
    r, v = ic.plummer(n, 300 + n)
    move = np.sort(rng.choice(n, size=nact, replace=False)).astype(np.uint32)
    f0 = np.zeros((n, 6))
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.set_active(move)
        h.update_acc_jrk(nact)
        f = h.download()["f"]
    fo = oracle.update_acc_jrk(move, p6(r, v), pm_of(r), E2, f0)
    check_forces(f, fo, move, p6(r, v), pm_of(r), E2, f"update n={n} nact={nact} fp64")
    rest = np.setdiff1d(np.arange(n), move)
    assert np.all(f[rest] == 0.0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,nact", [(262144, 8192), (300000, 20000)])
def test_update_large_active_sets_within_tolerance(oracle, n, nact, precision):
    """Active sets of 8192 or more over more than 256 j tiles: the plan's
    chunks of 4 tiles and the finish pass's 4 lanes per particle.  A seeded
    sample of 512 of the active particles against the oracle (each
    particle's sum is independent of the others'), and every particle
    outside the active set keeps zero forces."""
    rng = np.random.default_rng(n + nact)
    r, v = ic.plummer(n, 700 + n)
    move = np.sort(rng.choice(n, size=nact, replace=False)).astype(np.uint32)
    with Hermite4HIP(r, v, E2, precision=precision) as h:
        h.set_active(move)
        h.update_acc_jrk(nact)
        f = h.download()["f"]
    sample = np.sort(rng.choice(move, size=512, replace=False)).astype(np.uint32)
    fo = oracle.update_acc_jrk(sample, p6(r, v), pm_of(r), E2, np.zeros((n, 6)))
    label = f"update n={n} nact={nact} {precision} (512 sampled)"
    if precision == "fp64":
        check_forces(f, fo, sample, p6(r, v), pm_of(r), E2, label)
    else:
        check_mixed(f, fo, sample, label)
    rest = np.setdiff1d(np.arange(n), move)
    assert np.all(f[rest] == 0.0)


def test_autotune_table_is_valid_and_keeps_results_within_tolerance(oracle):
    """h4h_autotune picks a (workgroup share, lane-group width) per log2(nact)
    bucket on the context's own particles; forces computed under the table
    stay within the FP64 bound, set_tuning(None) restores the default rule
    and invalid widths are refused."""
    from gravidy_amd._lib import H4HError
    n = 8192
    r, v = ic.plummer(n, 818)
    rng = np.random.default_rng(5)
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        table = h.autotune(reps=1)
        assert set(table) == set(range(14)), table  # nact = 1, 2, ..., 8192
        for b, (g, w) in table.items():
            assert w in (4, 16, 64, 256) and 64 <= g <= 65536, (b, g, w)
        for nact in (3, 100, 2000, n):
            move = np.sort(rng.choice(n, size=nact, replace=False)).astype(np.uint32)
            h.set_active(move)
            h.update_acc_jrk(nact)
            f = h.download()["f"]
            fo = oracle.update_acc_jrk(move, p6(r, v), pm_of(r), E2, np.zeros((n, 6)))
            check_forces(f, fo, move, p6(r, v), pm_of(r), E2, f"tuned nact={nact}")
        h.set_tuning(None)
        with pytest.raises(H4HError):
            h.set_tuning({3: (512, 8)})


def test_update_after_autotune_takes_every_work_item(oracle):
    """At N = 65536 a full update has more work items than workgroups, so
    the update kernel hands them out from its work queue.  The autotuner
    launches the update many times without a finish pass in between; the
    queue must still start from zero for every launch: forces computed
    after it, for new positions, match the oracle."""
    n = 65536
    r, v = ic.plummer(n, 9)
    r2, v2 = ic.plummer(n, 10)
    move = np.arange(n, dtype=np.uint32)
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.autotune(reps=1)
        h.upload(r2, v2)
        h.set_active(move)
        h.update_acc_jrk(n)
        f = h.download()["f"]
    fo = oracle.update_acc_jrk(move, p6(r2, v2), pm_of(r2), E2, np.zeros((n, 6)))
    check_forces(f, fo, move, p6(r2, v2), pm_of(r2), E2, "after autotune n=65536")


# --- initial timesteps ----------------------------------------------------

def _exact_power_inputs():
    """|a| values that make ETA*(|a|/1) land exactly on, and one ulp either
    side of, powers of two (|j| = 1)."""
    eta = float(np.float32(0.01))
    rows = []
    for k in range(-12, -2):
        target = 2.0 ** k
        x = target / eta
        for _ in range(64):
            y = eta * x
            if y == target:
                break
            x = np.nextafter(x, np.inf if y < target else -np.inf)
        for a in (x, np.nextafter(x, np.inf), np.nextafter(x, -np.inf)):
            rows.append([a, 0, 0, 1.0, 0, 0])
    rows.append([1.0, 0, 0, 0, 0, 0])   # |j| = 0
    rows.append([0, 0, 0, 1.0, 0, 0])   # |a| = 0
    rows.append([0, 0, 0, 0, 0, 0])     # 0/0
    return np.array(rows, dtype=np.float64)


def test_init_dt_is_bitwise(oracle):
    r, v = ic.plummer(2048, 5)
    fo = oracle.init_acc_jrk(p6(r, v), pm_of(r), E2)
    extra = _exact_power_inputs()
    f = np.concatenate([fo, extra])
    n = f.shape[0]
    r2 = np.concatenate([r, np.tile(r[:1], (extra.shape[0], 1))])
    v2 = np.concatenate([v, np.tile(v[:1], (extra.shape[0], 1))])
    with Hermite4HIP(r2, v2, E2, precision="exact") as h:
        h.upload(r2, v2, f)
        ctime = h.init_dt(1.0e10, itime=0.0)
        s = h.download()
    to, dto, co = oracle.init_dt(f, itime=0.0)
    bad = np.nonzero(s["dt"] != dto)[0]
    assert bad.size == 0, [(i, f[i], s["dt"][i], dto[i]) for i in bad[:5]]
    assert bits_equal(s["t"], to)
    assert ctime == co


# --- energy ---------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 8, 1000, 4096])
def test_energy_matches(oracle, n):
    if n == 8:
        r = np.array([[x, y, z, 1.0] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
        v = np.zeros((8, 4))
    else:
        r, v = ic.plummer(n, 400 + n)
    with Hermite4HIP(r, v, E2) as h:
        ek, ep = h.get_energy()
    eko, epo = oracle.get_energy(r, v)
    assert ek == pytest.approx(eko, rel=1e-13, abs=1e-300)
    assert ep == pytest.approx(epo, rel=1e-13, abs=1e-300)


def test_energy_of_coincident_particles_is_minus_infinity(oracle):
    """The energy is unsoftened: two particles at one position give -inf
    in the reference's arithmetic (division by sqrt(0)) and on the GPU."""
    r = np.array([[0.5, 0.25, -1.0, 0.5], [0.5, 0.25, -1.0, 0.25], [1.0, 0.0, 0.0, 0.25]])
    v = np.zeros((3, 4))
    with Hermite4HIP(r, v, E2) as h:
        ek, ep = h.get_energy()
    eko, epo = oracle.get_energy(r, v)
    assert epo == -np.inf and ep == -np.inf
    assert ek == eko == 0.0


# --- whole runs -------------------------------------------------------------

def test_plummer1024_run_exact_is_bitwise(oracle):
    """Plummer N=1024, eps=1e-4, eta=0.01, dt in [2^-23, 2^-3], to t=1:
    identical block-timestep schedule and bit-identical final state."""
    r, v = ic.plummer(1024, 12345)
    ref = oracle.integrate(r, v, E2, 1.0)
    with Hermite4HIP(r, v, E2, precision="exact") as h:
        run = h.integration(1.0)
        s = h.download()
    assert np.array_equal(run["step_time"], ref["step_time"])
    assert np.array_equal(run["step_nact"], ref["step_nact"])
    assert run["interactions"] == ref["interactions"]
    assert bits_equal(s["r"], ref["r"]) and bits_equal(s["v"][:, :3], ref["v"][:, :3])
    assert bits_equal(s["dt"], ref["dt"]) and bits_equal(s["t"], ref["t"])
    assert np.array_equal(run["energy_time"], ref["energy_time"])
    np.testing.assert_allclose(run["energy"], ref["energy"], rtol=1e-13)


def test_device_loop_matches_host_loop(oracle):
    """h4h_run (graph-replayed device loop, and the same loop launched
    without graphs) and one h4h_step call per block step give the same
    schedule and bit-identical state, and the device loop honours max_steps
    and the output threshold."""
    r, v = ic.plummer(2048, 77)
    out = {}
    for mode, graphs in (("run", True), ("step", True), ("run", False)):
        with Hermite4HIP(r, v, E2, precision="fp64") as h:
            h.set_graphs(graphs)
            res = h.integration(0.25, mode=mode)
            out[mode, graphs] = (res, h.download())
    (ra, sa) = out["run", True]
    for key in (("step", True), ("run", False)):
        rb, sb = out[key]
        assert np.array_equal(ra["step_time"], rb["step_time"]), key
        assert np.array_equal(ra["step_nact"], rb["step_nact"]), key
        assert bits_equal(sa["r"], sb["r"]) and bits_equal(sa["dt"], sb["dt"]), key
        np.testing.assert_array_equal(ra["energy"], rb["energy"])
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.init_acc_jrk()
        atime = h.init_dt()
        info, st, sn = h.run(0.0, atime, 0.25, 1.0, 5)
        assert info["steps"] == 5 and info["stopped"] == 3
        assert np.array_equal(st, ra["step_time"][:5]) and np.array_equal(sn, ra["step_nact"][:5])
        info, st, _ = h.run(info["itime"], info["atime"], 0.25, 0.0625, 1 << 16)
        assert info["stopped"] == 1 and st[-1] >= 0.0625 and np.all(st[:-1] < 0.0625)
        assert np.array_equal(st, ra["step_time"][5:5 + len(st)])


@pytest.mark.parametrize("n,piece", [(4096, 1), (4096, 31), (4096, 33), (1024, 1), (1024, 7),
                                     (1024, 33)])
def test_device_loop_in_pieces_is_bitwise(n, piece):
    """A device loop cut into runs of `piece` block steps by max_steps
    (runs that end inside or at the end of a captured chunk; N = 1024 with
    one-workgroup steps in between), with one h4h_step among them, gives
    the schedule and the state of one uninterrupted run of the same steps,
    bit for bit: the loop tail of each run's last step is made before the
    run returns, and the next run starts from it."""
    r, v = ic.plummer(n, 303)
    total = 100
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.init_acc_jrk()
        atime = h.init_dt()
        info, st_ref, sn_ref = h.run(0.0, atime, 1.0, 1.0, total)
        assert info["steps"] == total
        ref = h.download()
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.init_acc_jrk()
        atime = h.init_dt()
        itime, done, st, sn = 0.0, 0, [], []
        while done < total:
            if done == 50:
                nact, nxt = h.step(atime)
                itime, atime = atime, nxt
                st.append(itime)
                sn.append(nact)
                done += 1
                continue
            info, s1, n1 = h.run(itime, atime, 1.0, 1.0, min(piece, total - done))
            assert info["steps"] == min(piece, total - done) and info["stopped"] == 3
            itime, atime = info["itime"], info["atime"]
            st += list(s1)
            sn += list(n1)
            done += info["steps"]
        got = h.download()
    assert np.array_equal(np.array(st), st_ref) and np.array_equal(np.array(sn), sn_ref)
    for key in ("r", "v", "f", "t", "dt", "a2", "a3"):
        assert bits_equal(got[key], ref[key]), key


@pytest.mark.parametrize("small_path", [64, 0])
def test_plummer1024_run_fp64_schedule_and_energy_error_match(oracle, small_path):
    """Fast FP64 kernel over the same run, with the one-workgroup step for
    small active sets (the default, 64) and without it (0): the same
    block-timestep schedule as the oracle (ITIME and nact of every block
    step), and a cumulative energy error at t_end (CumE, Logger.cpp:328)
    equal to the oracle's to 2 significant digits."""
    r, v = ic.plummer(1024, 12345)
    ref = oracle.integrate(r, v, E2, 1.0)
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.set_small_path(small_path)
        run = h.integration(1.0)
        small = h.stats()["small_steps"]
    assert (small > 0) == (small_path > 0)
    cum = abs((run["energy"][-1] - run["energy"][0]) / run["energy"][0])
    cum_ref = abs((ref["energy"][-1] - ref["energy"][0]) / ref["energy"][0])
    print(f"fp64 N=1024 small_path={small_path}: steps={len(run['step_time'])}/"
          f"{len(ref['step_time'])} ({small} in one workgroup) CumE={cum:.6e} "
          f"oracle CumE={cum_ref:.6e}")
    assert np.array_equal(run["step_time"], ref["step_time"])
    assert np.array_equal(run["step_nact"], ref["step_nact"])
    assert run["interactions"] == ref["interactions"]
    assert float(f"{cum:.1e}") == float(f"{cum_ref:.1e}") or abs(cum - cum_ref) <= 0.05 * cum_ref


# --- mixed precision ---------------------------------------------------------

def check_mixed(f, fo, idx, label):
    ra = np.linalg.norm(f[idx, :3] - fo[idx, :3], axis=1) / np.linalg.norm(fo[idx, :3], axis=1)
    rj = np.linalg.norm(f[idx, 3:] - fo[idx, 3:], axis=1) / np.linalg.norm(fo[idx, 3:], axis=1)
    rms_a = float(np.sqrt((ra ** 2).mean()))
    rms_j = float(np.sqrt((rj ** 2).mean()))
    print(f"{label}: acc rel err rms {rms_a:.3e} median {np.median(ra):.3e} "
          f"p99 {np.percentile(ra, 99):.3e} max {ra.max():.3e}; jerk rms {rms_j:.3e} "
          f"median {np.median(rj):.3e} p99 {np.percentile(rj, 99):.3e} max {rj.max():.3e}")
    assert rms_a <= MIXED_ACC_RMS
    assert rms_j <= MIXED_JERK_RMS
    assert ra.max() <= MIXED_ACC_MAX and rj.max() <= MIXED_JERK_MAX


@pytest.mark.parametrize("n,nact", [(1024, 1024), (8192, 8192), (65536, 3), (65536, 37),
                                    (65536, 300), (65536, 5000)])
def test_update_mixed_within_tolerance(oracle, n, nact):
    rng = np.random.default_rng(n + 7 * nact)
    r, v = ic.plummer(n, 500 + n)
    move = np.sort(rng.choice(n, size=nact, replace=False)).astype(np.uint32)
    with Hermite4HIP(r, v, E2, precision="mixed") as h:
        h.set_active(move)
        h.update_acc_jrk(nact)
        f = h.download()["f"]
    fo = oracle.update_acc_jrk(move, p6(r, v), pm_of(r), E2, np.zeros((n, 6)))
    check_mixed(f, fo, move, f"update n={n} nact={nact} mixed")


def test_set_precision_between_updates_matches_fresh_contexts():
    """h4h_set_precision re-encodes the predictor's pair weights (a double
    for the FP64 kernel, float bits for the mixed one): after every switch
    the update is bit-identical to that of a context created with the new
    precision."""
    n, nact = 4096, 700
    rng = np.random.default_rng(11)
    r, v = ic.plummer(n, 77)
    move = np.sort(rng.choice(n, size=nact, replace=False)).astype(np.uint32)

    def fresh(prec):
        with Hermite4HIP(r, v, E2, precision=prec) as h:
            h.set_active(move)
            h.update_acc_jrk(nact)
            return h.download()["f"]

    want = {p: fresh(p) for p in ("fp64", "mixed", "exact")}
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        h.set_active(move)
        for prec in ("mixed", "fp64", "exact", "mixed", "exact", "fp64"):
            h.set_precision(prec)
            h.update_acc_jrk(nact)
            f = h.download()["f"]
            assert bits_equal(f[move], want[prec][move]), prec


def test_plummer1024_run_mixed_energy_drift(oracle):
    """Plummer N=1024 to t=1 with the mixed kernel: the cumulative energy
    error at t_end stays within 2x of the FP64 path's."""
    r, v = ic.plummer(1024, 12345)
    runs = {}
    for prec in ("fp64", "mixed"):
        with Hermite4HIP(r, v, E2, precision=prec) as h:
            run = h.integration(1.0)
        e = run["energy"]
        runs[prec] = (abs((e[-1] - e[0]) / e[0]), len(run["step_time"]))
    print(f"N=1024 CumE fp64={runs['fp64'][0]:.4e} ({runs['fp64'][1]} steps) "
          f"mixed={runs['mixed'][0]:.4e} ({runs['mixed'][1]} steps)")
    assert runs["mixed"][0] <= 2.0 * runs["fp64"][0]


# --- ensembles -----------------------------------------------------------------

@pytest.mark.parametrize("graphs", [True, False])
def test_ensemble_exact_clusters_match_oracle(oracle, graphs):
    """Six uneven clusters in one context, exact mode, to t=0.25: every
    cluster's schedule length, final state and energies equal the oracle's
    run of that cluster alone (graph-replayed and directly launched loop)."""
    from gravidy_amd import EnsembleHIP
    sizes = [256, 300, 512, 777, 1000, 1024]
    systems = [ic.plummer(n, 900 + n) for n in sizes]
    with EnsembleHIP(systems, E2, precision="exact") as ens:
        ens.set_graphs(graphs)
        res = ens.integration(0.25)
        states = ens.download()
    assert list(res["energy_time"]) == [0.0, 0.125, 0.25]
    for c, (r, v) in enumerate(systems):
        ref = oracle.integrate(r, v, E2, 0.25)
        assert res["steps"][c] == len(ref["step_time"])
        assert res["interactions"][c] == ref["interactions"]
        assert bits_equal(states[c]["r"], ref["r"]), c
        assert bits_equal(states[c]["v"][:, :3], ref["v"][:, :3]), c
        assert bits_equal(states[c]["dt"], ref["dt"]) and bits_equal(states[c]["t"], ref["t"])
        np.testing.assert_allclose(res["energy"][:, c], ref["energy"], rtol=1e-13)


@pytest.mark.parametrize("sizes,t_end", [([2, 300, 3, 5, 257], 0.125), ([1, 64], 2.0 ** -16)])
def test_ensemble_with_tiny_clusters_matches_oracle(oracle, sizes, t_end):
    """Clusters of 1, 2, 3 and 5 particles beside larger ones.  A single
    particle has no force, so its steps are NaN-derived and normalize_dt
    keeps it at the smallest step (2^-23), as in the reference."""
    from gravidy_amd import EnsembleHIP
    systems = [ic.plummer(n, 1100 + n) for n in sizes]
    with EnsembleHIP(systems, E2, precision="exact") as ens:
        res = ens.integration(t_end, interval=t_end)
        states = ens.download()
    for c, (r, v) in enumerate(systems):
        ref = oracle.integrate(r, v, E2, t_end, interval=t_end)
        assert res["steps"][c] == len(ref["step_time"]), c
        assert bits_equal(states[c]["r"], ref["r"]), c
        assert bits_equal(states[c]["dt"], ref["dt"]) and bits_equal(states[c]["t"], ref["t"]), c
        np.testing.assert_allclose(res["energy"][:, c], ref["energy"], rtol=1e-13, atol=1e-300)


def test_tune_launch_leaves_the_state_and_results_unchanged():
    """h4h_tune_launch runs its trial steps on a copy of the state: the
    state after tuning is bit-identical to the state before, and an exact
    run after tuning equals an untuned one."""
    from gravidy_amd import EnsembleHIP
    systems = [ic.plummer(n, 700 + n) for n in (300, 1024, 2000, 513)]
    runs = []
    for tune in (False, True):
        with EnsembleHIP(systems, E2, precision="exact") as ens:
            if tune:
                ens.init()
                before = ens.download()
                choice = ens.tune_launch(32)
                after = ens.download()
                for a, b in zip(before, after):
                    for key in ("r", "v", "f", "t", "dt"):
                        assert bits_equal(a[key], b[key]), key
                assert choice["workgroups"] in (256, 512, 1024, 2048, 4096)
                assert choice["max_jsplit"] in (16, 64, 256, 1024)
                assert choice["iw_fill16"] in (8, 12, 15) and choice["ms_per_step"] > 0
            res = ens.integration(0.125)
            runs.append((res, ens.download()))
    with EnsembleHIP(systems, E2, precision="exact") as ens:
        ens.set_launch_choice(choice["workgroups"], choice["max_jsplit"], choice["iw_fill16"])
        runs.append((ens.integration(0.125), ens.download()))
        from gravidy_amd._lib import H4HError
        for bad in ((0, 64, 12), (1024, 0, 12), (1024, 64, 0), (1024, 64, 17)):
            with pytest.raises(H4HError):
                ens.set_launch_choice(*bad)
    (ra, sa) = runs[0]
    for rb, sb in runs[1:]:
        assert np.array_equal(ra["steps"], rb["steps"])
        for a, b in zip(sa, sb):
            assert bits_equal(a["r"], b["r"]) and bits_equal(a["dt"], b["dt"])
    with Hermite4HIP(*systems[2], E2, precision="fp64") as h:
        h.init_acc_jrk()
        h.init_dt()
        s0 = h.download()
        h.tune_launch(32)
        s1 = h.download()
        assert bits_equal(s0["r"], s1["r"]) and bits_equal(s0["f"], s1["f"])


def test_ensemble_of_one_equals_single_context():
    from gravidy_amd import EnsembleHIP
    r, v = ic.plummer(3000, 31)
    with Hermite4HIP(r, v, E2, precision="fp64") as h:
        run = h.integration(0.125)
        s1 = h.download()
    with EnsembleHIP([(r, v)], E2, precision="fp64") as ens:
        res = ens.integration(0.125)
        s2 = ens.download()[0]
    assert res["steps"][0] == len(run["step_time"])
    assert bits_equal(s1["r"], s2["r"]) and bits_equal(s1["dt"], s2["dt"])
    np.testing.assert_array_equal(res["energy"][:, 0], run["energy"])


def test_ensemble_fp64_energy_errors_match_oracle(oracle):
    """FP64 kernel, 5 clusters of 1k-4k to t=0.25: per-cluster cumulative
    energy error agrees with the oracle's to 2 significant digits."""
    from gravidy_amd import EnsembleHIP
    sizes = [1024, 1500, 2048, 3000, 4096]
    systems = [ic.plummer(n, 700 + n) for n in sizes]
    with EnsembleHIP(systems, E2, precision="fp64") as ens:
        res = ens.integration(0.25)
    for c, (r, v) in enumerate(systems):
        ref = oracle.integrate(r, v, E2, 0.25)
        e, eo = res["energy"][:, c], ref["energy"]
        cum = abs((e[-1] - e[0]) / e[0])
        cum_ref = abs((eo[-1] - eo[0]) / eo[0])
        print(f"cluster {c} n={sizes[c]}: CumE {cum:.4e} oracle {cum_ref:.4e} "
              f"steps {res['steps'][c]}/{len(ref['step_time'])}")
        assert abs(cum - cum_ref) <= 0.05 * cum_ref + 1e-15


def test_ensemble_mixed_energy_drift_within_2x_fp64():
    """Mixed kernel in an ensemble (5 clusters of 1k-4k to t=0.25): each
    cluster's cumulative energy error is within 2x its FP64 one, and every
    cluster reaches t_end."""
    from gravidy_amd import EnsembleHIP
    sizes = [1024, 1500, 2048, 3000, 4096]
    systems = [ic.plummer(n, 900 + n) for n in sizes]
    cum = {}
    for precision in ("fp64", "mixed"):
        with EnsembleHIP(systems, E2, precision=precision) as ens:
            res = ens.integration(0.25)
        e = res["energy"]
        assert list(res["energy_time"]) == [0.0, 0.125, 0.25]
        cum[precision] = np.abs((e[-1] - e[0]) / e[0])
    for c, n in enumerate(sizes):
        print(f"cluster {c} n={n}: CumE fp64 {cum['fp64'][c]:.4e} mixed {cum['mixed'][c]:.4e}")
        assert cum["mixed"][c] <= 2.0 * cum["fp64"][c] + 1e-12
