"""Host-side mirror of GraviDy's Hermite4 integrator interface over the
HIP backend (libh4hip.so).

`Hermite4HIP` keeps the reference's method names and call order
(Hermite4.hpp:48-88, cpu/Hermite4CPU.hpp:51-69): init_acc_jrk, init_dt,
find_particles_to_move, save_old_acc_jrk, predicted_pos_vel,
update_acc_jrk, correction_pos_vel, next_integration_time, get_energy and
integration().  All state stays on the GPU; host arrays are produced only
by download().
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, PRECISIONS

# common.hpp:122,129
D_TIME_MIN = 1.1920928955078125e-07
D_TIME_MAX = 0.125
# common.hpp:114,119 (float 0.01)
ETA_DEFAULT = float(np.float32(0.01))


def softening_e2(eps):
    """Softening squared as NbodySystem.cpp:50 computes it (float eps,
    float product)."""
    e = np.float32(eps)
    return float(np.float32(e * e))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _opt(a, shape):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class Hermite4HIP:
    """One N-body system resident on the current HIP device.

    r: (n,4) positions + mass (double4 h_r); v: (n,3|4) velocities.
    e2: softening squared (use softening_e2(eps) for the reference's value).
    """

    def __init__(self, r, v, e2, eta=ETA_DEFAULT, precision="fp64",
                 dt_min=D_TIME_MIN, dt_max=D_TIME_MAX):
        self.lib = _lib.load()
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if r.ndim != 2 or r.shape[1] != 4:
            raise ValueError("r must be (n,4): x, y, z, m")
        n = r.shape[0]
        if v.shape == (n, 3):
            v = np.concatenate([v, np.zeros((n, 1))], axis=1)
        if v.shape != (n, 4):
            raise ValueError("v must be (n,3) or (n,4)")
        self.n = n
        self.e2 = float(e2)
        self.eta = float(np.float32(eta))
        self.dt_min = float(dt_min)
        self.dt_max = float(dt_max)
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.precision = prec
        self._ctx = ctypes.c_void_p()
        check("h4h_create", self.lib.h4h_create(ctypes.byref(self._ctx), n, self.e2,
                                                self.eta, self.dt_min, self.dt_max, prec))
        self.upload(r, v)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self._ctx:
            self.lib.h4h_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- configuration ---------------------------------------------------
    def set_precision(self, precision):
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check("h4h_set_precision", self.lib.h4h_set_precision(self._ctx, prec))
        self.precision = prec

    def set_launch(self, workgroups=0, max_jsplit=0):
        check("h4h_set_launch", self.lib.h4h_set_launch(self._ctx, workgroups, max_jsplit))

    def autotune(self, reps=2):
        """Runs h4h_autotune on this system; returns {bucket: (grid, iw)}."""
        g = np.zeros(32, np.uint32)
        w = np.zeros(32, np.uint32)
        u = ctypes.POINTER(ctypes.c_uint32)
        check("h4h_autotune", self.lib.h4h_autotune(self._ctx, int(reps), g.ctypes.data_as(u),
                                                    w.ctypes.data_as(u)))
        return {b: (int(g[b]), int(w[b])) for b in range(32) if g[b] or w[b]}

    def set_tuning(self, table):
        """table: {bucket: (grid, iw)} as autotune returns, or None."""
        u = ctypes.POINTER(ctypes.c_uint32)
        if table is None:
            check("h4h_set_tuning", self.lib.h4h_set_tuning(self._ctx, None, None))
            return
        g = np.zeros(32, np.uint32)
        w = np.zeros(32, np.uint32)
        for b, (gg, ww) in table.items():
            g[int(b)] = gg
            w[int(b)] = ww
        check("h4h_set_tuning", self.lib.h4h_set_tuning(self._ctx, g.ctypes.data_as(u),
                                                        w.ctypes.data_as(u)))

    def set_timing(self, enable=True):
        check("h4h_set_timing", self.lib.h4h_set_timing(self._ctx, 1 if enable else 0))

    def tune_launch(self, steps=64):
        """Launch autotuner (h4h_tune_launch) over the next `steps` device
        block steps from the current state (left unchanged); returns the
        installed choice."""
        c = _lib.H4HLaunchChoice()
        check("h4h_tune_launch", self.lib.h4h_tune_launch(self._ctx, int(steps), ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def set_launch_choice(self, workgroups, max_jsplit, iw_fill16):
        """Installs a launch (as tune_launch returns it)."""
        c = _lib.H4HLaunchChoice(int(workgroups), int(max_jsplit), int(iw_fill16), 0.0)
        check("h4h_set_launch_choice", self.lib.h4h_set_launch_choice(self._ctx, ctypes.byref(c)))

    def set_graphs(self, enable=True):
        """Device loops replay captured graphs (default) or launch directly."""
        check("h4h_set_graphs", self.lib.h4h_set_graphs(self._ctx, 1 if enable else 0))

    def set_small_path(self, max_active=64):
        """Block steps with 1..max_active (at most 256) active particles of a
        system of at most 1024 run in one workgroup (0: every step on the
        multi-workgroup kernels)."""
        check("h4h_set_small_path", self.lib.h4h_set_small_path(self._ctx, int(max_active)))

    def stats(self, reset=False):
        s = _lib.H4HStats()
        check("h4h_get_stats", self.lib.h4h_get_stats(self._ctx, ctypes.byref(s), int(reset)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    # -- transfers -------------------------------------------------------
    def upload(self, r, v, f=None, t=None, dt=None):
        n = self.n
        r = _opt(r, (n, 4))
        v = np.ascontiguousarray(v, dtype=np.float64)
        if v.shape == (n, 3):
            v = np.concatenate([v, np.zeros((n, 1))], axis=1)
        v = _opt(v, (n, 4))
        f = _opt(f, (n, 6))
        t = _opt(t, (n,))
        dt = _opt(dt, (n,))
        check("h4h_upload", self.lib.h4h_upload(
            self._ctx, _dp(r), _dp(v), None if f is None else _dp(f),
            None if t is None else _dp(t), None if dt is None else _dp(dt)))

    def download(self):
        n = self.n
        out = {
            "r": np.zeros((n, 4)), "v": np.zeros((n, 4)), "f": np.zeros((n, 6)),
            "t": np.zeros(n), "dt": np.zeros(n), "old": np.zeros((n, 6)),
            "a2": np.zeros((n, 3)), "a3": np.zeros((n, 3)),
        }
        check("h4h_download", self.lib.h4h_download(
            self._ctx, _dp(out["r"]), _dp(out["v"]), _dp(out["f"]), _dp(out["t"]),
            _dp(out["dt"]), _dp(out["old"]), _dp(out["a2"]), _dp(out["a3"])))
        return out

    def download_predictor(self):
        p = np.zeros((self.n, 6))
        pm = np.zeros(self.n, np.float32)
        check("h4h_download_predictor", self.lib.h4h_download_predictor(
            self._ctx, _dp(p), pm.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p, pm

    def upload_predictor(self, p, pm):
        p = _opt(p, (self.n, 6))
        pm = np.ascontiguousarray(pm, dtype=np.float32)
        check("h4h_upload_predictor", self.lib.h4h_upload_predictor(
            self._ctx, _dp(p), pm.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def set_active(self, move):
        move = np.ascontiguousarray(move, dtype=np.uint32)
        check("h4h_set_active", self.lib.h4h_set_active(
            self._ctx, move.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), move.size))
        return move.size

    def get_active(self):
        move = np.zeros(self.n, np.uint32)
        k = ctypes.c_uint32()
        check("h4h_get_active", self.lib.h4h_get_active(
            self._ctx, move.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(k)))
        return move[:k.value].copy()

    # -- the Hermite4 interface --------------------------------------------
    def init_acc_jrk(self):
        check("h4h_init_acc_jrk", self.lib.h4h_init_acc_jrk(self._ctx))

    def init_dt(self, atime=1.0e10, eta=ETA_DEFAULT, itime=0.0):
        c = ctypes.c_double(atime)
        check("h4h_init_dt", self.lib.h4h_init_dt(self._ctx, ctypes.byref(c),
                                                  float(np.float32(eta)), float(itime)))
        return c.value

    def find_particles_to_move(self, itime):
        k = ctypes.c_uint32()
        check("h4h_find_particles_to_move",
              self.lib.h4h_find_particles_to_move(self._ctx, float(itime), ctypes.byref(k)))
        return k.value

    def save_old_acc_jrk(self):
        check("h4h_save_old_acc_jrk", self.lib.h4h_save_old_acc_jrk(self._ctx))

    def predicted_pos_vel(self, itime):
        check("h4h_predicted_pos_vel", self.lib.h4h_predicted_pos_vel(self._ctx, float(itime)))

    def update_acc_jrk(self, nact):
        check("h4h_update_acc_jrk", self.lib.h4h_update_acc_jrk(self._ctx, int(nact)))

    def correction_pos_vel(self, itime, nact):
        check("h4h_correction_pos_vel",
              self.lib.h4h_correction_pos_vel(self._ctx, float(itime), int(nact)))

    def next_integration_time(self):
        a = ctypes.c_double()
        check("h4h_next_integration_time",
              self.lib.h4h_next_integration_time(self._ctx, ctypes.byref(a)))
        return a.value

    def step(self, itime):
        """Fused block step; returns (nact, next time)."""
        k = ctypes.c_uint32()
        a = ctypes.c_double()
        check("h4h_step", self.lib.h4h_step(self._ctx, float(itime), ctypes.byref(k),
                                            ctypes.byref(a)))
        return k.value, a.value

    def run(self, itime, atime, t_end, t_stop, max_steps, record=True):
        """Block steps on the device until a step with ITIME >= t_stop is
        done, ITIME >= t_end at the loop test, or max_steps steps (h4h_run).
        Returns (info dict, step times, step nacts)."""
        io = _lib.H4HRunInfo(float(itime), float(atime), 0, 0, 0, 0, 0)
        st = np.zeros(max_steps) if record else None
        sn = np.zeros(max_steps, np.uint32) if record else None
        check("h4h_run", self.lib.h4h_run(
            self._ctx, float(t_end), float(t_stop), int(max_steps), ctypes.byref(io),
            None if st is None else _dp(st),
            None if sn is None else sn.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        info = {k: getattr(io, k) for k, _ in io._fields_}
        k = info["steps"]
        return info, (st[:k] if record else None), (sn[:k] if record else None)

    def get_energy(self):
        ek = ctypes.c_double()
        ep = ctypes.c_double()
        check("h4h_get_energy", self.lib.h4h_get_energy(self._ctx, ctypes.byref(ek),
                                                        ctypes.byref(ep)))
        return ek.value, ep.value

    def synchronize(self):
        check("h4h_synchronize", self.lib.h4h_synchronize(self._ctx))

    def integration(self, t_end, interval=0.125, eta_s=ETA_DEFAULT, max_steps=None,
                    on_output=None, chunk_steps=65536, mode="run", checkpoint=None,
                    restart=None):
        """Hermite4CPU::integration's loop (Hermite4CPU_integration.cpp:38-130)
        with every kernel on the GPU; the block steps between two output
        times run on the device (h4h_run).  integration_time and
        interval_time are float as in the reference (NbodySystem.hpp:69-70)
        and the interval test ITIME >= interval * k is evaluated in float.
        mode="step" drives the same loop one h4h_step call per block step.
        checkpoint: path of a binary checkpoint (gravidy_amd.checkpoint)
        written at every output time and at the end; restart: continue from
        such a checkpoint bit for bit (same e2, eta, bounds, precision and
        interval) instead of starting at t = 0.
        Returns the schedule (ITIME and nact per block step), the energies
        at the start and at every output time, and energy_ini (the energy at
        t = 0, carried through restarts)."""
        if mode == "step":
            if checkpoint or restart:
                raise ValueError("checkpoints need mode='run'")
            return self._integration_steps(t_end, interval, eta_s, max_steps, on_output)
        t_end32 = float(np.float32(t_end))
        iv32 = np.float32(interval)
        steps_t, steps_n = [], []
        if restart:
            c = self._read_restart(restart, float(iv32))
            itime, atime = c["itime"], c["atime"]
            output_factor = c["output_factor"]
            interactions, nsteps, iterations = c["interactions"], c["nsteps"], c["iterations"]
            energy_ini = c["energy_ini"]
            ek, ep = self.get_energy()
            e_time, e_val = [itime], [ek + ep]
        else:
            itime = 0.0
            self.init_acc_jrk()
            atime = self.init_dt(1.0e10, eta_s, itime)
            ek, ep = self.get_energy()
            energy_ini = ek + ep
            e_time, e_val = [itime], [ek + ep]
            if on_output:
                on_output(self, itime)
            output_factor = 1
            interactions = nsteps = iterations = 0
        done = 0
        while itime < t_end32:
            budget = chunk_steps if max_steps is None else min(chunk_steps, max_steps - done)
            if budget <= 0:
                break
            t_stop = float(np.float32(iv32 * np.float32(output_factor)))
            info, st, sn = self.run(itime, atime, t_end32, t_stop, budget)
            steps_t.append(st)
            steps_n.append(sn)
            done += info["steps"]
            iterations += info["steps"]
            interactions += info["interactions"]
            nsteps += info["nsteps"]
            if info["steps"]:
                itime, atime = info["itime"], info["atime"]
            if info["stopped"] == _lib.H4H_STOP_OUTPUT:
                ek, ep = self.get_energy()
                e_time.append(itime)
                e_val.append(ek + ep)
                if on_output:
                    on_output(self, itime)
                output_factor += 1
                if checkpoint:
                    self._write_checkpoint(checkpoint, itime, atime, float(iv32), output_factor,
                                           energy_ini, iterations, nsteps, interactions)
            elif info["stopped"] == _lib.H4H_STOP_END:
                break
        if checkpoint:
            self._write_checkpoint(checkpoint, itime, atime, float(iv32), output_factor,
                                   energy_ini, iterations, nsteps, interactions)
        return {
            "step_time": np.concatenate(steps_t) if steps_t else np.zeros(0),
            "step_nact": (np.concatenate(steps_n) if steps_n else np.zeros(0)).astype(np.uint32),
            "energy_time": np.array(e_time), "energy": np.array(e_val),
            "energy_ini": energy_ini, "interactions": interactions,
        }

    def _write_checkpoint(self, path, itime, atime, interval, output_factor, energy_ini,
                          iterations, nsteps, interactions):
        from .checkpoint import write_checkpoint
        st = self.download()
        st.update({"n": self.n, "e2": self.e2, "eta": self.eta, "dt_min": self.dt_min,
                   "dt_max": self.dt_max, "precision": self.precision, "itime": itime,
                   "atime": atime, "interval": interval, "output_factor": output_factor,
                   "snapshot_number": output_factor, "energy_ini": energy_ini,
                   "iterations": iterations, "nsteps": nsteps, "interactions": interactions})
        write_checkpoint(path, st)

    def _read_restart(self, path, interval):
        from .checkpoint import CheckpointError, read_checkpoint
        c = read_checkpoint(path)
        for key, mine in (("n", self.n), ("e2", self.e2), ("eta", self.eta),
                          ("dt_min", self.dt_min), ("dt_max", self.dt_max),
                          ("precision", self.precision), ("interval", interval)):
            if c[key] != mine:
                raise CheckpointError(f"restart: the checkpoint's {key} ({c[key]!r}) differs "
                                      f"from this run's ({mine!r})")
        self.upload(c["r"], c["v"], c["f"], c["t"], c["dt"])
        return c

    def _integration_steps(self, t_end, interval, eta_s, max_steps, on_output):
        t_end32 = np.float32(t_end)
        iv32 = np.float32(interval)
        itime = 0.0
        self.init_acc_jrk()
        atime = self.init_dt(1.0e10, eta_s, itime)
        ek, ep = self.get_energy()
        e_time, e_val = [itime], [ek + ep]
        if on_output:
            on_output(self, itime)
        steps_t, steps_n = [], []
        output_factor = 1
        interactions = 0
        while itime < t_end32:
            if max_steps is not None and len(steps_t) >= max_steps:
                break
            itime = atime
            nact, atime = self.step(itime)
            interactions += nact * self.n
            steps_t.append(itime)
            steps_n.append(nact)
            if itime >= float(np.float32(iv32 * np.float32(output_factor))):
                ek, ep = self.get_energy()
                e_time.append(itime)
                e_val.append(ek + ep)
                if on_output:
                    on_output(self, itime)
                output_factor += 1
        return {
            "step_time": np.array(steps_t), "step_nact": np.array(steps_n, np.uint32),
            "energy_time": np.array(e_time), "energy": np.array(e_val),
            "interactions": interactions,
        }
