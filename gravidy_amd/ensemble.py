"""Batched ensembles: many independent clusters of uneven size integrated
together in one context (h4h_create_ensemble / h4h_ensemble_*).  Each
cluster follows the reference's own loop (Hermite4CPU_integration.cpp:
38-130) with its own block times; the device advances every live cluster
by one block step per step launch."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, PRECISIONS
from .hermite4 import D_TIME_MAX, D_TIME_MIN, ETA_DEFAULT


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class EnsembleHIP:
    """systems: list of (r4, v4) per cluster ((n_c,4) each)."""

    def __init__(self, systems, e2, eta=ETA_DEFAULT, precision="fp64",
                 dt_min=D_TIME_MIN, dt_max=D_TIME_MAX):
        self.lib = _lib.load()
        self.sizes = np.array([r.shape[0] for r, _ in systems], dtype=np.uint32)
        self.K = len(systems)
        self.bounds = np.concatenate([[0], np.cumsum(self.sizes.astype(np.int64))])
        r = np.ascontiguousarray(np.concatenate([np.asarray(r, float) for r, _ in systems]))
        v = np.concatenate([np.asarray(v, float) for _, v in systems])
        if v.shape[1] == 3:
            v = np.concatenate([v, np.zeros((v.shape[0], 1))], axis=1)
        v = np.ascontiguousarray(v)
        self.n = r.shape[0]
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._ctx = ctypes.c_void_p()
        check("h4h_create_ensemble", self.lib.h4h_create_ensemble(
            ctypes.byref(self._ctx), self.K,
            self.sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), float(e2),
            float(np.float32(eta)), float(dt_min), float(dt_max), prec))
        check("h4h_upload", self.lib.h4h_upload(self._ctx, _dp(r), _dp(v), None, None, None))

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

    def set_small_path(self, max_active=256):
        """When every cluster has at most 1024 particles, each cluster makes
        its block steps with 1..max_active (at most 256; the default for an
        ensemble) active particles in a workgroup of its own (0: every step
        on the multi-workgroup kernels)."""
        check("h4h_set_small_path", self.lib.h4h_set_small_path(self._ctx, int(max_active)))

    def set_launch(self, workgroups=0, max_jsplit=0):
        check("h4h_set_launch", self.lib.h4h_set_launch(self._ctx, workgroups, max_jsplit))

    def stats(self, reset=False):
        s = _lib.H4HStats()
        check("h4h_get_stats", self.lib.h4h_get_stats(self._ctx, ctypes.byref(s), int(reset)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def init(self, eta_s=ETA_DEFAULT):
        check("h4h_ensemble_init", self.lib.h4h_ensemble_init(self._ctx, float(np.float32(eta_s))))

    def run(self, t_end, t_stop, max_steps=1 << 62):
        info = (_lib.H4HClusterInfo * self.K)()
        check("h4h_ensemble_run", self.lib.h4h_ensemble_run(
            self._ctx, float(t_end), float(t_stop), int(max_steps), info))
        return [{k: getattr(x, k) for k, _ in x._fields_} for x in info]

    def energy(self):
        ek = np.zeros(self.K)
        ep = np.zeros(self.K)
        check("h4h_ensemble_energy", self.lib.h4h_ensemble_energy(self._ctx, _dp(ek), _dp(ep)))
        return ek, ep

    def download(self):
        n = self.n
        r = np.zeros((n, 4))
        v = np.zeros((n, 4))
        f = np.zeros((n, 6))
        t = np.zeros(n)
        dt = np.zeros(n)
        check("h4h_download", self.lib.h4h_download(self._ctx, _dp(r), _dp(v), _dp(f), _dp(t),
                                                    _dp(dt), None, None, None))
        out = []
        for c in range(self.K):
            a, b = self.bounds[c], self.bounds[c + 1]
            out.append({"r": r[a:b], "v": v[a:b], "f": f[a:b], "t": t[a:b], "dt": dt[a:b]})
        return out

    def integration(self, t_end, interval=0.125, eta_s=ETA_DEFAULT):
        """Every cluster runs the reference's loop to t_end with outputs at
        interval * k (float arithmetic as there).  Returns per-cluster
        energies at t = 0 and at each output time, step counts,
        interactions and the block time of each cluster's last step."""
        t_end32 = float(np.float32(t_end))
        iv32 = np.float32(interval)
        self.init(eta_s)
        ek, ep = self.energy()
        energies = [ek + ep]
        times = [0.0]
        steps = np.zeros(self.K, np.int64)
        inter = np.zeros(self.K, np.int64)
        itime = np.zeros(self.K)
        k = 1
        while True:
            t_stop = float(np.float32(iv32 * np.float32(k)))
            info = self.run(t_end32, t_stop)
            steps += [x["steps"] for x in info]
            inter += [x["interactions"] for x in info]
            itime = np.array([x["itime"] if x["steps"] else itime[c] for c, x in enumerate(info)])
            if all(x["steps"] == 0 for x in info):
                break
            if any(x["stopped"] == _lib.H4H_STOP_OUTPUT for x in info):
                ek, ep = self.energy()
                energies.append(ek + ep)
                times.append(max(x["itime"] for x in info))
                k += 1
            if all(x["itime"] >= t_end32 for x in info):
                break
        return {"energy": np.array(energies), "energy_time": np.array(times),
                "steps": steps, "interactions": inter, "itime": itime}
