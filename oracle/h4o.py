"""ctypes wrapper over oracle/libh4oracle.so (the CPU restatement of
GraviDy's Hermite4CPU FP64 path).  Test infrastructure only.

Array layouts are those documented at the top of h4_oracle.c:
r, v (n,4) float64; f (n,6); p (n,6); pm (n,) float32; a2, a3 (n,3).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libh4oracle.so")

# common.hpp:122,129
D_TIME_MIN = 1.1920928955078125e-07
D_TIME_MAX = 0.125
# common.hpp:114,119 -- ETA_S / ETA_N are float 0.01
ETA_DEFAULT = float(np.float32(0.01))


def softening_e2(eps):
    """e2 exactly as NbodySystem.cpp:50 computes it: float eps squared in
    float (op.softening is float, OptionsParser.hpp:63), stored as double."""
    e = np.float32(eps)
    return float(np.float32(e * e))


_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_u = ctypes.POINTER(ctypes.c_uint32)


class _Record(ctypes.Structure):
    _fields_ = [
        ("step_time", _d),
        ("step_nact", _u),
        ("energy_time", _d),
        ("energy", _d),
        ("max_steps", ctypes.c_uint64),
        ("max_energy", ctypes.c_uint32),
        ("n_energy", ctypes.c_uint32),
        ("interactions", ctypes.c_uint64),
        ("nsteps", ctypes.c_uint64),
    ]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise RuntimeError(f"oracle library {_LIB} is not built (run make -C oracle)")
    lib = ctypes.CDLL(_LIB)
    u32, dbl, flt = ctypes.c_uint32, ctypes.c_double, ctypes.c_float
    lib.h4o_force_calculation.argtypes = [_d, _d, flt, dbl, _d]
    lib.h4o_init_acc_jrk.argtypes = [u32, _d, _f, dbl, _d]
    lib.h4o_update_acc_jrk.argtypes = [u32, u32, _u, _d, _f, dbl, _d]
    lib.h4o_predicted_pos_vel.argtypes = [u32, dbl, _d, _d, _d, _d, _d, _f]
    lib.h4o_find_particles_to_move.argtypes = [u32, dbl, _d, _d, _u]
    lib.h4o_find_particles_to_move.restype = u32
    lib.h4o_next_integration_time.argtypes = [u32, _d, _d]
    lib.h4o_next_integration_time.restype = dbl
    lib.h4o_init_dt_value.argtypes = [dbl, dbl, dbl]
    lib.h4o_init_dt_value.restype = dbl
    lib.h4o_init_dt.argtypes = [u32, _d, flt, dbl, _d, _d, _d, dbl, dbl]
    lib.h4o_save_old_acc_jrk.argtypes = [u32, _u, _d, _d]
    lib.h4o_get_timestep_normal.argtypes = [_d, _d, dbl, _d, flt]
    lib.h4o_get_timestep_normal.restype = dbl
    lib.h4o_normalize_dt.argtypes = [dbl, dbl, dbl, dbl, dbl]
    lib.h4o_normalize_dt.restype = dbl
    lib.h4o_correction_pos_vel.argtypes = [u32, dbl, u32, _u, _d, _d, _d, _d, _d,
                                           _d, _d, _d, _d, flt, dbl, dbl]
    lib.h4o_get_energy.argtypes = [u32, _d, _d, _d, _d]
    lib.h4o_integrate.argtypes = [u32, _d, _d, _d, _d, _d, _d, _d, dbl, flt, flt,
                                  flt, flt, dbl, dbl, ctypes.POINTER(_Record)]
    lib.h4o_integrate.restype = ctypes.c_int64
    lib.h4o_integrate_at.argtypes = [u32, _d, _d, _d, _d, _d, _d, _d, dbl, flt, flt, dbl,
                                     flt, flt, dbl, dbl, ctypes.POINTER(_Record)]
    lib.h4o_integrate_at.restype = ctypes.c_int64
    lib.h4o_block_steps.argtypes = [u32, _d, _d, _d, _d, _d, _d, _d, dbl, flt, _d,
                                    ctypes.c_uint64, dbl, dbl, ctypes.POINTER(_Record)]
    lib.h4o_block_steps.restype = ctypes.c_int64
    lib.h4o_num_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _c(a, dtype, shape=None):
    a = np.ascontiguousarray(a, dtype=dtype)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class Oracle:
    """Function-per-method view of the oracle; every method mirrors one
    reference method (see the citations in h4_oracle.c)."""

    def __init__(self):
        self.lib = load()

    def num_threads(self):
        return int(self.lib.h4o_num_threads())

    def force_calculation(self, pi, pj, mj, e2, fi=None):
        pi = _c(pi, np.float64, (6,))
        pj = _c(pj, np.float64, (6,))
        fi = np.zeros(6) if fi is None else _c(fi, np.float64, (6,)).copy()
        self.lib.h4o_force_calculation(_p(pi, ctypes.c_double), _p(pj, ctypes.c_double),
                                       float(np.float32(mj)), float(e2), _p(fi, ctypes.c_double))
        return fi

    def init_acc_jrk(self, p, pm, e2):
        n = p.shape[0]
        p = _c(p, np.float64, (n, 6))
        pm = _c(pm, np.float32, (n,))
        f = np.zeros((n, 6))
        self.lib.h4o_init_acc_jrk(n, _p(p, ctypes.c_double), _p(pm, ctypes.c_float),
                                  float(e2), _p(f, ctypes.c_double))
        return f

    def update_acc_jrk(self, move, p, pm, e2, f):
        n = p.shape[0]
        move = _c(move, np.uint32)
        p = _c(p, np.float64, (n, 6))
        pm = _c(pm, np.float32, (n,))
        f = _c(f, np.float64, (n, 6)).copy()
        self.lib.h4o_update_acc_jrk(n, move.size, _p(move, ctypes.c_uint32),
                                    _p(p, ctypes.c_double), _p(pm, ctypes.c_float),
                                    float(e2), _p(f, ctypes.c_double))
        return f

    def predicted_pos_vel(self, itime, t, r, v, f):
        n = r.shape[0]
        t = _c(t, np.float64, (n,))
        r = _c(r, np.float64, (n, 4))
        v = _c(v, np.float64, (n, 4))
        f = _c(f, np.float64, (n, 6))
        p = np.zeros((n, 6))
        pm = np.zeros(n, np.float32)
        self.lib.h4o_predicted_pos_vel(n, float(itime), _p(t, ctypes.c_double),
                                       _p(r, ctypes.c_double), _p(v, ctypes.c_double),
                                       _p(f, ctypes.c_double), _p(p, ctypes.c_double),
                                       _p(pm, ctypes.c_float))
        return p, pm

    def find_particles_to_move(self, itime, t, dt):
        n = t.shape[0]
        t = _c(t, np.float64, (n,))
        dt = _c(dt, np.float64, (n,))
        move = np.zeros(max(n, 1), np.uint32)
        k = self.lib.h4o_find_particles_to_move(n, float(itime), _p(t, ctypes.c_double),
                                                _p(dt, ctypes.c_double), _p(move, ctypes.c_uint32))
        return move[:k].copy()

    def next_integration_time(self, t, dt):
        n = t.shape[0]
        t = _c(t, np.float64, (n,))
        dt = _c(dt, np.float64, (n,))
        return float(self.lib.h4o_next_integration_time(n, _p(t, ctypes.c_double),
                                                         _p(dt, ctypes.c_double)))

    def init_dt_value(self, x, dt_min=D_TIME_MIN, dt_max=D_TIME_MAX):
        return float(self.lib.h4o_init_dt_value(float(x), float(dt_min), float(dt_max)))

    def init_dt(self, f, eta=ETA_DEFAULT, itime=0.0, ctime=1.0e10,
                dt_min=D_TIME_MIN, dt_max=D_TIME_MAX):
        n = f.shape[0]
        f = _c(f, np.float64, (n, 6))
        t = np.zeros(n)
        dt = np.zeros(n)
        c = ctypes.c_double(ctime)
        self.lib.h4o_init_dt(n, ctypes.byref(c), float(np.float32(eta)), float(itime),
                             _p(f, ctypes.c_double), _p(t, ctypes.c_double),
                             _p(dt, ctypes.c_double), float(dt_min), float(dt_max))
        return t, dt, c.value

    def get_timestep_normal(self, a2i, a3i, dti, fi, eta=ETA_DEFAULT):
        a2i = _c(a2i, np.float64, (3,))
        a3i = _c(a3i, np.float64, (3,))
        fi = _c(fi, np.float64, (6,))
        return float(self.lib.h4o_get_timestep_normal(_p(a2i, ctypes.c_double),
                                                       _p(a3i, ctypes.c_double), float(dti),
                                                       _p(fi, ctypes.c_double),
                                                       float(np.float32(eta))))

    def normalize_dt(self, new_dt, old_dt, t, dt_min=D_TIME_MIN, dt_max=D_TIME_MAX):
        return float(self.lib.h4o_normalize_dt(float(new_dt), float(old_dt), float(t),
                                                float(dt_min), float(dt_max)))

    def correction_pos_vel(self, itime, move, p, f, old, a2, a3, r, v, t, dt,
                           eta=ETA_DEFAULT, dt_min=D_TIME_MIN, dt_max=D_TIME_MAX):
        """Returns updated copies (a2, a3, r, v, t, dt)."""
        n = r.shape[0]
        move = _c(move, np.uint32)
        p = _c(p, np.float64, (n, 6))
        f = _c(f, np.float64, (n, 6))
        old = _c(old, np.float64, (n, 6))
        a2 = _c(a2, np.float64, (n, 3)).copy()
        a3 = _c(a3, np.float64, (n, 3)).copy()
        r = _c(r, np.float64, (n, 4)).copy()
        v = _c(v, np.float64, (n, 4)).copy()
        t = _c(t, np.float64, (n,)).copy()
        dt = _c(dt, np.float64, (n,)).copy()
        D = ctypes.c_double
        self.lib.h4o_correction_pos_vel(n, float(itime), move.size, _p(move, ctypes.c_uint32),
                                        _p(p, D), _p(f, D), _p(old, D), _p(a2, D), _p(a3, D),
                                        _p(r, D), _p(v, D), _p(t, D), _p(dt, D),
                                        float(np.float32(eta)), float(dt_min), float(dt_max))
        return a2, a3, r, v, t, dt

    def get_energy(self, r, v):
        n = r.shape[0]
        r = _c(r, np.float64, (n, 4))
        v = _c(v, np.float64, (n, 4))
        ek = ctypes.c_double()
        ep = ctypes.c_double()
        self.lib.h4o_get_energy(n, _p(r, ctypes.c_double), _p(v, ctypes.c_double),
                                ctypes.byref(ek), ctypes.byref(ep))
        return ek.value, ep.value

    def integrate(self, r, v, e2, t_end, interval=0.125, eta_s=ETA_DEFAULT,
                  eta_n=ETA_DEFAULT, dt_min=D_TIME_MIN, dt_max=D_TIME_MAX,
                  max_steps=10_000_000, max_energy=4096, itime0=0.0):
        """Runs Hermite4CPU::integration (Hermite4CPU_integration.cpp:38-130)
        from ITIME = itime0 (0, or the start time of a resumed run).
        Returns a dict with the final state and the recorded schedule."""
        n = r.shape[0]
        r = _c(r, np.float64, (n, 4)).copy()
        v = _c(v, np.float64, (n, 4)).copy()
        f = np.zeros((n, 6))
        t = np.zeros(n)
        dt = np.zeros(n)
        a2 = np.zeros((n, 3))
        a3 = np.zeros((n, 3))
        step_time = np.zeros(max_steps)
        step_nact = np.zeros(max_steps, np.uint32)
        e_time = np.zeros(max_energy)
        e_val = np.zeros(max_energy)
        rec = _Record(_p(step_time, ctypes.c_double), _p(step_nact, ctypes.c_uint32),
                      _p(e_time, ctypes.c_double), _p(e_val, ctypes.c_double),
                      max_steps, max_energy, 0, 0, 0)
        D = ctypes.c_double
        steps = self.lib.h4o_integrate_at(n, _p(r, D), _p(v, D), _p(f, D), _p(t, D), _p(dt, D),
                                          _p(a2, D), _p(a3, D), float(e2),
                                          float(np.float32(eta_s)), float(np.float32(eta_n)),
                                          float(itime0), float(np.float32(t_end)),
                                          float(np.float32(interval)), float(dt_min),
                                          float(dt_max), ctypes.byref(rec))
        if steps < 0:
            raise RuntimeError(f"oracle integration exceeded max_steps={max_steps}")
        return {
            "r": r, "v": v, "f": f, "t": t, "dt": dt, "a2": a2, "a3": a3,
            "step_time": step_time[:steps].copy(), "step_nact": step_nact[:steps].copy(),
            "energy_time": e_time[:rec.n_energy].copy(), "energy": e_val[:rec.n_energy].copy(),
            "interactions": int(rec.interactions), "nsteps": int(rec.nsteps),
        }

    def block_steps(self, state, e2, atime, steps, eta_n=ETA_DEFAULT, dt_min=D_TIME_MIN,
                    dt_max=D_TIME_MAX):
        """`steps` block steps of Hermite4CPU::integration's loop body
        (Hermite4CPU_integration.cpp:83-99) from state = dict of r, v (n,4),
        f (n,6), t, dt (n,) and a2, a3 (n,3), updated in place, starting at
        ITIME = atime.  Returns (next block time, step times, step nacts,
        interactions)."""
        n = state["r"].shape[0]
        for key, shape in (("r", (n, 4)), ("v", (n, 4)), ("f", (n, 6)), ("t", (n,)),
                           ("dt", (n,)), ("a2", (n, 3)), ("a3", (n, 3))):
            a = state.get(key)
            if a is None:
                state[key] = np.zeros(shape)
            elif a.shape != shape or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"{key}: need a C-contiguous float64 array of shape {shape}")
        step_time = np.zeros(max(1, steps))
        step_nact = np.zeros(max(1, steps), np.uint32)
        rec = _Record(_p(step_time, ctypes.c_double), _p(step_nact, ctypes.c_uint32),
                      None, None, steps, 0, 0, 0, 0)
        D = ctypes.c_double
        at = ctypes.c_double(atime)
        k = self.lib.h4o_block_steps(n, _p(state["r"], D), _p(state["v"], D), _p(state["f"], D),
                                     _p(state["t"], D), _p(state["dt"], D), _p(state["a2"], D),
                                     _p(state["a3"], D), float(e2), float(np.float32(eta_n)),
                                     ctypes.byref(at), int(steps), float(dt_min), float(dt_max),
                                     ctypes.byref(rec))
        return at.value, step_time[:k].copy(), step_nact[:k].copy(), int(rec.interactions)
