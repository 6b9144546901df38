"""Initial conditions: seeded Plummer spheres and the reference's IC file
format.

File format (NbodySystem::read_input_file, NbodySystem.cpp:79-207): one
particle per line, 8 whitespace-separated columns `id m x y z vx vy vz`;
lines whose first non-blank character is '#' are comments, and the first
comment line holding a "time" token followed by a value (e.g. the
snapshot header "# Time: 0.1250") sets the start time.  Masses are parsed
as float (`(float)strtod`, NbodySystem.cpp:174), positions and velocities
as double.  A data line without exactly 8 columns stops the read with an
error, as in the reference.

Plummer sampling follows Aarseth, Henon & Wielen (1974, A&A 37, 183):
radius from the inverse cumulative mass M(r) = r^3 (1+r^2)^(-3/2), speed
by von Neumann rejection on g(q) = q^2 (1-q^2)^(7/2), isotropic
directions; then scaled to Henon units (G = M = 1, E = -1/4) by r *= 3pi/16
and v /= sqrt(3pi/16), and shifted to the centre-of-mass frame.
"""
import math

import numpy as np

PLUMMER_SCALE = 3.0 * math.pi / 16.0


def _isotropic(rng, mag):
    n = mag.shape[0]
    u = rng.random(n)
    w = rng.random(n)
    z = (1.0 - 2.0 * u) * mag
    rxy = np.sqrt(np.maximum(mag * mag - z * z, 0.0))
    phi = 2.0 * math.pi * w
    return np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1)


def plummer(n, seed, rmax=10.0, masses=None, com=True):
    """n-body Plummer sphere in Henon units.

    seed: integer seed of numpy's PCG64 generator (bit-stable across
    platforms).  rmax: truncation radius in Plummer radii (radii are
    redrawn until r < rmax).  masses: optional (n,) masses (default 1/n);
    they are rounded to float32 as the reference stores them.
    Returns (r4, v4): (n,4) positions + mass, (n,4) velocities + 0.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    rng = np.random.Generator(np.random.PCG64(seed))
    rad = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x1 = rng.random(todo.size)
        with np.errstate(divide="ignore"):
            r = 1.0 / np.sqrt(np.power(x1, -2.0 / 3.0) - 1.0)
        ok = np.isfinite(r) & (r < rmax)
        rad[todo[ok]] = r[ok]
        todo = todo[~ok]
    pos = _isotropic(rng, rad)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.random(todo.size)
        x5 = rng.random(todo.size)
        ok = 0.1 * x5 < x4 * x4 * np.power(1.0 - x4 * x4, 3.5)
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    vesc = math.sqrt(2.0) * np.power(1.0 + rad * rad, -0.25)
    vel = _isotropic(rng, q * vesc)
    pos *= PLUMMER_SCALE
    vel /= math.sqrt(PLUMMER_SCALE)
    if masses is None:
        m = np.full(n, 1.0 / n)
    else:
        m = np.asarray(masses, dtype=np.float64)
        if m.shape != (n,):
            raise ValueError("masses must have shape (n,)")
    m = m.astype(np.float32).astype(np.float64)
    if com:
        mt = m.sum()
        pos -= (m[:, None] * pos).sum(axis=0) / mt
        vel -= (m[:, None] * vel).sum(axis=0) / mt
    r4 = np.concatenate([pos, m[:, None]], axis=1)
    v4 = np.concatenate([vel, np.zeros((n, 1))], axis=1)
    return r4, v4


def salpeter_masses(n, seed, mlo=0.1, mhi=100.0, alpha=2.35, total=1.0):
    """n masses with dN/dm ~ m^-alpha on [mlo, mhi] (inverse-CDF sampling),
    rescaled to sum to `total`, rounded to float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.random(n)
    a1 = 1.0 - alpha
    m = np.power(mlo ** a1 + u * (mhi ** a1 - mlo ** a1), 1.0 / a1)
    m *= total / m.sum()
    return m.astype(np.float32).astype(np.float64)


def plummer_salpeter_bh(n, seed, bh_mass=0.01, eps_rmax=10.0):
    """n Salpeter-mass Plummer stars (total mass 1) plus one central
    particle of mass bh_mass at rest at the origin: n+1 bodies."""
    m = salpeter_masses(n, seed + 1)
    r4, v4 = plummer(n, seed, rmax=eps_rmax, masses=m)
    bh_r = np.array([[0.0, 0.0, 0.0, float(np.float32(bh_mass))]])
    bh_v = np.zeros((1, 4))
    return np.concatenate([bh_r, r4]), np.concatenate([bh_v, v4])


def log_uniform_sizes(count, seed, nmin=1024, nmax=32768):
    """Cluster sizes N_k = round(exp(U[ln nmin, ln nmax]))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.random(count)
    return np.rint(np.exp(math.log(nmin) + u * (math.log(nmax) - math.log(nmin)))).astype(np.int64)


def write_ic(path, r4, v4, ids=None, time=None):
    """Writes the reference's 8-column format; doubles with %.17g so strtod
    round-trips them, masses with %.9g so (float)strtod round-trips."""
    n = r4.shape[0]
    ids = np.arange(n) if ids is None else ids
    with open(path, "w") as fh:
        if time is not None:
            fh.write(f"# Time: {time:.17g}\n")
        for i in range(n):
            fh.write("%d %.9g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                ids[i], np.float32(r4[i, 3]), r4[i, 0], r4[i, 1], r4[i, 2],
                v4[i, 0], v4[i, 1], v4[i, 2]))


class ICFormatError(ValueError):
    pass


def read_ic(path):
    """Reads an IC or snapshot file as NbodySystem::read_input_file does.
    Returns (ids, r4, v4, time) with time None if no "time" header."""
    ids, rows = [], []
    time = None
    with open(path) as fh:
        for line in fh:
            stripped = line.lstrip(" ")
            if stripped.startswith("#"):
                if time is None:
                    toks = stripped.replace("#", "").replace(":", "").lower().split()
                    for k, tok in enumerate(toks):
                        if tok == "time" and k + 1 < len(toks):
                            time = float(toks[k + 1])
                            break
                continue
            toks = line.split()
            if not toks:
                # an empty line has 0 columns: the reference stops there
                raise ICFormatError(f"{path}: wrong line format: only eight numbers per line are allowed")
            if len(toks) != 8:
                raise ICFormatError(f"{path}: wrong line format: only eight numbers per line are allowed")
            ids.append(int(toks[0]))
            rows.append([float(np.float32(float(toks[1])))] + [float(x) for x in toks[2:]])
    a = np.array(rows, dtype=np.float64).reshape(-1, 7)
    r4 = np.concatenate([a[:, 1:4], a[:, 0:1]], axis=1)
    v4 = np.concatenate([a[:, 4:7], np.zeros((a.shape[0], 1))], axis=1)
    return np.array(ids, dtype=np.int64), r4, v4, time
