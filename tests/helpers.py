"""Shared test helpers: seeded synthetic inputs in the reference's one-tile layout."""

import numpy as np

from oracle import pic_oracle as O


def fp32(a):
    """Round to FP32 and back, so the FP64 oracle sees exactly what the GPU sees."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def random_particles(p: O.Params, S=2, P=64, seed=0, vmax=0.3, inactive_frac=0.0, margin=0.0):
    """Uniform positions inside the domain (optionally kept `margin` cells from
    the faces), velocities with |v_i| < vmax * d_i / dt (sub-cell moves)."""
    rng = np.random.default_rng(seed)
    x = np.zeros((S, P, 3))
    u = np.zeros((S, P, 3))
    for a in range(3):
        half = p.wind[a] / 2
        lo, hi = -half + margin * p.d[a], half - margin * p.d[a]
        x[..., a] = rng.uniform(lo, hi, size=(S, P)) if p.N[a] > 1 else rng.uniform(-0.4, 0.4, (S, P)) * p.d[a]
        u[..., a] = rng.uniform(-vmax, vmax, size=(S, P)) * p.d[a] / p.dt
    active = rng.random((S, P)) >= inactive_frac
    return fp32(x), fp32(u), active


def species(S=2, charge=(-1.0, 1.0), mass=(1.0, 1836.0), weight=(0.5, 0.5), update_x=True, update_u=True):
    return O.Species(
        charge=np.asarray(charge[:S], dtype=np.float64),
        mass=np.asarray(mass[:S], dtype=np.float64),
        weight=np.asarray(weight[:S], dtype=np.float64),
        update_x=np.full((S, 3), update_x, dtype=bool),
        update_u=np.full((S, 3), update_u, dtype=bool),
    )


def random_field(p: O.Params, seed, scale=1.0, refresh=True):
    rng = np.random.default_rng(seed)
    f = fp32(rng.standard_normal(p.shape) * scale)
    return O.refresh(f, p) if refresh else f


def interior(f, g):
    return f[g:-g, g:-g, g:-g]


def rel_err(a, b):
    """Norm-wise relative error max|a-b| / max|b| (fields contain zeros)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)
