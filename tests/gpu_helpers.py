"""Converters between the oracle's NumPy layout and the GPU tensors."""

import numpy as np
import torch

from oracle import pic_oracle as O

TOL = 1.0e-5        # FP32 kernels vs the FP64 oracle, norm-wise relative error


def oracle_params(static, dyn):
    return O.Params(
        N=(int(dyn.Nx), int(dyn.Ny), int(dyn.Nz)), wind=(dyn.x_wind, dyn.y_wind, dyn.z_wind),
        d=(dyn.dx, dyn.dy, dyn.dz), g=static.guard_cells, dt=dyn.dt, C=dyn.C, eps=dyn.eps, mu=dyn.mu,
        alpha=dyn.alpha, shape_factor=static.shape_factor, relativistic=static.relativistic,
        particle_pusher=static.particle_pusher, current_deposition=static.current_deposition,
        current_filter=static.current_filter, bc=static.boundary_conditions,
        particle_bc=static.particle_boundary_conditions, electrostatic=static.electrostatic)


def oracle_species(cfg):
    return O.Species(np.asarray(cfg.charge, float), np.asarray(cfg.mass, float), np.asarray(cfg.weight, float),
                     np.asarray(cfg.update_x, bool), np.asarray(cfg.update_u, bool))


def gpu_field(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float32)).cuda().view(1, 1, 1, *a.shape)


def host(t):
    t = t.detach()
    if t.dim() == 6:
        t = t[0, 0, 0]
    return t.double().cpu().numpy()


def host_particles(parts):
    x = parts.x[0, 0, 0].double().cpu().numpy()
    u = parts.u[0, 0, 0].double().cpu().numpy()
    a = parts.active[0, 0, 0].cpu().numpy().astype(bool)
    return x, u, a


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)
