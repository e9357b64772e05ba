"""Reverse-mode gradients of time_loop_electrodynamic (Esirkepov branch).

The reference has no gradient code; this module differentiates the step it
defines (evolve.py:16-103) with hand-written adjoint kernels
(``pic_adjoint_yee``, ``pic_adjoint_particles``) behind
``torch.autograd.Function``:

* ``AdjointRun`` advances the allocation-free runner from (x0, v0, E0, B0),
  keeping a checkpoint every ``checkpoint_every`` steps, and evaluates the
  loss = final field energy U_E + U_B in FP64 (utils.py:108-187).
* ``AdjointRun.backward`` walks the segments backwards: it recomputes the
  states of one segment from its checkpoint, then applies the adjoint step to
  each state in reverse order.  Memory is one checkpoint per segment plus one
  segment of states.
* ``FieldEnergyLoss`` wraps both so that torch autograd chains the gradient
  into whatever produced x0 and v0 (drift velocity, perturbation amplitude,
  per-particle momenta; see ``two_stream_state``).

Scope: periodic field and particle boundaries, no PML, the Esirkepov current
(the branch ``current_deposition == "esirkepov"``), any pusher, shape factor
and reduced axes.  Adjoint quantities are FP64; the forward state is the FP32
state of the forward kernels.
"""

from __future__ import annotations

import ctypes

import math

import torch

from . import _native as N
from ._tensors import pic_particles, soa_view
from .particles.particle_class import TiledParticles
from .runner import ElectrodynamicRunner
from .utils import compute_energy


def _check_supported(static, dyn):
    if static.current_deposition != "esirkepov":
        raise NotImplementedError("the adjoint covers the Esirkepov branch of time_loop_electrodynamic")
    if static.electrostatic:
        raise NotImplementedError("the adjoint covers the electrodynamic step")


class AdjointRun:
    """Forward with checkpoints, then reverse sweep with recomputation.
    min_blocks is passed through to pic_adjoint_particles, whose first-order
    kernels choose their own occupancy (the argument stays for callers that
    set it)."""

    def __init__(self, static_parameters, dynamic_parameters, species_config, active, external_fields=None,
                 checkpoint_every=20, min_blocks=8):
        _check_supported(static_parameters, dynamic_parameters)
        self.static, self.dynamic = static_parameters, dynamic_parameters
        self.cfg = species_config
        self.active = active
        self.external = external_fields
        self.k = max(1, int(checkpoint_every))
        self.min_blocks = int(min_blocks)
        self.runner = None
        self.steps = 0
        self.generation = 0      # counts forward runs; a loss's backward needs the run it came from
        self.final = None

    # ------------------------------------------------------------ helpers
    def _fields(self, E0, B0):
        from .workloads import empty_fields
        f = list(empty_fields(self.static, self.dynamic, device=E0[0].device))
        f[0] = tuple(e.reshape(f[0][0].shape).float() for e in E0)
        f[1] = tuple(b.reshape(f[1][0].shape).float() for b in B0)
        if self.external is not None:
            f[5] = self.external
        return tuple(f)

    def _snapshot(self):
        r = self.runner
        k = r.cur
        return (r.pos[k].clone(), r.vel[k].clone(), r.act[k].clone(), [c.clone() for c in r.E[r.ecur]],
                [c.clone() for c in r.B])

    def _restore(self, snap):
        r = self.runner
        pos, vel, act, E, B = snap
        k = r.cur
        r.pos[k].copy_(pos)
        r.vel[k].copy_(vel)
        r.act[k].copy_(act)
        for c in range(3):
            r.E[r.ecur][c].copy_(E[c])
            r.B[c].copy_(B[c])

    def _interior(self, t):
        g = self.runner.desc.g
        n = self.runner.desc.n
        return t[g:g + n[0], g:g + n[1], g:g + n[2]]

    # ------------------------------------------------------------ forward
    def forward(self, x0, v0, E0, B0, steps):
        """x0, v0: (3, S, P); E0, B0: 3 ghosted components.  Returns the FP64
        loss tensor U_E + U_B after `steps` steps."""
        S, P = x0.shape[1], x0.shape[2]
        parts = TiledParticles(x=soa_view(x0.float().contiguous()), u=soa_view(v0.float().contiguous()),
                               active=self.active.reshape(1, 1, 1, S, P))
        self.runner = ElectrodynamicRunner(parts, self.cfg, self._fields(E0, B0), self.static, self.dynamic,
                                           sort_every=0)
        d = self.runner.desc
        if any(d.bc[a] != 0 or d.pbc[a] != 0 for a in range(3)):
            raise NotImplementedError("the adjoint covers periodic field and particle boundaries")
        self.steps = int(steps)
        self.ckpt = {}
        for n in range(self.steps):
            if n % self.k == 0:
                self.ckpt[n] = self._snapshot()
            self.runner.step()
        self.generation += 1
        # the final fields seed the adjoint; kept apart from the runner, whose
        # buffers the reverse sweep's recomputation overwrites
        r = self.runner
        self.final = ([c.clone() for c in r.E[r.ecur]], [c.clone() for c in r.B])
        return self.loss()

    def loss(self):
        parts, fields = self.runner.state()
        ue, ub, _ = compute_energy(parts, fields[0], fields[1], self.static, self.dynamic, self.cfg)
        return ue + ub

    # ----------------------------------------------------------- backward
    def backward(self):
        """Returns (lam_x0, lam_v0, lam_E0, lam_B0): FP64 gradients of the loss
        with respect to x0, v0 (3, S, P) and the interior of E0, B0 (3, Nx, Ny, Nz)."""
        r = self.runner
        d = r.desc
        dev = r.device
        s = N.stream_handle()
        dV = d.d[0] * d.d[1] * d.d[2]
        E_end, B_end = self.final
        lamE = torch.stack([self._interior(c).double() * (d.eps * dV) for c in E_end]).contiguous()
        lamB = torch.stack([self._interior(c).double() * (dV / d.mu) for c in B_end]).contiguous()
        lamJ = torch.empty_like(lamE)
        work = torch.empty((6,) + tuple(lamE.shape[1:]), dtype=torch.float64, device=dev)
        lx = torch.zeros((3, r.S, r.P), dtype=torch.float64, device=dev)
        lv = torch.zeros_like(lx)
        defer = torch.empty(1 + r.S * r.P, dtype=torch.int32, device=dev)
        scratch = torch.empty(7 * r.S * r.P, dtype=torch.float64, device=dev)
        p3 = lambda t: N.P3(*[t[c].data_ptr() for c in range(3)])   # noqa: E731
        w6 = (N.c_vp * 6)(*[work[c].data_ptr() for c in range(6)])
        LE, LB, LJ, LX, LV = p3(lamE), p3(lamB), p3(lamJ), p3(lx), p3(lv)
        # one segment of recomputed states, allocated once
        k = min(self.k, self.steps)
        gshape = tuple(r.J[0].shape)
        seg_x = torch.empty((k, 3, r.S, r.P), dtype=torch.float32, device=dev)
        seg_v = torch.empty_like(seg_x)
        # periodic particle boundaries never change the active set: one copy serves every step
        seg_a = r.act[r.cur].clone()
        seg_E = torch.empty((k, 3) + gshape, dtype=torch.float32, device=dev)
        seg_B = torch.empty_like(seg_E)
        # C-ABI arguments of every segment slot, built once (the reverse sweep is
        # launch-bound at small sizes)
        lib = N.lib()
        seg_pp = [ctypes.byref(pic_particles([seg_x[m, 0], seg_x[m, 1], seg_x[m, 2]],
                                             [seg_v[m, 0], seg_v[m, 1], seg_v[m, 2]], seg_a)) for m in range(k)]
        seg_pE = [p3(seg_E[m]) for m in range(k)]
        seg_pB = [p3(seg_B[m]) for m in range(k)]
        dref, spref, dfr, scr = ctypes.byref(d), ctypes.byref(r.sp), N.ptr(defer), N.ptr(scratch)
        adj_yee, adj_part = lib.pic_adjoint_yee, lib.pic_adjoint_particles
        starts = sorted(self.ckpt)
        for si in reversed(range(len(starts))):
            n0 = starts[si]
            n1 = starts[si + 1] if si + 1 < len(starts) else self.steps
            self._restore(self.ckpt[n0])
            for m in range(n1 - n0):
                seg_x[m].copy_(r.pos[r.cur])
                seg_v[m].copy_(r.vel[r.cur])
                torch.stack(r.E[r.ecur], out=seg_E[m])
                torch.stack(r.B, out=seg_B[m])
                r.step()
            for m in reversed(range(n1 - n0)):
                N.check(adj_yee(dref, LE, LB, LJ, w6, s), "pic_adjoint_yee")
                N.check(adj_part(dref, spref, seg_pp[m], seg_pE[m], seg_pB[m], r.Ee, r.Be, LJ, LX, LV, LE, LB,
                                 self.min_blocks, dfr, scr, s), "pic_adjoint_particles")
        return lx, lv, lamE, lamB


class FieldEnergyLoss(torch.autograd.Function):
    """loss = U_E + U_B after `steps` electrodynamic steps from (x0, v0, E0, B0);
    E0 and B0 are (3, Nx+2g, Ny+2g, Nz+2g) stacks with refreshed ghosts, and
    their gradient is returned on the interior (ghost entries zero)."""

    @staticmethod
    def forward(ctx, x0, v0, E0, B0, run: AdjointRun, steps: int):
        loss = run.forward(x0.detach(), v0.detach(), tuple(E0.detach()), tuple(B0.detach()), steps)
        ctx.run = run
        ctx.generation = run.generation
        ctx.meta = (x0.dtype, v0.dtype, E0.dtype, B0.dtype, E0.shape)
        return loss.detach().clone()

    @staticmethod
    def backward(ctx, grad):
        if ctx.run.generation != ctx.generation:
            raise RuntimeError("this AdjointRun was run forward again after the loss was computed; its "
                               "checkpoints now belong to the later run (use one AdjointRun per loss)")
        lx, lv, lE, lB = ctx.run.backward()
        xd, vd, ed, bd, gshape = ctx.meta
        g = grad.to(torch.float64)
        d = ctx.run.runner.desc
        gi = [slice(d.g, d.g + d.n[a]) for a in range(3)]
        gE = torch.zeros(gshape, dtype=torch.float64, device=lE.device)
        gB = torch.zeros(gshape, dtype=torch.float64, device=lB.device)
        gE[:, gi[0], gi[1], gi[2]] = lE
        gB[:, gi[0], gi[1], gi[2]] = lB
        return (g * lx).to(xd), (g * lv).to(vd), (g * gE).to(ed), (g * gB).to(bd), None, None


def field_energy_loss(x0, v0, E0, B0, run: AdjointRun, steps: int):
    return FieldEnergyLoss.apply(x0, v0, E0, B0, run, steps)


# ------------------------------------------------------------------ inputs
def _inverse_cdf(u, W, amp, m):
    """x' in [0, W) with (x' + amp sin(k x') / k) = u, k = 2 pi m / W (Newton)."""
    k = 2 * math.pi * m / W
    y = u.clone()
    for _ in range(12):
        y = y - (y + (amp / k) * torch.sin(k * y) - u) / (1 + amp * torch.cos(k * y))
    return y


def two_stream_state(static, dynamic, ppc, drift, amp, vth=0.01, seed=0, mode=1, device="cuda"):
    """Two counter-streaming beams (+drift, -drift along x), `ppc` particles per
    cell each, Maxwellian spread vth, density 1 + amp cos(2 pi mode x / W) along x.
    drift and amp may be 0-d tensors that require grad: the returned FP64
    x0, v0 (3, 2, P) carry their derivatives (positions through the implicit
    derivative of the inverse CDF, dx/damp = -(sin(k x)/k) / (1 + amp cos(k x)))."""
    N_ = (int(dynamic.Nx), int(dynamic.Ny), int(dynamic.Nz))
    d = (float(dynamic.dx), float(dynamic.dy), float(dynamic.dz))
    W = (float(dynamic.x_wind), float(dynamic.y_wind), float(dynamic.z_wind))
    drift = torch.as_tensor(drift, dtype=torch.float64, device=device)
    amp = torch.as_tensor(amp, dtype=torch.float64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    ncell = N_[0] * N_[1] * N_[2]
    P = ppc * ncell
    cells = torch.arange(ncell, device=device, dtype=torch.int64)
    ci = (cells // (N_[1] * N_[2]), (cells // N_[2]) % N_[1], cells % N_[2])
    xs, vs = [], []
    for beam in range(2):
        comps_x, comps_v = [], []
        for a in range(3):
            base = ci[a].repeat_interleave(ppc).to(torch.float64)
            frac = torch.rand(P, generator=gen, device=device, dtype=torch.float64)
            if N_[a] == 1:
                pos = (frac - 0.5) * d[a] * 0.8
            else:
                u = (base + frac) * d[a]
                if a == 0:
                    k = 2 * math.pi * mode / W[0]
                    y = _inverse_cdf(u, W[0], amp.detach(), mode)
                    dy = -(torch.sin(k * y) / k) / (1 + amp.detach() * torch.cos(k * y))
                    u = y + dy * (amp - amp.detach())
                pos = u - W[a] / 2
            noise = vth * torch.randn(P, generator=gen, device=device, dtype=torch.float64)
            vel = noise + (drift if beam == 0 else -drift) if a == 0 else noise
            comps_x.append(pos)
            comps_v.append(vel)
        xs.append(torch.stack(comps_x))
        vs.append(torch.stack(comps_v))
    x0 = torch.stack(xs, dim=1)
    v0 = torch.stack(vs, dim=1)
    return x0, v0
