"""Batched ensembles of independent simulations (uneven grids and particle counts).

``EnsembleRunner`` advances every member with the semantics of
``time_loop_electrodynamic``'s Esirkepov branch (evolve.py:16-103), using one
persistent launch per phase for the whole ensemble:

1. ``pic_fill_f32`` zeroes the batch J arrays;
2. ``pic_batch_fused_push_esirkepov`` walks a work list of (member, block)
   items with a fixed grid of persistent blocks (gather + push + Esirkepov +
   position update + boundaries, the single-simulation block function);
3. ``pic_batch_field_step`` runs one block per member for the J fold/refresh
   and the three Yee passes.

``capture`` records one step into a HIP graph (torch.cuda.CUDAGraph) that
``replay`` launches; the ops allocate nothing and use no memset, so the graph
replays the same three kernels.  ``autotune`` picks the particles per block
(256 x rounds) and the persistent grid size from timed steps.

Members must have periodic field and particle boundaries, alpha = 1 and no
external fields (the two-stream ensembles this targets).
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _native as N
from ._tensors import as_field3, make_desc, make_species, particle_planes, pic_particles, soa_view
from .particles.particle_class import TiledParticles


class PicBatchSim(ctypes.Structure):
    _fields_ = [("desc", N.PicDesc), ("species", N.PicSpecies), ("particle_offset", ctypes.c_int64),
                ("field_offset", ctypes.c_int64)]


THREADS = 256


class EnsembleRunner:
    def __init__(self, members, rounds=4, grid=None, device="cuda"):
        """members: list of (particles, species_config, fields, static, dynamic)
        with the reference's argument conventions (fields is the 8-tuple)."""
        if not members:
            raise ValueError("empty ensemble")
        self.device = torch.device(device)
        self.n = len(members)
        table = (PicBatchSim * self.n)()
        p_off, f_off = 0, 0
        self.layout = []
        for s, (parts, cfg, fields, static, dyn) in enumerate(members):
            if static.current_deposition != "esirkepov" or static.electrostatic:
                raise NotImplementedError("ensembles run the electrodynamic Esirkepov step")
            if fields[6] is not None:
                raise NotImplementedError("ensembles run without PML (pml_state must be None)")
            desc = make_desc(static, dyn)
            xp, up, act = particle_planes(parts)
            S, P = xp[0].shape
            sp = make_species(cfg, P)
            ctypes.memmove(ctypes.addressof(table[s].desc), ctypes.addressof(desc), ctypes.sizeof(N.PicDesc))
            ctypes.memmove(ctypes.addressof(table[s].species), ctypes.addressof(sp), ctypes.sizeof(N.PicSpecies))
            table[s].particle_offset = p_off
            table[s].field_offset = f_off
            nf = desc.dims[0] * desc.dims[1] * desc.dims[2]
            self.layout.append((p_off, S, P, f_off, tuple(desc.dims), desc))
            p_off += S * P
            f_off += nf
        self.slots, self.felems = p_off, f_off
        rc = N.lib().pic_batch_check(ctypes.addressof(table), self.n, self.slots, self.felems)
        N.check(rc, "pic_batch_check (periodic boundaries, alpha = 1, consistent slices are required)")
        self._table = table
        raw = bytes(bytearray(table))
        self.table = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
        dev = self.device
        self.x = [torch.empty((3, self.slots), dtype=torch.float32, device=dev) for _ in range(2)]
        self.u = [torch.empty((3, self.slots), dtype=torch.float32, device=dev) for _ in range(2)]
        self.act = [torch.zeros(self.slots, dtype=torch.uint8, device=dev) for _ in range(2)]
        self.E = [torch.zeros((3, self.felems), dtype=torch.float32, device=dev) for _ in range(2)]
        self.B = torch.zeros((3, self.felems), dtype=torch.float32, device=dev)
        self.Bh = torch.zeros_like(self.B)
        self.J = torch.zeros_like(self.B)
        for s, (parts, cfg, fields, static, dyn) in enumerate(members):
            po, S, P, fo, dims, desc = self.layout[s]
            xp, up, act = particle_planes(parts)
            nf = dims[0] * dims[1] * dims[2]
            for a in range(3):
                self.x[0][a, po:po + S * P].copy_(xp[a].reshape(-1))
                self.u[0][a, po:po + S * P].copy_(up[a].reshape(-1))
                self.E[0][a, fo:fo + nf].copy_(as_field3(fields[0][a], "E").reshape(-1))
                self.B[a, fo:fo + nf].copy_(as_field3(fields[1][a], "B").reshape(-1))
            self.act[0][po:po + S * P].copy_(act.reshape(-1).to(torch.uint8))
        self.cur = 0
        self.graph = None
        props = torch.cuda.get_device_properties(self.device)
        self.cus = props.multi_processor_count
        self.set_launch(rounds, grid if grid is not None else 4 * self.cus)

    # ----------------------------------------------------------- launch shape
    def set_launch(self, rounds, grid):
        if rounds not in (2, 4, 8):
            raise ValueError("rounds must be 2, 4 or 8")
        self.rounds, self.grid = int(rounds), int(grid)
        per = THREADS * self.rounds
        items = []
        for s, (po, S, P, fo, dims, desc) in enumerate(self.layout):
            for b in range(math.ceil(S * P / per)):
                items.append((s, b))
        self.work = torch.tensor(items, dtype=torch.int32, device=self.device).reshape(-1)
        self.n_work = len(items)
        self.graph = None

    # ----------------------------------------------------------------- step
    def _pp(self, k):
        return pic_particles([self.x[k][0], self.x[k][1], self.x[k][2]], [self.u[k][0], self.u[k][1], self.u[k][2]],
                             self.act[k])

    def step(self):
        s = N.stream_handle()
        src, dst = self.cur, 1 - self.cur
        N.call("pic_fill_f32", N.ptr(self.J), self.J.numel(), 0.0, s)
        tab = ctypes.c_void_p(self.table.data_ptr())
        N.call("pic_batch_fused_push_esirkepov", tab, N.ptr(self.work), self.n_work, self.rounds, self.grid,
               ctypes.byref(self._pp(src)), N.p3(self.E[src]), N.p3(self.B),
               ctypes.byref(self._pp(dst)), N.p3(self.J), s)
        N.call("pic_batch_field_step", tab, self.n, N.p3(self.E[src]), N.p3(self.B), N.p3(self.Bh), N.p3(self.J),
               N.p3(self.E[dst]), s)
        self.cur = dst

    def capture(self):
        """Record two steps (one per ping-pong parity) into a HIP graph."""
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            self.step()
            self.step()          # warm both parities outside the graph
        torch.cuda.current_stream(self.device).wait_stream(stream)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        cur0 = self.cur
        with torch.cuda.graph(g):
            self.step()
            self.step()
        self.cur = cur0
        self.graph = g
        return g

    def replay(self):
        """Two steps from the captured graph."""
        if self.graph is None:
            self.capture()
        self.graph.replay()

    # ------------------------------------------------------------ autotune
    def autotune(self, rounds=(2, 4, 8), grids=None, reps=3):
        """Fastest (rounds, grid) over timed eager steps; the steps advance the
        ensemble (their results are the ensemble's next states)."""
        grids = grids or (2 * self.cus, 4 * self.cus, 8 * self.cus)
        times = {}
        for r in rounds:
            for gsz in grids:
                self.set_launch(r, gsz)
                self.step()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(reps):
                    self.step()
                b.record()
                torch.cuda.synchronize()
                times[(r, gsz)] = a.elapsed_time(b) / reps
        best = min(times, key=times.get)
        self.set_launch(*best)
        return best, times

    # --------------------------------------------------------------- views
    def member(self, s):
        """(TiledParticles, E, B, J) views of member s's current state."""
        po, S, P, fo, dims, desc = self.layout[s]
        k = self.cur
        xs = self.x[k][:, po:po + S * P].reshape(3, S, P)
        us = self.u[k][:, po:po + S * P].reshape(3, S, P)
        act = self.act[k][po:po + S * P].reshape(S, P).bool()
        nf = dims[0] * dims[1] * dims[2]
        f = lambda t: tuple(t[c, fo:fo + nf].reshape(dims) for c in range(3))   # noqa: E731
        parts = TiledParticles(x=soa_view(xs), u=soa_view(us), active=act.view(1, 1, 1, S, P))
        return parts, f(self.E[k]), f(self.B), f(self.J)

    @property
    def particles(self):
        return sum(S * P for (_, S, P, _, _, _) in self.layout)
