"""Allocation-free stepping of time_loop_electrodynamic for long runs.

``ElectrodynamicRunner`` performs exactly the call sequence of
``evolve.time_loop_electrodynamic`` (Esirkepov or direct branch) on
preallocated ping-pong buffers: particle SoA planes alternate between two
sets, E alternates between two buffers and B returns to its own buffer after
the two half steps.  Optional cell re-sorting every ``sort_every`` steps keeps
particles cache-local; with the tile-ordered fused kernels (exact grids) the
re-bin stops at its permutation and the next step's fused kernel reads
the particles through it (the permutation is applied on its own only when
the particles are read before that step, through ``cur`` or ``state()``).
Per-phase HIP-event timing is available for the
benchmark (events are recorded on the stream the kernels run on).
"""

from __future__ import annotations

import torch

from . import _native as N
from ._tensors import as_field3, check_fields, make_desc, make_species, particle_planes, pic_particles, soa_view
from .deposition.Esirkepov import finish_current
from .deposition.J_from_rhov import deposit_direct_into, finish_direct_current
from .particles.particle_class import TiledParticles
from .solvers.first_order_yee import _check_pml, _pml_args, yee_b, yee_e
from .sorting import key_bits
from .utils import is_all_zero


class ElectrodynamicRunner:
    PHASES = ("particles", "current_fold", "yee")

    def __init__(self, particles, species_config, fields, static_parameters, dynamic_parameters, sort_every=0,
                 tile=None, rounds=4, min_blocks=1):
        """tile (4 or 8): keep particles in tile-major cell order (sorted once
        here, then every ``sort_every`` steps) and use the tile-ordered kernels:
        the fused push + Esirkepov kernel (3-D first order); in the
        J_from_rhov branch the fused push + direct-deposit kernel on exact 3-D
        first-order grids, else the tiled direct deposit."""
        E, B, J, rho, phi, external_fields, pml_state, overflow = fields
        _check_pml(pml_state)
        if static_parameters.particle_pusher not in ("boris", "higuera_cary"):
            raise ValueError(f"Unknown particle_pusher: {static_parameters.particle_pusher}")
        self.static, self.dynamic = static_parameters, dynamic_parameters
        self.desc = make_desc(static_parameters, dynamic_parameters)
        self.esirkepov = static_parameters.current_deposition == "esirkepov"
        self.filter = static_parameters.current_filter
        xp, up, act = particle_planes(particles)
        self.S, self.P = xp[0].shape
        self.sp = make_species(species_config, self.P)
        dev = xp[0].device
        self.device = dev
        self.pos = [torch.empty((3, self.S, self.P), dtype=torch.float32, device=dev) for _ in range(2)]
        self.vel = [torch.empty((3, self.S, self.P), dtype=torch.float32, device=dev) for _ in range(2)]
        self.act = [torch.empty((self.S, self.P), dtype=torch.bool, device=dev) for _ in range(2)]
        for a in range(3):
            self.pos[0][a].copy_(xp[a])
            self.vel[0][a].copy_(up[a])
        self.act[0].copy_(act)
        self._cur = 0
        # a tile re-bin whose permutation the next fused step applies while it
        # reads (self.perm; the particles still sit in the set self._cur)
        self._pending = False
        self.E = [[as_field3(c, "E").clone() for c in E], [as_field3(c, "E").clone() for c in E]]
        self.B = [as_field3(c, "B").clone() for c in B]
        self.Bt = [b.clone() for b in self.B]
        self.J = [as_field3(c, "J").clone() for c in J]
        self.scratch = [b.clone() for b in self.B]
        self.ecur = 0
        ext_E = [as_field3(c, "external E") for c in external_fields[0]]
        ext_B = [as_field3(c, "external B") for c in external_fields[1]]
        self._ext = (ext_E, ext_B)
        # the runner's kernels are the FP32 ones (float64 states step through evolve.time_loop_electrodynamic)
        if xp[0].dtype != torch.float32:
            raise ValueError("ElectrodynamicRunner runs float32 particles and fields")
        check_fields(self.desc, *self.E[0], *self.B, *self.J, *ext_E, *ext_B, dtype=torch.float32)
        self.Ee = None if all(is_all_zero(c) for c in ext_E) else N.p3(ext_E)
        self.Be = None if all(is_all_zero(c) for c in ext_B) else N.p3(ext_B)
        self.pml_b = self.pml_e = None
        if pml_state is not None:      # the runner advances its own copies of the PML memories in place
            b_mem, prof3, b_new = _pml_args(self.desc, pml_state, 1)
            e_mem, _, e_new = _pml_args(self.desc, pml_state, 0)
            self.pml_b, self.pml_e = (b_mem, prof3), (e_mem, prof3)
            pml_state = (e_new, b_new, pml_state[2])
        self.rest = (rho, phi, external_fields, pml_state, overflow)
        self.sort_every = int(sort_every)
        self.steps_done = 0
        self.events = None
        self.keys = None
        self.sorts_done = 0
        self.tile, self.rounds, self.min_blocks, self.starts = None, int(rounds), int(min_blocks), None
        # launch shape of the fused direct kernel: 4 resident blocks per CU
        # (two slots per thread) at tile 4, one at tile 8
        self.direct_min_blocks = 1
        self.sort_lookahead = None      # tile re-sort keys at x + t u; None: t = sort_every * dt / 2
        self.tile_ws, self.direct_fused = None, False
        self.defer_permute = False
        if tile is not None:
            self.set_tile(tile)

    # -------------------------------------------------------------- helpers
    @property
    def cur(self):
        """Index of the particle set that holds the current particles (in
        their sorted order: a deferred permutation is applied first)."""
        self._materialize()
        return self._cur

    def _materialize(self):
        if self._pending:
            self._pending = False
            src, dst = self._cur, 1 - self._cur
            N.call("pic_permute_particles", self._pp(src), self._pp(dst), N.ptr(self.perm), self.S * self.P,
                   N.stream_handle())
            self._cur = dst

    def _pp(self, k):
        return pic_particles([self.pos[k][0], self.pos[k][1], self.pos[k][2]],
                             [self.vel[k][0], self.vel[k][1], self.vel[k][2]], self.act[k])

    def enable_timing(self, on=True):
        self.events = {ph: [] for ph in self.PHASES} if on else None

    def _mark(self, phase):
        if self.events is None:
            return None
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record()
        self.events[phase].append(ev)
        return ev

    def phase_ms(self):
        """Mean duration (ms) of each phase over the recorded steps."""
        out = {}
        for ph, evs in (self.events or {}).items():
            if evs:
                out[ph] = sum(a.elapsed_time(b) for a, b in evs) / len(evs)
        return out

    # ----------------------------------------------------------------- step
    def step(self):
        d, s = self.desc, N.stream_handle()
        src, dst = self._cur, 1 - self._cur
        # the fused tile kernels read through a pending permutation; anything else
        # (or a switch to the strided comparison kernel since the re-bin) applies it first
        if self._pending and not (self.tile and (self.direct_fused or (self.esirkepov and d.variant != 1))):
            self._materialize()
            src, dst = self._cur, 1 - self._cur
        E, Et = self.E[self.ecur], self.E[1 - self.ecur]
        if self.esirkepov:
            for c in self.J:
                N.call("pic_fill_f32", N.ptr(c), c.numel(), 0.0, s)
            ev = self._mark("particles")
            if self.tile and self._pending:
                N.call("pic_fused_push_esirkepov_tiled_gather", d, self.sp, self._pp(src), N.ptr(self.perm), N.p3(E),
                       N.p3(self.B), self.Ee, self.Be, self._pp(dst), N.p3(self.J), self.tile, N.ptr(self.starts),
                       self.rounds, self.min_blocks, N.ptr(self.tile_ws), s)
                self._pending = False
            elif self.tile:
                N.call("pic_fused_push_esirkepov_tiled", d, self.sp, self._pp(src), N.p3(E), N.p3(self.B), self.Ee,
                       self.Be, self._pp(dst), N.p3(self.J), self.tile, N.ptr(self.starts), self.rounds,
                       self.min_blocks, N.ptr(self.tile_ws), s)
            else:
                N.call("pic_fused_push_esirkepov", d, self.sp, self._pp(src), N.p3(E), N.p3(self.B), self.Ee,
                       self.Be, self._pp(dst), N.p3(self.J), s)
            if ev:
                ev[1].record()
            ev = self._mark("current_fold")
            finish_current(d, self.J, s)
        elif self.tile and self.direct_fused:
            for c in self.J:
                N.call("pic_fill_f32", N.ptr(c), c.numel(), 0.0, s)
            ev = self._mark("particles")
            if self._pending:
                N.call("pic_fused_push_direct_tiled_gather", d, self.sp, self._pp(src), N.ptr(self.perm), N.p3(E),
                       N.p3(self.B), self.Ee, self.Be, self._pp(dst), N.p3(self.J), self.tile, N.ptr(self.starts),
                       self.direct_min_blocks, N.ptr(self.tile_ws), s)
                self._pending = False
            else:
                N.call("pic_fused_push_direct_tiled", d, self.sp, self._pp(src), N.p3(E), N.p3(self.B), self.Ee,
                       self.Be, self._pp(dst), N.p3(self.J), self.tile, N.ptr(self.starts), self.direct_min_blocks,
                       N.ptr(self.tile_ws), s)
            if ev:
                ev[1].record()
            ev = self._mark("current_fold")
            finish_direct_current(d, self.J, self.filter, s, scratch=self.scratch)
        else:
            ev = self._mark("particles")
            uo = [self.vel[dst][0], self.vel[dst][1], self.vel[dst][2]]
            N.call("pic_push", d, self.sp, self._pp(src), N.p3(E), N.p3(self.B), self.Ee, self.Be, N.p3(uo), s)
            self.act[dst].copy_(self.act[src])
            half = pic_particles([self.pos[src][0], self.pos[src][1], self.pos[src][2]], uo, self.act[dst])
            xd = [self.pos[dst][0], self.pos[dst][1], self.pos[dst][2]]
            N.call("pic_advance_positions", d, self.sp, half, 0.5, N.p3(xd), s)
            N.call("pic_particle_boundaries", d, self.sp, self._pp(dst), s)
            if ev:
                ev[1].record()
            ev = self._mark("current_fold")
            deposit_direct_into(d, self.sp, self._pp(dst), self.J, self.filter, s, scratch=self.scratch,
                                tile=self.tile, starts=self.starts)
            N.call("pic_advance_positions", d, self.sp, self._pp(dst), 0.5, N.p3(xd), s)
            N.call("pic_particle_boundaries", d, self.sp, self._pp(dst), s)
        if ev:
            ev[1].record()
        ev = self._mark("yee")
        yee_b(d, E, self.B, self.Bt, False, s, scratch=self.scratch, pml=self.pml_b)
        yee_e(d, E, self.Bt, self.J, Et, s, scratch=self.scratch, pml=self.pml_e)
        yee_b(d, Et, self.Bt, self.B, True, s, scratch=self.scratch, pml=self.pml_b)
        if ev:
            ev[1].record()
        self._cur = dst
        self.ecur = 1 - self.ecur
        self.steps_done += 1
        if self.sort_every and self.steps_done % self.sort_every == 0:
            self.sort()

    def autotune(self, reps=3):
        """Pick the fastest launch shape of the fused kernel on the current
        state: the slot-ordered kernel's variants, or in tile mode the tiled
        kernel's slots per thread.  The timed launches write the inactive
        particle set and J, which the next step overwrites."""
        from .autotune import tune_fused, tune_tiled
        if not self.esirkepov:
            return None, {}
        src, dst = self.cur, 1 - self.cur
        E = self.E[self.ecur]
        if self.tile:
            best, times = tune_tiled(self.desc, self.sp, self._pp(src), E, self.B, self.Ee, self.Be, self._pp(dst),
                                     self.J, self.tile, self.starts, self.tile_ws, reps=reps)
            self.rounds, self.min_blocks = best
            return best, times
        best, times = tune_fused(self.desc, self.sp, self._pp(src), E, self.B, self.Ee, self.Be, self._pp(dst),
                                 self.J, reps=reps)
        self.desc.variant = best
        return best, times

    def _sort_buffers(self):
        n = self.S * self.P
        dev = self.device
        if getattr(self, "perm", None) is None:
            self.perm = torch.empty(n, dtype=torch.int32, device=dev)
        if self.tile:
            nbytes = int(N.lib().pic_tile_rebin_workspace_bytes(self.desc, self.sp, self.tile))
            if getattr(self, "rebin_ws", None) is None or self.rebin_ws.numel() < nbytes:
                self.rebin_ws = torch.empty(max(1, nbytes), dtype=torch.uint8, device=dev)
        elif getattr(self, "keys", None) is None:
            self.keys = torch.empty(n, dtype=torch.int32, device=dev)
            self.keys_sorted = torch.empty(n, dtype=torch.int32, device=dev)
            self.sort_ws = torch.empty(max(1, int(N.lib().pic_sort_workspace_bytes(n))), dtype=torch.uint8,
                                       device=dev)

    def set_tile(self, tile):
        """Switch to the tile-ordered kernel with `tile`^3-cell tiles (None:
        the slot-ordered kernel) and sort the particles accordingly."""
        self._materialize()
        if tile is None:
            self.tile, self.starts, self.tile_ws, self.direct_fused = None, None, None, False
            self.defer_permute = False
            return
        d = self.desc
        if self.esirkepov and not (d.shape_factor == 1 and all(d.n[a] > 1 for a in range(3))):
            raise ValueError("the tile-ordered Esirkepov kernel needs the 3-D first-order path")
        if tile not in (4, 8):
            raise ValueError("tile must be 4 or 8")
        self.tile = int(tile)
        self.direct_min_blocks = 4 if self.tile == 4 else 1
        # direct branch: one fused tile kernel on exact 3-D first-order grids with cells equal to the node
        # spacing, the separate push / advance / boundary / tiled-deposit kernels otherwise
        self.direct_fused = (not self.esirkepov and d.shape_factor == 1 and all(d.n[a] > 1 for a in range(3))
                             and bool(N.lib().pic_tile_exact_grid(d))
                             and all(d.d[a] == d.sp_c[a] for a in range(3)))
        # a re-bin leaves its gather to the next step's fused kernel, which reads
        # the particles through perm (both branches' fused kernels on exact grids)
        self.defer_permute = self.direct_fused or (self.esirkepov and d.variant != 1
                                                   and bool(N.lib().pic_tile_exact_grid(d)))
        tiles = int(N.lib().pic_tile_count(d, self.tile))
        self.starts = torch.empty(self.S * (tiles + 1) + 1, dtype=torch.int32, device=self.device)
        # the tiled kernel's list of particles for its general route (1 + slots int32)
        self.tile_ws = torch.zeros(1 + self.S * self.P, dtype=torch.int32, device=self.device)
        self.sort()

    def sort(self):
        """Re-order the current particle buffers (into the other set): in
        tile mode a counting re-bin into tile-key order (the order inside one
        cell is unspecified; results do not depend on it), otherwise the
        stable sort into the reference's cell order.  perm[i] is the previous
        slot of slot i.  With defer_permute the re-bin stops at perm and the
        next step (or the first access to ``cur``) applies it."""
        self._sort_buffers()
        s = N.stream_handle()
        src, dst = self.cur, 1 - self.cur
        n = self.S * self.P
        if self.tile:
            # keys at the position half a re-sort interval ahead: the order is then
            # freshest in the middle of the interval instead of at its start
            ahead = self.sort_lookahead
            if ahead is None:
                ahead = 0.5 * self.sort_every * float(self.desc.dt)
            defer = self.defer_permute
            N.call("pic_tile_rebin", self.desc, self.sp, self._pp(src), self.tile, ahead,
                   None if defer else self._pp(dst), N.ptr(self.starts), N.ptr(self.perm), N.ptr(self.rebin_ws),
                   self.rebin_ws.numel(), s)
            if defer:
                self._pending = True
                self.sorts_done += 1
                return
        else:
            ncell = self.desc.n[0] * self.desc.n[1] * self.desc.n[2]
            N.call("pic_cell_keys", self.desc, self.sp, self._pp(src), N.ptr(self.keys), s)
            N.call("pic_sort_permutation", N.ptr(self.keys), N.ptr(self.keys_sorted), N.ptr(self.perm), n,
                   key_bits(self.S, ncell), N.ptr(self.sort_ws), self.sort_ws.numel(), s)
            N.call("pic_permute_particles", self._pp(src), self._pp(dst), N.ptr(self.perm), n, s)
        self._cur = dst
        self.sorts_done += 1

    def state(self):
        """(TiledParticles, fields 8-tuple) views of the current buffers."""
        k = self.cur
        parts = TiledParticles(x=soa_view(self.pos[k]), u=soa_view(self.vel[k]),
                               active=self.act[k].view(1, 1, 1, self.S, self.P))
        v6 = lambda t: t.view(1, 1, 1, *t.shape)   # noqa: E731
        rho, phi, ext, pml, ovf = self.rest
        fields = (tuple(v6(c) for c in self.E[self.ecur]), tuple(v6(c) for c in self.B), tuple(v6(c) for c in self.J),
                  rho, phi, ext, pml, ovf)
        return parts, fields
