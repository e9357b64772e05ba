"""Batched ensembles: every member advanced by the persistent batch kernels
matches the same member advanced alone (single-simulation runner, whose step
is itself checked against the oracle), and a HIP-graph replay matches eager
steps."""

import numpy as np
import pytest
import torch

from oracle import pic_oracle as O
from pypic3d_amd import workloads
from pypic3d_amd.batch import EnsembleRunner
from pypic3d_amd.runner import ElectrodynamicRunner
from tests.gpu_helpers import TOL, host, host_particles, oracle_params, oracle_species, rel_err

pytestmark = pytest.mark.gpu


def _members(seed=3):
    specs = [((24, 1, 1), 20), ((16, 12, 1), 9), ((10, 1, 1), 33), ((6, 5, 4), 5), ((9, 7, 1), 17),
             ((20, 6, 1), 2)]
    out = []
    for m, (N_, ppc) in enumerate(specs):
        static, dyn = workloads.make_parameters(N_, cfl=0.5, name=f"m{m}")
        w = workloads.plasma_frequency_weight(2 * ppc, dyn.dt, 0.2)
        sp = [workloads.SpeciesSpec(ppc, weight=w, vth=(0.02,) * 3, drift=(0.15, 0, 0), perturb=0.01),
              workloads.SpeciesSpec(ppc, weight=w, vth=(0.02,) * 3, drift=(-0.15, 0, 0), perturb=0.01)]
        parts, cfg = workloads.plasma(static, dyn, sp, seed=seed + m)
        out.append((parts, cfg, workloads.empty_fields(static, dyn), static, dyn))
    return out


def _clone_member(mem):
    parts, cfg, fields, static, dyn = mem
    from pypic3d_amd.particles.particle_class import TiledParticles
    p2 = TiledParticles(x=parts.x.clone(), u=parts.u.clone(), active=parts.active.clone())
    f2 = tuple(tuple(c.clone() for c in f) if isinstance(f, tuple) and f and torch.is_tensor(f[0]) else f
               for f in fields)
    return p2, cfg, f2, static, dyn


@pytest.mark.parametrize("rounds,grid", [(4, 64), (2, 7), (8, 1000)])
def test_batch_matches_single_simulation_runs(rounds, grid):
    members = _members()
    batch = EnsembleRunner([_clone_member(m) for m in members], rounds=rounds, grid=grid)
    singles = [ElectrodynamicRunner(*_clone_member(m)) for m in members]
    for _ in range(3):
        batch.step()
        for r in singles:
            r.step()
    torch.cuda.synchronize()
    for s, r in enumerate(singles):
        bp, bE, bB, bJ = batch.member(s)
        sp_, sf = r.state()
        xb, ub, ab = host_particles(bp)
        xs, us, as_ = host_particles(sp_)
        np.testing.assert_array_equal(ab, as_)
        assert np.max(np.abs(xb - xs)) <= 1e-6 * max(float(members[s][4].x_wind), 1.0)
        assert rel_err(ub, us) <= TOL
        for c in range(3):
            assert rel_err(host(bJ[c]), host(sf[2][c])) <= TOL, (s, "J", c)
            assert rel_err(host(bE[c]), host(sf[0][c])) <= TOL, (s, "E", c)
            assert rel_err(host(bB[c]), host(sf[1][c])) <= TOL, (s, "B", c)


def test_batch_first_step_matches_oracle():
    members = _members(seed=8)
    batch = EnsembleRunner([_clone_member(m) for m in members])
    states = []
    for parts, cfg, fields, static, dyn in members:
        x, u, a = host_particles(parts)
        p = oracle_params(static, dyn)
        z = [np.zeros(p.shape) for _ in range(3)]
        st = O.State(x, u, a, z, [c.copy() for c in z])
        states.append(O.time_loop_electrodynamic(st, oracle_species(cfg), p))
    batch.step()
    torch.cuda.synchronize()
    for s, want in enumerate(states):
        bp, bE, bB, bJ = batch.member(s)
        xb, ub, ab = host_particles(bp)
        np.testing.assert_array_equal(ab, want.active)
        assert rel_err(ub, want.u) <= TOL
        for c in range(3):
            assert rel_err(host(bJ[c]), want.J[c]) <= TOL, (s, "J", c)
            assert rel_err(host(bE[c]), want.E[c]) <= TOL, (s, "E", c)


def test_graph_replays_match_eager_steps():
    members = _members(seed=11)
    eager = EnsembleRunner([_clone_member(m) for m in members])
    graphed = EnsembleRunner([_clone_member(m) for m in members])
    graphed.capture()          # advances two steps outside the graph first
    eager.step()
    eager.step()
    for rep in range(3):
        graphed.replay()
        eager.step()
        eager.step()
        torch.cuda.synchronize()
        for s in range(len(members)):
            gp, gE, gB, gJ = graphed.member(s)
            ep, eE, eB, eJ = eager.member(s)
            xg, ug, ag = host_particles(gp)
            xe, ue, ae = host_particles(ep)
            np.testing.assert_array_equal(ag, ae)
            assert rel_err(ug, ue) <= TOL, (rep, s)
            for c in range(3):
                assert rel_err(host(gE[c]), host(eE[c])) <= TOL, (rep, s, c)
