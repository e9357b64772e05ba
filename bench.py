"""Throughput of the PIC hot path (push + Esirkepov + Yee) on one MI355X.

Workload (default): periodic 256^3 grid, warm isotropic electron plasma at
vth = 0.05c, 32 particles per cell (536,870,912 particles, SoA FP32), first
order (CIC) Esirkepov current, relativistic Boris, Yee B/2 -> E -> B/2,
omega_p dt = 0.1, cfl 0.5.  A step is one time_loop_electrodynamic call
(evolve.py:16-103) executed by the allocation-free runner on tile-sorted
particles (4^3-cell tiles by default), with a tile re-bin every
--sort-every steps (default 6; counted in the timed region).

Prints ONE JSON line.  value = particle-steps/s over the timed steps.
roofline: the fused particle kernel (gather + Boris + Esirkepov + position
update + boundaries), algorithmic bytes 50 per particle (read x,y,z,vx,vy,vz
and the active byte, write them back) + 36 per cell (E, B gather reads and the
J write), divided by the mean duration of the particle phase from HIP events
recorded on the stream the kernel runs on (the phase is the fused kernel plus
its rare-list kernel, about 0.15 ms; the launch after each re-bin reads the
particles through the re-bin's permutation, 4 more bytes per particle, and
that permutation is never applied on its own).  traffic: calibrated FETCH_SIZE +
WRITE_SIZE bytes per launch from the kernel's committed rocprofv3 summary in
profiles/ for this workload (null when there is none).
cpu_baseline: the FP64 NumPy oracle (oracle/) on a bounded sample of the same
workload, its particles split over one worker process per host core.
two_stream_fp64: the 64-cell FP64 two-stream configuration, 500 steps, on
the GPU's FP64 path and on the oracle (one core).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s (push+Esirkepov+Yee) 256³·32ppc; HBM GB/s; adjoint/fwd cost"


def measured_traffic(kernel, n, ppc):
    """HBM bytes per launch of `kernel` on this workload from its committed
    rocprofv3 summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{n}cubed_{ppc}ppc.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1].split()[0]), os.path.relpath(path, ROOT)
    return None, None


_POOL_STATE = {}


def _oracle_chunk(bounds):
    """One worker's share of an oracle step: push, raw Esirkepov deposit,
    position update and particle boundaries for slots [lo, hi) (the inputs
    are inherited from the parent at fork); returns the raw J and a checksum
    of the new positions."""
    from oracle import pic_oracle as O
    lo, hi = bounds
    st, sp, p = _POOL_STATE["st"], _POOL_STATE["sp"], _POOL_STATE["p"]
    x, u, act = st.x[:, lo:hi], st.u[:, lo:hi], st.active[:, lo:hi]
    un = O.particle_push(x, u, act, sp, st.E, st.B, p)
    J = O.esirkepov_raw(x, un, act, sp, p)
    xn = O.update_positions(x, un, act, sp, p.dt)
    xn, un, act, _ = O.refresh_particles(xn, un, act, p)
    return J, float(xn.sum())


def cpu_baseline(n=48, ppc=32):
    """The FP64 NumPy oracle on the host's cores: one electrodynamic
    Esirkepov step of an n^3 x ppc warm plasma (the benchmark workload at a
    bounded size), the particles split over one worker process per core
    (push, deposit, position update, boundaries), the J partial sums folded
    and the Yee updates done in the parent."""
    import multiprocessing as mp
    import numpy as np
    from oracle import pic_oracle as O
    cores = min(16, len(os.sched_getaffinity(0)))
    p = O.Params(N=(n, n, n), wind=(float(n),) * 3, g=2, dt=0.5 / 3, current_deposition="esirkepov")
    rng = np.random.default_rng(20261202)
    P = n ** 3 * ppc
    x = np.stack([rng.uniform(-n / 2, n / 2, (1, P)) for _ in range(3)], -1).astype(np.float32).astype(np.float64)
    u = (rng.standard_normal((1, P, 3)) * 0.05).astype(np.float32).astype(np.float64)
    sp = O.Species(np.array([-1.0]), np.array([1.0]), np.array([0.36 / 32]), np.ones((1, 3), bool),
                   np.ones((1, 3), bool))
    E = [O.refresh((rng.standard_normal(p.shape) * 0.01), p) for _ in range(3)]
    B = [O.refresh((rng.standard_normal(p.shape) * 0.01), p) for _ in range(3)]
    _POOL_STATE.update(st=O.State(x, u, np.ones((1, P), bool), E, B), sp=sp, p=p)
    edges = np.linspace(0, P, cores + 1).astype(int)
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        t0 = time.perf_counter()
        parts = pool.map(_oracle_chunk, list(zip(edges[:-1], edges[1:])))
        J = [sum(pj[0][c] for pj in parts) for c in range(3)]
        J = [O.refresh(O.fold(c, p, p.particle_bc), p, p.particle_bc) for c in J]
        B1 = O.update_B(E, B, p)
        E1 = O.update_E(E, B1, J, p)
        O.update_B(E1, B1, p, do_filter=True)
        dt = time.perf_counter() - t0
    _POOL_STATE.clear()
    return {"value": P / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": f"FP64 NumPy oracle, one electrodynamic Esirkepov step of a {n}^3 x {ppc} ppc warm plasma "
                      f"({P} particles) split over {cores} worker processes, {dt:.1f} s"}


def two_stream_fp64(steps=500, oracle=True):
    """The 64-cell two-stream configuration in FP64 for `steps` steps: the
    GPU's FP64 path (time_loop_electrodynamic on float64 tensors) and the FP64
    NumPy oracle on one core, as particle-steps/s."""
    import numpy as np
    import torch
    from oracle import pic_oracle as O
    from pypic3d_amd import workloads
    from pypic3d_amd.evolve import time_loop_electrodynamic
    static, dyn, specs = workloads.config_two_stream_1d()
    parts, cfg = workloads.plasma(static, dyn, specs, seed=20261202, dtype=torch.float64)
    fields = workloads.empty_fields(static, dyn, dtype=torch.float64)
    n_part = int(parts.active.numel())
    x = parts.x[0, 0, 0].cpu().numpy()
    u = parts.u[0, 0, 0].cpu().numpy()
    a = parts.active[0, 0, 0].cpu().numpy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        parts, fields = time_loop_electrodynamic(parts, cfg, fields, static, dyn)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    out = {"steps": steps, "particles": n_part, "gpu_ms_per_step": 1e3 * gpu_s / steps,
           "gpu_particle_steps_per_s": n_part * steps / gpu_s}
    if oracle:
        p = O.Params(N=(64, 1, 1), wind=(float(dyn.x_wind), float(dyn.y_wind), float(dyn.z_wind)), g=2,
                     dt=float(dyn.dt))
        sp = O.Species(np.asarray(cfg.charge), np.asarray(cfg.mass), np.asarray(cfg.weight),
                       np.asarray(cfg.update_x), np.asarray(cfg.update_u))
        z = [np.zeros(p.shape) for _ in range(3)]
        st = O.State(x, u, a, z, [c.copy() for c in z])
        t0 = time.perf_counter()
        for _ in range(steps):
            st = O.time_loop_electrodynamic(st, sp, p)
        cpu_s = time.perf_counter() - t0
        out.update(oracle_cores=1, oracle_particle_steps_per_s=n_part * steps / cpu_s,
                   oracle_s=cpu_s)
    return out


def _timed(fn, reps):
    import torch
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def extras(n, ppc, steps, seed):
    """The other deposition paths of the same workload: time_loop_electrodynamic
    with J_from_rhov (filter "none" and "bilinear"), compute_rho, and the
    electrostatic CG per iteration (FP32 vectors, FP64 dots)."""
    import torch
    from pypic3d_amd import _native as N
    from pypic3d_amd import workloads
    from pypic3d_amd.deposition.rho import deposit_rho_into
    from pypic3d_amd.runner import ElectrodynamicRunner
    from pypic3d_amd.solvers.electrostatic_yee import last_solve_info, poisson_cg_into

    out = {}
    n_cells = n ** 3
    for filt in ("none", "bilinear"):
        static, dyn, species = workloads.config_warm_3d(n=n, ppc=ppc, vth=0.05, current_deposition="j_from_rhov")
        static = static._replace(current_filter=filt)
        parts, cfg = workloads.plasma(static, dyn, species, seed=seed)
        r = ElectrodynamicRunner(parts, cfg, workloads.empty_fields(static, dyn), static, dyn, sort_every=4, tile=4)
        del parts
        ms = _timed(r.step, steps)
        npart = r.S * r.P
        r.enable_timing(True)
        for _ in range(2):
            r.step()
        torch.cuda.synchronize()
        ph = r.phase_ms()
        r.enable_timing(False)
        # particle phase (the fused direct tile kernel on this exact grid): particle planes read and
        # written (50 B/particle), E and B gathered (24 B/cell), J written (12 B/cell)
        out[f"j_from_rhov_{filt}"] = {"ms_per_step": ms, "particle_steps_per_s": npart / (ms * 1e-3),
                                      "fused_kernel": bool(r.direct_fused), "phases_ms": ph,
                                      "particle_phase_algorithmic_gbs":
                                          (50 * npart + 36 * n_cells) / (ph["particles"] * 1e-3) / 1e9}
        if filt == "none":
            rho = torch.empty_like(r.J[0])
            ms_rho = _timed(lambda: deposit_rho_into(r.desc, r.sp, r._pp(r.cur), rho, "none", N.stream_handle(),
                                                     tile=r.tile, starts=r.starts), 3)
            out["rho"] = {"ms": ms_rho, "particles_per_s": npart / (ms_rho * 1e-3),
                          "algorithmic_bytes": 13 * npart + 4 * n_cells}
            phi = torch.zeros_like(rho)
            iters = 50

            def cg():
                phi.zero_()
                poisson_cg_into(r.desc, rho, phi, 0.0, iters, N.stream_handle(), rtol=0.0)
            ms_cg = _timed(cg, 2)
            k = last_solve_info["iterations"] or iters
            out["poisson_cg"] = {"iterations": k, "ms_per_iteration": ms_cg / k,
                                 "hbm_gbs_at_40B_per_cell_iter": 40 * n_cells * k / (ms_cg * 1e-3) / 1e9}
        del r
        torch.cuda.empty_cache()
    out["yee_pml"] = yee_pml(n)
    return out


def yee_pml(n, layers=16, reps=20):
    """The three Yee updates of one step on the n^3 grid without and with a PML
    of `layers` cells on all six walls (algorithmic bytes: 120 B/cell without;
    with it each of the three updates also reads its six ADE memories and the
    three profiles, 36 B/cell, and writes a memory where its axis' profile is
    nonzero)."""
    import torch
    from pypic3d_amd import _native as N
    from pypic3d_amd import workloads
    from pypic3d_amd._tensors import make_desc
    from pypic3d_amd.boundary_conditions.PML import initialize_tiled_pml_state, load_pml_from_toml
    from pypic3d_amd.solvers.first_order_yee import _pml_args, yee_b, yee_e

    static, dyn = workloads.make_parameters((n, n, n), cfl=0.5)
    desc = make_desc(static, dyn)
    f = workloads.empty_fields(static, dyn)
    E = [c[0, 0, 0] for c in f[0]]
    B = [c[0, 0, 0] for c in f[1]]
    J = [c[0, 0, 0] for c in f[2]]
    E2 = [c.clone() for c in E]
    Bt = [c.clone() for c in B]
    walls = [{"wall": w, "thickness": layers} for w in ("-x", "+x", "-y", "+y", "-z", "+z")]
    _, _, _, _, prof = load_pml_from_toml(walls, static, dyn)
    state = initialize_tiled_pml_state(static, dyn, prof)
    pb, pe = _pml_args(desc, state, 1), _pml_args(desc, state, 0)
    s = N.stream_handle()
    out = {}
    for name, pml_b, pml_e in (("none", None, None), ("six_walls", pb[:2], pe[:2])):
        def step():
            yee_b(desc, E, B, Bt, False, s, pml=pml_b)
            yee_e(desc, E, Bt, J, E2, s, pml=pml_e)
            yee_b(desc, E2, Bt, B, True, s, pml=pml_b)
        ms = _timed(step, reps)
        cells = n ** 3
        if name == "none":
            bytes_ = 120 * cells
        else:
            g = desc.g
            inside = [int((t[0, 0, 0, g:-g, g:-g, g:-g] != 0).sum()) for t in state[2]]
            writes = 4 * sum(inside[a] for a in (1, 2, 2, 0, 0, 1))
            bytes_ = 120 * cells + 3 * (36 * cells + writes)
        out[name] = {"ms_per_step": ms, "hbm_gbs": bytes_ / (ms * 1e-3) / 1e9}
    out["layers"] = layers
    return out


def ensemble(n_members, steps, seed):
    """Batched two-stream ensemble against sequential single-simulation runs
    of the same members (both advanced from identical initial states)."""
    import torch
    from pypic3d_amd import workloads
    from pypic3d_amd.batch import EnsembleRunner
    from pypic3d_amd.runner import ElectrodynamicRunner

    members = workloads.config_two_stream_ensemble(n_members=n_members, seed=seed)
    batch = EnsembleRunner(members)
    npart = batch.particles
    (rounds, grid), tune = batch.autotune()
    batch.capture()
    reps = max(1, steps // 2)
    batch.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        batch.replay()
    b.record()
    torch.cuda.synchronize()
    t_batch = a.elapsed_time(b) / (2 * reps)
    del batch
    torch.cuda.empty_cache()
    singles = [ElectrodynamicRunner(*m) for m in members]
    for r in singles:
        r.step()
    torch.cuda.synchronize()
    a.record()
    for _ in range(steps):
        for r in singles:
            r.step()
    b.record()
    torch.cuda.synchronize()
    t_seq = a.elapsed_time(b) / steps

    def graph_of(runners):
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for r in runners:
                r.step()
                r.step()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            for r in runners:       # two steps per member: the replay advances the state
                r.step()
                r.step()
        g.replay()
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            g.replay()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / (2 * reps)

    # the same sequential launches recorded in one graph (no host launch cost)
    t_seq_graph = graph_of(singles)
    del singles
    torch.cuda.empty_cache()
    # the batched kernels launched for one member at a time, also in one graph
    solo = [EnsembleRunner([m], rounds=rounds, grid=grid) for m in members]
    del members
    t_solo_graph = graph_of(solo)
    del solo
    torch.cuda.empty_cache()
    return {"members": n_members, "particles": npart, "rounds": rounds, "grid": grid,
            "batched_ms_per_step": t_batch, "sequential_ms_per_step": t_seq,
            "sequential_graph_ms_per_step": t_seq_graph,
            "batched_kernels_one_member_at_a_time_graph_ms_per_step": t_solo_graph,
            "speedup_vs_sequential": t_seq / t_batch, "speedup_vs_sequential_graph": t_seq_graph / t_batch,
            "speedup_vs_batched_kernels_one_member_at_a_time": t_solo_graph / t_batch,
            "batched_particle_steps_per_s": npart / (t_batch * 1e-3)}


def adjoint(n=64, ppc=32, steps=200, every=20, seed=20261202):
    """Differentiable two-stream (n x n x 1, two beams x ppc): forward with a
    checkpoint every `every` steps, then the adjoint sweep (recompute + reverse)
    for d(U_E + U_B)/d(drift, amplitude, per-particle momenta)."""
    import torch
    from pypic3d_amd import workloads
    from pypic3d_amd.adjoint import AdjointRun, field_energy_loss, two_stream_state
    from pypic3d_amd.particles.particle_class import SpeciesConfig
    import numpy as np

    static, dyn = workloads.make_parameters((n, n, 1), cfl=0.5)
    w = workloads.plasma_frequency_weight(2 * ppc, dyn.dt, 0.1)
    cfg = SpeciesConfig(charge=np.array([-1.0, -1.0]), mass=np.array([1.0, 1.0]), weight=np.array([w, w]),
                        update_x=np.ones((2, 3), bool), update_u=np.ones((2, 3), bool))
    drift = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(1e-3, dtype=torch.float64, requires_grad=True)
    x0, v0 = two_stream_state(static, dyn, ppc, drift, amp, vth=0.01, seed=seed, device="cuda")
    g = int(static.guard_cells)
    shape = (3, n + 2 * g, n + 2 * g, 1 + 2 * g)
    E0 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    B0 = torch.zeros_like(E0)
    act = torch.ones((2, x0.shape[2]), dtype=torch.bool, device="cuda")
    run = AdjointRun(static, dyn, cfg, act, checkpoint_every=every)
    # plain forward cost: the same steps without checkpoints
    run.forward(x0.detach(), v0.detach(), tuple(E0), tuple(B0), 2)
    r = run.runner
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = field_energy_loss(x0, v0, E0, B0, run, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gd, ga, gv = torch.autograd.grad(loss, [drift, amp, v0])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"grid": [n, n, 1], "particles": int(x0.shape[1] * x0.shape[2]), "steps": steps,
            "checkpoint_every": every, "forward_ms_per_step": 1e3 * fwd,
            "forward_with_checkpoints_ms_per_step": 1e3 * (t1 - t0) / steps,
            "adjoint_ms_per_step": 1e3 * (t2 - t1) / steps, "adjoint_over_forward": (t2 - t1) / steps / fwd,
            "loss": float(loss.detach()), "dloss_ddrift": float(gd), "dloss_damplitude": float(ga),
            "max_abs_dloss_dv0": float(gv.abs().max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=32)
    ap.add_argument("--sort-every", type=int, default=6)
    ap.add_argument("--tile", type=int, default=4, choices=(0, 4, 8),
                    help="cells per tile edge of the tile-ordered fused kernel (0: slot-ordered kernel)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--variant", type=int, default=None,
                    help="use this fused-kernel launch shape instead of autotuning (tile mode: slots per thread)")
    ap.add_argument("--min-blocks", type=int, default=None, choices=(1, 3, 4),
                    help="tile mode with --variant: resident-block target of the fused kernel")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble comparison")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the J_from_rhov / rho / CG measurements of the same workload")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only run a plain device copy of MiB MiB five times (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU (domain-decomposed) path is "
                                   "design-only, so --gpus must be 1"}))
        sys.exit(2)

    import torch
    from pypic3d_amd import _native as N
    from pypic3d_amd import workloads
    from pypic3d_amd.runner import ElectrodynamicRunner

    torch.cuda.set_device(0)
    static, dyn, species = workloads.config_warm_3d(n=args.n, ppc=args.ppc, vth=0.05)
    parts, cfg = workloads.plasma(static, dyn, species, seed=20261202)
    fields = workloads.empty_fields(static, dyn)
    runner = ElectrodynamicRunner(parts, cfg, fields, static, dyn, sort_every=args.sort_every,
                                  tile=args.tile or None)
    del parts
    torch.cuda.synchronize()
    n_particles = runner.S * runner.P
    n_cells = args.n ** 3
    print(f"[bench] {n_particles} particles on {args.n}^3, warmup {args.warmup}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        runner.step()
    torch.cuda.synchronize()
    if args.variant is not None:
        if runner.tile:
            runner.rounds = args.variant
            runner.min_blocks = args.min_blocks or runner.min_blocks
            variant = (runner.rounds, runner.min_blocks)
        else:
            runner.desc.variant = variant = args.variant
        tune_ms = {}
    else:
        variant, tune_ms = runner.autotune() if not args.no_autotune else \
            ((runner.rounds, runner.min_blocks) if runner.tile else 0, {})
    print(f"[bench] autotune fused variants (ms): {tune_ms} -> {variant}", file=sys.stderr, flush=True)
    runner.step()
    torch.cuda.synchronize()

    runner.enable_timing(True)
    sorts0 = runner.sorts_done
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        runner.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    phases = runner.phase_ms()
    sorts = runner.sorts_done - sorts0
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s; phases {phases}", file=sys.stderr, flush=True)

    if runner.tile:
        exact = bool(N.lib().pic_tile_exact_grid(runner.desc)) and runner.desc.variant != 1
        kernel_name = "fused_tile_exact_kernel" if exact else "fused_tiled_kernel"
    else:
        kernel_name = "fused_push_esirkepov_v2"
    traffic, traffic_src = measured_traffic(kernel_name, args.n, args.ppc)
    ms_per_step = 1e3 * elapsed / args.steps
    value = n_particles * args.steps / elapsed
    kern_ms = phases.get("particles")
    kern_bytes = 50 * n_particles + 36 * n_cells
    if runner.tile and runner.defer_permute:
        # the launch after each re-bin also reads the 4-byte source slot of every particle
        kern_bytes += 4 * n_particles * sorts // args.steps
    achieved = kern_bytes / (kern_ms * 1e-3) / 1e9 if kern_ms else None
    step_bytes = 50 * n_particles + 156 * n_cells
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"warm isotropic electron plasma {args.n}^3 x {args.ppc} ppc, vth 0.05c, periodic, "
                        "CIC Esirkepov, relativistic Boris, Yee",
            "particles": n_particles, "cells": n_cells, "sort_every": args.sort_every,
            "resorted_particles_per_step": n_particles * sorts / args.steps,
            "tile": runner.tile, "fused_variant": variant, "autotune_ms": tune_ms,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic, "traffic_source": traffic_src, "kernel_ms": kern_ms,
            "algorithmic_bytes": kern_bytes,
        },
        "step_hbm_gbs": step_bytes * args.steps / elapsed / 1e9,
        "phase_ms": phases,
    }
    if not args.no_extras:
        del runner
        torch.cuda.empty_cache()
        print("[bench] J_from_rhov, rho and CG on the same workload", file=sys.stderr, flush=True)
        out["other_paths"] = extras(args.n, args.ppc, 3, 20261202)
    if not args.no_ensemble:
        print("[bench] batched ensemble vs sequential single-simulation launches", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble(args.ensemble_members, 20, 20261202)
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        print("[bench] adjoint step cost (differentiable two-stream)", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint()
        torch.cuda.empty_cache()
    if not args.no_extras:
        print("[bench] FP64 two-stream, 500 steps (GPU FP64 path and oracle)", file=sys.stderr, flush=True)
        out["two_stream_fp64"] = two_stream_fp64(oracle=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (NumPy oracle sample on the host cores)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
