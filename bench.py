#!/usr/bin/env python3
"""Benchmark of the MI355X GPGMP hot path; prints one JSON line.

Headline workload (value): the plain stochastic inhomogeneous drift-diffusion
solver on the 256x256 Schnakenberg configuration (2 species, 4 reactions,
D_s(x) = D0_s (1 + 0.5 sin(2 pi x/L)), drift (0.02, 0), initial Poisson(200)/(75)).
A step is one host-loop iteration of DiffusionModel::oclSolve: the SSA over the
iteration's dt plus the due X/Y diffusion sweeps; the default K = 10^4 steps is
the whole configured run.  value = cells * K / wall time (cell-steps/s) with the
state resident in HBM.

Also reported (extra): the batched ensemble (512 realizations, autotuned launch
shape, HIP-graph chunks) against the plain per-realization solver, and the CPU
oracle timed on a bounded sample of the headline workload.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 100]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-steps/s at 1 GPU, batched vs plain; HBM GB/s; adjoint grad/s at 1e-4 rel"


def _host_info():
    """CPU model, thread count and the OpenCL devices of the host (the reference's kernels
    would need a CPU OpenCL device; ROCm's OpenCL exposes GPU agents only)."""
    import subprocess

    info = {"cpu_model": None, "cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
    except Exception as e:  # noqa: BLE001
        info["cpu_model"] = "unknown (%s)" % type(e).__name__
    try:
        out = subprocess.run(["clinfo"], capture_output=True, text=True, timeout=60).stdout
        types = [l.split(":", 1)[1].strip() for l in out.splitlines() if "Device Type:" in l]
        plats = [l.split(":", 1)[1].strip() for l in out.splitlines() if l.strip().startswith("Platform Name:")]
        cpu = [t for t in types if "CPU" in t]
        info["opencl"] = {"platforms": sorted(set(plats)), "device_types": types,
                          "cpu_device": bool(cpu),
                          "reference_on_cpu_opencl": "available" if cpu else
                          "unavailable: no CPU OpenCL device (clinfo lists %s)" % (", ".join(types) or "no device")}
    except Exception as e:  # noqa: BLE001
        info["opencl"] = {"reference_on_cpu_opencl": "unavailable: clinfo failed (%s)" % type(e).__name__}
    return info


def cpu_baseline(steps, single_steps):
    """The CPU oracle (a restatement of the reference kernels, OpenMP over cells) on the
    first `steps` iterations of the headline workload with every host thread, and on the
    first `single_steps` iterations with one thread."""
    from oracle import oracle
    from inchman_amd import configs

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = configs.schnakenberg_model(256)
    st = configs.schnakenberg_state(256)
    t0 = time.perf_counter()
    oracle.run(m, st, steps, 5, nthreads=threads)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    oracle.run(m, st, single_steps, 5, nthreads=1)
    d1 = time.perf_counter() - t1
    out = {"value": 65536 * steps / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
           "sample": "first %d host-loop iterations of the 256x256 Schnakenberg run (%.1f s)" % (steps, dt),
           "single_thread": {"value": 65536 * single_steps / d1, "unit": "cell-steps/s", "cores": 1,
                             "sample": "first %d host-loop iterations of the same run (%.1f s)" % (single_steps, d1)}}
    out.update(_host_info())
    return out


TIMING_EVERY = 128  # time 4 of every 128 steps: events on every launch add ~22% dispatch overhead
GRAPH_CHUNK = 128   # HIP-graph chunk of the timed run (the timing windows run eagerly)

# Counters of the timed launches (the 4 counted steps of every window of TIMING_EVERY in the
# timed region), per dispatch, by timing slot (1 SSA launch, 2 Y sweeps, 4 X sweeps + busy
# cells), for the bench windows that were profiled: SQ_INSTS_VALU (wave64 VALU instructions,
# summed over the chip) and HBM bytes (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE KiB, the factors a
# plain device copy reads).  Keyed by the number of timed steps (the timed region always
# starts from the initial state, so the window is fixed by it).  Commands, raw counters and
# the launch selection: profiles/gmp_step_schnakenberg256_counters.md.
WINDOW_COUNTERS = {
    20: {"valu": {1: 7734881.5, 2: 2255003.0, 4: 3276716.0},
         "hbm_bytes": {1: 1847424, 2: 1635872, 4: 2938256},
         "source": "rocprofv3 --pmc SQ_INSTS_VALU / FETCH_SIZE / WRITE_SIZE over bench.py --steps 20 --warmup 5: "
                   "the 4 counted launches of each kernel (timed iterations 1..4)"},
    10000: {"valu": {1: 1779384.7, 2: 8454541.7, 4: 9159980.5},
            "hbm_bytes": {1: 1670766, 2: 1666009, 4: 1668537},
            "source": "rocprofv3 --pmc SQ_INSTS_VALU / FETCH_SIZE / WRITE_SIZE over bench.py --steps 10000 --warmup 100: "
                      "the 316 counted launches of each kernel (timed iterations t with t % 128 in 1..4)"},
}

# MI355X's VALU issue peak: 256 CUs x 4 SIMDs, one wave64 32-bit VALU instruction per SIMD
# every 4 cycles (16 lanes), 2.4 GHz
VALU_PEAK = 256 * 4 * 2.4e9 / 4


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def plain_run(solver, st, seed0, steps, timing, chunk=GRAPH_CHUNK):
    import torch

    solver.synchronizeStateBufferToDevice(st)
    solver.run(0, seed0)
    solver.kernelTimes(on=timing)
    # the graphs the run replays are captured here, ahead of the timed region (nothing runs)
    solver.prepareRun(steps, chunk)
    solver.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # HIP-graph chunks: each chunk's timing window runs eagerly with its events, the rest of
    # the chunk replays a graph
    solver.runContinue(steps, chunk)
    solver.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, n, b = solver.kernelTimes(on=False)
    return wall, ms, n, b


def ensemble_extra(steps=5000, sample=32, chunk=50):
    """The batched ensemble on its configuration (512 realizations of the plain GMP network,
    sides log-uniform in [64, 512], seeds s_r = Threefry(r, 0xBA7C), `steps` host-loop
    iterations), autotuned launch shape, HIP-graph chunks captured ahead; against the plain
    solver run realization by realization on a stratified sample (every 512/sample-th
    realization in size order) for the same number of steps."""
    import numpy as np
    import torch
    from inchman_amd import Ensemble, configs

    n = 512
    sides = configs.ensemble_sides(n)
    seeds = configs.ensemble_seeds(n)
    m = configs.plain_gmp_model(32)
    ens = Ensemble(m, sides, sides, seeds)
    states = [configs.ensemble_state(sd, r) for r, sd in enumerate(sides)]
    for r, st in enumerate(states):
        ens.setState(r, st)
    block, cpt = ens.autotune(10)
    for r, st in enumerate(states):
        ens.setState(r, st)
    ens.reset()
    ens.prepareRun(steps, chunk)
    ens.synchronize()
    cells = ens.totalCells()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(steps, chunk)
    ens.synchronize()
    wall = time.perf_counter() - t0
    batched = cells * steps / wall
    idx = np.argsort(sides, kind="stable")[n // (2 * sample)::n // sample]
    tot, cs = 0.0, 0
    for r in idx:
        p = configs.plain_gmp_model(int(sides[r])).createSolver()
        p.synchronizeStateBufferToDevice(states[r])
        p.run(0, int(seeds[r]))
        p.prepareRun(steps, chunk)
        p.synchronize()
        t0 = time.perf_counter()
        p.runContinue(steps, chunk)
        p.synchronize()
        tot += time.perf_counter() - t0
        cs += int(sides[r]) ** 2 * steps
    plain = cs / tot
    return {"workload": "512 realizations of the plain GMP network, sides 64..512 (sum %d cells), %d steps, "
                        "seeds Threefry(r, 0xBA7C)" % (cells, steps),
            "batched_cell_steps_per_s": batched, "batched_wall_s": wall, "plain_cell_steps_per_s": plain,
            "plain_sample": "%d realizations (every %dth by size, sides %d..%d, sum %d cells), %d steps each"
                            % (len(idx), n // sample, int(sides[idx].min()), int(sides[idx].max()),
                               int((sides[idx].astype(np.int64) ** 2).sum()), steps),
            "batched_over_plain": batched / plain, "launch": {"block": block, "cells_per_thread": cpt},
            "graph_chunk": chunk}


def rdd_extra(reps=3):
    """Differentiable reaction-diffusion, 1024x1024, 2000 steps, checkpoints every 100:
    forward, adjoint (segment recompute + reverse sweep) and grad/s."""
    import torch
    from inchman_amd import configs
    from inchman_amd.rdd import ReactionDiffusion

    prob = configs.rdd_problem(1024)
    W = H = 1024
    steps = prob["steps"]
    r = ReactionDiffusion(W, H, steps, prob["dt"], prob["h"], prob["D0"], prob["k"], 100)
    dev = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    D, u0, v0, ut, vt = (dev(prob[k]) for k in ("D", "u0", "v0", "ut", "vt"))
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fw, bw = [], []
    for it in range(reps + 1):
        u, v = u0.clone(), v0.clone()
        torch.cuda.synchronize()
        e[0].record()
        r.forward(D, u, v, save=True)
        e[1].record()
        lu = ((2.0 / u.numel()) * (u - ut)).contiguous()
        lv = ((2.0 / u.numel()) * (v - vt)).contiguous()
        torch.cuda.synchronize()
        e[1].record()
        r.backward(D, lu, lv)
        e[2].record()
        torch.cuda.synchronize()
        if it:  # the first repetition is the warmup
            fw.append(e[0].elapsed_time(e[1]))
            bw.append(e[1].elapsed_time(e[2]))
    f, b = min(fw), min(bw)
    # four fused steps per launch: u, v, D read once with a 4-cell halo ((64+8)x(16+8) per
    # 64x16 tile) and u, v written once per four steps
    moved = (12.0 * (72 * 24) / (64 * 16) + 8.0) / 4
    return {"workload": "Schnakenberg RDD 1024x1024, 2000 fp32 Euler steps, checkpoints every 100",
            "forward_ms": f, "backward_ms": b, "adjoint_over_forward": b / f, "grad_per_s": 1e3 / (f + b),
            "forward_cell_steps_per_s": W * H * steps / (f * 1e-3),
            "forward_algorithmic_GBps": 20.0 * W * H * steps / (f * 1e-3) / 1e9,
            "forward_algorithmic_note": "20 B per cell-step: u, v, D read and u, v written by an unfused step",
            "forward_steps_per_launch": 4,
            "forward_moved_bytes_per_cell_step": moved,
            "forward_moved_GBps": moved * W * H * steps / (f * 1e-3) / 1e9}


def det_diffusion_extra(side=8192, reps=20):
    """The reference's deterministic stencils (homogeneous diffusion of 2 species; the
    inhomogeneous drift-diffusion X and Y sweeps of one species) on a side x side fp32
    grid, time-stepped by ping-pong between two state buffers of 2 x 256 MiB (larger than
    the caches): achieved HBM GB/s from the algorithmic bytes (homogeneous: read + write
    8 B per cell-species; inhomogeneous: state, D, r read and state written, 16 B per
    cell) over HIP-event kernel time on the launching stream."""
    import ctypes as C
    import torch
    from inchman_amd._lib import call

    A = side * side
    g = torch.Generator(device="cuda").manual_seed(3)
    buf = [torch.rand(2 * A, device="cuda", generator=g) * 100 for _ in range(2)]
    D = torch.rand(A, device="cuda", generator=g) * 0.9 + 0.1
    r = (torch.rand(A, device="cuda", generator=g) - 0.5) * 0.6
    Dh = (C.c_float * 2)(0.6, 0.2)
    st = torch.cuda.current_stream()
    sh = int(st.cuda_stream)
    out = {"workload": "deterministic stencils, %dx%d fp32, ping-pong time steps" % (side, side)}

    def timed(fn, nbytes):
        for k in range(4):
            fn(k)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(st)
        for k in range(reps):
            fn(k)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"ms_per_call": ms, "GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}

    out["homogeneous_2_species"] = timed(
        lambda k: call("gpgmp_det_hom_diffuse_device", buf[k & 1].data_ptr(), buf[(k + 1) & 1].data_ptr(), side, side,
                       2, Dh, 0.05, 1.0, sh), 8.0 * 2 * A)
    for axis, name in ((0, "inhomogeneous_x"), (1, "inhomogeneous_y")):
        out[name] = timed(lambda k: call("gpgmp_det_inhom_sweep_device", buf[k & 1].data_ptr(),
                                         buf[(k + 1) & 1].data_ptr(), side, side, 0, axis, 0.1, D.data_ptr(),
                                         r.data_ptr(), 1.0, sh), 16.0 * A)
    return out


def rk4_moments_extra(side=1024, dt=0.01, reps=20, moment_side=8192):
    """RK4 reactions (deterministic_performReaction_rk4) of the Schnakenberg network in
    concentration form on side x side cells over dt; the X moments of one species
    (inhomogeneous_computeMomentsX + reduceMomentsBlocks) of an int32 2-species state of
    moment_side x moment_side cells (4 B read per cell of the species)."""
    import numpy as np
    import torch
    from inchman_amd import configs
    from inchman_amd._lib import call

    m = configs.schnakenberg_model(side)
    cm = m.compile()
    rng = np.random.default_rng(3)
    st = np.stack([rng.uniform(50, 250, (side, side)), rng.uniform(20, 120, (side, side))]).astype(np.float32)
    x = torch.from_numpy(st).cuda()
    sh = int(torch.cuda.current_stream().cuda_stream)
    call("gpgmp_det_rk4_device", cm.handle, x.data_ptr(), dt, sh)  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call("gpgmp_det_rk4_device", cm.handle, x.data_ptr(), dt, sh)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {"rk4": {"workload": "RK4 reactions, Schnakenberg (concentration form), %dx%d cells, dt %g"
                   % (side, side, dt), "ms_per_call": t * 1e3 / reps, "cells_per_s": side * side * reps / t,
                   "algorithmic_GBps": 16.0 * side * side * reps / t / 1e9}}
    W = H = moment_side
    state = torch.randint(0, 1000, (2, H, W), dtype=torch.int32, device="cuda")
    sums = torch.zeros(2, dtype=torch.int64, device="cuda")
    call("gpgmp_moments_x_device", state.data_ptr(), W, H, 1, sums.data_ptr(), sh)  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call("gpgmp_moments_x_device", state.data_ptr(), W, H, 1, sums.data_ptr(), sh)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    n = state[1].to(torch.int64)
    ok = sums.tolist() == [int((n * torch.arange(W, device="cuda")[None, :]).sum()), int(n.sum())]
    out["moments_x"] = {"workload": "X moments of one species, %dx%d int32 cells" % (W, H),
                        "ms_per_call": t * 1e3 / reps, "GBps": 4.0 * W * H * reps / t / 1e9,
                        "hbm_frac": 4.0 * W * H * reps / t / 1e9 / HBM_PEAK_GBS, "exact": ok}
    return out


def aqss_extra(side=1024, dt=0.1, reps=5, cpu_side=64):
    """alpha-QSS reactions (deterministic_performReaction_aqss) of the Schnakenberg network
    in concentration form on side x side cells (u ~ U(50, 250), v ~ U(20, 120)) over dt,
    with the reference's corrector factor and with the fixed one; the oracle's float
    instance (one thread) on a cpu_side x cpu_side sample beside it."""
    import ctypes as C
    import numpy as np
    import torch
    from inchman_amd import configs
    from inchman_amd._lib import AQSS_PADE_FIX, call
    from oracle import oracle

    m = configs.schnakenberg_model(side)
    cm = m.compile()
    rng = np.random.default_rng(2)
    st = np.stack([rng.uniform(50, 250, (side, side)), rng.uniform(20, 120, (side, side))]).astype(np.float32)
    src = torch.from_numpy(st).cuda()
    x = src.clone()
    sh = int(torch.cuda.current_stream().cuda_stream)
    out = {"workload": "alpha-QSS, Schnakenberg (concentration form), %dx%d cells, dt %g" % (side, side, dt)}
    for name, flags in (("reference_corrector", 0), ("fixed_corrector", AQSS_PADE_FIX)):
        x.copy_(src)
        call("gpgmp_det_aqss_device", cm.handle, x.data_ptr(), dt, flags, sh)  # warmup
        torch.cuda.synchronize()
        t = 0.0
        for _ in range(reps):
            x.copy_(src)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            call("gpgmp_det_aqss_device", cm.handle, x.data_ptr(), dt, flags, sh)
            torch.cuda.synchronize()
            t += time.perf_counter() - t0
        out[name] = {"ms_per_call": t * 1e3 / reps, "cells_per_s": side * side * reps / t}
    tbl = oracle.table_from_model(m)
    t0 = time.perf_counter()
    oracle.det_aqss(st[:, :cpu_side, :cpu_side], tbl, dt, np.float32, False)
    cpu = time.perf_counter() - t0
    out["cpu_oracle_1_thread_cells_per_s"] = cpu_side * cpu_side / cpu
    return out


def individuals_extra(n=1 << 20, side=1024, steps=20, cpu_steps=2):
    """A particle-based species (IndividualSolver): n particles with their own
    diffusivities/drifts on a side x side reflective grid; each host-loop iteration does
    the due X/Y hop sweeps, each followed by the stable sort, cell offsets, per-cell
    constants and the time-step minimum (read back to the host, as the reference does)."""
    import numpy as np
    import torch
    from inchman_amd.individuals import IndividualSpecies

    rng = np.random.default_rng(8)
    cells = rng.integers(0, side * side, n).astype(np.uint32)
    props = np.stack([rng.uniform(0.1, 1.0, n), rng.uniform(0.1, 1.0, n), rng.uniform(-0.2, 0.2, n),
                      rng.uniform(-0.2, 0.2, n)]).astype(np.float32)
    g = IndividualSpecies(side, side, float(side), n)
    g.setParticles(cells, props)
    g.run(2, 1)  # warmup
    g.setParticles(cells, props)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.run(steps, 21)
    wall = time.perf_counter() - t0
    sw = g.sweeps()
    out = {"workload": "individual species, %d particles, %dx%d reflective cells, %d iterations" % (n, side, side, steps),
           "sweeps": sw, "particle_hops_per_s": n * sw / wall, "iterations_per_s": steps / wall, "wall_s": wall}
    from oracle import oracle
    p = oracle.ind_params(side, side, float(side))
    t1 = time.perf_counter()
    oracle.ind_run(p, cells, np.arange(n, dtype=np.uint32), props, cpu_steps, 21)
    cpu = time.perf_counter() - t1
    g.setParticles(cells, props)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g.run(cpu_steps, 21)
    gpu = time.perf_counter() - t2
    out["cpu_oracle_1_thread_iterations_per_s"] = cpu_steps / cpu
    out["gpu_iterations_per_s_same_sample"] = cpu_steps / gpu
    out["maintain"] = maintain_extra(side)
    return out


def maintain_extra(side=1024, calls=20):
    """Population maintenance after a reaction step (IndividualSolver::maintainPopulation):
    the per-cell goal alternates between two independent Poisson(1) fields on side x side
    cells (about 2^20 particles, ~0.37 of them killed and as many created per call), goals
    resident on the device, a newIndividuals rule drawing one RANDOM per new particle."""
    import numpy as np
    import torch
    from inchman_amd.individuals import IndividualSpecies

    A = side * side
    rng = np.random.default_rng(12)
    goals = [rng.poisson(1.0, A).astype(np.int32) for _ in range(2)]
    n0 = int(goals[1].sum())
    cells = np.repeat(np.arange(A, dtype=np.uint32), goals[1])
    props = np.ones((4, n0), np.float32)
    cap = 3 * A
    g = IndividualSpecies(side, side, float(side), cap)
    g.setParticles(cells, props)
    dev = [torch.from_numpy(x).cuda() for x in goals]
    rule = ((2, 0, 0, 0), (0.5, 0, 0, 0), (1.0, 0, 0, 0))
    for k in range(2):
        g.maintainPopulation(dev[k & 1], 100 + k, rule)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(calls):
        g.maintainPopulation(dev[k & 1], 200 + k, rule)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the oracle on one call of the same workload
    from oracle import oracle
    p = oracle.ind_params(side, side, float(side))
    keys = np.concatenate([cells, np.full(cap - n0, 0xFFFFFFFF, np.uint32)])
    vals = np.arange(cap, dtype=np.uint32)
    pr = np.zeros((4, cap), np.float32)
    t1 = time.perf_counter()
    st, *_ = oracle.ind_maintain(p, keys, vals, n0, pr, goals[0], 7, oracle.new_rule(*rule))
    cpu = time.perf_counter() - t1
    assert st == 0
    return {"workload": "maintainPopulation, %dx%d cells, goals alternating between two Poisson(1) fields "
                        "(~%d particles), RANDOM rule" % (side, side, n0),
            "ms_per_call": wall * 1e3 / calls, "cells_per_s": A * calls / wall,
            "cpu_oracle_1_thread_ms_per_call": cpu * 1e3}


def particles_extra(n=10_000_000, side=1024, steps=1000):
    """10^7 particles on a 1024x1024 periodic domain: move, stable radix sort, offsets,
    in-cell A+A->0 (p = 0.05 per pair per step)."""
    import numpy as np
    import torch
    from inchman_amd.particles import ParticleSystem

    rng = np.random.default_rng(4)
    x = rng.uniform(0, side, n).astype(np.float32)
    y = rng.uniform(0, side, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    ps = ParticleSystem(side, D=1.0, dt=0.01, vortex_amp=0.1, p_annihilate=0.05, seed=17, capacity=n)
    out = {"workload": "10^7 particles, 1024x1024 periodic cells, %d steps" % steps,
           "note": "the first step after set() sorts an unsorted list (radix step); later steps run on the "
                   "list sorted by cell (counts of the moves to the 3x3 neighbours and one scatter); an "
                   "untimed step of a small list first loads every kernel's code",
           "algorithmic_bytes": {"sorted_step_per_particle": 52, "sorted_step_per_cell": 160,
                                 "radix_pass_per_particle": 32, "note": "sorted step: move reads x, y, id "
                                 "and writes x2, y2, code (24 B), the scatter reads x2, y2, id, code and "
                                 "writes x, y, id (28 B); per cell: the move counts (36 B), the cell pass "
                                 "(cnt 36 B in, pre 32 + record 16 + survivors 4 B out), the offsets and "
                                 "records (16 B), the scatter's record and pre reads (~20 B)"}}
    ps.set(x[:200000], y[:200000], ids[:200000])
    ps.step(0, 3)
    ps.count()
    # (first step timed, k steps from it, repetitions): from the unsorted list; steps 1..10 from
    # the list the first step sorted (that step untimed); the whole decaying run.  The 10-step
    # windows last about 2 ms, so each is timed 5 times from the same state and the median kept.
    for label, first, k, reps in (("first_10_steps", 0, 10, 5), ("sorted_10_steps", 1, 10, 5),
                                  ("full_run", 0, steps, 1)):
        walls = []
        for _ in range(reps):
            ps.set(x, y, ids)
            if first:
                ps.step(0, first)
            p0 = ps.processed()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ps.step(first, k)
            n_end = ps.count()
            walls.append(time.perf_counter() - t0)
            done = ps.processed() - p0
        wall = sorted(walls)[len(walls) // 2]
        out[label] = {"particle_steps_per_s": done / wall, "particle_steps": done, "wall_s": wall,
                      "live_at_end": n_end}
        if reps > 1:
            out[label]["wall_s_samples"] = walls
        if label == "sorted_10_steps":
            # algorithmic HBM bytes of the sorted steps / wall time, against the 8 TB/s peak
            b = 52.0 * done + 160.0 * side * side * k
            out[label]["algorithmic_GBps"] = b / wall / 1e9
            out[label]["hbm_frac"] = b / wall / 8e12
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=3000)
    ap.add_argument("--cpu-single-steps", type=int, default=60)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--only-extra", default=None,
                    help="run one extra workload (ensemble, rdd, det_diffusion, particles, individuals, aqss, "
                         "rk4_moments) and print its JSON line; for profiling")
    ap.add_argument("--particle-steps", type=int, default=1000, help="steps of the particle extra's full run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver runs on a single MI355X "
              "(multi-GPU sharding is not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    from inchman_amd import configs

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.only_extra:
        fns = {"ensemble": ensemble_extra, "rdd": rdd_extra, "det_diffusion": det_diffusion_extra,
               "particles": lambda: particles_extra(steps=args.particle_steps), "individuals": individuals_extra,
               "aqss": aqss_extra, "rk4_moments": rk4_moments_extra}
        print(json.dumps({args.only_extra: fns[args.only_extra]()}), flush=True)
        return
    m = configs.schnakenberg_model(256)
    st = configs.schnakenberg_state(256)
    seed0 = 5
    solver = m.createSolver()
    # warmup: W iterations from the initial state
    solver.synchronizeStateBufferToDevice(st)
    solver.run(args.warmup, seed0)
    solver.synchronize()
    # timed: K iterations from the initial state; dispatch-tied HIP events on the solver's
    # stream around every launch of 4 consecutive steps out of every 128 (after one more
    # step with uncounted events: events right after untimed launches read skewed)
    wall, ms, n, b = plain_run(solver, st, seed0, args.steps, timing=TIMING_EVERY)
    cells = 256 * 256
    value = cells * args.steps / wall
    # the step kernels (index 3 is the hot-cell continuation), by total time in the run
    dom = int(max((0, 1, 2, 4), key=lambda i: ms[i]))
    avg_ms = ms[dom] / max(n[dom], 1)
    achieved = (b[dom] / max(n[dom], 1)) / (avg_ms * 1e-3) / 1e9
    # kernel names as the profiler prints them (template arguments: species, mode, SSA or
    # sweep lanes per cell, sweep scheme), for the default launch shape
    g_ssa, g_sweep = solver.getLaunch()[2:4]
    assert (g_ssa, g_sweep) == (1, 0), "names below assume the default lanes per cell (1, 0)"
    names = {0: "k_gmp_step<2, 0, 1, 0> (homogeneous)",
             1: "k_gmp_step<2, 1, 1, 0> (gather of Y outflow + SSA)",
             2: "k_gmp_step<2, 2, 8, 0> (gather of X outflow + Y sweeps)",
             3: "k_gmp_hot<2, 1> (deferred busy cells; not launched on this shape)",
             4: "k_gmp_xs_hot<2, 8> (X sweeps + deferred busy cells)"}
    # VALU issue against its peak, from the counters of this window's timed launches (the
    # kernels are RNG-bound: one Threefry-2x32-20 per SSA event and per particle pair); the
    # HBM roofline of the same launches beside it
    win = WINDOW_COUNTERS.get(args.steps)
    valu = {}
    for i in (1, 2, 4):
        if win and i in win["valu"] and n[i]:
            valu[i] = win["valu"][i] / (ms[i] / n[i] * 1e-3)
    frac = {i: v / VALU_PEAK for i, v in valu.items()}
    bad = {names[i]: f for i, f in frac.items() if f > 1.0}
    if bad:
        progress("VALU fraction above 1 (counter window does not match the timed launches): %s" % bad)
    roof = {"bound": "valu", "kernel": names[dom],
            "achieved": valu.get(dom), "peak": VALU_PEAK, "unit": "wave64 VALU instructions/s",
            "frac": frac.get(dom),
            "traffic": win["hbm_bytes"].get(dom) if win else None,
            "avg_launch_ms": avg_ms, "launches_timed": int(n[dom]),
            "instructions_per_launch": win["valu"].get(dom) if win else None,
            "counters": (win["source"] if win else
                         "no counter profile of a %d-step timed window: VALU fraction not reported" % args.steps),
            "hbm": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "algorithmic_bytes_per_launch": b[dom] / max(n[dom], 1),
                    "traffic_bytes_per_launch": win["hbm_bytes"].get(dom) if win else None},
            "consistent": not bad}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "stochastic inhomogeneous drift-diffusion, Schnakenberg, 256x256, "
                               "reference-exact Threefry streams",
                   "grid": [256, 256], "species": 2, "reactions": 4, "iterations": args.steps},
        "roofline": roof,
    }
    extra = {"kernel_avg_launch_ms": {names[i]: ms[i] / n[i] for i in range(5) if n[i]},
             "kernel_valu_frac": {names[i]: f for i, f in frac.items()},
             "kernel_hbm_frac": {names[i]: (b[i] / n[i]) / (ms[i] / n[i] * 1e-3) / 1e9 / HBM_PEAK_GBS
                                 for i in range(5) if n[i]},
             "kernel_launches_timed": {names[i]: int(n[i]) for i in range(5) if n[i]},
             "kernel_timing": "every launch of 4 consecutive steps out of every %d of the timed region"
                               % TIMING_EVERY}
    # the state handed over from and back to host memory (the boundary's host-buffer calls,
    # 2 x 65536 int32 each way), timed on their own after the run and added to its wall time:
    # the rate a caller that keeps the state on the host would see; not the headline value
    import torch as _t
    _t.cuda.synchronize()
    t0 = time.perf_counter()
    solver.synchronizeStateBufferToDevice(st)
    solver.synchronize()
    t1 = time.perf_counter()
    host_state = solver.synchronizeStateBufferFromDevice()
    t2 = time.perf_counter()
    extra["pcie_inclusive"] = {"cell_steps_per_s": cells * args.steps / (wall + (t2 - t0)),
                               "copy_in_s": t1 - t0, "copy_out_s": t2 - t1,
                               "note": "the timed run's wall time plus one state copy from and one to host memory"}
    del host_state
    progress("headline done: %.4g cell-steps/s" % value)
    if args.cpu_steps > 0:
        # same sample as the CPU baseline, for a like-for-like ratio
        w2, _, _, _ = plain_run(solver, st, seed0, args.cpu_steps, timing=False)
        extra["gpu_on_cpu_sample_cell_steps_per_s"] = cells * args.cpu_steps / w2
    if not args.no_extra:
        for key, fn in (("ensemble", ensemble_extra), ("rdd", rdd_extra), ("det_diffusion", det_diffusion_extra),
                        ("particles", lambda: particles_extra(steps=args.particle_steps)),
                        ("individuals", individuals_extra),
                        ("aqss", aqss_extra), ("rk4_moments", rk4_moments_extra)):
            extra[key] = fn()
            progress(key + " done")
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, args.cpu_single_steps)
        progress("cpu baseline done")
    out["extra"] = extra
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
