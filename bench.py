#!/usr/bin/env python3
"""Benchmark: all-atom AMBER pair energy+force throughput of the gradient-step
loop on one MI355X.

Workload (one "step" = one pass of the hot path over the whole batch): a
300-residue synthetic globule antigen (~2.5k heavy atoms) and 1,024 binders
of 20 residues (~167 atoms each); per step every candidate atom's LJ +
Coulomb + hydrophobic energy and force against the antigen (FP32 pair terms,
TwoSum-compensated, FP64 per-candidate totals), then the gradient update of
every residue; scoring with top-64 selection every 10 steps. Inputs are
resident in HBM before the timed region.

value = in-cutoff atom pairs evaluated (energy + force) per second, counted
exactly by replaying the same deterministic trajectory with the counting
kernel variant after the timed run.

Prints one JSON line. Only one GPU is supported (--gpus > 1 exits).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "de-novo-antigen-binder_amd"))
sys.path.insert(0, str(ROOT))

FLOP_PER_PAIR_EF = 40.0   # algorithmic FLOP per in-cutoff pair, energy + force
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 peak (vector and matrix)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """Worker threads of the CPU baselines: RAYON_NUM_THREADS when set, else
    one per logical CPU (Rayon's default)."""
    return int(os.environ.get("RAYON_NUM_THREADS", "0")) or os.cpu_count() or 1


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def sample_states(pop, params, steps, every, start):
    """The loop's states every `every` steps (and the last), stepped with the
    per-step API, which reproduces the loop's trajectory bit for bit; each
    state's in-cutoff pairs counted on the GPU with the reference predicate."""
    pop.set_coords(*start)
    states, pairs = [], []
    for it in range(steps + 1):
        if it % every == 0 or it == steps:
            states.append(pop.coords())
            pairs.append(pop.pair_count()[0])
        if it < steps:
            pop.step(params, it)
    return states, pairs


def cpu_baseline(w, states, state_pairs, budget_s=12.0):
    """The oracle's threaded C restatement of the reference's CPU path (grid
    E+F over all antigen atoms, src/energy.rs:322-398 pattern with AMBER
    per-atom parameters, parallel over candidates as the Rayon loop of
    src/diffusion.rs:1021-1024), timed on this host over the states the GPU
    loop itself passes through, round robin, for about budget_s seconds."""
    from oracle import oracle as O
    from dnab.structure import AtomCloud
    threads = host_threads()
    ag = w.antigens[0].atoms
    off = w.pop.cand_atom_off()
    a = w.pop.atoms
    clouds = [AtomCloud(x, y, z, a.charge, a.r_min_half, a.epsilon, a.hydrophobic, a.atom_type)
              for x, y, z in states]
    O.baseline_grid_ef(ag, clouds[0], off, w.cutoff, threads, forces=True)  # warm-up
    passes, pairs, t0 = 0, 0, time.perf_counter()
    while True:
        k = passes % len(clouds)
        O.baseline_grid_ef(ag, clouds[k], off, w.cutoff, threads, forces=True)
        pairs += state_pairs[k]
        passes += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or passes >= 100000:
            break
    # the reference's own energy pattern (brute force over all antigen atoms), first state
    reps_b, t0 = 0, time.perf_counter()
    while True:
        O.baseline_energy_brute(ag, clouds[0], off, w.cutoff, threads)
        reps_b += 1
        dtb = time.perf_counter() - t0
        if dtb >= budget_s / 3 or reps_b >= 1000:
            break
    return {"value": pairs / dt, "unit": "atom-pairs/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": (f"{passes} full-batch grid E+F passes over the {len(clouds)} states the "
                       f"GPU loop passes through (every 20th step; {min(state_pairs)}.."
                       f"{max(state_pairs)} in-cutoff pairs each) in {dt:.1f} s on {threads} "
                       f"threads; brute-force energy pattern on the start state: "
                       f"{state_pairs[0] * reps_b / dtb:.3g} pairs/s")}


class AllatomScorer:
    """The oracle's f32 allatom_energy (interaction_energy_atoms_breakdown
    plus disulfide_energy, src/diffusion.rs:281-283) of fixed candidates,
    scored by `threads` host threads: the C structs are built once, so each
    call releases the GIL for the whole scoring (ctypes), as the reference's
    Rayon workers score candidates independently."""

    def __init__(self, ag_atoms, proteins, cutoff):
        import ctypes as C
        import numpy as np
        from oracle import oracle as O
        self.C, self.O, self.cutoff = C, O, cutoff
        self.a, self.ka = O.cloud(ag_atoms)
        self.ea, self.kea = O.ext(ag_atoms.atom_type)
        self.items = []
        for p in proteins:
            b, kb = O.cloud(p.atoms)
            eb, keb = O.ext(p.atoms.atom_type)
            sg = [np.ascontiguousarray(v, np.float32)
                  for v in O.sg_positions(p.atoms.x, p.atoms.y, p.atoms.z, p.res_start,
                                          p.amino_acid, 0, len(p.amino_acid))]
            self.items.append((b, kb, eb, keb, sg))

    def rate(self, threads, budget_s):
        """Scorings per second over about budget_s seconds, and the count."""
        import threading
        import numpy as np
        C, lib = self.C, self.O.lib
        counts = [0] * threads
        deadline = time.perf_counter() + budget_s

        def work(t):
            out = np.zeros(5, np.float32)
            op = out.ctypes.data_as(C.POINTER(C.c_float))
            k, n = t, 0
            while time.perf_counter() < deadline:
                b, _, eb, _, sg = self.items[k % len(self.items)]
                lib.or_energy_full_f32(C.byref(self.a), C.byref(b), C.byref(self.ea),
                                       C.byref(eb), self.cutoff, op)
                lib.or_disulfide_f32(*[v.ctypes.data_as(C.POINTER(C.c_float)) for v in sg],
                                     len(sg[0]))
                k += threads
                n += 1
            counts[t] = n

        t0 = time.perf_counter()
        ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        el = time.perf_counter() - t0
        return sum(counts) / el, sum(counts), el


def copy_calibration(mib):
    """A device-to-device copy of `mib` MiB through the HIP runtime (its copy
    kernel reads and writes exactly that many bytes): the reference point
    that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def measured_traffic(step_kernels):
    """HBM bytes per launch of the dominant kernel of the c2 loop (the fused
    trajectory kernel, or the two pair kernels of a step), from the committed
    FETCH_SIZE / WRITE_SIZE profile, or None."""
    f = ROOT / "profiles" / ("pair_traffic_c2.txt" if step_kernels else "trajectory_traffic_c2.txt")
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split()[1])
    return None


def traj_kernel_time(ctx):
    """Event-timed milliseconds of the fused loop's launches so far: the
    step launches and the final evaluation's."""
    return ctx.kernel_time("trajectory")[0] + ctx.kernel_time("trajectory_final")[0]


def measured_valu(step_kernels):
    """VALU issue of the dominant kernel of the c2 loop from the committed
    SQ_INSTS_VALU profile: (fraction of the wave64 issue peak, wave
    instructions per in-cutoff pair), or None."""
    f = ROOT / "profiles" / "valu_c2.txt"
    if not f.exists():
        return None
    key = "pair_kernels" if step_kernels else "trajectory"
    vals = {}
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0].startswith(key + "_valu_"):
            vals[parts[0][len(key) + 6:]] = float(parts[1])
    return vals or None


def saturated_pair_rate(ctx, n_cand, reps=20):
    """The pair kernels fed enough work: the c2 antigen with n_cand binders at
    their starting positions (surface contact), energies + forces evaluated
    reps times; in-cutoff pairs counted exactly once by the counting
    variant, launches timed with HIP events on the context stream."""
    import dnab
    from dnab import workloads as W
    w = W.config2(n_cand=n_cand, seed=77)
    ags, pop = W.device_objects(ctx, w)
    n_in, n_tested = pop.pair_count()
    tuning, _ = pop.autotune(reps=3)
    pop.evaluate(forces=True)                      # warm-up
    ctx.synchronize()
    ctx.set_timing(True)
    k0, n0 = ctx.kernel_time("pair_ef")
    for _ in range(reps):
        pop.evaluate(forces=True)
    ctx.synchronize()
    k1, n1 = ctx.kernel_time("pair_ef")
    ctx.set_timing(False)
    avg_ms = (k1 - k0) / max(1, n1 - n0)
    achieved = n_in * FLOP_PER_PAIR_EF / (avg_ms * 1e-3) / 1e12
    pop.close()
    for a in ags:
        a.close()
    return {"candidates": n_cand, "pairs_in_cutoff_per_launch": int(n_in),
            "pairs_tested_per_launch": int(n_tested), "kernel_avg_ms": avg_ms,
            "achieved": achieved, "frac": achieved / PEAK_FP32_TFLOPS,
            "pairs_per_s": n_in / (avg_ms * 1e-3), "pair_tuning": tuning}


def saturated_trajectory_rate(ctx, n_cand, steps=10):
    """The fused trajectory kernel fed enough work: the c2 antigen with n_cand
    binders at their starting positions (surface contact, where most pairs
    are), `steps` gradient steps; in-cutoff pairs counted by a replay with
    the counting variant, the launch timed with HIP events."""
    import numpy as np
    import dnab
    from dnab import workloads as W
    w = W.config2(n_cand=n_cand, seed=77)
    ags, pop = W.device_objects(ctx, w)
    x0, y0, z0 = pop.coords()

    def params(count):
        return dnab.langevin_params(n_steps=steps, score_every=0, top_k=64,
                                    restraint_radius=w.restraint_radius[0],
                                    count_pairs=1 if count else 0)
    pop.reset_pair_stats()
    pop.run(params(True))
    n_in, n_tested = pop.pair_stats()
    pop.set_coords(x0, y0, z0)
    tuning = pop.autotune_trajectory(params(False), trial_steps=steps)
    ctx.synchronize()
    ctx.set_timing(True)
    k0 = traj_kernel_time(ctx)
    t0 = time.perf_counter()
    pop.run(params(False))
    ctx.synchronize()
    wall_fused = time.perf_counter() - t0
    ms = traj_kernel_time(ctx) - k0
    ctx.set_timing(False)
    achieved = n_in * FLOP_PER_PAIR_EF / (ms * 1e-3) / 1e12
    # the per-step kernels on the same trajectory (the paths agree bit for
    # bit): their pair kernels' event time over the same evaluations
    xf = pop.coords()
    pop.set_coords(x0, y0, z0)
    pop.autotune(reps=3)
    pop.set_coords(x0, y0, z0)
    ctx.synchronize()
    ctx.set_timing(True)
    p0 = ctx.kernel_time("pair_ef")[0]
    pop.run(dnab.langevin_params(n_steps=steps, score_every=0, top_k=64,
                                 restraint_radius=w.restraint_radius[0], step_kernels=1,
                                 use_graph=0))
    ctx.synchronize()
    pair_ms = ctx.kernel_time("pair_ef")[0] - p0
    ctx.set_timing(False)
    same = all(np.array_equal(u, v) for u, v in zip(xf, pop.coords()))
    # the whole loop on each engine, wall clock (the per-step kernels
    # replayed from a graph, warmed up once)
    step_params = dnab.langevin_params(n_steps=steps, score_every=0, top_k=64,
                                       restraint_radius=w.restraint_radius[0], step_kernels=1,
                                       use_graph=1)
    pop.set_coords(x0, y0, z0)
    pop.run(step_params)
    pop.set_coords(x0, y0, z0)
    ctx.synchronize()
    t0 = time.perf_counter()
    pop.run(step_params)
    ctx.synchronize()
    wall_step = time.perf_counter() - t0
    same = same and all(np.array_equal(u, v) for u, v in zip(xf, pop.coords()))
    pop.close()
    for a in ags:
        a.close()
    return {"candidates": n_cand, "steps": steps, "evaluations": steps + 1,
            "pairs_in_cutoff": int(n_in), "pairs_tested": int(n_tested),
            "kernel_ms": ms, "ms_per_evaluation": ms / (steps + 1), "achieved": achieved,
            "frac": achieved / PEAK_FP32_TFLOPS, "pairs_per_s": n_in / (ms * 1e-3),
            "trajectory_tuning": tuning, "wall_ms": wall_fused * 1e3,
            "step_kernels_same_trajectory": {
                "pair_kernel_ms": pair_ms, "pairs_per_s": n_in / (pair_ms * 1e-3),
                "wall_ms": wall_step * 1e3, "bit_identical": bool(same)}}


def loop_leg(ctx, w, trial_steps=20):
    """One workload's gradient-step loop (w.n_steps steps, scored every
    w.score_every, top w.top_k kept) on both engines from the same start
    state: the fused trajectory kernel with its autotuned shape, and the
    per-step kernels (autotuned work split, graph replay). In-cutoff pairs
    are counted by a replay with the counting variant; the fused rate uses
    the event time of its launches (steps and final evaluation)."""
    import numpy as np
    import dnab
    from dnab import workloads as W
    ags, pop = W.device_objects(ctx, w)
    x0 = pop.coords()

    def params(n, count=False, step_kernels=0):
        return dnab.langevin_params(n_steps=n, score_every=w.score_every, top_k=w.top_k,
                                    restraint_radius=w.restraint_radius[0], with_noise=w.with_noise,
                                    count_pairs=1 if count else 0, step_kernels=step_kernels,
                                    use_graph=1)
    trial = min(w.n_steps, trial_steps)
    tuning = pop.autotune_trajectory(params(trial), trial_steps=trial)
    pop.set_coords(*x0)
    pop.reset_pair_stats()
    pop.run(params(w.n_steps, count=True))
    n_in, n_tested = pop.pair_stats()
    xc = pop.coords()

    def timed(step_kernels, reps=2):
        best = None
        for _ in range(reps):
            pop.set_coords(*x0)
            ctx.synchronize()
            ctx.set_timing(not step_kernels)
            k0 = traj_kernel_time(ctx)
            t0 = time.perf_counter()
            idx, _ = pop.run(params(w.n_steps, step_kernels=step_kernels))
            ctx.synchronize()
            dt = time.perf_counter() - t0
            k = traj_kernel_time(ctx) - k0
            ctx.set_timing(False)
            if best is None or dt < best[0]:
                best = (dt, k, idx)
        return best
    dt_f, k_f, idx_f = timed(0)
    xf = pop.coords()
    pop.set_coords(*x0)
    pair_tuning, _ = pop.autotune(reps=3)
    dt_s, _, idx_s = timed(1)
    xs = pop.coords()
    same = (all(np.array_equal(a, b) for a, b in zip(xf, xs)) and np.array_equal(idx_f, idx_s) and
            all(np.array_equal(a, b) for a, b in zip(xf, xc)))
    achieved = n_in * FLOP_PER_PAIR_EF / (k_f * 1e-3) / 1e12
    out = {"workload": w.description, "antigen_atoms": [int(a.n_atoms()) for a in w.antigens],
           "candidates": int(w.pop.n_cand), "candidate_atoms": int(len(w.pop.atoms)),
           "cutoff_A": w.cutoff, "steps": w.n_steps, "evaluations": w.n_steps + 1,
           "score_every": w.score_every, "pairs_in_cutoff": int(n_in), "pairs_tested": int(n_tested),
           "tested_per_in_cutoff": n_tested / max(n_in, 1),
           "fused": {"wall_ms": dt_f * 1e3, "kernel_ms": k_f, "pairs_per_s": n_in / dt_f,
                     "kernel_pairs_per_s": n_in / (k_f * 1e-3), "achieved_tflops": achieved,
                     "frac": achieved / PEAK_FP32_TFLOPS, "trajectory_tuning": tuning},
           "step_kernels": {"wall_ms": dt_s * 1e3, "pairs_per_s": n_in / dt_s,
                            "pair_tuning": pair_tuning, "graph": True},
           "paths_bit_identical": bool(same)}
    pop.close()
    for a in ags:
        a.close()
    return out


def probe_pair_counts(ctx, ag_xyz, probe_xyz, cell, rc):
    """In-cutoff (candidate atom, antigen atom) pairs under the reference's
    FP32 predicate (src/energy.rs:284): all of them, and those the
    reference's 27-cell probe (src/spatial.rs:44-58) lists on a grid of
    `cell` over the antigen atoms."""
    import numpy as np
    import dnab
    A = np.ascontiguousarray(ag_xyz, np.float32)
    P = np.ascontiguousarray(probe_xyz, np.float32)

    def keep(d):
        r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        return ~((r2 > np.float32(rc * rc)) | (r2 < np.float32(0.25)))
    n_all = 0
    for k in range(0, len(P), 2048):
        n_all += int(keep(P[k:k + 2048, None, :] - A[None, :, :]).sum())
    g = dnab.SpatialHashGrid(ctx, cell)
    g.build(A[:, 0], A[:, 1], A[:, 2])
    off, idx = g.query_neighbors(P[:, 0], P[:, 1], P[:, 2])
    pi = np.repeat(np.arange(len(P)), np.diff(off.astype(np.int64)))
    found = int(keep(P[pi] - A[idx.astype(np.int64)]).sum())
    g.close()
    return n_all, found


def c1_leg(ctx, reps=3):
    """run_allatom at the plain method's shape: a 16-residue peptide antigen,
    64 binders of 12 residues on a ring, 600 refinement iterations with the
    Monte-Carlo moves, no broad phase (the reference's path without a GPU),
    grid cells of 8 A; and the in-cutoff pairs the reference's 27-cell probe
    misses with 8 A cells (10 A cutoff) against 10 A cells, on the start
    binders and on the results."""
    import numpy as np
    import dnab
    from dnab import workloads as W
    w = W.config1()
    ag = w.antigens[0]
    dag = dnab.DeviceAntigen(ctx, ag, cutoff=w.cutoff, cell_size=w.cell_size)
    kw = dict(ab_length=12, population=64, cpu_iterations=600, top_n=64, n_broad=0, seed=11)
    dnab.run_allatom(ctx, dag, **kw)  # warm-up
    best, res = 1e9, None
    for _ in range(reps):
        t0 = time.perf_counter()
        res = dnab.run_allatom(ctx, dag, **kw)
        best = min(best, time.perf_counter() - t0)
    A = np.stack([ag.atoms.x, ag.atoms.y, ag.atoms.z], 1)
    a = w.pop.atoms
    start = np.stack([a.x, a.y, a.z], 1)
    final = np.concatenate([np.stack([p.atoms.x, p.atoms.y, p.atoms.z], 1) for p, _, _ in res])
    missed = {}
    for name, P in (("start", start), ("results", final)):
        n_all, f8 = probe_pair_counts(ctx, A, P, 8.0, w.cutoff)
        _, f10 = probe_pair_counts(ctx, A, P, 10.0, w.cutoff)
        missed[name] = {"pairs_in_cutoff": n_all, "found_cell_8": f8, "missed_cell_8": n_all - f8,
                        "found_cell_10": f10, "missed_cell_10": n_all - f10}
    dag.close()
    e = [r[1] for r in res]
    return {"workload": w.description + ", run_allatom: 600 iterations with moves, no broad phase",
            "antigen_atoms": int(ag.n_atoms()), "ms_per_call": best * 1e3,
            "binder_iterations_per_s": 64 * 600 / best, "best_energy": e[0],
            "energies_sorted": bool(all(x <= y for x, y in zip(e, e[1:]))),
            "missed_pairs": missed}


def interface_positions(ab, ag):
    """is_interface (src/filters.rs:96-105): Calpha within 8 A of an antigen Calpha."""
    import numpy as np
    p, q = ab.ca_pos(), ag.ca_pos()
    d = p[:, None, :] - q[None, :, :]
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return [int(r) for r in np.nonzero((r2 < np.float32(64.0)).any(axis=1))[0]]


def allatom_energy_bench(ctx, w, pop, coords, reps=10, cpu_budget_s=3.0):
    """allatom_energy (src/diffusion.rs:281-283: the five pair terms of
    interaction_energy_atoms_breakdown plus the disulfide term) of every
    binder of the workload at its start state, one breakdown evaluation of the
    population per call. The CPU baseline is the oracle's f32 restatement of
    the same sum on all host threads, over the same binders."""
    import numpy as np
    pop.set_coords(*coords)
    pairs, _ = pop.pair_count()
    pop.evaluate(forces=False, breakdown=True)     # warm-up
    ctx.synchronize()
    ctx.set_timing(True)
    k0 = ctx.kernel_time("pair_breakdown")
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        pop.evaluate(forces=False, breakdown=True)
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    k1 = ctx.kernel_time("pair_breakdown")
    ctx.set_timing(False)
    dt = float(np.median(times))
    terms = pop.terms()
    out = {
        "workload": ("allatom_energy (LJ, Coulomb, hydrophobic, H-bond, EEF1 solvation, disulfide) "
                     "of the 1,024 binders at their start state"),
        "candidates": int(pop.n_cand),
        "pairs_in_cutoff": int(pairs),
        "ms_per_call": dt * 1e3,
        "kernel_ms": (k1[0] - k0[0]) / reps,
        "value": pop.n_cand / dt,
        "unit": "candidates/s",
        "pairs_per_s": pairs / dt,
        "mean_energy": float(terms.sum(1).mean()),
    }
    threads = host_threads()
    rate, n, el = AllatomScorer(w.antigens[0].atoms, candidate_proteins(w, coords),
                                w.cutoff).rate(threads, cpu_budget_s)
    out["cpu_baseline"] = {
        "value": rate, "unit": "candidates/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
        "sample": f"{n} f32 allatom_energy scorings of the workload's binders in {el:.1f} s "
                  f"on {threads} threads"}
    out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def candidate_proteins(w, coords):
    """The workload's binders as AtomProteins at the given coordinates."""
    import numpy as np
    from dnab.structure import AtomProtein
    p = w.pop
    off = p.cand_atom_off()
    x, y, z = coords
    prots = []
    for c in range(p.n_cand):
        r0, r1 = int(p.cand_res_off[c]), int(p.cand_res_off[c + 1])
        cl = p.candidate_cloud(c)
        a0, a1 = int(off[c]), int(off[c + 1])
        cl.x, cl.y, cl.z = x[a0:a1].copy(), y[a0:a1].copy(), z[a0:a1].copy()
        prots.append(AtomProtein(cl, (p.res_atom_off[r0:r1 + 1] - off[c]).astype(np.uint32),
                                 p.res_amino_acid[r0:r1].copy(),
                                 (p.res_ca_atom[r0:r1] - off[c]).astype(np.uint32)))
    return prots


def maturation_bench(ctx, w, ag_dev, coords, reps=5, cpu_budget_s=4.0):
    """Batched affinity maturation (dnab_affinity_maturation) of every binder
    the loop refined, at its interface positions, as src/main.rs:341-347 runs
    it per result. Timed: the C call alone, inputs packed beforehand. The CPU
    baseline restates the reference's cost: one f32 allatom_energy (the
    oracle's C restatement of interaction_energy_atoms_breakdown plus the
    disulfide term) per trial the reference's scan makes, scored on all host
    threads."""
    import numpy as np
    from dnab.api import MaturationCall
    from dnab.structure import AtomProtein, population_from_proteins
    p = w.pop
    off = p.cand_atom_off()
    x, y, z = coords
    ag = w.antigens[0]
    abs_ = candidate_proteins(w, coords)
    muts = [interface_positions(prot, ag) for prot in abs_]
    npos = sum(len(m) for m in muts)
    call = MaturationCall(ctx, ag_dev, population_from_proteins(abs_), muts)
    call.run()                                     # warm-up
    phases = ("maturation_host_build", "maturation_host_upload", "maturation_host_device",
              "maturation_host_assemble", "pair_breakdown", "maturation_scan",
              "population_create_pack", "population_create_upload", "population_create_sort")
    ctx.set_timing(True)
    k0 = {n: ctx.kernel_time(n) for n in phases}
    times = []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        call.run()
        times.append(time.perf_counter() - t0)
    k1 = {n: ctx.kernel_time(n) for n in phases}
    ctx.set_timing(False)
    dt = float(np.median(times))
    prots, e, _, _ = call.results()
    # trials the reference's scan makes: up to the first improving type, else 19
    trials = 0
    for a, m in enumerate(muts):
        old, new = abs_[a].amino_acid, prots[a].amino_acid
        for r in m:
            if new[r] != old[r]:
                trials += [t for t in range(20) if t != old[r]].index(int(new[r])) + 1
            else:
                trials += 19
    out = {
        "workload": ("affinity maturation (greedy point-mutation scan scored by allatom_energy) "
                     "of the 1,024 binders after the 200-step loop, at their interface positions"),
        "antibodies": len(abs_),
        "antibodies_with_positions": sum(1 for m in muts if m),
        "positions": npos,
        "residues_scored": int(call.host.n_res + 20 * npos),
        "reference_trials": trials,
        "ms_per_call": dt * 1e3,
        "value": npos / dt,
        "unit": "positions/s",
        "phase_ms": {n: (k1[n][0] - k0[n][0]) / reps for n in phases},
        "mutations_kept": int(sum(int((a.amino_acid != b.amino_acid).sum())
                                  for a, b in zip(abs_, prots))),
    }
    # CPU baseline: f32 allatom_energy per trial, all host threads
    sample = [a for a, m in zip(abs_, muts) if m] or abs_[:1]
    threads = host_threads()
    rate, n_sc, el = AllatomScorer(ag.atoms, sample, w.cutoff).rate(threads, cpu_budget_s)
    cpu_val = npos / (trials / rate) if trials else None
    out["cpu_baseline"] = {
        "value": cpu_val, "unit": "positions/s", "cores": threads, "kind": "port",
        "cpu": cpu_model(),
        "sample": (f"{n_sc} f32 allatom_energy scorings of the workload's antibodies with "
                   f"interface positions in {el:.1f} s on {threads} threads "
                   f"({rate:.0f} per s), times the {trials} trials the reference's scan makes")}
    out["speedup_vs_cpu"] = (out["value"] / cpu_val) if cpu_val else None
    return out


def refine_bench(ctx, w, ag_dev, coords, survivors, cpu_budget_s=5.0):
    """dnab_refine: allatom_diffusion_step with its Monte-Carlo moves
    (src/diffusion.rs:290-416), all iterations in one launch. Two workloads
    of run_allatom: the refinement phase (the loop's 64 best binders, 600
    iterations with noise, src/diffusion.rs:1083-1095) and the broad-sampling
    step count with the moves on all 1,024 binders (200 iterations, no noise,
    src/diffusion.rs:1017-1024). CPU baseline: the oracle's C port of the
    same step, threaded over candidates like the reference's Rayon loop,
    timed on a bounded sample."""
    import numpy as np
    import dnab
    from dnab.api import RefineCall, mc_params
    from dnab.structure import AtomProtein, population_from_proteins
    prots = candidate_proteins(w, coords)
    out = {}
    for name, cands, steps, noise, total in (
            ("refinement_64x600", [prots[int(i)] for i in survivors[:64]], 600, 1, 600),
            ("broad_1024x200", prots, 200, 0, 800)):
        lp = dnab.langevin_params(n_steps=steps, with_noise=noise, schedule_offset=0,
                                  schedule_total=total, restraint_radius=w.restraint_radius[0],
                                  seed=2024)
        host = population_from_proteins(cands)
        call = RefineCall(ctx, ag_dev, host, lp, mc_params(), energies=False)
        call.run()                                       # warm-up (and the kernel's first load)
        ctx.set_timing(True)
        k0 = ctx.kernel_time("refine")
        ctx.synchronize()
        t0 = time.perf_counter()
        call.run()
        dt = time.perf_counter() - t0
        k1 = ctx.kernel_time("refine")
        ctx.set_timing(False)
        cs = len(cands) * steps
        out[name] = {"candidates": len(cands), "iterations": steps, "noise": bool(noise),
                     "ms_per_call": dt * 1e3, "kernel_ms": k1[0] - k0[0],
                     "value": cs / dt, "unit": "candidate-iterations/s"}
    # CPU baseline: the C port on the host cores, 64 candidates, a bounded number of iterations
    from oracle import oracle as O
    import ctypes as C
    from dnab import _native as N
    cg = []
    for a in range(20):
        q, e, sg, h = C.c_float(), C.c_float(), C.c_float(), C.c_uint8()
        N.lib.dnab_residue_cg_params(a, C.byref(q), C.byref(e), C.byref(sg), C.byref(h))
        cg.append((q.value, e.value, sg.value, h.value))
    tables = O.refine_tables(np.array(cg, np.float32))
    cands = [prots[int(i)] for i in survivors[:64]]
    host = population_from_proteins(cands)
    slots = np.zeros((host.n_res, 14, 3), np.float32)
    for g in range(host.n_res):
        b, e = int(host.res_atom_off[g]), int(host.res_atom_off[g + 1])
        slots[g, :e - b] = np.stack([host.atoms.x[b:e], host.atoms.y[b:e], host.atoms.z[b:e]], 1)
    aa = host.res_amino_acid.copy()
    rot = np.zeros(host.n_res, np.uint8)
    ag = w.antigens[0]
    threads = int(os.environ.get("RAYON_NUM_THREADS", "0")) or os.cpu_count() or 1
    prm = dict(step=0.08, max_disp=2.0, restraint_k=0.02, restraint_radius=w.restraint_radius[0],
               with_noise=1, seed_lo=2024, seed_hi=0, mut_p=0.08, rot_p=0.12, bb_p=0.04,
               max_torsion=15.0, iface_sq=64.0, moves=3, center=ag.ca_center_of_mass())
    done, t0 = 0, time.perf_counter()
    while True:
        k = 10
        temps = [O.temperature(done + i, 600) for i in range(k)]
        sigs = [O.sigma(done + i, 600) for i in range(k)]
        O.refine_run(tables, ag.ca_residue_cloud(), ag.amino_acid, host.cand_res_off, aa, rot,
                     slots, temps, sigs, done, prm, threads)
        done += k
        el = time.perf_counter() - t0
        if el >= cpu_budget_s or done >= 600:
            break
    out["cpu_baseline"] = {
        "value": 64 * done / el, "unit": "candidate-iterations/s", "cores": threads,
        "kind": "port",
        "sample": (f"the oracle's C port of the step with its moves, 64 candidates x {done} "
                   f"iterations of the refinement phase in {el:.1f} s on {threads} threads")}
    out["speedup_vs_cpu"] = out["refinement_64x600"]["value"] / out["cpu_baseline"]["value"]
    return out


def prescreen_bench(ctx, args):
    """Learned prescreen at the config-5 shape: train on exact energies of
    shell binders around antigens of 100..1,000 residues, then screen a pool
    of 65,536 candidates against a fresh 300-residue antigen; compare the
    pipeline (map features + fused MLP + top-n + exact re-score of the kept
    set + top-64) with exact scoring of the whole pool, and the network's kept
    set with the map energy's (the screen without the network)."""
    import numpy as np
    import dnab
    from dnab import workloads as W
    t0 = time.perf_counter()
    per = max(1024, args.prescreen_train // 10)
    ds = dnab.Dataset(ctx, capacity=10 * per)
    for k, n_res in enumerate(range(100, 1001, 100)):
        r = W.globule_radius(n_res)
        ag = W.globule_antigen(n_res, r, 5000 + k)
        pop_h = W.shell_population(per, 20, ag.ca_center_of_mass(), r + W.SHELL_GAP, r,
                                   (5 << 32) | (k << 24))
        dag = dnab.DeviceAntigen(ctx, ag, cutoff=10.0, cell_size=10.0, restraint_radius=r + W.SHELL_GAP)
        dpop = dnab.DevicePopulation(ctx, [dag], pop_h)
        ds.add(dpop)
        dpop.close()
        dag.close()
        log(f"[bench] prescreen training data: antigen {n_res} res, {len(ds)} rows")
    t_data = time.perf_counter() - t0
    mlp = dnab.Mlp(ctx, seed=1)
    t0 = time.perf_counter()
    steps = max(200, 2 * len(ds) // 512)
    hist = mlp.train(ds, n_steps=steps, batch=512, lr=1e-3)
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    ds.close()
    # screening pool
    w = W.config2(n_cand=args.prescreen_pool, seed=55)
    ags, pop = W.device_objects(ctx, w)
    t0 = time.perf_counter()
    ags[0].build_maps()
    ctx.synchronize()
    t_maps = time.perf_counter() - t0
    dnab.exact_topk(pop, 64)            # warm-up
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        idx_e, e_e = dnab.exact_topk(pop, 64)
    t_exact = (time.perf_counter() - t0) / reps
    top = set(idx_e.tolist())
    S = dnab.map_energy(pop)
    mlp.predict(pop)                    # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        mlp.predict(pop)
    t_predict = (time.perf_counter() - t0) / reps
    out = {"pool": w.pop.n_cand, "train_rows": 10 * per, "train_steps": steps,
           "train_s": round(t_train, 2), "data_s": round(t_data, 2),
           "maps_build_ms": t_maps * 1e3, "maps_bytes": int(ags[0].map_info()["bytes"]),
           "loss_first": float(np.mean(hist[:50])), "loss_last": float(np.mean(hist[-50:])),
           "exact_cand_per_s": w.pop.n_cand / t_exact, "exact_ms": t_exact * 1e3,
           "features_mlp_ms": t_predict * 1e3,
           "recall_at_64": {}, "recall_at_64_map_energy": {}, "pipeline_cand_per_s": {},
           "pipeline_ms": {}}
    best = None
    for keep in (512, 1024, 2048):
        mlp.prescreen(pop, keep, 64)    # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            idx_p, _ = mlp.prescreen(pop, keep, 64)
        t_p = (time.perf_counter() - t0) / reps
        rec = len(set(idx_p.tolist()) & top) / 64
        rec_map = len(set(np.argsort(S, kind="stable")[:keep].tolist()) & top) / 64
        out["recall_at_64"][str(keep)] = rec
        out["recall_at_64_map_energy"][str(keep)] = rec_map
        out["pipeline_cand_per_s"][str(keep)] = w.pop.n_cand / t_p
        out["pipeline_ms"][str(keep)] = t_p * 1e3
        if best is None and rec >= 0.95:
            best = w.pop.n_cand / t_p
            out["speedup_vs_exact_at_95pct_recall"] = t_exact / t_p
    out["screened_cand_per_s_at_95pct_recall"] = best
    # per-stage kernel times of one screen (keep 512, HIP events), and the
    # fused MLP kernel against the bf16 MFMA peak (2 M (32 256 + 2 256^2 + 256) FLOP)
    names = ("map_features", "mlp_infer", "trajectory_subset")
    ctx.set_timing(True)
    k0 = {n: ctx.kernel_time(n) for n in names}
    mlp.prescreen(pop, 512, 64)
    k1 = {n: ctx.kernel_time(n) for n in names}
    ctx.set_timing(False)
    out["stage_kernel_ms"] = {n: k1[n][0] - k0[n][0] for n in names}
    ms = out["stage_kernel_ms"]["mlp_infer"]
    if ms > 0:
        flop = 2.0 * w.pop.n_cand * (32 * 256 + 2 * 256 * 256 + 256)
        out["mlp_infer"] = {"kernel_ms": ms, "tflops": flop / (ms * 1e-3) / 1e12,
                            "frac_bf16_mfma_peak": flop / (ms * 1e-3) / 2.5e15}
    pop.close()
    for a in ags:
        a.close()
    mlp.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph", type=int, default=1,
                    help="per-step kernel path: replay the step loop as a hipGraph")
    ap.add_argument("--step-kernels", type=int, default=0,
                    help="time the per-step kernel path (1) instead of the fused trajectory kernel (0)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--n-cand", type=int, default=1024)
    ap.add_argument("--no-prescreen", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-saturated", action="store_true")
    ap.add_argument("--no-maturation", action="store_true")
    ap.add_argument("--no-allatom", action="store_true")
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-loops", action="store_true", help="skip the c3 / c4 loop legs")
    ap.add_argument("--no-c1", action="store_true", help="skip the c1 run_allatom leg")
    ap.add_argument("--tuning", default="",
                    help="fixed pair work split 'survivors,chunks,waves' instead of autotuning")
    ap.add_argument("--traj-tuning", default="",
                    help="fixed fused-kernel shape 'waves,cand_per_group,balance,cells,chunks"
                         "[,rebalance]' instead of autotuning")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied device-to-device first (reference point for traffic counters)")
    ap.add_argument("--prescreen-train", type=int, default=1_000_000,
                    help="training rows (candidate, exact energy) for the prescreen")
    ap.add_argument("--prescreen-pool", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(candidates shard with no exchange but a top-k gather; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import dnab
    from dnab import workloads as W

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    ctx = dnab.GpuContext.try_init(0)
    if ctx is None:
        print("no gfx950 GPU visible", file=sys.stderr)
        sys.exit(1)
    info = ctx.device_info()
    log(f"[bench] device {info}")
    t = time.perf_counter()
    w = W.config2(n_cand=args.n_cand)
    log(f"[bench] workload built in {time.perf_counter() - t:.1f}s: {w.description}, "
        f"{w.antigens[0].n_atoms()} antigen atoms, {len(w.pop.atoms)} candidate atoms")
    ags, pop = W.device_objects(ctx, w)
    x0, y0, z0 = pop.coords()

    def params(n, count=False, step_kernels=args.step_kernels):
        return dnab.langevin_params(n_steps=n, score_every=w.score_every, top_k=w.top_k,
                                    restraint_radius=w.restraint_radius[0],
                                    use_graph=args.graph, count_pairs=1 if count else 0,
                                    step_kernels=step_kernels)

    def timed_run(step_kernels):
        """Wall time of the loop from the initial state, and the event-timed
        duration of its dominant kernel (per launch)."""
        pop.set_coords(x0, y0, z0)
        ctx.synchronize()
        name = "pair_ef" if step_kernels else "trajectory"
        # the fused loop is one launch: its events cost nothing; graph
        # replays carry no per-kernel events
        ctx.set_timing(not step_kernels or not args.graph)
        k0 = ctx.kernel_time(name)
        f0 = ctx.kernel_time("trajectory_final")[0]
        t0 = time.perf_counter()
        idx, _ = pop.run(params(args.steps, step_kernels=step_kernels))
        ctx.synchronize()
        dt = time.perf_counter() - t0
        k1 = ctx.kernel_time(name)
        f1 = ctx.kernel_time("trajectory_final")[0]
        ctx.set_timing(False)
        if step_kernels and args.graph:  # per-kernel time from an eager pass
            pop.set_coords(x0, y0, z0)
            ctx.set_timing(True)
            k0 = ctx.kernel_time(name)
            pop.run(dnab.langevin_params(n_steps=args.steps, score_every=w.score_every,
                                         top_k=w.top_k, restraint_radius=w.restraint_radius[0],
                                         use_graph=0, step_kernels=1))
            k1 = ctx.kernel_time(name)
            ctx.set_timing(False)
        launches = max(1, k1[1] - k0[1])
        # the fused loop's total includes its final evaluation's launch
        total = k1[0] - k0[0] + (0.0 if step_kernels else f1 - f0)
        return dt, (k1[0] - k0[0]) / launches, idx, total, launches

    # warm-up (untimed) of both paths; the autotuner picks the per-step pair
    # kernels' work split on the state the warm-up leaves (results do not
    # depend on it)
    if args.warmup:
        pop.run(params(args.warmup, step_kernels=1))
        pop.set_coords(x0, y0, z0)
        pop.run(params(args.warmup, step_kernels=0))
    if args.tuning:
        pop.set_tuning(*(int(v) for v in args.tuning.split(",")))
        tuning = pop.tuning()
        log(f"[bench] fixed pair work split {tuning}")
    else:
        tuning, tuning_ms = pop.autotune(reps=3)
        log(f"[bench] autotuned pair work split {tuning} ({tuning_ms * 1e3:.1f} us per evaluation)")
    if args.traj_tuning:
        v = [int(t) for t in args.traj_tuning.split(",")]
        pop.set_trajectory_tuning(*v)
        traj_tuning = pop.trajectory_tuning()
        log(f"[bench] fixed trajectory shape {traj_tuning}")
    else:
        # the fused loop's shape, timed from the start state the timed run uses
        pop.set_coords(x0, y0, z0)
        t = time.perf_counter()
        traj_tuning = pop.autotune_trajectory(params(min(args.steps, 50), step_kernels=0),
                                              trial_steps=min(args.steps, 50))
        log(f"[bench] autotuned trajectory shape {traj_tuning} in {time.perf_counter() - t:.1f}s")
    other = 1 - args.step_kernels
    dt, avg_ms, idx, total_ms, launches = timed_run(args.step_kernels)
    x1, y1, z1 = pop.coords()
    dt_other, avg_other, idx_other, _, _ = timed_run(other)
    xo, yo, zo = pop.coords()
    paths_identical = bool(np.array_equal(xo, x1) and np.array_equal(idx_other, idx))
    log(f"[bench] {'step kernels' if args.step_kernels else 'fused'}: {dt * 1e3:.2f} ms; "
        f"{'step kernels' if other else 'fused'}: {dt_other * 1e3:.2f} ms; identical {paths_identical}")

    # exact pair count: replay the same deterministic trajectory, counting
    pop.set_coords(x0, y0, z0)
    pop.reset_pair_stats()
    idx2, _ = pop.run(params(args.steps, count=True))
    n_in, n_tested = pop.pair_stats()
    xc, yc, zc = pop.coords()
    replay_ok = bool(np.array_equal(xc, x1) and np.array_equal(idx2, idx))
    n_evals = args.steps + 1            # K steps + the final scoring evaluation
    pairs_per_eval = n_in / n_evals
    if args.step_kernels:
        achieved = pairs_per_eval * FLOP_PER_PAIR_EF / (avg_ms * 1e-3) / 1e12
        kernel = "k_pair_chunks + k_pair_rest (energy + forces), per step"
    else:
        # the loop's launches together run every evaluation
        achieved = n_in * FLOP_PER_PAIR_EF / (total_ms * 1e-3) / 1e12
        kernel = ("k_trajectory (every step and the final evaluation, in launches of "
                  "rebalance_steps steps)")
    step_ms = {"fused": (dt_other if args.step_kernels else dt) * 1e3 / args.steps,
               "step_kernels": (dt if args.step_kernels else dt_other) * 1e3 / args.steps}

    value = n_in / dt
    out = {
        "metric": "atom-pair energy+force evals/s",
        "value": value,
        "unit": "atom-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"exact batched scoring at GPU-phase shape: 300-residue globule "
                         f"antigen vs {w.pop.n_cand:,} x 20-residue binders, {args.steps} "
                         f"gradient steps, scored every {w.score_every}, top-{w.top_k} kept"),
            "antigen_atoms": int(w.antigens[0].n_atoms()),
            "candidates": int(w.pop.n_cand),
            "candidate_atoms": int(len(w.pop.atoms)),
            "cutoff_A": w.cutoff,
            "path": "step kernels" if args.step_kernels else "fused trajectory kernel",
            "graph": bool(args.graph) if args.step_kernels else None,
            "ms_per_step_by_path": step_ms,
            "paths_bit_identical": paths_identical,
            "pairs_in_cutoff": int(n_in),
            "pairs_tested": int(n_tested),
            "evaluations": n_evals,
            "replay_identical": replay_ok,
            "pair_tuning": tuning,
            "trajectory_tuning": traj_tuning,
            "trajectory_tuning_source": "fixed" if args.traj_tuning else "autotuned",
        },
        "roofline": {
            "bound": "valu",
            "engine": "FP32 VALU (157.3 TFLOP/s, AMD's FP32 vector figure, assumes packed FP32)",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic(args.step_kernels),
            "traffic_source": ("rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE of the same kernel on "
                               "this workload, scaled by a device copy of known size: " +
                               ("profiles/pair_traffic_c2.txt" if args.step_kernels
                                else "profiles/trajectory_traffic_c2.txt")),
            "valu_issue": measured_valu(args.step_kernels),
            "valu_issue_source": ("rocprofv3 --pmc SQ_INSTS_VALU of the same kernel on this workload: "
                                  "profiles/valu_c2.txt (fraction of 1,024 SIMDs x 2,400 MHz / 4 cycles "
                                  "per wave64 instruction)"),
            "kernel": kernel,
            "kernel_avg_ms": avg_ms,
            "kernel_launches": launches,
            "kernel_total_ms": total_ms,
            "flop_per_pair": FLOP_PER_PAIR_EF,
        },
    }
    if not args.no_allatom:
        log("[bench] allatom_energy of the binders at their start state")
        out["allatom_energy"] = allatom_energy_bench(ctx, w, pop, (x0, y0, z0))
        log(f"[bench] allatom_energy: {out['allatom_energy']['ms_per_call']:.3f} ms per call")
    if not args.no_maturation:
        log("[bench] batched affinity maturation of the refined binders")
        out["maturation"] = maturation_bench(ctx, w, ags[0], (x1, y1, z1))
        log(f"[bench] maturation: {out['maturation']['ms_per_call']:.2f} ms per call, "
            f"{out['maturation']['positions']} positions")
    if not args.no_refine:
        log("[bench] refinement with Monte-Carlo moves")
        out["refine"] = refine_bench(ctx, w, ags[0], (x1, y1, z1), idx)
        log(f"[bench] refine: 64 x 600 in {out['refine']['refinement_64x600']['ms_per_call']:.1f} ms, "
            f"1024 x 200 in {out['refine']['broad_1024x200']['ms_per_call']:.1f} ms")
    if not args.no_saturated:
        log("[bench] kernels with enough work (16,384 binders in contact)")
        out["roofline"]["saturated"] = {
            "trajectory": saturated_trajectory_rate(ctx, 16384),
            "pair_kernels": saturated_pair_rate(ctx, 16384)}
    if not args.no_loops:
        log("[bench] gradient-step loops at the c3 and c4 shapes")
        loops = {"c2": {"pairs_per_s": value, "frac": achieved / PEAK_FP32_TFLOPS,
                        "tested_per_in_cutoff": n_tested / max(n_in, 1),
                        "fused_ms_per_step": step_ms["fused"],
                        "step_kernels_ms_per_step": step_ms["step_kernels"]}}
        for name in ("config3", "config4"):
            t = time.perf_counter()
            leg = loop_leg(ctx, getattr(W, name)())
            loops[name.replace("config", "c")] = leg
            log(f"[bench] {name}: fused {leg['fused']['wall_ms']:.2f} ms, step kernels "
                f"{leg['step_kernels']['wall_ms']:.2f} ms, {leg['pairs_in_cutoff']:,} pairs, "
                f"identical {leg['paths_bit_identical']} ({time.perf_counter() - t:.1f}s)")
        out["loops"] = loops
    if not args.no_c1:
        log("[bench] run_allatom at the c1 shape")
        out["c1_run_allatom"] = c1_leg(ctx)
        log(f"[bench] c1 run_allatom: {out['c1_run_allatom']['ms_per_call']:.1f} ms per call")
    if not args.no_precision:
        log("[bench] precision study at the config-4 shape")
        wp = W.config4(n_cand=1024)
        ags4, pop4 = W.device_objects(ctx, wp)
        out["precision"] = dict(workload=wp.description, antigen_atoms=int(wp.antigens[0].n_atoms()),
                                **dnab.precision_study(pop4))
        pop4.close()
        for a in ags4:
            a.close()
    if not args.no_prescreen:
        log("[bench] learned prescreen")
        out["prescreen"] = prescreen_bench(ctx, args)
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline on the loop's own states")
        states, state_pairs = sample_states(pop, params(args.steps), args.steps, 20, (x0, y0, z0))
        if not np.array_equal(states[-1][0], x1):
            raise RuntimeError("the per-step API did not reproduce the loop's trajectory")
        cb = cpu_baseline(w, states, state_pairs, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
