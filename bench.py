#!/usr/bin/env python3
"""Benchmark of the explicit spectral-element time step on one MI355X.

Workload: 512 x 512 elements of degree 4 (2049^2 GLL nodes), per-cell random
material (Vp ~ U[4000,6000], Vs = Vp/sqrt(3), rho ~ U[2500,3000], seed 2),
two Ricker sources (15 Hz), 16 receivers, dt = 0.2 CFL.  A step is one
explicit Verlet step of the whole mesh with receiver sampling; the inputs
are resident in HBM when the timed region starts.

Prints one JSON line: GLL node-updates/s of the FP64 path (the reference's
arithmetic; the headline `value`), with the FP32+Kahan and FP32 paths beside
it, the accuracy of FP32 against FP64 over the full 2000-step run, the
roofline of the dominant kernel (here and at twice the mesh width, whose
state exceeds the 256 MB Infinity Cache) and the reference algorithm timed
on the host CPU (oracle/: one core and all cores, with and without the
per-step receiver file), plus the ensemble and surrogate workloads.  The
timed advances are three repetitions of --steps steps after --warmup steps,
reported as their median; their buffers are sized beforehand, so none of
them allocates.  Other legs (one JSON line each): --variants MX P STEPS times
every element-kernel variant at MX x MX elements of degree P; --accuracy
N1,N2,... reports the FP32 modes' receiver error against FP64 as the run
grows; --split times the element and update kernels of the split schedule
apart (also part of the default line).  --gpus N > 1 exits: one GPU only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "GLL node-updates/s (FP32 vs FP64) at 512² p4; ensemble solves/s; surrogate err"
# algorithmic HBM bytes per node and step of the fused step kernel:
# read u, v, 1/M (or M); write u, v  (+ compensation read/write)
BYTES_PER_NODE = {0: 2 * 8 + 2 * 8 + 8 + 2 * 8 + 2 * 8,   # FP64: 72
                  1: 2 * 4 + 2 * 4 + 4 + 2 * 4 + 2 * 4,   # FP32: 36
                  2: 2 * 4 + 2 * 4 + 4 + 2 * 4 + 2 * 4 + 2 * 2 * 4}  # FP32 + Kahan: 52
CANONICAL_B_NODE = {0: 152, 1: 76, 2: 92}  # two-pass scheme bytes/node/step
# split schedule, algorithmic HBM bytes per node and step of each kernel:
# element kernel reads u, writes f; update kernel reads f, u, v, M (+ the
# compensation) and writes u, v (+ the compensation)
SPLIT_ELEMENT_BYTES = {0: 2 * 8 + 2 * 8, 1: 2 * 4 + 2 * 4, 2: 2 * 4 + 2 * 4}
SPLIT_UPDATE_BYTES = {0: 3 * 2 * 8 + 8 + 2 * 2 * 8, 1: 3 * 2 * 4 + 4 + 2 * 2 * 4, 2: 3 * 2 * 4 + 4 + 2 * 2 * 4 + 2 * 2 * 4}


def log(msg):
    """Progress on stderr (the result is the one JSON line on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))


# receiver record columns: ux uy vx vy ax ay curl
FIELDS = {"u": [0, 1], "v": [2, 3], "a": [4, 5], "curl": [6]}


def field_errors(tr, ref):
    return {k: rel_l2(tr[:, :, c], ref[:, :, c]) for k, c in FIELDS.items()}


def stream_copy_gbs(nbytes=1 << 30, reps=20):
    """Device copy bandwidth, (read + write bytes) / time, of two plain
    kernels over 1 GiB: the runtime's copy (b.copy_(a), a blit kernel) and an
    elementwise kernel (b = 1 * a), the better measure of what HBM delivers."""
    import torch

    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)

    def rate(f):
        for _ in range(3):
            f()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return 2 * nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9

    out = {"blit_copy": rate(lambda: b.copy_(a)), "elementwise_copy": rate(lambda: torch.mul(a, 1.0, out=b))}
    del a, b
    torch.cuda.empty_cache()
    return out


def measured_traffic(mx, p, prec, signature):
    """HBM bytes per launch of the step kernel, from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rows = json.load(f).get("rows", [])
    for r in rows:
        if r.get("mx") == mx and r.get("p") == p and r.get("precision") == prec and signature in r.get("kernel", ""):
            return r.get("traffic_bytes")
    return None


# (elements in x, elements in y, minimum waves per SIMD, node-phase prefetch
# level, lanes per element) of the register-blocked kernel's schedules, in the
# library's variant order (SE2W_KERNEL_REG + k)
REG_SCHEDULES = [(32, 4, 0, 4, 1), (64, 2, 0, 4, 1), (32, 2, 0, 2, 1), (32, 4, 0, 0, 2), (16, 4, 0, 4, 2),
                 (16, 8, 0, 4, 2), (16, 4, 0, 2, 2), (16, 4, 0, 0, 2)]
# (elements in x, elements in y, minimum waves per SIMD, node-load placement,
# lanes per element) of the FP64 register kernel's schedules
# (SE2W_KERNEL_REG64 + k)
REG64_SCHEDULES = [(32, 4, 3, 0, 2), (32, 4, 0, 0, 2), (16, 8, 0, 4, 2), (32, 4, 3, 4, 2), (32, 4, 2, 0, 1),
                   (64, 2, 2, 0, 1), (32, 4, 2, 4, 1), (16, 4, 1, 4, 2), (32, 4, 1, 4, 2)]
KERNEL_REG64 = 12


def kernel_signature(variant, p, mode):
    """The step kernel's template signature as rocprofv3 prints it."""
    if variant == 0:
        return f"tile_kernel<{p}, {mode}, true, 256>"
    if variant == 9:
        return f"mfma_kernel<{p}, {mode}>"
    if variant == 10:
        return f"line_kernel<{p}, {mode}>"
    if variant == 11:
        return f"tile_kernel<{p}, {mode}, true, 512>"
    if variant >= KERNEL_REG64:
        ex, ey, w, pf, lpe = REG64_SCHEDULES[variant - KERNEL_REG64]
        return f"reg64_kernel<{p}, {ex}, {ey}, {w}, {pf}, {lpe}>"
    ex, ey, w, pf, lpe = REG_SCHEDULES[variant - 1]
    return f"reg_kernel<{p}, {mode}, true, {ex}, {ey}, {w}, {pf}, {lpe}>"


def ens_kernel_signature(variant, p, mode):
    """The batched (ensemble) step kernel's template signature."""
    if variant == 0:
        return f"ens_tile_kernel<{p}, {mode}, 256>"
    if variant == 9:
        return f"ens_mfma_kernel<{p}, {mode}>"
    if variant == 10:
        return f"ens_line_kernel<{p}, {mode}>"
    if variant == 11:
        return f"ens_tile_kernel<{p}, {mode}, 512>"
    if variant >= KERNEL_REG64:
        ex, ey, w, pf, lpe = REG64_SCHEDULES[variant - KERNEL_REG64]
        return f"ens_reg64_kernel<{p}, {ex}, {ey}, {w}, {pf}, {lpe}>"
    ex, ey, w, pf, lpe = REG_SCHEDULES[variant - 1]
    return f"ens_reg_kernel<{p}, {mode}, {ex}, {ey}, {w}, {pf}, {lpe}>"


def kernel_name(variant, p, mode):
    what = ("LDS-staged tile" if variant == 0 else "matrix-core contractions" if variant == 9 else
            "line contractions" if variant == 10 else "LDS-staged tile, 512 threads" if variant == 11 else
            "FP64 %s per element in the reference's operation order, %dx%d elements" %
            (("lane", "lane pair")[REG64_SCHEDULES[variant - KERNEL_REG64][4] - 1],
             *REG64_SCHEDULES[variant - KERNEL_REG64][:2]) if variant >= KERNEL_REG64 else
            "%s per element, %dx%d elements" % (("thread", "lane pair")[REG_SCHEDULES[variant - 1][4] - 1],
                                                *REG_SCHEDULES[variant - 1][:2]))
    return f"{kernel_signature(variant, p, mode)}: fused element operator + Verlet update, {what}"


def run_precision(S, SC, sc, prec, steps, warmup, full_steps, profile_steps, kernel=None, reps=None):
    """`reps` timed advances of `steps` steps each (default 3, or 9 for runs
    under 200 steps, whose few milliseconds see host jitter) after `warmup` untimed
    steps; the run buffers are sized for the longest advance beforehand, and
    the timed advances are checked to allocate nothing.  Returns the median
    wall time, the per-launch kernel time and (full_steps > 0) the receiver
    rows of a separate full-length run."""
    if reps is None:
        reps = 3 if steps >= 200 else 9
    b = SC.build(sc, precision=prec)
    sv = b.solver
    if kernel is not None:
        sv.ctx.set_kernel(kernel)
    sv.reserve(max(steps, warmup, profile_steps, 1))
    caps = sv.ctx.reserved()
    if warmup > 0:
        sv.advance(warmup, b.dt)
    walls, event_ms = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        sv.advance(steps, b.dt)
        walls.append(time.perf_counter() - t0)
        event_ms.append(sv.ctx.timing()[0])
    if sv.ctx.reserved() != caps:
        raise RuntimeError("a timed advance grew the run buffers")
    # per-launch kernel time (HIP events on the library's stream)
    sv.ctx.set_profiling(True)
    sv.advance(profile_steps, b.dt)
    _, kern = sv.ctx.timing()
    sv.ctx.set_profiling(False)
    variant = sv.ctx.kernel()
    sv.close()
    out = {"wall_s": float(np.median(walls)), "wall_s_reps": walls, "event_ms": float(np.median(event_ms)),
           "event_ms_reps": event_ms,
           "kernel_ms": float(kern[0]), "dt": b.dt, "variant": variant}
    if full_steps > 0:
        bf = SC.build(sc, precision=prec)
        if kernel is not None:
            bf.solver.ctx.set_kernel(kernel)
        _, tr = bf.solver.advance(full_steps, bf.dt)
        out["trace"] = tr
        bf.solver.close()
    return out


def split_schedule_bench(S, SC, sc, precs, steps=100, warmup=10):
    """Element kernel and update kernel of the split schedule timed apart
    (HIP events around each launch on the library's stream), each against
    the HBM peak with its own algorithmic bytes per node."""
    nodes = sc.nnodes
    names = {0: "fp64", 1: "fp32", 2: "fp32k"}
    out = {}
    for prec in precs:
        log(f"split schedule, precision {prec}")
        b = SC.build(sc, precision=prec, schedule=S.SCHED_SPLIT)
        sv = b.solver
        sv.advance(warmup, b.dt)
        sv.ctx.set_profiling(True)
        sv.advance(steps, b.dt)
        run_ms, kern = sv.ctx.timing()
        variant = sv.ctx.kernel()
        sv.close()
        row = {"steps": steps, "ms_per_step": run_ms / steps,
               "element_kernel": kernel_signature(variant, sc.p, prec).replace(", true", ", false")
                                 .replace("reg64_kernel<", "reg64_element_kernel<"),
               "update_kernel": f"update_kernel<{sc.p}, {prec}>",
               "element_kernel_ms": float(kern[0]), "update_kernel_ms": float(kern[1]),
               "element_bytes_per_node": SPLIT_ELEMENT_BYTES[prec], "update_bytes_per_node": SPLIT_UPDATE_BYTES[prec]}
        for k, nb in (("element", SPLIT_ELEMENT_BYTES[prec]), ("update", SPLIT_UPDATE_BYTES[prec])):
            ms = row[k + "_kernel_ms"]
            gbs = nb * nodes / (ms * 1e-3) / 1e9 if ms > 0 else None
            row[k + "_GBs"] = gbs
            row[k + "_frac"] = gbs / HBM_PEAK_GBS if gbs else None
        out[names[prec]] = row
    return out


def cpu_info():
    """(CPU model, logical CPUs of the host, CPUs this process may run on)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_rate(sc, dt, threads, seconds_target, io_dir=None):
    """node-updates/s of the reference algorithm (dense FP64 element loop,
    oracle/ref_cpu.c) on `threads` OpenMP threads.  Setup (mesh, tables,
    mass) is excluded: a 1-step run and an n-step run are timed and the rate
    is taken over their difference of n - 1 steps.  io_dir: every step appends
    its receiver row to the receiver file there (open, write, close), as the
    reference does; otherwise no receiver output (-ignore_receiver_output)."""
    from oracle import oracle as O

    def timed(n):
        pr = O.Problem(sc, dt=dt, nthreads=threads, nt=n, rec_dir=io_dir)
        t0 = time.perf_counter()
        pr.run(max_steps=n, want_state=False, want_trace=False)
        return time.perf_counter() - t0

    t1 = timed(1)
    t3 = timed(3)
    per_step = max((t3 - t1) / 2, 1e-4)
    n = int(max(4, min(1000, 1 + seconds_target / per_step)))
    tn = timed(n)
    el = max(tn - t1, 1e-9)
    return sc.nnodes * (n - 1) / el, n - 1, el


def config1_gpu(S, SC):
    """The reference's own CPU-sized case (config 1: 64x64 elements, p = 6,
    point force, 1000 steps at 0.2 CFL) in FP64 on the GPU: wall time of
    the whole 1000-step run (setup excluded) and its receiver rows."""
    sc = SC.config1()
    b = SC.build(sc, precision=S.FP64)
    b.solver.reserve(b.nsteps)
    t0 = time.perf_counter()
    _, tr = b.solver.advance(b.nsteps, b.dt)
    t = time.perf_counter() - t0
    variant = b.solver.ctx.kernel()
    b.solver.close()
    return {"sc": sc, "dt": b.dt, "steps": b.nsteps, "seconds": t, "trace": tr,
            "kernel": kernel_signature(variant, sc.p, S.FP64)}


def cpu_baseline(sc, dt, seconds_target=8.0, config1=None):
    """The reference algorithm on the host's cores: one core, and all the
    cores this process may use (at most 16, the box's share), compute only
    and with the reference's per-step receiver file append.  With config1
    (the GPU's FP64 run of config 1), also the whole config-1 run on all
    cores, and whether the GPU's receiver rows equal its bits."""
    import tempfile

    model, nproc, usable = cpu_info()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = max(1, min(usable, env_threads or usable, 16))
    one, n1, el1 = cpu_rate(sc, dt, 1, seconds_target)
    allc, na, ela = cpu_rate(sc, dt, cores, seconds_target)
    with tempfile.TemporaryDirectory() as d:
        allio, nio, elio = cpu_rate(sc, dt, cores, seconds_target, io_dir=d)
    return {"value": allc, "unit": "node-updates/s", "cores": cores, "kind": "port",
            "sample": f"{na} steps of the {sc.mx}x{sc.my} p={sc.p} workload (FP64 dense reference element loop, "
                      f"oracle/ref_cpu.c, {cores} OpenMP threads, no receiver output): n-step run minus 1-step run "
                      f"= {ela:.1f} s",
            "single_core": {"value": one, "steps": n1, "seconds": el1},
            "all_cores": {"value": allc, "threads": cores, "steps": na, "seconds": ela},
            "all_cores_with_receiver_io": {"value": allio, "threads": cores, "steps": nio, "seconds": elio,
                                           "what": "each step appends its 16-station row to the receiver file "
                                                   "(fopen/fprintf/fclose), as se2wave.c:2904-2981 does"},
            "cpu_model": model, "nproc": nproc, "usable_cpus": usable,
            "config1": config1_cpu(config1, cores) if config1 else None}


def config1_cpu(c1, cores):
    """Config 1 in full on the host (the oracle, all cores) against the GPU's
    FP64 run of it."""
    from oracle import oracle as O

    sc = c1["sc"]
    pr = O.Problem(sc, dt=c1["dt"], nt=c1["steps"], nthreads=cores)
    t0 = time.perf_counter()
    _, tr, _, _, _ = pr.run(want_state=False)
    t = time.perf_counter() - t0
    return {"workload": f"{sc.mx}x{sc.my} elements, p={sc.p}, point force, {c1['steps']} steps (the whole run)",
            "gpu_fp64_s": c1["seconds"], "gpu_kernel": c1["kernel"], "cpu_fp64_s": t, "cpu_threads": cores,
            "speedup": t / c1["seconds"],
            "receiver_rows_bit_identical": bool(np.array_equal(tr, c1["trace"]))}


def ensemble_bench(S, SC, n=256, nt=1500):
    """The uneven ensemble workload: n instances of mx, my ~ U{64..192}, p = 6,
    FP32, per-cell random materials, one Ricker source, 8 receivers, nt steps
    each; all advanced together by one batched launch per step (HIP graph),
    against the same n instances solved one after another.  Beside it, the
    plain FP64 method on the same instances, batched the same way (each
    member bit-identical to its own FP64 solve): its throughput, and the
    receiver error of every FP32 member against it."""
    scs = [SC.config3_instance(i, nt=nt) for i in range(n)]
    nodes = sum(sc.nnodes for sc in scs)
    builts = [SC.build(sc, precision=S.FP32) for sc in scs]
    ens = S.Ensemble([b.solver for b in builts])
    best, _ = ens.autotune(steps=10, reps=2)
    for b in builts:
        b.solver.reserve(nt)
    log(f"ensemble: {n} members, batched run")
    dts = [b.dt for b in builts]
    t0 = time.perf_counter()
    res = ens.advance(nt, dts)
    t_ens = time.perf_counter() - t0
    run_ms, step_ms = ens.timing()
    tiles, grid = ens.launch()
    ens.close()
    for b in builts:
        b.solver.close()
    t_seq, errs = 0.0, []
    log("ensemble: sequential solves")
    for i, sc in enumerate(scs):
        b = SC.build(sc, precision=S.FP32)
        b.solver.reserve(nt)
        t0 = time.perf_counter()
        _, tr = b.solver.advance(nt, b.dt)
        t_seq += time.perf_counter() - t0
        errs.append(rel_l2(tr[:, :, :4], res[i][1][:, :, :4]))
        b.solver.close()
    # the plain FP64 method on the same instances, batched
    log(f"ensemble: {n} members in FP64, batched run")
    builts = [SC.build(sc, precision=S.FP64) for sc in scs]
    ens = S.Ensemble([b.solver for b in builts])
    best64, _ = ens.autotune(steps=5, reps=2)
    for b in builts:
        b.solver.reserve(nt)
    t0 = time.perf_counter()
    res64 = ens.advance(nt, dts)
    t_ens64 = time.perf_counter() - t0
    _, step_ms64 = ens.timing()
    ens.close()
    for b in builts:
        b.solver.close()
    e64 = {"u": [], "v": []}
    for i in range(n):
        tr = res64[i][1]
        e64["u"].append(rel_l2(res[i][1][:, :, 0:2], tr[:, :, 0:2]))
        e64["v"].append(rel_l2(res[i][1][:, :, 2:4], tr[:, :, 2:4]))
    return {"workload": f"{n} instances, mx, my ~ U{{64..192}}, p=6, FP32, per-cell random material, "
                        f"1 Ricker source (f0 ~ U[5,20] Hz), 8 receivers, {nt} steps each",
            "instances": n, "nodes": nodes, "steps": nt,
            "solves_per_s": n / t_ens, "sequential_solves_per_s": n / t_seq, "speedup_vs_sequential": t_seq / t_ens,
            "ensemble_s": t_ens, "sequential_s": t_seq, "batched_step_ms": step_ms,
            "node_updates_per_s": nodes * nt / t_ens,
            "kernel": ens_kernel_signature(best, 6, 1),
            "tiles": tiles, "grid": grid,
            "max_rel_l2_u_v_vs_sequential": max(errs),
            "fp64": {"solves_per_s": n / t_ens64, "ensemble_s": t_ens64, "batched_step_ms": step_ms64,
                     "node_updates_per_s": nodes * nt / t_ens64, "kernel": ens_kernel_signature(best64, 6, 0)},
            "receiver_rel_l2_vs_fp64": {"members": len(e64["u"]),
                                        "u_max": max(e64["u"]), "u_median": float(np.median(e64["u"])),
                                        "v_max": max(e64["v"]), "v_median": float(np.median(e64["v"]))}}


def beyond_cache_bench(S, SC, mx=1024, p=4, precs=(0, 2, 1), steps=100, warmup=10):
    """The fused step kernel at mx x mx elements of degree p, where the state
    (FP64: 72 B per node, 1.2 GB at 1024^2 p4) is several times the 256 MB
    Infinity Cache, so that the kernel's bytes come from HBM: mean kernel
    time (HIP events around `steps` back-to-back launches) against the HBM
    peak with the kernel's algorithmic bytes per node."""
    sc = SC.config2(nt=steps + warmup, mx=mx, p=p)
    names = {0: "fp64", 1: "fp32", 2: "fp32k"}
    out = {"mx": mx, "p": p, "nodes": sc.nnodes}
    for prec in precs:
        log(f"beyond-cache leg, {mx}x{mx} p={p}, precision {prec}")
        b = SC.build(sc, precision=prec)
        sv = b.solver
        sv.reserve(steps)
        sv.advance(warmup, b.dt)
        sv.ctx.set_profiling(True)
        sv.advance(steps, b.dt)
        _, kern = sv.ctx.timing()
        variant = sv.ctx.kernel()
        sv.close()
        ms = float(kern[0])
        gbs = BYTES_PER_NODE[prec] * sc.nnodes / (ms * 1e-3) / 1e9
        out[names[prec]] = {"kernel": kernel_signature(variant, p, prec), "kernel_ms": ms,
                            "bytes_per_node": BYTES_PER_NODE[prec],
                            "algorithmic_bytes_per_launch": BYTES_PER_NODE[prec] * sc.nnodes,
                            "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                            "node_updates_per_s": sc.nnodes / (ms * 1e-3)}
    return out


def surrogate_bench(S, SC, ntrain=4096, ntest=512, epochs=200, f0=None, timing=True, n_fp64=8):
    """The learned surrogate: training data from the batched solver (96^2,
    p=4, FP32), the MLP trained on the GPU, its held-out error against the
    solver and its time per answer against the solver's time per solve.  The
    batched FP32 solves that make the data are checked against FP64 solves
    of n_fp64 of the held-out instances.  f0: the Ricker peak frequency of
    the sources (default: the surrogate's own, 1 Hz)."""
    from se2dr_fork_amd import surrogate as SG

    f0 = SG.F0 if f0 is None else f0
    log(f"surrogate ({f0:g} Hz): generating the training set")
    tr = SG.generate(range(ntrain), log=log, f0=f0)
    te = SG.generate(range(100000, 100000 + ntest), log=log, f0=f0)
    log("surrogate: training")
    sg = SG.Surrogate(f0=f0)
    losses = sg.fit(tr, epochs=epochs, lr=1e-3)
    train_ms = sg.mlp.timing()[0]
    sg.predict(te.x)  # warm-up
    t0 = time.perf_counter()
    yp = sg.predict(te.x)
    t_wall = time.perf_counter() - t0
    fwd_ms = sg.mlp.timing()[1]
    sg.close()
    # the plain solver one instance at a time (what the surrogate replaces)
    nseq = 8 if timing else 0
    t_seq = 0.0
    for k in range(nseq):
        sc, _ = SG.instance(100000 + k, f0=f0)
        b = SC.build(sc, precision=S.FP32)
        b.solver.reserve(SG.STRIDE * SG.NSAMP)
        t0 = time.perf_counter()
        b.solver.advance(SG.STRIDE * SG.NSAMP, b.dt)
        t_seq += time.perf_counter() - t0
        b.solver.close()
    # the batched FP32 solves that made the data against FP64 solves
    e64 = []
    for k in range(n_fp64):
        sc, _ = SG.instance(100000 + k, f0=f0)
        b = SC.build(sc, precision=S.FP64)
        _, tr64 = b.solver.advance(SG.STRIDE * SG.NSAMP, b.dt)
        b.solver.close()
        e64.append(SG.rel_l2(te.y[k], SG.seismograms(tr64)))
    per_solve_batched = te.seconds / ntest
    per_answer = t_wall / ntest
    d = ((te.x[:, None, :2] - tr.x[None, :, :2]) ** 2).sum(-1)
    out = {"workload": f"{ntrain} training + {ntest} held-out FP32 solves, 96x96 elements, p=4, {f0:g} Hz Ricker point "
                        f"force, Vp = 5000 (1 + d), d ~ N(0, 5%) on 8x8 blocks; 4 receivers x (vx, vy) x 512 samples",
           "f0_hz": f0,
            "network": "66-1024-1024-1024-4096 GELU, BF16 MFMA / FP32 accumulate, Adam lr 1e-3 cosine-decayed, batch 256, "
                       f"{epochs} epochs, mirror-symmetry augmentation (4x)",
            "held_out_rel_l2": SG.rel_l2(yp, te.y),
            "held_out_median_sample_rel_l2": float(np.median(SG.per_sample_rel_l2(yp, te.y))),
            "mean_predictor_rel_l2": SG.rel_l2(np.broadcast_to(tr.y.mean(axis=0), te.y.shape), te.y),
            "nearest_source_predictor_rel_l2": SG.rel_l2(tr.y[np.argmin(d, axis=1)], te.y),
            "final_training_loss": float(losses[-1]),
            "data_generation_s": tr.seconds + te.seconds, "training_s": train_ms / 1e3,
            "surrogate_s_per_answer_wall": per_answer, "surrogate_device_ms_for_held_out": fwd_ms,
            "solver_s_per_solve_batched": per_solve_batched,
            "speedup_vs_batched_solver": per_solve_batched / per_answer,
            "training_solves_rel_l2_vs_fp64": {"solves": n_fp64, "max": max(e64) if e64 else None,
                                               "median": float(np.median(e64)) if e64 else None}}
    if nseq:
        per_solve_seq = t_seq / nseq
        out["solver_s_per_solve_sequential"] = per_solve_seq
        out["speedup_vs_sequential_solver"] = per_solve_seq / per_answer
    return out


def kernel_variant_bench(S, SC, mx, p, steps, precisions=("fp32",), full_steps=0):
    """Every element-kernel variant built for (p, precision) on the benchmark's
    material model at mx x mx elements: mean kernel time over `steps` steps
    (HIP events), node-updates/s, and with full_steps > 0 the receiver error
    of each variant after full_steps steps against the FP64 run."""
    sc = SC.config2(nt=max(steps, full_steps, 1), mx=mx, p=p) if p != 8 or mx != 2048 else SC.config5(nt=max(steps, full_steps, 1))
    near = 0
    if full_steps > 0 and p == 8 and mx == 2048:
        # config 5's receiver line sees no wave in its 500 steps: add one
        # receiver beside each source (sources peaking at t = 0) for the check
        sc = SC.with_near_source_receivers(sc)
        near = len(sc.src_xy)
    hip, _ = S.libraries()
    out = {"mx": mx, "p": p, "nodes": sc.nnodes, "steps": steps, "variants": []}
    ref = ref_state = None
    if full_steps > 0:
        b = SC.build(sc, precision=S.FP64)
        t0 = time.perf_counter()
        _, ref = b.solver.advance(full_steps, b.dt)
        out["fp64_s"] = time.perf_counter() - t0
        ref_state = b.solver.ctx.get_state()[:2]
        out["simulated_s"] = full_steps * b.dt
        b.solver.close()
        log(f"variants: FP64 reference run of {full_steps} steps in {out['fp64_s']:.1f} s")
    for name in precisions:
        prec = S.PRECISIONS[name]
        for v in range(hip.se2w_kernel_count()):
            b = SC.build(sc, precision=prec)
            try:
                b.solver.ctx.set_kernel(v)
            except S.Se2wError:
                b.solver.close()
                continue
            b.solver.ctx.set_profiling(True)
            b.solver.advance(steps, b.dt)
            _, kern = b.solver.ctx.timing()
            b.solver.ctx.set_profiling(False)
            row = {"precision": name, "variant": v, "kernel": kernel_signature(v, p, prec),
                   "kernel_ms": float(kern[0]), "node_updates_per_s": sc.nnodes / (kern[0] * 1e-3)}
            if ref is not None:
                b.solver.close()
                b = SC.build(sc, precision=prec)
                b.solver.ctx.set_kernel(v)
                _, tr = b.solver.advance(full_steps, b.dt)
                row["receiver_rel_l2_vs_fp64"] = field_errors(tr[:, : tr.shape[1] - near], ref[:, : ref.shape[1] - near])
                if near:
                    row["near_source_receiver_rel_l2_vs_fp64"] = field_errors(tr[:, -near:], ref[:, -near:])
                su, sv, _ = b.solver.ctx.get_state()
                row["state_rel_l2_vs_fp64"] = {"u": rel_l2(su, ref_state[0]), "v": rel_l2(sv, ref_state[1])}
                del su, sv
            out["variants"].append(row)
            log(f"variants: {name} variant {v} ({row['kernel']}) {row['kernel_ms']:.4f} ms")
            b.solver.close()
    return out


def accuracy_study(S, SC, nsteps_list, mx=512, p=4):
    """Receiver error of the FP32 modes against FP64 on the benchmark
    workload as the run grows: one run per precision of max(nsteps_list)
    steps, errors over the first n steps of the traces for each n."""
    n_max = max(nsteps_list)
    sc = SC.config2(nt=n_max, mx=mx, p=p)
    traces = {}
    for name in ("fp64", "fp32", "fp32k"):
        log(f"accuracy: {name}, {n_max} steps")
        b = SC.build(sc, precision=S.PRECISIONS[name])
        _, traces[name] = b.solver.advance(n_max, b.dt)
        dt = b.dt
        b.solver.close()
    out = {"mx": mx, "p": p, "dt": dt, "rows": []}
    for n in sorted(nsteps_list):
        row = {"steps": n, "simulated_s": n * dt}
        for name in ("fp32", "fp32k"):
            row[name] = field_errors(traces[name][:n], traces["fp64"][:n])
        out["rows"].append(row)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed steps (the workload's run length)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--full-steps", type=int, default=2000, help="length of the accuracy runs (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--mx", type=int, default=512)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy", action="store_true", help="skip the device copy-bandwidth measurement")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the split schedule (element and update kernels timed apart)")
    ap.add_argument("--split", action="store_true",
                    help="only the split schedule leg (one JSON line)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate measurement")
    ap.add_argument("--surrogate-frequencies", default="",
                    help="only the surrogate at each of these Ricker frequencies (Hz, comma-separated): held-out "
                         "error of a network trained on that frequency's solves (one JSON line)")
    ap.add_argument("--variants", nargs=3, type=int, metavar=("MX", "P", "STEPS"),
                    help="only time every element-kernel variant at MX x MX elements of degree P for STEPS steps "
                         "(2048 8 STEPS: the layered 2048^2 p=8 workload) and print that as JSON")
    ap.add_argument("--accuracy", default="",
                    help="comma-separated step counts: receiver error of FP32 and FP32+Kahan against FP64 "
                         "on the benchmark workload after each")
    ap.add_argument("--variant-precisions", default="fp32,fp32k",
                    help="with --variants: the precision modes to time (fp64, fp32, fp32k)")
    ap.add_argument("--variants-check", type=int, default=0,
                    help="with --variants: receiver error of each variant after this many steps against FP64")
    ap.add_argument("--ensemble-steps", type=int, default=1500)
    ap.add_argument("--precisions", default="fp64,fp32k,fp32",
                    help="the first one is the headline (value, dtype, roofline)")
    ap.add_argument("--no-beyond-cache", action="store_true",
                    help="skip the leg at twice the mesh width, whose state exceeds the Infinity Cache")
    ap.add_argument("--kernel", choices=["auto", "lds", "reg"], default="auto",
                    help="element-kernel variant of the FP32 modes (FP64 takes its autotuned variant)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU (the step would shard by element-aligned "
                                   "blocks with one interface-force exchange per step, not built), so --gpus "
                                   "must be 1"}))
        sys.exit(2)
    import se2dr_fork_amd as S
    from se2dr_fork_amd import scenarios as SC

    if args.variants:
        mx, p, steps = args.variants
        print(json.dumps(kernel_variant_bench(S, SC, mx, p, steps, tuple(args.variant_precisions.split(",")),
                                              args.variants_check)))
        return
    if args.surrogate_frequencies:
        rows = [surrogate_bench(S, SC, f0=float(f), timing=False) for f in args.surrogate_frequencies.split(",")]
        print(json.dumps({"surrogate_by_frequency": rows}))
        return
    if args.accuracy:
        print(json.dumps(accuracy_study(S, SC, [int(x) for x in args.accuracy.split(",")], mx=args.mx, p=args.p)))
        return

    sc = SC.config2(nt=max(args.full_steps, 1), mx=args.mx, p=args.p)
    precs = [S.PRECISIONS[x] for x in args.precisions.split(",")]
    if args.split:
        out = {"mx": args.mx, "p": args.p, "nodes": sc.nnodes,
               "split_schedule": split_schedule_bench(S, SC, sc, precs, steps=args.profile_steps)}
        if not args.no_copy:
            out["stream_copy_GBs"] = stream_copy_gbs()
        print(json.dumps(out))
        return
    res = {}
    for prec in precs:
        log(f"step kernel, precision {prec}")
        kern = None if (args.kernel == "auto" or prec == S.FP64) else {"lds": S.KERNEL_LDS, "reg": S.KERNEL_REG}[args.kernel]
        res[prec] = run_precision(S, SC, sc, prec, args.steps, args.warmup, args.full_steps, args.profile_steps, kern)
    head = precs[0]
    nodes = sc.nnodes
    rates = {p: nodes * args.steps / r["wall_s"] for p, r in res.items()}
    value = rates[head]
    ms_step = 1e3 * res[head]["wall_s"] / args.steps
    kern_ms = res[head]["kernel_ms"]
    achieved = BYTES_PER_NODE[head] * nodes / (kern_ms * 1e-3) / 1e9 if kern_ms > 0 else None
    errs = {}
    if S.FP64 in res and args.full_steps > 0:
        ref = res[S.FP64]["trace"]
        for p in precs:
            if p != S.FP64:
                errs[p] = field_errors(res[p]["trace"], ref)
    names = {0: "fp64", 1: "fp32", 2: "fp32k"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f64", 1: "f32", 2: "f32"}[head],
        "data": "synthetic",
        "config": {"workload": f"se2wave explicit SEM step, {args.mx}x{args.mx} elements, p={args.p} "
                               f"({nodes} GLL nodes), per-cell random material, 2 Ricker sources, "
                               f"16 receivers, {names[head]} arithmetic, fused step kernel",
                   "mx": args.mx, "my": args.mx, "p": args.p, "nodes": nodes},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic(args.mx, args.p, names[head],
                                                 kernel_signature(res[head]["variant"], args.p, head)),
                     "kernel": kernel_name(res[head]["variant"], args.p, head),
                     "bytes_per_node": BYTES_PER_NODE[head], "kernel_ms": kern_ms},
        "node_updates_per_s": {names[p]: r for p, r in rates.items()},
        "ms_per_step_by_precision": {names[p]: 1e3 * r["wall_s"] / args.steps for p, r in res.items()},
        "wall_s_reps": {names[p]: r["wall_s_reps"] for p, r in res.items()},
        "gpu_event_ms_reps": {names[p]: r["event_ms_reps"] for p, r in res.items()},
        "kernel_ms": {names[p]: r["kernel_ms"] for p, r in res.items()},
        "kernels": {names[p]: kernel_signature(r["variant"], args.p, p) for p, r in res.items()},
        "roofline_by_precision": {
            names[p]: {"bytes_per_node": BYTES_PER_NODE[p],
                       "achieved_GBs": BYTES_PER_NODE[p] * nodes / (r["kernel_ms"] * 1e-3) / 1e9,
                       "frac": BYTES_PER_NODE[p] * nodes / (r["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
            for p, r in res.items() if r["kernel_ms"] > 0},
        # the two-pass scheme's bytes (76 / 92 / 152 B per node) at the wall
        # rate: only for the 70 % of 1.05e11 node-updates/s target check, not
        # a roofline fraction of the fused kernel (which moves fewer bytes)
        "target_check": {names[p]: {"canonical_bytes_per_node": CANONICAL_B_NODE[p],
                                    "canonical_frac_at_wall_rate": rates[p] * CANONICAL_B_NODE[p] / (HBM_PEAK_GBS * 1e9),
                                    "fp32_target_node_updates_per_s": 0.7 * HBM_PEAK_GBS * 1e9 / CANONICAL_B_NODE[1]}
                         for p in rates},
        "receiver_rel_l2_vs_fp64": {names[p]: e for p, e in errs.items()},
        "receiver_error_steps": args.full_steps,
        "algorithmic_bytes_per_launch": BYTES_PER_NODE[head] * nodes,
    }
    if not args.no_beyond_cache:
        out["beyond_cache"] = beyond_cache_bench(S, SC, mx=2 * args.mx, p=args.p)
    if not args.no_split:
        out["split_schedule"] = split_schedule_bench(S, SC, sc, precs)
    if not args.no_copy:
        out["stream_copy_GBs"] = stream_copy_gbs()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(S, SC, args.ensemble_members, args.ensemble_steps)
    if not args.no_surrogate:
        out["surrogate"] = surrogate_bench(S, SC)
    if not args.no_cpu:
        log("config 1 on the GPU (FP64)")
        c1 = config1_gpu(S, SC)
        log("host CPU baseline")
        out["cpu_baseline"] = cpu_baseline(sc, res[head]["dt"], config1=c1)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
