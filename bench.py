"""Throughput of the LULESH Lagrange leap-frog cycle on one MI355X.

Workload: the Sedov blast at s=192 (7,077,888 elements, 7,189,057 nodes),
FP64, the first `--steps` cycles (default 500) from the initial state.  A
step is one cycle of the reference's driver loop: timeIncrement! +
lagrangeLeapFrog.  The domain is built on the device and stays resident in
HBM; warm-up cycles run on the same context, which is then reset to the
initial state before the timed region.

Prints ONE JSON line: zone-cycles/s over the timed cycles (the production
path: the captured one-cycle hipGraph replayed), the roofline of the
dominant kernel (algorithmic bytes per launch / mean launch time from HIP
events recorded on the context's stream over a second timed run of the
same cycles, launched one by one) and the CPU baseline (the oracle, one
thread, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "lulesh.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFS = 78.6   # MI355X FP64 (vector and matrix), with fused multiply-add
# FP64 flop per element of k_nodal (stress + FB hourglass + node update),
# counted by the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters at s=192:
# 12.37 GFLOP per launch over 7,077,888 elements
NODAL_FLOP_PER_ZONE = 1748


def alg_bytes(E, N):
    """Algorithmic HBM bytes per launch: per element (E) and node (N) of the
    work the kernel does, the footprint of the reference algorithm for those
    rows -- each array element a phase needs counted once, 8 B per double,
    4 B per index or mask, the per-corner force round trip of the
    atomic-free corner gather at 3 doubles per corner written and read,
    dead stores excluded.  k_nodal does the force phase (rows 3-6 of the
    path: nodelist 32E, positions/velocities 48N, p q v volo ss elemMass
    48E, corner forces written 192E) and the node update (row 7: corner list
    32E, corner forces read 192E, start/count 8N, nodalMass 8N, positions/
    velocities read 48N and written 48N); k_kin rows 8-10; k_qeos rows
    11-14; k_dt reads the k_qeos block minima."""
    return {
        "k_nodal": (320 * E + 48 * N) + (224 * E + 112 * N),
        "k_kin": 128 * E + 48 * N,
        "k_qeos": 196 * E,
        "k_dt": 16 * ((E + 255) // 256),
    }


def footprint_bytes(E, N):
    """Bytes each launch of this build's kernels must move at least: the
    corner forces never leave LDS and connectivity is computed, not read.
    k_nodal: positions/velocities read 48N and written 48N, nodalMass 8N,
    p q v volo ss elemMass 48E; k_kin: positions/velocities 48N, volo v 16E,
    ten outputs 80E; k_qeos: elemBC 4, delv x3 24, delx x3 24, ten inputs
    80, five outputs 40 (per element)."""
    return {
        "k_nodal": 48 * E + 104 * N,
        "k_kin": 96 * E + 48 * N,
        "k_qeos": 172 * E,
        "k_dt": 16 * ((E + 255) // 256),
    }


def bench_adjoint(nx=48, cycles=200, every=20, snapshots=10):
    """Adjoint/forward wall-time ratio: the gradient of the total internal
    energy after `cycles` cycles (hand-written adjoint kernels, checkpoints
    every `every` cycles, including its own forward pass) over the plain
    forward run of the same cycles, both from the initial state.  Beside it
    the memory-bounded binomial schedule with `snapshots` saved states and
    the store-everything end (one segment: every cycle's slot kept from the
    forward pass), same gradient bit for bit, with each scheme's memory."""
    import numpy as np
    from lulesh_amd import Domain, LuleshProblem, grad_run, grad_run_revolve, run
    d = Domain(LuleshProblem(cycles, True, nx, 11, 1, 1))
    # warm-up: the forward graph is captured and the gradient's checkpoint and
    # segment buffers are allocated (first use) outside the timed runs
    run(d, 5)
    d.reset()
    grad_run(d, cycles, every)
    d.reset()
    t0 = time.perf_counter()
    run(d, cycles)
    fwd = time.perf_counter() - t0
    d.reset()
    t0 = time.perf_counter()
    g = grad_run(d, cycles, every)
    grad = time.perf_counter() - t0
    d.reset()
    t0 = time.perf_counter()
    gr = grad_run_revolve(d, cycles, snapshots)
    grad_rv = time.perf_counter() - t0
    # every cycle's slot kept from the forward pass: no replay at all
    d.reset()
    grad_run(d, cycles, cycles)  # first use allocates the slots
    d.reset()
    t0 = time.perf_counter()
    ga = grad_run(d, cycles, cycles)
    grad_all = time.perf_counter() - t0
    E, N = d.numElem, d.numNode
    d.close()
    state, slot = 8 * (6 * N + 5 * E), 8 * (6 * N + 14 * E)
    mem_two = -(-cycles // every) * state + (every + 1) * slot
    mem_rv = min(snapshots, cycles) * state + 2 * slot
    return {"config": f"s={nx}, {cycles} cycles, checkpoint every {every}",
            "forward_ms": round(fwd * 1e3, 3), "gradient_ms": round(grad * 1e3, 3),
            "ratio": round(grad / fwd, 3), "loss": g["loss"],
            "checkpoint_MB": round(mem_two / 1e6, 1),
            "revolve": {"snapshots": snapshots, "gradient_ms": round(grad_rv * 1e3, 3),
                        "ratio": round(grad_rv / fwd, 3), "cycles_run": gr["forward_cycles"],
                        "checkpoint_MB": round(mem_rv / 1e6, 1),
                        "same_gradient": bool(np.array_equal(gr["grad_e"], g["grad_e"]))},
            "store_all": {"checkpoint_every": cycles, "gradient_ms": round(grad_all * 1e3, 3),
                          "ratio": round(grad_all / fwd, 3),
                          "checkpoint_MB": round((state + (cycles + 1) * slot) / 1e6, 1),
                          "same_gradient": bool(np.array_equal(ga["grad_e"], g["grad_e"]))}}


def bench_batched(n=64, seed=0, cap=5000):
    """Batched speedup: 64 uneven Sedov instances (the seeded ragged
    workload) run to stop time (at most `cap` cycles each) one after another
    on the GPU, over the same instances in one batch."""
    from lulesh_amd import Batch, c5_instances
    nx, scale, hg, nr = c5_instances(n, seed)
    lone = [Batch([nx[i]], [scale[i]], [hg[i]], [nr[i]]) for i in range(n)]
    for b in lone:
        b.run(2)
        b.views[0].reset()
    t0 = time.perf_counter()
    lone_cycles = [int(b.run(cap, check=False)[0]) for b in lone]
    seq = time.perf_counter() - t0
    lone_err = [b.errors()[0] for b in lone]
    for b in lone:
        b.close()
    batch = Batch(nx, scale, hg, nr)
    batch.run(2)
    batch.views[0].reset()  # re-initialises every instance of the batch
    t0 = time.perf_counter()
    cyc = batch.run(cap, check=False)
    bt = time.perf_counter() - t0
    errs = batch.errors()
    batch.close()
    if [int(c) for c in cyc] != lone_cycles or errs != lone_err:
        raise RuntimeError("batched cycle counts or errors differ from the sequential runs")
    return {"config": f"{n} Sedov instances (seed {seed}), s in [{min(nx)}, {max(nx)}], "
                      f"{sum(v ** 3 for v in nx)} elements, each to stoptime, {cap} cycles or "
                      f"its first reference error",
            "sequential_ms": round(seq * 1e3, 2), "batched_ms": round(bt * 1e3, 2),
            "speedup": round(seq / bt, 3), "cycles_min": min(lone_cycles),
            "cycles_max": max(lone_cycles), "zone_cycles": int(sum(c * v ** 3 for c, v in zip(lone_cycles, nx))),
            "reached_stoptime": int(sum(c < cap and e == 0 for c, e in zip(lone_cycles, lone_err))),
            "stopped_by_error": int(sum(e != 0 for e in lone_err))}


def bench_batch_cycles(k, n=64, seed=0):
    """The batched workload's instances advanced k cycles together, all of
    them still running (the first instance to stop does so at cycle 463)."""
    from lulesh_amd import Batch, c5_instances
    nx, scale, hg, nr = c5_instances(n, seed)
    b = Batch(nx, scale, hg, nr)
    b.run(2)
    t0 = time.perf_counter()
    cyc = b.run(2 + k)
    wall = time.perf_counter() - t0
    b.close()
    if not all(int(c) == k for c in cyc):
        raise RuntimeError("an instance stopped early")
    return {"config": f"{n} Sedov instances (seed {seed}), {sum(v ** 3 for v in nx)} elements, "
                      f"cycles 3-{2 + k}", "ms_per_cycle": round(wall * 1e3 / k, 4)}


STATE_FIELDS = ("e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd")
STATE_SCALARS = ("time", "deltatime", "cycle", "dtcourant", "dthydro")


def cpu_baseline(nx, cycles, nthreads, state=None, pin=None):
    """The oracle (the CPU restatement of the reference) timed over `cycles`
    cycles of the driver loop at s=nx: from `state` (fields and scalars of a
    running domain) when given, else from the initial state.  With `pin` the
    process runs on that one core for the measurement.  Returns the rate,
    the seconds and the oracle's end state (e, x) for a cross-check."""
    import oracle as O
    od = O.OracleDomain(nx, O.QUIRKS_REFERENCE, nthreads)
    if state is not None:
        for name, val in state["fields"].items():
            od.set_field(name, val)
        for name, val in state["scalars"].items():
            od.set(name, val)
    E = od.get("numElem")
    old = os.sched_getaffinity(0)
    if pin is not None:
        os.sched_setaffinity(0, {pin})
    try:
        t0 = time.perf_counter()
        for _ in range(cycles):
            rc = od.cycle_once()
            if rc:
                raise RuntimeError(f"oracle error {rc}")
        dt = time.perf_counter() - t0
    finally:
        if pin is not None:
            os.sched_setaffinity(0, old)
    end = {"e": od.field("e"), "x": od.field("x"), "cycle": od.get("cycle")}
    od.close()
    return E * cycles / dt, dt, end


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--cpu-cycles", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--only", choices=("adjoint", "batched"), default=None,
                    help="run only the adjoint-ratio or the batched-speedup measurement and print "
                         "its JSON object (for profiling them on their own)")
    ap.add_argument("--batch-cycles", type=int, default=0, metavar="K",
                    help="only advance the 64-instance batched workload K cycles with every "
                         "instance live, after 2 warm-up cycles, and print its time per cycle "
                         "(for profiling the batched kernels on their own)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the "
                         "reference point for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(5):
            b.copy_(a)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 5
        print(json.dumps({"calibrate_copy_bytes": n * 8, "copy_ms": round(ms, 4),
                          "copy_GB_s": round(2 * n * 8 / (ms * 1e-3) / 1e9, 1)}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs one single-GPU domain; multi-GPU (--gpus > 1) is not "
              "part of it.", file=sys.stderr)
        sys.exit(2)
    if args.only:
        print(json.dumps(bench_adjoint() if args.only == "adjoint" else bench_batched()))
        return
    if args.batch_cycles:
        print(json.dumps(bench_batch_cycles(args.batch_cycles)))
        return

    from lulesh_amd import Domain, LuleshProblem, run

    nx = args.nx
    dom = Domain(LuleshProblem(args.steps, True, nx, 11, 1, 1))
    E, N = dom.numElem, dom.numNode
    if args.warmup > 0:
        run(dom, args.warmup)
    # timed region 1: the production path (the captured one-cycle hipGraph
    # replayed, loop test on the device) -> value
    dom.reset()
    t0 = time.perf_counter()
    ncyc = run(dom, args.steps)
    wall = time.perf_counter() - t0
    if ncyc != args.steps:
        raise RuntimeError(f"ran {ncyc} cycles, expected {args.steps}")
    value = E * ncyc / wall
    # timed region 2: the same cycles launched one by one with HIP events
    # around every kernel on the context's stream -> per-kernel durations
    dom.reset()
    dom.set_timing(True)
    t0 = time.perf_counter()
    ncyc2 = run(dom, args.steps)
    wall_events = time.perf_counter() - t0
    kt = dom.kernel_times()
    dom.set_timing(False)
    if ncyc2 != args.steps:
        raise RuntimeError(f"ran {ncyc2} cycles, expected {args.steps}")

    ab = alg_bytes(E, N)
    fb = footprint_bytes(E, N)
    kernels = {}
    for name, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        avg = ms / cnt
        gbs = ab[name] / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
        fgbs = fb[name] / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
        kernels[name] = {"avg_ms": round(avg, 5), "launches": cnt,
                         "alg_bytes": ab[name], "GB_s": round(gbs, 1),
                         "footprint_bytes": fb[name], "footprint_GB_s": round(fgbs, 1)}
    dominant = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    dk = kernels[dominant]
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as fh:
                traffic = json.load(fh).get(dominant)
        except Exception:
            traffic = None
    # achieved: the algorithmic bytes of the work (the reference algorithm's
    # footprint, corner round trip included) over the launch time -- the
    # rate a kernel streaming those bytes would need; traffic: the bytes
    # this kernel moves (copy-calibrated FETCH_SIZE + WRITE_SIZE), far below
    # them for k_nodal, whose corner forces stay in LDS
    roofline = {"bound": "hbm", "achieved": dk["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dk["GB_s"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dominant, "footprint_GB_s": dk["footprint_GB_s"],
                "basis": "algorithmic bytes of the unfused reference path (corner-force round "
                         "trip included) per launch time; the kernel itself moves `traffic` bytes"}

    # k_nodal is bound by FP64 issue, not HBM: the same kernel against the
    # FP64 peak (its arithmetic is unfused mul/add for bitwise parity, so
    # half of that peak is the reachable ceiling)
    nk = kernels.get("k_nodal")
    roofline_fp64 = None
    if nk:
        tf = NODAL_FLOP_PER_ZONE * E / (nk["avg_ms"] * 1e-3) / 1e12
        roofline_fp64 = {"bound": "fp64-valu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFS,
                         "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFS, 4),
                         "kernel": "k_nodal"}

    adjoint = None if args.no_adjoint else bench_adjoint()
    batched = None if args.no_batched else bench_batched()

    cpu = cpu_par = None
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        import numpy as np
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        # the sample starts where the timed GPU cycles ended (nodes moving,
        # the blast under way), not at the initial state where nothing moves
        c0 = dom.cycle
        state = {"fields": {f: dom.get(f) for f in STATE_FIELDS},
                 "scalars": {n: dom.get(n) for n in STATE_SCALARS}}
        allowed = sorted(os.sched_getaffinity(0))
        pin = allowed[0]
        v, secs, end = cpu_baseline(nx, args.cpu_cycles, args.cpu_threads, state, pin)
        # the same cycles on the GPU must end bit for bit where the oracle did
        run(dom, c0 + args.cpu_cycles)
        same = (dom.cycle == end["cycle"] and
                np.array_equal(dom.get("e").view(np.uint64), end["e"].view(np.uint64)) and
                np.array_equal(dom.get("x").view(np.uint64), end["x"].view(np.uint64)))
        cpu = {"value": round(v, 1), "unit": "zone-cycles/s", "cores": args.cpu_threads,
               "kind": "port",
               "sample": f"CPU oracle (C++ restatement), s={nx}, cycles {c0 + 1}-{c0 + args.cpu_cycles} "
                         f"of the same workload started from the GPU's state after cycle {c0}, "
                         f"{secs:.1f} s, {args.cpu_threads} thread(s) pinned to core {pin}",
               "sample_bitwise_equal_to_gpu": bool(same)}
        # the same sample over the host's cores (OpenMP; bit-identical results)
        nthr = max(1, min(16, len(allowed)))
        vp, secs_p, _ = cpu_baseline(nx, args.cpu_cycles, nthr, state)
        cpu_par = {"value": round(vp, 1), "unit": "zone-cycles/s", "cores": nthr,
                   "sample": f"same sample, {secs_p:.1f} s, {nthr} OpenMP threads"}

    out = {
        "metric": "FP64 zone-cycles/s at s=192; adjoint/forward time ratio; batched speedup",
        "value": round(value, 1),
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": ncyc,
        "warmup": args.warmup,
        "ms_per_step": round(wall / ncyc * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Sedov blast initial state built on the device",
        "config": {"workload": f"Sedov s={nx} first {ncyc} cycles", "nx": nx, "elements": E,
                   "nodes": N, "cycles": ncyc, "quirks": "reference"},
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_par,
        # verifyAndWriteFinalOutput's timing lines (src/utils.jl:128-129,161-162)
        "reference_report": {"grind_time_us_per_zone_cycle": round(wall * 1e6 / ncyc / E, 8),
                             "fom_z_per_s": round(1000.0 / (wall * 1e6 / ncyc / E), 3)},
        "kernels": kernels,
        # the second timed region (kernels launched one by one, HIP events
        # around each): its ms per cycle, for comparison with ms_per_step
        "ms_per_step_event_timed": round(wall_events / ncyc2 * 1e3, 4),
        "adjoint": adjoint,
        "batched": batched,
        "alg_bytes_per_cycle": sum(ab.values()),
        "footprint_bytes_per_cycle": sum(fb.values()),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
