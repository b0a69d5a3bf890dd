"""Benchmark: element-steps/s of the batched FP32-bulk Cosserat-rod step on the
4096-rod ragged batch (n log-uniform in [16, 1024], each rod at its own
stable dt) on one MI355X, beside the fused FP64 step, the plain (op-by-op)
FP64 step and the CPU oracle.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 100] [--chunk 100]

Prints ONE JSON line.  `value` is the FP32-bulk element-steps/s over the
timed steps with the state resident in HBM (element-steps = sum of rod
element counts x steps; ghost slots are not counted), replayed as hipGraph
chunks of min(--chunk, --steps) steps per launch.

`roofline` prices the dominant kernel (the fused FP32-bulk step) by what
bounds it: it holds the rod state on chip for all steps of a launch, so HBM
is touched once per launch and VALU issue bounds it.  `achieved` is its VALU
instructions per launch (the committed rocprofv3 counter summary's
instructions per wave and step x waves x steps per launch) over its average
launch duration, measured with HIP events on the batch stream around the
timed region; `peak` is one wave64 VALU instruction per CU per cycle.
`traffic` and `hbm_measured` are its HBM bytes per launch from the counter
summary (FETCH_SIZE x 2 + WRITE_SIZE, calibrated on a plain device copy);
`equivalent_streaming` is what a kernel streaming the state every step
would move.  `roofline_streaming` prices the same kernel at one step per
launch (state through HBM every step), an HBM-bound launch.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "element-steps/s batched FP32-bulk vs plain FP64 at ≤1e-6 rel err; p99 latency"
HBM_PEAK_GBS = 8000.0


def b32(n):
    return 360 * n + 72


def b64(n):
    return 432 * n + 48


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_block(wl):
    from elasticapp_amd.workloads import build_block
    return build_block(wl)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_fp32_ragged4096_pmc.md")
CONTACT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_fp32_rods_on_plane_pmc.md")


def _summary_lines(path, tag, nfields):
    """(lanes, epl, spl, key=value dict) of the machine-readable lines `tag
    lanes=L epl=P spl=S ...` of a counter summary under profiles/."""
    out = []
    try:
        with open(path) as f:
            for line in f:
                w = line.strip().strip("`").split()
                if len(w) == nfields and w[0] == tag:
                    kv = dict(x.split("=", 1) for x in w[1:])
                    out.append((int(kv["lanes"]), int(kv["epl"]), int(kv["spl"]), kv))
    except (OSError, ValueError, KeyError):
        pass
    return out


def _closest(lines, lanes, epl, spl):
    """The line of this tile shape measured at `spl` steps per launch, else at
    the nearest step count; None without one for the shape."""
    cand = [(abs(s - spl), s, kv) for (l, e, s, kv) in lines if (l, e) == (lanes, epl)]
    if not cand:
        return None
    cand.sort(key=lambda c: c[0])
    return cand[0][1], cand[0][2]


def measured_traffic(lanes, epl, spl, tag="traffic", path=TRAFFIC_FILE, exact=True):
    """HBM bytes per launch of the FP32-bulk kernel at this tile shape, from
    the rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines of the form
    `traffic lanes=L epl=P spl=S bytes_per_launch=B` for the ragged batch,
    `traffic_contact ...` for the rods on a plane), or None.  With
    exact=False the line of the nearest steps per launch is taken (one pass
    over state and constants per launch whatever its steps) and the result is
    (bytes, spl measured at)."""
    c = _closest(_summary_lines(path, tag, 5), lanes, epl, spl)
    if c is None or (exact and c[0] != spl):
        return None
    b = float(c[1]["bytes_per_launch"])
    return b if exact else (b, c[0])


# VALU issue peak of MI355X: 256 CUs x 2.4 GHz, each CU issuing one wave64
# VALU instruction per cycle (4 SIMDs of 16 lanes, 4 cycles per wave64
# instruction), which is the 78.6 TFLOP/s FP64 vector peak counted in
# instructions (64 lanes x 2 flops per FMA)
VALU_PEAK_GINSTR = 256 * 2.4


def measured_valu(lanes, epl, spl, exact=True, tag="valu", path=TRAFFIC_FILE):
    """(waves per launch, VALU instructions per wave and step) of the
    FP32-bulk kernel at this tile shape on the ragged batch, from the
    rocprofv3 SQ counter summary (`valu lanes=L epl=P spl=S waves=W
    insts_per_wave_step=I`; `valu_contact ...` for the rods on a plane), or
    None.  With exact=False the line of the nearest steps per launch is taken
    and its step count returned third."""
    c = _closest(_summary_lines(path, tag, 6), lanes, epl, spl)
    if c is None or (exact and c[0] != spl):
        return None
    r = (int(c[1]["waves"]), float(c[1]["insts_per_wave_step"]))
    return r if exact else r + (c[0],)


def persistent_roofline(n, kms, launches, steps, lanes, epl, spl, traffic_tag="traffic", valu_tag="valu",
                        path=TRAFFIC_FILE):
    """The roofline of FP32-bulk launches that hold the state on chip for
    their steps: HBM is touched once per launch (state in and out, constants
    in), so what bounds a launch is VALU issue: its VALU instructions
    (counter summary, per wave and step) over its event-timed duration,
    against one wave64 VALU instruction per CU per cycle.  The measured HBM
    traffic and the bytes a kernel streaming the state every step would move
    are reported beside it."""
    steps_launch = steps / launches
    bytes_alg = float(b32(n).sum()) * steps_launch
    traffic = measured_traffic(lanes, epl, spl, tag=traffic_tag, path=path, exact=False)
    valu = measured_valu(lanes, epl, spl, exact=False, tag=valu_tag, path=path)
    roofline = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINSTR, "unit": "G wave64 VALU instr/s",
                "frac": None, "traffic": traffic[0] if traffic else None,
                "kernel": "k_fused<float,%d,%d> (FP32-bulk)" % (epl, lanes // epl),
                "kernel_ms_per_launch": kms, "launches": launches, "steps_per_launch": steps_launch}
    if valu is not None:
        waves, ipws, valu_spl = valu
        g = waves * ipws * steps_launch / (kms * 1e-3) / 1e9
        roofline.update(achieved=g, frac=g / VALU_PEAK_GINSTR, insts_per_wave_step=ipws, waves_per_launch=waves,
                        counters_at_steps_per_launch=valu_spl)
    if traffic is not None:
        hbm = traffic[0] / (kms * 1e-3) / 1e9
        roofline["hbm_measured"] = {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS,
                                    "bytes_per_launch": traffic[0], "counters_at_steps_per_launch": traffic[1]}
    roofline["equivalent_streaming"] = {
        "bytes_per_launch": bytes_alg, "GB_per_s": bytes_alg / (kms * 1e-3) / 1e9, "unit": "GB/s",
        "note": "B32(n) = 360 n + 72 bytes per rod-step times the launch's steps: what a kernel streaming the "
                "state through HBM every step would move; not a fraction of peak (the state stays on chip)"}
    return roofline


def stream_copy_gbs(torch, nbytes=2 << 30, reps=10):
    """Device-to-device copy rate (read + write bytes / s) of a 2 GiB buffer:
    the measured HBM ceiling beside the 8 TB/s peak."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    fn()
    sync()
    return time.perf_counter() - t0


def cpu_model_name():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(wl, budget_s=12.0, n_sample=512):
    """The oracle (FP64 CPU restatement) on a bounded sample of the same batch:
    every k-th rod of the ragged batch (n_sample rods spread over the whole
    length distribution) packed into blocks of <= 1024 elements, one
    std::thread per host core of one GPU's share, timed over enough steps to
    fill ~budget_s."""
    from elasticapp_amd.workloads import Workload
    from tests.gpu_helpers import oracle_system
    threads = min(16, len(os.sched_getaffinity(0)))  # the host share of one GPU
    stride = max(1, len(wl.n_elems) // n_sample)
    idx = list(range(0, len(wl.n_elems), stride))[:n_sample]
    pos = {r: k for k, r in enumerate(idx)}  # end forces follow their rods into the sample's numbering
    sub = Workload("cpu_sample", [wl.n_elems[i] for i in idx], [wl.states[i] for i in idx], wl.dt, 0,
                   forces=[(pos[f[0]],) + tuple(f[1:]) for f in wl.forces if f[0] in pos])
    sysm = oracle_system(sub, max_block_elems=1024)
    elems = sum(sub.n_elems)
    t = time.perf_counter()
    sysm.step(wl.dt, 5, threads)
    probe = time.perf_counter() - t
    steps = int(max(10, min(50000, budget_s / max(probe / 5, 1e-6))))
    t = time.perf_counter()
    sysm.step(wl.dt, steps, threads)
    el = time.perf_counter() - t
    return {"value": elems * steps / el, "unit": "element-steps/s", "cores": threads, "kind": "port",
            "sample": "every %dth rod of the batch (%d rods, %d elements, %.0f%% of its elements) in blocks of "
                      "<=1024 elements, %d FP64 steps, %.1f s on %d threads of %s (oracle built -O3 "
                      "-march=x86-64-v3 -ffp-contract=off)"
                      % (stride, len(idx), elems, 100.0 * elems / sum(wl.n_elems), steps, el, threads,
                         cpu_model_name())}


def contact_leg(args, torch):
    """Rods on a plane under gravity with friction: 100,000 rods x 128
    elements, FP32-bulk fused steps in hipGraph chunks, beside a few plain
    FP64 steps."""
    from elasticapp_amd import FP32_BULK
    from elasticapp_amd.workloads import rods_on_plane
    wl = rods_on_plane(steps=args.contact_steps)
    n = np.array(wl.n_elems)
    elems = int(n.sum())
    blk = make_block(wl)

    def sync():
        blk.synchronize()
        torch.cuda.synchronize()

    blk.set_fused_config(0, 0, args.chunk)
    blk.autotune(FP32_BULK, args.chunk)
    blk.step_fused_graph(args.chunk, args.chunk, FP32_BULK)
    sync()
    el = timed(lambda: blk.step_fused_graph(args.contact_steps, args.chunk, FP32_BULK), sync)
    kms, launches = blk.last_kernel_ms()
    lanes, epl, spl = blk.fused_config()
    flags = blk.check_finite()
    assert not flags.any(), "non-finite state in %d contact rods" % int(flags.sum())
    z = blk.get_field("Position")[2]
    zmin = float(min(z[s0:s0 + k + 1].min() for s0, k in zip(blk.starts, n)))  # rod nodes, not ghosts
    v = elems * args.contact_steps / el
    roofline = persistent_roofline(n, kms, launches, args.contact_steps, lanes, epl, spl, traffic_tag="traffic_contact",
                                   valu_tag="valu_contact", path=CONTACT_TRAFFIC_FILE)
    log("rods on a plane: %d rods x %d, %.3f s for %d steps -> %.3e element-steps/s, VALU issue %s"
        % (len(n), n[0], el, args.contact_steps, v,
           "%.2f" % roofline["frac"] if roofline["frac"] is not None else "(no counter summary for this shape)"))
    # FP32-bulk against fused FP64 over the same steps (each rod's positions
    # relative to its length), before the plain steps below move blk on
    from elasticapp_amd import FP64
    b64c = make_block(wl)
    b64c.set_fused_config(0, 0, args.chunk)
    b64c.step_fused_graph(args.chunk + args.contact_steps, args.chunk, FP64)
    x32, x64 = blk.get_field("Position"), b64c.get_field("Position")
    b64c.close()
    Lc = np.array([st["ReferenceElementLength"].sum() for st in wl.states])
    cerr = np.array([np.abs(x32[:, s0:s0 + k + 1] - x64[:, s0:s0 + k + 1]).max() for s0, k in zip(blk.starts, n)]) / Lc
    log("rods on a plane: fp32-bulk vs fp64 after %d steps: max |dx|/L %.3e (median %.3e)"
        % (args.chunk + args.contact_steps, cerr.max(), np.median(cerr)))
    # plain FP64 steps continue from the fused state on the same batch
    blk.step(wl.dt, 2)
    sync()
    elp = timed(lambda: blk.step(wl.dt, args.contact_plain_steps), sync)
    blk.close()
    vp = elems * args.contact_plain_steps / elp
    log("rods on a plane, plain fp64: %.3e element-steps/s" % vp)
    cpu = None if args.no_cpu else _cpu_rate(wl, 1024, 5.0, min(16, len(os.sched_getaffinity(0))))
    return {"cpu_baseline": cpu,
            "workload": "rods_on_plane: %d rods x %d elements, L 1 m, r 0.01 m, rho 1000, E 1e6, rest curvature "
                        "N(0,0.5), gravity -9.81 z, plane z=0 k 1e3 c 10 mu 0.3 ct 10" % (len(n), n[0]),
            "elements": elems, "dt": wl.dt, "steps": args.contact_steps, "value": v, "unit": "element-steps/s",
            "ms_per_step": el * 1e3 / args.contact_steps, "tile_lanes": lanes, "elements_per_thread": epl,
            "steps_per_launch": spl, "min_z": zmin,
            "fp32_bulk_position_err_vs_fp64": {"max_over_rods": float(cerr.max()), "median_over_rods": float(np.median(cerr)),
                                               "steps": args.chunk + args.contact_steps},
            "roofline": roofline,
            "fp64_plain": {"value": vp, "steps": args.contact_plain_steps}, "ratio_fp32bulk_vs_plain_fp64": v / vp}


def _cpu_rate(wl, n_rods, budget_s, threads):
    """Oracle element-steps/s on the first n_rods of a workload, blocks of <=
    1024 elements, about budget_s seconds of stepping."""
    from elasticapp_amd.workloads import Workload
    from tests.gpu_helpers import oracle_system
    sub = Workload("cpu_sample", wl.n_elems[:n_rods], wl.states[:n_rods], wl.dt, 0,
                   forces=wl.forces[:n_rods] if wl.forces else [], clamps=wl.clamps[:n_rods] if wl.clamps else [],
                   gravity=wl.gravity, contact=wl.contact)
    sysm = oracle_system(sub, max_block_elems=1024)
    elems = sum(sub.n_elems)
    t = time.perf_counter()
    sysm.step(wl.dt, 5, threads)
    probe = time.perf_counter() - t
    steps = int(max(10, min(50000, budget_s / max(probe / 5, 1e-7))))
    t = time.perf_counter()
    sysm.step(wl.dt, steps, threads)
    el = time.perf_counter() - t
    return {"value": elems * steps / el, "unit": "element-steps/s", "cores": threads, "kind": "port",
            "sample": "%d rods (%d elements), %d FP64 steps, %.1f s" % (n_rods, elems, steps, el)}


def small_batches_leg(args, torch, cpu):
    """The Timoshenko cantilever (1 rod x 100, 50,000 steps) and the uniform
    batch (1024 rods x 64, 10,000 steps): FP64 fused and plain rates, the
    fused FP64 error against the oracle for the cantilever, the roofline
    fraction of the uniform batch, and the oracle timed beside them."""
    from elasticapp_amd import FP32_BULK, FP64
    from elasticapp_amd.workloads import build_block, timoshenko_cantilever, uniform_1024x64
    from tests.gpu_helpers import oracle_system, rel_state_error
    threads = min(16, len(os.sched_getaffinity(0)))
    out = {}
    wl = timoshenko_cantilever(steps=50000)
    blk = build_block(wl)
    sync = lambda: (blk.synchronize(), torch.cuda.synchronize())
    blk.step_fused(1000, FP64)
    blk = build_block(wl)
    el = timed(lambda: blk.step_fused(wl.steps, FP64), sync)
    sysm = oracle_system(wl, max_block_elems=1024)
    t = time.perf_counter()
    sysm.step(wl.dt, wl.steps, 1)
    cpu_el = time.perf_counter() - t
    err = rel_state_error(blk, sysm)
    plain = build_block(wl)
    elp = timed(lambda: plain.step(wl.dt, 1000), lambda: (plain.synchronize(), torch.cuda.synchronize()))
    # FP32-bulk over the same steps: its tip against FP64, relative to the rod length
    b32c = build_block(wl)
    b32c.step_fused(wl.steps, FP32_BULK)
    L0 = float(wl.states[0]["ReferenceElementLength"].sum())
    err32 = float(np.abs(b32c.get_rod("Position", 0) - blk.get_rod("Position", 0)).max() / L0)
    b32c.close()
    out["timoshenko_cantilever"] = {"steps": wl.steps, "fp64_fused": 100 * wl.steps / el, "fp64_plain": 100 * 1000 / elp,
                            "unit": "element-steps/s", "max_rel_err_vs_oracle": max(err.values()),
                            "fp32_bulk_position_err_vs_fp64": err32,
                            "tip": blk.get_rod("Position", 0)[:, -1].tolist(),
                            "cpu_baseline": {"value": 100 * wl.steps / cpu_el, "unit": "element-steps/s", "cores": 1,
                                             "kind": "port", "sample": "the whole run, 1 thread"}}
    log("timoshenko cantilever (100 elements): fp64 fused %.3e, plain %.3e element-steps/s, max rel err vs oracle %.1e, "
        "fp32-bulk position error %.2e" % (out["timoshenko_cantilever"]["fp64_fused"], out["timoshenko_cantilever"]["fp64_plain"],
                                            max(err.values()), err32))
    wl = uniform_1024x64(steps=10000)
    n = np.array(wl.n_elems)
    blk = build_block(wl)
    blk.set_fused_config(0, 0, 100)
    blk.autotune(FP64, 100)
    sync = lambda: (blk.synchronize(), torch.cuda.synchronize())
    el = timed(lambda: blk.step_fused_graph(wl.steps, 100, FP64), sync)
    kms, launches = blk.last_kernel_ms()
    eq = float(b64(n).sum()) * (wl.steps // launches) / (kms * 1e-3) / 1e9
    elems = int(n.sum())
    # FP32-bulk over the same steps: its rate and its position error against FP64
    b32_ = build_block(wl)
    b32_.set_fused_config(0, 0, 100)
    b32_.autotune(FP32_BULK, 100)  # trial launches from a snapshot, restored afterwards
    sync32 = lambda: (b32_.synchronize(), torch.cuda.synchronize())
    el32 = timed(lambda: b32_.step_fused_graph(wl.steps, 100, FP32_BULK), sync32)
    x32, x64 = b32_.get_field("Position"), blk.get_field("Position")
    L = np.array([st["ReferenceElementLength"].sum() for st in wl.states])
    err32 = max(float(np.abs(x32[:, s0:s0 + k + 1] - x64[:, s0:s0 + k + 1]).max() / L[i])
                for i, (s0, k) in enumerate(zip(blk.starts, n)))
    b32_.close()
    blk.close()
    out["uniform_1024x64"] = {"steps": wl.steps, "fp64_fused": elems * wl.steps / el,
                              "fp32_bulk": elems * wl.steps / el32, "unit": "element-steps/s",
                              "fp32_bulk_position_err_vs_fp64": err32,
                              "fp64_fused_kernel_ms_per_launch": kms, "fp64_fused_launches": launches,
                              "fp64_equivalent_streaming_gbs": eq,
                              "note": "the fused FP64 launch holds the state on chip for 100 steps, so its "
                                      "equivalent streaming bytes (432 n + 48 per rod-step) over the launch time are "
                                      "no HBM fraction",
                              "cpu_baseline": None if not cpu else _cpu_rate(wl, 1024, 5.0, threads)}
    log("1024 x 64 uniform rods: fp64 fused %.3e, fp32-bulk %.3e element-steps/s, fp32-bulk position error %.2e"
        % (out["uniform_1024x64"]["fp64_fused"], out["uniform_1024x64"]["fp32_bulk"], err32))
    return out


def _served_error_vs_fp64(wl, dts, sample):
    """FP32-bulk served results against FP64: the sampled requests' rods
    stepped again in one FP64 fused batch, each at its own dt for its
    request's steps (step budgets), positions compared per rod relative to
    its length.  `sample` = [(rods, steps, [Position of each rod])]."""
    from elasticapp_amd import FP64
    from elasticapp_amd.workloads import Workload, build_block
    rods = [int(r) for rr, _, _ in sample for r in rr]
    steps = [h for rr, h, _ in sample for _ in rr]
    got = [x for _, _, xs in sample for x in xs]
    force = {f[0]: f for f in wl.forces}
    sub = Workload("served_sample", [wl.n_elems[r] for r in rods], [wl.states[r] for r in rods], wl.dt, 0,
                   forces=[(i,) + tuple(force[r][1:]) for i, r in enumerate(rods) if r in force])
    blk = build_block(sub)
    blk.set_rod_dt(np.array([dts[r] for r in rods]))
    for i, h in enumerate(steps):
        blk.set_step_budget(i, h)
    blk.step_fused_graph(max(steps), 100, FP64)
    err = np.empty(len(rods))
    for i, r in enumerate(rods):
        L = float(wl.states[r]["ReferenceElementLength"].sum())
        err[i] = np.abs(got[i] - blk.get_rod("Position", i)).max() / L
    blk.close()
    h = np.array(steps)
    short = err[h <= 10000]
    return {"max_over_rods": float(err.max()), "median_over_rods": float(np.median(err)), "rods": len(rods),
            "requests": len(sample), "max_steps": int(max(steps)),
            "max_over_rods_horizon_le_1e4": float(short.max()) if short.size else None,
            "rods_horizon_le_1e4": int(short.size),
            "note": "positions of every rod of every 40th served request (FP32-bulk, 500 steps per tick) against "
                    "the same rods stepped by the fused FP64 path, |dx| / rod length"}


def serving_leg(args, torch, capacity):
    """Continuous batching: Poisson arrivals of requests (1-64 template rods of
    32-512 elements, horizons of 1e3-1e5 steps) served by the FP32-bulk fused
    kernel from a pool of rod slots that grows on demand within the 200 GB
    HBM budget (erod_server_create_budgeted; requests admitted in order, a
    waiting request reserving its slot classes); latency = completion -
    arrival.  The arrival rate is fixed (`--serve-rate`, requests per
    second), so that p50 / p99 are comparable between runs and builds; the
    server's own capacity, measured first as the element-steps per second it
    advances on a backlog of requests, is reported beside it with the load the
    rate offers (`capacity` is the batched rate of the ragged batch, for
    comparison).  Every 40th request's result is checked against FP64."""
    from elasticapp_amd import FP32_BULK, Server
    from elasticapp_amd.workloads import build_block, serving_library, serving_requests
    wl, dts = serving_library()
    lib = build_block(wl)
    lib.set_rod_dt(dts)
    n_lib = np.array(wl.n_elems)
    reqs = serving_requests(args.serve_requests, len(n_lib))
    work = np.array([n_lib[r].sum() * h for r, h in reqs], dtype=np.float64)
    # mean work of the request distribution (sampled over many draws, not just this run's)
    mean_work = float(np.mean([n_lib[r].sum() * h for r, h in serving_requests(20000, len(n_lib), seed=99)]))
    capacities = [64, 128, 256, 512]
    budget = 200e9
    spt = args.serve_steps_per_tick
    # the server's capacity: a backlog of 3,000 requests, admitted as the
    # budget allows; 20 ticks timed after 3 that admit and grow the pool
    sat = Server.budgeted(lib, capacities, precision=FP32_BULK, steps_per_tick=spt, hbm_budget_bytes=budget)
    for i, (rods, h) in enumerate(serving_requests(3000, len(n_lib), seed=3006)):
        sat.submit(i, rods, h)
    for _ in range(3):
        for rid in sat.tick():
            sat.release(rid)
    t0 = time.perf_counter()
    k0, e0 = sat.stats()
    for _ in range(20):
        for rid in sat.tick():
            sat.release(rid)
    sat_wall = time.perf_counter() - t0
    k1, e1 = sat.stats()
    server_capacity = (e1 - e0) / sat_wall
    sat_active = sat.status()["active_elements"]
    sat.close()
    log("serving, server capacity: %.3e element-steps/s over %d ticks with %d active elements"
        % (server_capacity, k1 - k0, sat_active))
    lam = args.serve_rate
    rng = np.random.default_rng(3005)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, len(reqs)))
    srv = Server.budgeted(lib, capacities, precision=FP32_BULK, steps_per_tick=spt, hbm_budget_bytes=budget)
    lat = np.full(len(reqs), np.nan)
    nxt = done = ticks = non_finite = 0
    max_active = max_pool = 0
    sample = []
    t0 = time.perf_counter()
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            srv.submit(nxt, reqs[nxt][0], reqs[nxt][1])
            nxt += 1
        st = srv.status()
        if st["running"] == 0 and st["queued"] == 0 and nxt < len(reqs):
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
            continue
        for rid in srv.tick():
            lat[rid] = time.perf_counter() - t0 - arrivals[rid]
            x = srv.read(rid, 0, "Position")     # the result leaves the device
            non_finite += int(not np.isfinite(x).all())
            if rid % 40 == 0:
                rods, h = reqs[rid]
                sample.append((rods, h, [srv.read(rid, k, "Position") for k in range(len(rods))]))
            srv.release(rid)
            done += 1
        st = srv.status()
        max_active = max(max_active, st["active_elements"])
        max_pool = max(max_pool, st["pool_bytes"])
        ticks += 1
    wall = time.perf_counter() - t0
    info = srv.pool_info()
    srv.close()
    err = _served_error_vs_fp64(wl, dts, sample) if sample else None
    lib.close()
    thr = float(work.sum()) / wall
    out = {"workload": "%d requests, U{1..64} rods of %d templates (n logU[32,512], materials as the ragged batch, "
                       "own stable dt), horizons logU[1e3,1e5] steps, Poisson arrivals" % (len(reqs), len(n_lib)),
           "lambda_per_s": lam, "lambda": "fixed", "server_capacity_element_steps_per_s": server_capacity,
           "server_capacity_ticks": k1 - k0, "server_capacity_active_elements": sat_active,
           "ragged_batch_element_steps_per_s": capacity,
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "mean_s": float(np.mean(lat)), "max_s": float(np.max(lat)), "throughput": thr, "unit": "element-steps/s",
           "non_finite_results": non_finite,
           "utilization": thr / server_capacity, "offered_load_fraction": lam * mean_work / server_capacity,
           "wall_s": wall, "ticks": ticks, "steps_per_tick": spt,
           "admission": "HBM budget (pool grows on demand), in order, a waiting request reserving its classes",
           "slot_classes": capacities, "pool_growths": info["growths"], "pool_slots": info["slots"],
           "pool_bytes_max": max_pool, "active_elements_max": max_active, "hbm_budget_bytes": budget,
           "fp32_bulk_position_err_vs_fp64": err}
    log("serving: lambda %.1f/s, p50 %.3f s, p99 %.3f s, %.3e element-steps/s (%.0f%% of the server's capacity), "
        "%d ticks, pool grown %d times to %d slots (%.2f GB)"
        % (lam, out["p50_s"], out["p99_s"], thr, 100 * thr / server_capacity, ticks, info["growths"], info["slots"],
           max_pool / 1e9))
    if err:
        log("serving: fp32-bulk vs fp64 on %d rods of %d requests: max |dx|/L %.3e, median %.3e; "
            "%d rods with horizons up to 1e4 steps: max %s"
            % (err["rods"], err["requests"], err["max_over_rods"], err["median_over_rods"],
               err["rods_horizon_le_1e4"], "%.3e" % err["max_over_rods_horizon_le_1e4"]
               if err["max_over_rods_horizon_le_1e4"] is not None else "-"))
    return out


def run_steps(blk, n, spl, precision):
    """n fused steps of `spl` steps per launch: whole hipGraph chunks of spl
    steps, the rest (n % spl, fewer than one chunk) as one direct launch."""
    if n > 0:
        blk.step_fused_graph(n, spl, precision)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000, help="timed steps of the ragged batch")
    ap.add_argument("--warmup", type=int, default=100, help="untimed steps before them")
    ap.add_argument("--chunk", type=int, default=100,
                    help="steps per launch and per captured hipGraph replay (at most --steps)")
    ap.add_argument("--horizon", type=int, default=10000,
                    help="steps after which the FP32-bulk positions are compared with FP64 (at least warmup + steps)")
    ap.add_argument("--plain-steps", type=int, default=200, help="steps timed on the op-by-op FP64 path")
    ap.add_argument("--stream-steps", type=int, default=200, help="steps timed at one step per launch")
    ap.add_argument("--contact-steps", type=int, default=2000, help="steps of the contact batch (0: skip it)")
    ap.add_argument("--contact-plain-steps", type=int, default=20)
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests of the serving leg (0: skip it)")
    ap.add_argument("--serve-steps-per-tick", type=int, default=500, help="fused steps per server tick")
    ap.add_argument("--serve-rate", type=float, default=450.0, help="Poisson arrival rate of the serving leg (1/s)")
    ap.add_argument("--small-batches", type=int, default=1, help="1: also run the cantilever and the uniform batch")
    ap.add_argument("--tile", default="", help="LANES,EPL: fixed tile shape instead of the autotuner's choice")
    ap.add_argument("--tile64", default="", help="LANES,EPL[,SPLIT_LANES,SPLIT_EPL]: fixed fused FP64 plan")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rods", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: liberod has no multi-GPU path (rods are independent, so a batch "
              "would shard by rods with no exchange per step)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.chunk < 1 or args.contact_steps < 0:
        print("--steps and --chunk must be positive, --warmup and --contact-steps not negative", file=sys.stderr)
        sys.exit(2)
    # steps per launch of the timed region: a launch holds the rod state on
    # chip for its steps, so the timed steps are whole launches of at most
    # --chunk steps (fewer timed steps than a chunk: one launch of all of them)
    spl = min(args.chunk, args.steps)

    import torch
    from elasticapp_amd import FP32_BULK, FP64
    from elasticapp_amd.workloads import ragged_4096

    torch.cuda.init()
    dev_name = torch.cuda.get_device_name(0)
    # compute units this process sees: the persistent launches hold one
    # workgroup per CU, so their time goes by rounds of tiles over the CUs
    dev_cus = int(torch.cuda.get_device_properties(0).multi_processor_count)
    log("device", dev_name)
    wl = ragged_4096(n_rods=args.rods, steps=args.steps)
    n = np.array(wl.n_elems)
    elems = int(n.sum())
    log("workload ragged_4096: %d rods, %d elements, dt per rod %.3e .. %.3e; %d timed steps after %d warm-up steps, "
        "%d steps per launch" % (len(n), elems, float(np.min(wl.info["rod_dt"])), float(np.max(wl.info["rod_dt"])),
                                 args.steps, args.warmup, spl))

    copy_gbs = stream_copy_gbs(torch)
    log("device-to-device copy: %.0f GB/s" % copy_gbs)

    # ---- FP32-bulk: the timed steps ------------------------------------------------------
    # every rod steps with its own stable dt (the axial and rotational limits)
    rod_dt = np.asarray(wl.info["rod_dt"])
    blk32 = make_block(wl)
    blk32.set_rod_dt(rod_dt)
    if args.tile:
        lanes, epl = (int(x) for x in args.tile.split(","))
        blk32.set_fused_config(lanes, epl, spl)
        log("fixed shape (lanes, elements/thread, steps/launch) = %s" % (blk32.fused_config(),))
    else:
        blk32.set_fused_config(0, 0, spl)
        # one trial launch of `spl` steps per tile shape
        tune_ms = blk32.autotune(FP32_BULK, spl)
        log("autotuned shape (lanes, elements/thread, steps/launch) = %s, second group %s, %.4f ms/step"
            % (blk32.fused_config(), blk32.fused_split(), tune_ms))

    def sync32():
        blk32.synchronize()
        torch.cuda.synchronize()

    blk32.capture_fused_graph(spl, FP32_BULK)  # the capture is set-up, not part of the timed steps
    run_steps(blk32, args.warmup, spl, FP32_BULK)
    sync32()
    el32 = timed(lambda: run_steps(blk32, args.steps, spl, FP32_BULK), sync32)
    v32 = elems * args.steps / el32
    # dominant kernel: HIP events on the batch stream around the timed launches
    kms, launches = blk32.last_kernel_ms()
    log("fp32-bulk: %.4f s for %d steps -> %.3e element-steps/s; %d launches, %.4f ms/launch"
        % (el32, args.steps, v32, launches, kms))
    lanes, epl, _ = blk32.fused_config()
    split32 = blk32.fused_split()

    # ---- fused FP64 over the same steps ----------------------------------------------------
    blk64 = make_block(wl)
    blk64.set_rod_dt(rod_dt)
    if args.tile64:
        # a fixed plan (counter runs time each dispatch apart, which misleads the autotuner)
        t64 = [int(x) for x in args.tile64.split(",")]
        blk64.set_fused_config(t64[0], t64[1], spl)
        if len(t64) == 4:
            blk64.set_fused_split(t64[2], t64[3])
    else:
        blk64.set_fused_config(0, 0, spl)
        blk64.autotune(FP64, spl)

    def sync64():
        blk64.synchronize()
        torch.cuda.synchronize()

    blk64.capture_fused_graph(spl, FP64)
    run_steps(blk64, args.warmup, spl, FP64)
    sync64()
    el64 = timed(lambda: run_steps(blk64, args.steps, spl, FP64), sync64)
    v64 = elems * args.steps / el64
    plan64 = {"tile": list(blk64.fused_config()), "second_group": list(blk64.fused_split())}
    log("fp64 fused: %.4f s for %d steps -> %.3e element-steps/s (tiles %s, second group %s)"
        % (el64, args.steps, v64, plan64["tile"], plan64["second_group"]))

    # ---- both on to the error horizon, in launches of --chunk steps (timed: the
    # sustained rates of long runs) --------------------------------------------------------
    horizon = max(args.horizon, args.warmup + args.steps)
    rest = horizon - args.warmup - args.steps
    sustained = None
    if rest > 0:
        for blk, prec in ((blk32, FP32_BULK), (blk64, FP64)):
            lz, ez, _ = blk.fused_config()
            blk.set_fused_config(lz, ez, args.chunk)
        els32 = timed(lambda: run_steps(blk32, rest, args.chunk, FP32_BULK), sync32)
        kms_s32, launches_s32 = blk32.last_kernel_ms()
        els64 = timed(lambda: run_steps(blk64, rest, args.chunk, FP64), sync64)
        sustained = {"steps": rest, "steps_per_launch": args.chunk,
                     "fp32_bulk": elems * rest / els32, "fp64_fused": elems * rest / els64,
                     "fp32_bulk_kernel_ms_per_launch": kms_s32, "fp32_bulk_launches": launches_s32,
                     "unit": "element-steps/s",
                     "note": "the steps from the timed ones up to the error horizon, in erod_step_fused_graph "
                             "chunks of steps_per_launch steps (a chunk of one kernel launch goes out as "
                             "that launch, a chunk of several as a hipGraph replay)"}
        log("sustained over %d more steps at %d steps/launch: fp32-bulk %.3e, fp64 fused %.3e element-steps/s"
            % (rest, args.chunk, sustained["fp32_bulk"], sustained["fp64_fused"]))
    flags = blk32.check_finite()
    assert not flags.any(), "non-finite FP32-bulk state in %d rods" % int(flags.sum())
    # error of the FP32-bulk positions against FP64 after the same steps:
    # per rod max |x32 - x64| over its nodes, relative to the rod length
    x32 = blk32.get_field("Position")
    x64 = blk64.get_field("Position")
    lengths = np.array([st["ReferenceElementLength"].sum() for st in wl.states])
    rod_err = np.empty(len(n))
    for i, s0 in enumerate(blk64.starts):
        rod_err[i] = np.abs(x32[:, s0:s0 + n[i] + 1] - x64[:, s0:s0 + n[i] + 1]).max() / lengths[i]
    pos_err = float(rod_err.max())
    log("fp32-bulk position error vs fp64 after %d steps: max over rods |dx|/L = %.3e (median %.3e)"
        % (horizon, pos_err, float(np.median(rod_err))))
    blk64.close()

    # the same kernel and tile shape at one step per launch: the state goes
    # through HBM every step (after the error check: these steps advance blk32)
    blk32.set_fused_config(lanes, epl, 1)
    blk32.step_fused(10, FP32_BULK)
    blk32.step_fused(args.stream_steps, FP32_BULK)
    kms_s, launches_s = blk32.last_kernel_ms()
    bytes_s = float(b32(n).sum())
    achieved_s = bytes_s / (kms_s * 1e-3) / 1e9
    log("fp32-bulk at 1 step/launch: %.4f ms/launch -> %.0f GB/s algorithmic (%.3e element-steps/s)"
        % (kms_s, achieved_s, elems / (kms_s * 1e-3)))
    # the cost of the scheduled director re-orthonormalisation (one
    # Newton-Schulz step over every rod's directors: 72 B read and written
    # per element), against the sustained step time at one every 1,000 steps
    sync32()
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        blk32.reorthonormalize()
    sync32()
    ro_ms = (time.perf_counter() - t) * 1e3 / reps
    step_ms = (kms_s32 / args.chunk) if sustained else (kms / (args.steps / launches))
    reortho = {"ms_per_call": ro_ms, "every": 1000, "cost_fraction": ro_ms / (1000 * step_ms),
               "bytes_per_call": 144.0 * elems, "GB_per_s": 144.0 * elems / (ro_ms * 1e-3) / 1e9,
               "note": "host-timed over %d calls; scheduled on the fused paths by erod_set_reortho_interval" % reps}
    log("reorthonormalisation: %.4f ms per call, %.2f%% of the step time at one every 1,000 steps"
        % (ro_ms, 100 * reortho["cost_fraction"]))
    blk32.close()

    # ---- roofline of the timed launches ----------------------------------------------------
    roofline = persistent_roofline(n, kms, launches, args.steps, lanes, epl, spl)

    # ---- plain FP64 (one kernel per reference operator) ----------------------------------
    blkp = make_block(wl)
    blkp.step(wl.dt, 5)
    blkp.synchronize()
    elp = timed(lambda: blkp.step(wl.dt, args.plain_steps), lambda: (blkp.synchronize(), torch.cuda.synchronize()))
    vp = elems * args.plain_steps / elp
    log("plain fp64: %.3f s for %d steps -> %.3e element-steps/s" % (elp, args.plain_steps, vp))
    blkp.close()

    small = small_batches_leg(args, torch, not args.no_cpu) if args.small_batches else None
    contact = contact_leg(args, torch) if args.contact_steps > 0 else None
    serving = serving_leg(args, torch, v32) if args.serve_requests > 0 else None

    cpu = None if args.no_cpu else cpu_baseline(wl)
    if cpu:
        log("cpu oracle: %.3e element-steps/s on %d threads" % (cpu["value"], cpu["cores"]))

    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el32 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 state)",
        "data": "synthetic",
        "config": {"workload": "ragged_4096: 4096 rods, n log-uniform [16,1024] (seed 2002), L U[0.5,3] m, "
                               "r U[0.005,0.05] m, E logU[1e5,1e7], rho U[500,5000], damping U[0.01,0.5], "
                               "N(0,0.5) rest curvature, N(0,1e-2) velocities, N(0,1) N tip forces; "
                               "each rod its own stable dt (axial Courant and rotational limits)",
                   "rods": int(len(n)), "elements": elems,
                   "dt_per_rod": {"min": float(rod_dt.min()), "median": float(np.median(rod_dt)),
                                  "max": float(rod_dt.max())},
                   "steps_per_launch": spl, "graph_chunk": spl,
                   "tile_lanes": lanes, "elements_per_thread": epl, "second_group": list(split32),
                   "device": dev_name, "compute_units": dev_cus},
        "roofline": roofline,
        "reortho": reortho,
        "roofline_streaming": {"bound": "hbm", "achieved": achieved_s, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": achieved_s / HBM_PEAK_GBS,
                               "traffic": measured_traffic(lanes, epl, 1),
                               "kernel_ms_per_launch": kms_s, "launches": launches_s, "steps_per_launch": 1,
                               "algorithmic_bytes_per_launch": bytes_s,
                               "element_steps_per_s": elems / (kms_s * 1e-3)},
        "stream_copy_gbs": copy_gbs,
        "sustained": sustained,
        "cantilever_and_uniform": small,
        "rods_on_plane": contact,
        "serving": serving,
        "cpu_baseline": cpu,
        "fp64_fused": {"value": v64, "ms_per_step": el64 * 1e3 / args.steps,
                       "plan": {"tile (lanes, elements/thread, steps/launch)": plan64["tile"],
                                "second group (lanes, elements/thread, tiles)": plan64["second_group"]}},
        "fp64_plain": {"value": vp, "steps": args.plain_steps},
        "ratio_fp32bulk_vs_plain_fp64": v32 / vp,
        "ratio_fp32bulk_vs_fused_fp64": v32 / v64,
        "fp32_bulk_position_err_vs_fp64": {"value": pos_err, "median_over_rods": float(np.median(rod_err)),
                                           "metric": "max over rods of max |x32 - x64| / rod length",
                                           "after_steps": horizon},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
