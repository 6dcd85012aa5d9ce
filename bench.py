"""Device-load benchmark: BSIM3 instance-evaluations per second on one MI355X.

Workload: the seeded random CMOS logic netlist of 5,000 2-input NAND/NOR
gates (20,000 BSIM3 transistors, W/L within +-20% of 0.5u/0.13u, RC
interconnect of 1-3 segments per net).  A step is one full device load of
every BSIM3 instance, as the reference's Newton iteration performs it:
zero the outputs (the assembler's job), updateState, loadDAEVectors,
loadDAEMatrices (Master loops of N_DEV_MOSFET_B3.C:9222-9928), at a
transient iteration (charges, capacitances and limiter terms all live) with
seeded node voltages from U[-0.2, 1.5] V.  All vectors stay resident in HBM.

Prints one JSON line.  --gpus N with N > 1 exits: this build measures one
GPU (the load of one circuit is sharded across GPUs by graph partition,
which is not part of it).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes of one BSIM3 instance evaluation (full load):
# node voltages 4x8, previous store 5x8, store+state writes 10x8, stamp
# values 50x8 (F 3, Q 4, limiter 7, dF/dx 18, dQ/dx 16, rounded as counted
# for the reference data contract), indices (4 nodes + 34 stamp slots) x4,
# and the per-instance size-dependent + junction record (B3R_N + B3I_N
# doubles, read once per instance when every instance has its own record).
B3_SDP_FIELDS = 84
B3_GEO_FIELDS = 15
B3_BYTES_PER_EVAL = 32 + 40 + 80 + 400 + 152 + 8 * (B3_SDP_FIELDS + B3_GEO_FIELDS)
HBM_PEAK_GBS = 8000.0
# HBM bytes per BSIM3 eval launch (the staged split kernel the library picks
# for 20000 instances) on this workload, from the FETCH_SIZE
# and WRITE_SIZE counter passes scaled by a plain-copy calibration
# (profiles/bsim3_device_load_random_logic_5000g_hbm_traffic.md); valid for
# the default --gates only.
B3_EVAL_TRAFFIC_BYTES = 29.70e6
# FP64 operations per BSIM3 eval launch (the same kernel and workload), from
# the SQ_INSTS_VALU_{ADD,MUL,TRANS,FMA}_F64 counters (per-wave instruction
# counts; 64 lanes each, an FMA two operations): 64 (333575 + 345814 +
# 76987 + 2 x 556793)
B3_EVAL_FP64_FLOPS = 64 * (333575 + 345814 + 76987 + 2 * 556793)
FP64_PEAK_TFLOPS = 78.6
# Eval waves up to which the library's automatic choice is the LDS-staged
# split kernel (one eval wave per 64 instances; two eval waves per workgroup
# above 256), else the HBM-record kernel.
B3_STAGED_MAX_WAVES = 2 * 256


def b3_kernel_name(n_b3, variant=0):
    """The BSIM3 eval kernel a launch uses: the library's default rule for
    variant 0 (automatic), else the form the autotuner chose."""
    waves = (n_b3 + 63) // 64
    if variant == 0:
        variant = 1 if waves > B3_STAGED_MAX_WAVES else 3
    if variant == 1:
        return "bsim3_eval_kernel"
    if variant == 2:
        return "bsim3_eval_staged_kernel<false, 1>"
    return "bsim3_eval_staged_kernel<true, %d>" % (2 if waves > 256 else 1)


def autotune_choice(dev):
    """The autotuner's launch choice for a handle (xg_autotune_info)."""
    from xyce_amd._lib import lib
    ch = np.zeros(3, np.int32)
    us = np.zeros(6)
    lib().xg_autotune_info(dev.h, ch.ctypes.data, us.ctypes.data)
    names = {0: "auto", 1: "hbm", 2: "staged", 3: "staged_split"}
    return {"eval_form": names[int(ch[0])], "linear_in_bsim3_launch": bool(ch[1]),
            "source": ["default", "timed", "cached"][int(ch[2])],
            "us_per_launch": {"hbm": us[0], "hbm_linear_apart": us[1], "staged": us[2],
                              "staged_linear_apart": us[3], "staged_split": us[4],
                              "staged_split_linear_apart": us[5]},
            "variant": int(ch[0])}


def _finite(o):
    """Non-finite floats (an untimed mode's inf) as null, so the line is
    strict JSON."""
    if isinstance(o, dict):
        return {k: _finite(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_finite(v) for v in o]
    if isinstance(o, float) and not math.isfinite(o):
        return None
    return o


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


T_START = time.perf_counter()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--gates", type=int, default=5000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches only (for counter collection)")
    ap.add_argument("--mc-samples", type=int, default=256,
                    help="Monte-Carlo ring samples in the batched-transient leg")
    ap.add_argument("--mc-plain", type=int, default=16,
                    help="samples (evenly spaced) timed one circuit per launch")
    ap.add_argument("--mc-tstop", type=float, default=0.5e-9)
    ap.add_argument("--mc-dt", type=float, default=1e-12)
    ap.add_argument("--no-mc", action="store_true", help="skip the batched-transient leg")
    ap.add_argument("--mc-only", action="store_true",
                    help="run only the batched-transient leg and print its result "
                         "(for profiling that leg alone)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests in the serving leg (Poisson at --serve-rate)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-plain", type=int, default=200,
                    help="first requests of the same stream served one circuit per "
                         "launch (default: all of them)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-host-loads", action="store_true",
                    help="skip the host-pointer (PCIe-staged) load timing")
    ap.add_argument("--no-grid", action="store_true",
                    help="skip the power-grid (100k inverter) load timing")
    ap.add_argument("--no-big-tran", action="store_true",
                    help="skip the large single-circuit transients (5,000-gate logic "
                         "5 ns, power grid 2 ns)")
    ap.add_argument("--big-legs", default="logic,grid",
                    help="large single-circuit transients to run (logic, grid)")
    ap.add_argument("--big-span", type=float, default=1.0,
                    help="fraction of the logic transient's 5 ns to simulate; 1 is the "
                         "configured workload")
    ap.add_argument("--grid-span", type=float, default=0.0005,
                    help="fraction of the power grid's 2 ns to simulate (default 1 ps: "
                         "its 100,000 clock phases put a PULSE corner, where the step "
                         "control ends a step, every ~2.5 fs, so 2 ns takes ~8e5 steps)")
    ap.add_argument("--big-fixed", action="store_true",
                    help="large transients with fixed steps of --big-dt instead of the "
                         "reference's step control")
    ap.add_argument("--big-dt", type=float, default=10e-12,
                    help="fixed step of the large single-circuit transients")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 512 MiB device copy, the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def copy_calibration(torch):
    """A plain device-to-device copy of known size (512 MiB read, 512 MiB
    written), three times; its counter readings give the KiB the profiler
    reports per real HBM byte."""
    src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def graph_timing(torch, dev, stream, step, timed, args):
    """The same step captured once in a hipGraph and replayed: every kernel
    of the step still runs on every replay; only the per-launch host cost
    goes."""
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        dev.set_stream(torch.cuda.current_stream().cuda_stream)
        step()
    dev.set_stream(stream.cuda_stream)
    for _ in range(args.warmup):
        graph.replay()
    return timed(graph.replay, args.steps)


def batched_transient_leg(torch, args):
    """Monte-Carlo BSIM3 ring oscillators (stage counts U{11..301}, per-device
    VTH0 += N(0, 30 mV), TOX *= U[0.98, 1.02], seed 3): every sample's full
    transient (DC operating point, then fixed steps) in one batch, against
    the same samples one circuit per launch.  Setup (netlist analysis and
    upload) is outside both timings; state stays resident on the device."""
    from xyce_amd import circuits
    from xyce_amd.montecarlo import run_batched, run_plain
    parts = circuits.monte_carlo_rings(args.mc_samples, seed=3)
    probes = ["out0"]
    sync = torch.cuda.synchronize
    progress("batched transient: %d samples" % len(parts))
    b = run_batched(parts, args.mc_tstop, args.mc_dt, probes, every=10, sync=sync)
    progress("batched transient: %.3f s; plain form on %d samples"
             % (b["seconds"], args.mc_plain))
    res = {
        "workload": "%d Monte-Carlo BSIM3 ring oscillators, 11-301 stages, %d FETs; "
                    ".TRAN %.3g s at %.3g s steps after the DC operating point"
                    % (len(parts), sum(len(q.mos) for q in parts), args.mc_tstop, args.mc_dt),
        "solves_per_s": len(parts) / b["seconds"],
        "batch_latency_s": b["seconds"],
        "batch_newton_iterations": int(b["newton"]),
        "batch_step_failures": int(b["step_failures"]),
        "batch_host_reads": b["host_reads"],
        "fixed_steps_per_graph_launch": 8,
        "batch_setup_s": b["setup_s"],
    }
    if args.mc_plain <= 0:
        return res
    pick = np.unique(np.linspace(0, len(parts) - 1, args.mc_plain).round().astype(int))
    p = run_plain([parts[k] for k in pick], args.mc_tstop, args.mc_dt, probes, every=10,
                  sync=sync)
    progress("plain form: %.3f s" % float(p["seconds"].sum()))
    diff = float(np.abs(p["waves"] - b["waves"][pick]).max())
    plain_s = p["seconds"]
    return {
        "workload": "%d Monte-Carlo BSIM3 ring oscillators, 11-301 stages, %d FETs; "
                    ".TRAN %.3g s at %.3g s steps after the DC operating point"
                    % (len(parts), sum(len(q.mos) for q in parts), args.mc_tstop, args.mc_dt),
        "solves_per_s": len(parts) / b["seconds"],
        "plain_solves_per_s": len(pick) / float(plain_s.sum()),
        "speedup_vs_plain": (len(parts) / b["seconds"]) / (len(pick) / float(plain_s.sum())),
        "batch_latency_s": b["seconds"],
        "plain_latency_p50_s": float(np.percentile(plain_s, 50)),
        "plain_latency_p99_s": float(np.percentile(plain_s, 99)),
        "plain_samples_timed": int(len(pick)),
        "batch_newton_iterations": int(b["newton"]),
        "plain_newton_iterations_mean": float(p["newton"].mean()),
        "batch_step_failures": int(b["step_failures"]),
        "batch_host_reads": b["host_reads"],
        "fixed_steps_per_graph_launch": 8,
        "batch_setup_s": b["setup_s"],
        "max_waveform_diff_batch_vs_plain_V": diff,
    }


def power_grid_leg(torch, args):
    """The 500 x 500 resistive power-grid mesh driving 100,000 BSIM3
    inverters (200,000 FETs, 100,000 clock sources, 550,004 unknowns, seed
    5): one device load per Newton step (every device: BSIM3, resistors,
    capacitors, sources, and the ordered reduction into F, Q, the limiter
    vectors and both Jacobians) and the Jacobian assembly of a transient
    step, timed at full scale with seeded node voltages."""
    from xyce_amd import circuits
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad, _ptr
    progress("power grid: building the netlist")
    t0 = time.perf_counter()
    nl = circuits.power_grid_bsim3()
    nl.layout()
    dev = DeviceLoad(nl)
    setup = time.perf_counter() - t0
    progress("power grid: %d unknowns, %d CSR entries, setup %.1f s"
             % (nl.n_nodes, dev.nnz, setup))
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(13)
    x = rng.uniform(-0.2, 1.4, n1)
    x[-1] = 0.0
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    X = cu(x)
    SC, SN = cu(rng.uniform(-1e-14, 1e-14, nl.n_state)), cu(np.zeros(nl.n_state))
    TC, TN = cu(rng.uniform(-0.2, 1.4, nl.n_store)), cu(np.zeros(nl.n_store))
    out = torch.zeros(4 * n1 + 3 * dev.nnz, dtype=torch.float64, device="cuda")
    F, Q, FL, QL = (out[i * n1:(i + 1) * n1] for i in range(4))
    JF, JQ, J = (out[4 * n1 + i * dev.nnz:4 * n1 + (i + 1) * dev.nnz] for i in range(3))
    qn0, fprev, rhs = cu(np.zeros(n1)), cu(np.zeros(n1)), cu(np.zeros(n1))
    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    dev.set_output_mode(True)
    dev.set_solver_state(transient=1, newtonIter=1)
    dev.set_time(0.3e-9, 0.0)

    def load():
        dev.load_all(X, SC, TC, SN, TN, F, Q, FL, QL, JF, JQ)

    def assemble():
        check(lib().xg_tia_assemble(dev.h, 1, 2, 1e-12, *[_ptr(t) for t in (
            qn0, fprev, F, Q, FL, QL, JF, JQ, rhs, J)]), "assemble")

    def timed(fn, k):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(k):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    k = 30
    load_ms = timed(load, k)
    asm_ms = timed(assemble, k)
    eval_ms = timed(lambda: dev.update_state(X, SC, TC, SN, TN), k)
    n_b3 = len(nl.bsim3)
    progress("power grid: load %.3f ms, assembly %.3f ms" % (load_ms, asm_ms))
    res = {
        "workload": "500x500 resistive power-grid mesh + 100,000 BSIM3 inverters with "
                    "clock sources (200,000 FETs, %d unknowns, %d CSR entries, seed 5)"
                    % (nl.n_nodes, dev.nnz),
        "device_load_ms_per_newton_step": load_ms,
        "jacobian_assembly_ms": asm_ms,
        "newton_step_load_plus_assembly_ms": load_ms + asm_ms,
        "bsim3_evals_per_s_in_load": n_b3 / (load_ms * 1e-3),
        "device_eval_ms": eval_ms,
        "eval_roofline_frac": B3_BYTES_PER_EVAL * n_b3 / (eval_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "setup_s": setup,
        "linear_solve": "dense-front (multifrontal) LU of the one 550k-unknown "
                        "component: timed per Newton iteration in grid_transient",
    }
    dev.close()
    return res


def large_transient_leg(torch, args, which):
    """One large single circuit's whole transient on the device: "logic" =
    the 5,000-gate random BSIM3 logic (20,000 FETs, one 20k-unknown
    component; .TRAN 5 ns from a zero start with the supply ramped over
    50 ps), "grid" = the 500 x 500 power grid with 100,000 BSIM3 inverters
    (550,004 unknowns; DC operating point, then .TRAN 2 ns).  Steps from
    the reference's step control (local error estimate, breakpoints at the
    PULSE corners; --big-fixed: fixed steps of --big-dt).  Each Newton
    iteration is one graph: device load, ONESTEP
    assembly, the dense-front (multifrontal) LU of the component, update,
    norms and the convergence test.  Setup (netlist, analysis, upload,
    graph capture) is reported apart from the run."""
    from xyce_amd import circuits
    from xyce_amd._lib import lib
    from xyce_amd.transient import Transient
    if which == "logic":
        nl = circuits.random_logic_bsim3(n_gates=args.gates, seed=2, ramp=50e-12)
        full = 5e-9
        tstop, probes, uic = full * args.big_span, ["g0", "g1"], True
        name = ("random CMOS logic, %d NAND2/NOR2 gates, %d BSIM3 (seed 2), supply ramped "
                "over 50 ps" % (args.gates, 4 * args.gates))
    else:
        nl = circuits.power_grid_bsim3()
        full = 2e-9
        tstop, probes, uic = full * args.grid_span, ["out0", "g250_250"], False
        name = "500x500 power grid + 100,000 BSIM3 inverters (seed 5)"
    progress("large transient (%s): setup" % which)
    t0 = time.perf_counter()
    tr = Transient(nl)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    info = np.zeros(6, np.int64)
    lib().xg_lu_front_info(tr.dev.h, info.ctypes.data)
    progress("large transient (%s): %d unknowns, setup %.1f s, %d fronts, %d levels, "
             "%.3g flops per factorisation" % (which, nl.n_nodes, setup, info[1], info[2],
                                               info[5]))
    t0 = time.perf_counter()
    last = [t0]

    def report(step, n):
        now = time.perf_counter()
        if now - last[0] >= 20.0:
            last[0] = now
            progress("large transient (%s): step %d (%s), %.1f s" % (which, step, n, now - t0))
    counts = None
    if args.big_fixed:
        t, w = tr.run(tstop, args.big_dt, probes, every=10, uic=uic, progress=report)
    else:
        t, w, counts = tr.run_adaptive(tstop, args.big_dt, probes, uic=uic, progress=report)
    torch.cuda.synchronize()
    run = time.perf_counter() - t0
    tr.close()
    progress("large transient (%s): %.2f s, %d Newton iterations" % (which, run,
                                                                     tr.newton_total))
    steps = ("fixed steps of %.3g s" % args.big_dt if args.big_fixed else
             "the reference's step control (TSTEP %.3g s, RELTOL 1e-3, ABSTOL 1e-6)"
             % args.big_dt)
    return {
        "workload": name + "; .TRAN %.3g s, %s" % (tstop, steps),
        "step_control": None if counts is None else counts,
        "unknowns": nl.n_nodes,
        "solves_per_s": 1.0 / run,
        "run_s": run,
        "setup_s": setup,
        "steps": int(tr.n_steps),
        "newton_iterations": int(tr.newton_total),
        "step_failures": int(tr.step_failures),
        "ms_per_newton_iteration": 1e3 * run / max(tr.newton_total, 1),
        "fronts": {"count": int(info[1]), "levels": int(info[2]), "values": int(info[3]),
                   "largest": int(info[4]), "flops_per_factorisation": int(info[5])},
        "final_probe_V": [float(v) for v in w[:, -1]],
        # the PULSE corners the step control must land on: in the span run
        # and in the configured .TRAN
        "breakpoints": {"in_span": len(nl.breakpoints(tstop)),
                        "in_configured_tran": len(nl.breakpoints(full))},
    }


def host_loads_leg(nl, x, sta, sto, args):
    """The drop-in host-pointer mode (xg_update_state_host,
    xg_load_vectors_host, xg_load_matrices_host) on the same circuit and
    inputs: the caller's vectors and CSR value arrays stay in host memory
    (numpy), as the reference's N_LAS objects do, and every call stages its
    inputs and results over PCIe.  One Newton iteration's loads per step;
    the step time includes the transfers."""
    import ctypes
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad
    L = lib()
    progress("host-pointer loads")
    dev = DeviceLoad(nl)
    dev.set_solver_state(transient=1, newtonIter=1)
    dev.set_output_mode(True)
    n1 = nl.n_nodes + 1
    hs = [x.copy(), sta.copy(), sto.copy(), np.zeros(nl.n_state), sto.copy()]
    ho = [np.zeros(n1) for _ in range(4)] + [np.zeros(dev.nnz) for _ in range(2)]
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def step():
        check(L.xg_update_state_host(dev.h, *[p(a) for a in hs]), "update host")
        check(L.xg_load_vectors_host(dev.h, *[p(a) for a in ho[:4]]), "vectors host")
        check(L.xg_load_matrices_host(dev.h, *[p(a) for a in ho[4:]]), "matrices host")
    for _ in range(max(args.warmup, 3)):
        step()
    k = max(10, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(k):
        step()
    ms = (time.perf_counter() - t0) * 1e3 / k
    moved = 8 * (n1 + 2 * nl.n_state + 3 * nl.n_store + 4 * n1 + 2 * dev.nnz)
    dev.close()
    return {"ms_per_newton_iteration": ms,
            "evals_per_s": len(nl.bsim3) / (ms * 1e-3),
            "host_bytes_per_iteration": moved,
            "note": "host (numpy) arrays in and out of every call; PCIe-inclusive, "
                    "never the headline value"}


def serving_leg(args):
    """A Poisson stream of transient-simulation requests (seed 4): 60%
    BSIM3 rings of 11-101 stages (0.1 ns), 30% random BSIM3 logic of
    125-1,250 gates (500-5,000 FETs, 0.05 ns), 10% RC ladders of 10,000 nodes with 1,000
    diodes (0.05 ns), all at 1 ps steps.  Batched: continuous batching
    (micro-batches of one kind every 5 ms, up to 64 requests, 3 batches in
    flight).  Plain: the same stream (its first --serve-plain requests), one
    circuit per launch, one at a time.  Latency = completion - arrival,
    including the library setup and upload of each batch.  Each form starts
    with an empty analysed-plan cache (the process-wide reuse of analysed
    sub-circuits), so neither reuses the other's analyses."""
    from xyce_amd._lib import check, lib
    from xyce_amd.serving import poisson_requests, serve, summary

    def fresh_cache():
        check(lib().xg_lu_cache_set_limit(256 << 20), "plan cache reset")
    progress("serving: %d requests at %.0f/s" % (args.serve_requests, args.serve_rate))
    fresh_cache()
    reqs = poisson_requests(rate=args.serve_rate, n=args.serve_requests, seed=4,
                            logic_gates=(125, 1250))
    b = summary(serve(reqs, max_batch=64, workers=3, window=0.005))
    progress("serving: batched %.1f solves/s, p99 %.3f s" % (b["solves_per_s"],
                                                            b["latency_p99_s"]))
    res = {"workload": "Poisson %.0f req/s, %d requests: 60%% BSIM3 rings 11-101 stages, "
                       "30%% random BSIM3 logic 125-1250 gates, 10%% RC ladders 10k nodes + "
                       "1k diodes (seed 4)" % (args.serve_rate, args.serve_requests),
           "batched": b}
    if args.serve_plain > 0:
        plain = poisson_requests(rate=args.serve_rate, n=args.serve_requests, seed=4,
                                 logic_gates=(125, 1250))[:args.serve_plain]
        fresh_cache()
        p = summary(serve(plain, max_batch=1, workers=1, window=0.001))
        progress("serving: plain %.1f solves/s, p99 %.3f s" % (p["solves_per_s"],
                                                              p["latency_p99_s"]))
        res["plain"] = p
        res["plain_requests"] = args.serve_plain
    return res


def cpu_baseline(nl, x, sta, sto, budget_s):
    """The oracle's restatement of the reference's BSIM3 Master loops (array
    of structs, pointer read-modify-write stamping), built -O3 -march=native
    on this host, timed serially on one core and partitioned over this
    process's CPU share the way Xyce distributes devices over MPI ranks
    (fixed-order combine of the partitions' output copies).  The serial leg
    takes 3/4 of the wall budget and the all-cores leg 1/8 (about as much
    CPU time on 16 threads).  `value` is the all-cores rate."""
    from oracle.baseline import PartitionedBsim3, host_description, native_library
    L, _ = native_library()
    host = host_description()
    share = host["affinity"]
    if host.get("cgroup_cpus"):
        share = min(share, max(1, int(host["cgroup_cpus"])))
    # one GPU's share of a box's host cores
    cores = max(1, min(16, share))
    n = len(nl.bsim3)
    legs = {}
    for P in (1, cores):
        if P in legs:
            continue
        pb = PartitionedBsim3(nl, P, L)
        pb.set_solver_state(transient=1, newtonIter=1)
        sn, tn = np.zeros(nl.n_state), sto.copy()
        pb.run(x, sta.copy(), sto.copy(), sn, tn, 2)                 # warm up
        k = 3 if P == 1 else 20       # thread start-up is in every run
        el, _ = pb.run(x, sta.copy(), sto.copy(), sn, tn, k)
        share = 0.75 if P == 1 else 0.125
        reps = max(3, int(share * budget_s / max(el / k, 1e-6)))
        el, _ = pb.run(x, sta.copy(), sto.copy(), sn, tn, reps)
        pb.close()
        legs[P] = (n * reps / el, reps, el)
    v1, r1, e1 = legs[1]
    vp, rp, ep = legs[cores]
    return dict(value=vp, unit="BSIM3 instance-evals/s", cores=cores, kind="port",
                sample="%d full BSIM3 loads of the %d-instance workload in %.1f s on %d host "
                       "threads (instances in %d contiguous partitions, fixed-order combine); "
                       "serial: %d loads in %.1f s on one core; oracle restatement of the "
                       "reference Master loops, gcc -O3 -march=native on this host"
                       % (rp, n, ep, cores, cores, r1, e1),
                serial={"value": v1, "cores": 1}, host=host)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU; the device load of one circuit would shard by "
              "graph partition across GPUs, which this build does not implement "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    progress("importing torch")
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad

    if args.mc_only:
        print(json.dumps({"batched_transient": batched_transient_leg(torch, args)}),
              flush=True)
        return

    progress("building the %d-gate netlist" % args.gates)
    nl = circuits.random_logic_bsim3(n_gates=args.gates, seed=2)
    nl.layout()
    n_b3 = len(nl.bsim3)
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(11)
    x = rng.uniform(-0.2, 1.5, n1)
    x[-1] = 0.0
    sta = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto = rng.uniform(-0.2, 1.5, nl.n_store)

    dev = DeviceLoad(nl)
    tune = _finite(autotune_choice(dev))
    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    dev.set_solver_state(transient=1, newtonIter=1)
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    X, SC, TC = cu(x), cu(sta), cu(sto)
    SN, TN = cu(np.zeros(nl.n_state)), cu(sto)
    # F, Q, dFdxdVp, dQdxdVp and the dF/dx, dQ/dx CSR values in one buffer,
    # zeroed by one fill (the assembler's job before every load).
    out = torch.zeros(4 * n1 + 2 * dev.nnz, dtype=torch.float64, device="cuda")
    F, Q, FL, QL = (out[i * n1:(i + 1) * n1] for i in range(4))
    JF = out[4 * n1:4 * n1 + dev.nnz]
    JQ = out[4 * n1 + dev.nnz:]
    if args.copy_calib:
        copy_calibration(torch)

    # The step: the library's overwrite output mode writes every output
    # entry (contributions summed from +0.0, untouched entries set to +0.0),
    # which is the zeroing pass and the accumulating load in one.
    def step():
        dev.load_all(X, SC, TC, SN, TN, F, Q, FL, QL, JF, JQ)

    # The same with the reference's accumulate semantics: zero, then add.
    def step_accumulate():
        out.zero_()
        dev.load_all(X, SC, TC, SN, TN, F, Q, FL, QL, JF, JQ)

    progress("device load: warmup and timed steps")
    dev.set_output_mode(False)
    for _ in range(args.warmup):
        step_accumulate()
    torch.cuda.synchronize()

    def timed(fn, steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(torch.cuda.current_stream())
        for _ in range(steps):
            fn()
        e1.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, time.perf_counter() - t0

    ms_accumulate, _ = timed(step_accumulate, args.steps)
    dev.set_output_mode(True)
    for _ in range(args.warmup):
        step()
    ms_eager, wall_eager = timed(step, args.steps)
    ms_graph, wall_graph = float("inf"), None
    if not args.no_graph:
        ms_graph, wall_graph = graph_timing(torch, dev, stream, step, timed, args)
    # report the faster of the two launch modes; both run the whole step
    if ms_graph < ms_eager:
        ms_step, wall, launch = ms_graph, wall_graph, "hipGraph replay of the whole step"
    else:
        ms_step, wall, launch = ms_eager, wall_eager, "eager stream launches"

    # Kernel durations: back-to-back launches of one kernel between two
    # events on the library's stream, so the figure is the kernel's own
    # average duration (the inputs are unchanged between launches, so every
    # launch does the same work).
    n_k = min(args.steps, 50)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(n_k):
        dev.update_state_type("bsim3", X, SC, TC, SN, TN)
    b.record(stream)
    torch.cuda.synchronize()
    eval_ms = a.elapsed_time(b) / n_k
    a.record(stream)
    for _ in range(n_k):
        dev.load_vectors(F, Q, FL, QL)
        dev.load_matrices(JF, JQ)
    b.record(stream)
    torch.cuda.synchronize()
    reduce_ms = a.elapsed_time(b) / n_k
    achieved = B3_BYTES_PER_EVAL * n_b3 / (eval_ms * 1e-3) / 1e9

    value = n_b3 / (ms_step * 1e-3)
    res = {
        "metric": "BSIM3 instance-evals/s; batched transient solves/s + p99 latency vs plain",
        "value": value,
        "unit": "instance-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "random CMOS logic, %d NAND2/NOR2 gates, %d BSIM3 "
                               "instances, RC interconnect (seed 2); full device load "
                               "at a transient Newton iteration" % (args.gates, n_b3),
                   "n_nodes": nl.n_nodes, "nnz": dev.nnz, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": B3_EVAL_TRAFFIC_BYTES if args.gates == 5000 else None,
                     "kernel": b3_kernel_name(n_b3, tune["variant"]), "kernel_ms": eval_ms,
                     "bytes_per_eval": B3_BYTES_PER_EVAL,
                     # the same launch's FP64 VALU rate: counter operations
                     # over the measured duration, against the vector peak
                     "fp64": None if args.gates != 5000 else {
                         "achieved": B3_EVAL_FP64_FLOPS / (eval_ms * 1e-3) / 1e12,
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": B3_EVAL_FP64_FLOPS / (eval_ms * 1e-3) / 1e12
                         / FP64_PEAK_TFLOPS,
                         "flops_per_launch": B3_EVAL_FP64_FLOPS}},
        "autotune": tune,
        "wall_s_timed_region": wall,
        "launch": launch,
        "ms_per_step_eager": ms_eager,
        "ms_per_step_graph": ms_graph,
        "output_mode": "overwrite (output zeroing folded into the ordered reduction)",
        "ms_per_step_zero_then_accumulate": ms_accumulate,
        "reduction_ms": reduce_ms,
    }
    progress("device load: %.4f ms per step" % ms_step)
    if not args.no_host_loads:
        res["host_loads"] = host_loads_leg(nl, x, sta, sto, args)
    if not args.no_mc:
        res["batched_transient"] = batched_transient_leg(torch, args)
    if not args.no_grid:
        res["power_grid"] = power_grid_leg(torch, args)
    if not args.no_big_tran:
        for which in args.big_legs.split(","):
            res[which + "_transient"] = large_transient_leg(torch, args, which)
    if not args.no_serve:
        res["serving"] = serving_leg(args)
    if not args.no_cpu:
        progress("CPU baseline (%.0f s sample)" % args.cpu_seconds)
        sta_h, sto_h = sta.copy(), sto.copy()
        res["cpu_baseline"] = cpu_baseline(nl, x, sta_h, sto_h, args.cpu_seconds)
    print(json.dumps(_finite(res)), flush=True)


if __name__ == "__main__":
    main()
