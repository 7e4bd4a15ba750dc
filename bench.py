"""Benchmark of the explicit time step on one MI355X.

Default workload (--workload 3d-8m): the 3D box of the mixed-precision
configuration at full size -- 200 x 200 x 20 km, ~8.09 M tetrahedra / 1.42 M
nodes, elasto-visco-plastic Mohr-Coulomb (cohesion softening 44 -> 4 MPa
scaled per element, friction 30 -> 5 deg over plastic strain 0.1), per-element
log10 viscosity U[21,25] and cohesion N(44, 4.4) MPa, linear 273 -> 1573 K
temperature, 5 km weak slab, +-1 cm/yr side walls, dynamic-relaxation mass
scaling -- stepped in FP64; the mixed-precision step of the same model runs
beside it, and a short differentiated run of the adjoint configuration gives
the adjoint/forward time.
--workload 3d-fp64: 100 x 100 x 10 km, ~1.02 M tetrahedra (FP64).
--workload 3d-mixed: the 8.09 M-tet model with mixed precision primary.
--workload adjoint / ensemble: the differentiated run and the batched models.

Synthetic inputs are built on the host and resident in HBM before the timed
region.  A step is one pass of the explicit step over the whole mesh through
the fused six-pass schedule (des_run).  Prints one JSON line.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

# the CPU baselines' OpenMP placement, read when the OpenMP runtime loads
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "elem-steps/s & % HBM peak: FP64 vs mixed-prec step; adjoint/forward time"

# Algorithmic bytes of the fused passes per element and per node (3D, FP64):
# pass C reads conn, volume, volume_old, stress, strain, plstrain and writes
# stress, strain, plstrain, delta_plstrain, dpressure (256 B/element) and
# gathers coordinates, velocity, dvoldt and temperature (64 B/node).  Models
# with per-element viscosity and cohesion read them in pass C (+16 B/element).
PASS_BYTES_3D = {
    "pass_a": (72, 56), "pass_b": (72, 48), "pass_c": (256, 64),
    "pass_d": (32, 20), "pass_e": (200, 40), "pass_f": (112, 112),
}
# mixed precision: FP32 deviator, strain and rheology in passes C and E, FP64
# plastic strain; force contributions in 12-byte FP32 slots (E writes, F reads)
PASS_BYTES_3D_MIXED = dict(PASS_BYTES_3D, pass_c=(156, 64), pass_e=(112, 40), pass_f=(64, 112))
ELEM_PARAM_BYTES = {"fp64": 16, "mixed": 8}     # per-element viscosity and cohesion, pass C

# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per
# dispatch), each scaled by what the counter reads for a plain 1 GiB copy
# (FETCH_SIZE reads 0.5x, WRITE_SIZE 1.0x the bytes moved): (2 FETCH + WRITE)
# * 1024.  The commands and per-kernel tables are in profiles/.
TRAFFIC = {
    "3d-8m": {"k_pass_a": 818262353, "k_pass_b": 828603477, "k_pass_c": 2127069184, "k_pass_d": 375919251,
              "k_pass_e": 1821302892, "k_pass_f": 1230029760},
    "3d-fp64": {"k_pass_a": 92678592, "k_pass_b": 102013877, "k_pass_c": 250262697, "k_pass_d": 35105536,
                "k_pass_e": 216038982, "k_pass_f": 152488707},
    "3d-mixed": {"k_pass_a": 818268498, "k_pass_b": 828633387, "k_pass_c_mixed": 1433182104, "k_pass_d": 375282400,
                 "k_pass_e_mixed": 1044767783, "k_pass_f32": 737890883},
}

WORKLOADS = {
    "3d-8m": ("3d box 200x200x20 km, 8.09 M tets, per-element viscosity and cohesion, EVP, FP64 step "
              "(mixed-precision step beside it)", 8_000_000),
    "3d-fp64": ("3d-fp64 extension box 100x100x10 km, EVP Mohr-Coulomb", 1_000_000),
    "3d-mixed": ("3d box 200x200x20 km, 8.09 M tets, per-element viscosity and cohesion, EVP, "
                 "mixed-precision step", 8_000_000),
    "adjoint": ("adjoint box 50x50x10 km, maxwell, per-element log10 viscosity U[22,24]", None),
    "ensemble": ("ensemble of 256 box models (128 2D with 1k-50k triangles, 128 3D with 10k-200k "
                 "tetrahedra, log-uniform sizes, per-model extension rate, friction angle and "
                 "weak-seed position, seed 6), mixed precision", None),
}
MODEL_OF = {"3d-8m": "3d-mixed", "3d-fp64": "3d-fp64", "3d-mixed": "3d-mixed"}


def kernel_name(precision, p):
    """The kernel that runs fused pass p at this precision."""
    from dynearthsol_amd import abi
    if precision == abi.DES_MIXED and p in ("pass_c", "pass_e"):
        return f"k_{p}_mixed"
    if precision == abi.DES_MIXED and p == "pass_f":
        return "k_pass_f32"
    return "k_dt_partial+k_dt_final" if p == "dt" else "k_" + p


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baselines(model, nsteps, port_steps, runs=3):
    """The CPU step on this host over a bounded sample of the same workload:
    the reference's own step functions (oracle/_ref, OpenMP build; its
    rheology.cxx is serial because its OpenMP clause does not compile) and
    the oracle's C restatement (every function OpenMP-parallel).  Median of
    `runs` runs of nsteps (reference) and port_steps (restatement) steps each,
    from the same initial state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    host = {"cpu": cpu_model(), "nproc": os.cpu_count(), **env}
    try:
        from oracle import OracleState, ref_available, ref_lib
    except Exception as e:   # the checker is absent: report why
        return ({"value": None, "unit": "elem-steps/s", "cores": 0, "kind": "reference",
                 "sample": f"unavailable: {e}"}, None)
    base = OracleState(model)
    out = []
    if ref_available(3, omp=True):
        threads = ref_lib(3, omp=True).ref3_nthreads()
        secs = [base.copy().ref_run(nsteps, omp=True) for _ in range(runs)]
        med = statistics.median(secs)
        out.append({"value": model.nelem * nsteps / med, "unit": "elem-steps/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"median of {runs} runs of {nsteps} steps of the same {model.nelem}-element "
                              f"model from its initial state ({', '.join(f'{s:.2f}' for s in secs)} s): the "
                              f"reference's fields/geometry/bc/matprops with OpenMP ({threads} threads), "
                              f"rheology.cxx serial (its OpenMP clause does not compile)",
                    "host": host})
    else:
        out.append({"value": None, "unit": "elem-steps/s", "cores": 0, "kind": "reference",
                    "sample": "oracle/_ref not built"})
    secs = []
    for _ in range(runs):
        o = base.copy()
        t0 = time.perf_counter()
        o.step(port_steps)
        secs.append(time.perf_counter() - t0)
    med = statistics.median(secs)
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    out.append({"value": model.nelem * port_steps / med, "unit": "elem-steps/s", "cores": threads,
                "kind": "port",
                "sample": f"median of {runs} runs of {port_steps} steps of the same model "
                          f"({', '.join(f'{s:.2f}' for s in secs)} s): the oracle's C restatement of the "
                          f"step, every function OpenMP-parallel ({threads} threads)",
                "host": host})
    return out[0], out[1]


def timed(g, steps, precision):
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, precision)
    g.sync()
    return time.perf_counter() - t0


def rel(a, b):
    s = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / s if s else float(np.max(np.abs(a - b)))


def bench_ensemble(args):
    """All models in one batch (des_batch_run) against the same models run one
    at a time on their own contexts (des_run), both in mixed precision."""
    from dynearthsol_amd import abi, model as M
    from dynearthsol_amd.lib import DesBatch, DesModel
    t0 = time.perf_counter()
    models = M.config_ensemble(size_scale=args.ensemble_scale)
    n2 = sum(m.ndims == 2 for m in models)
    E = sum(m.nelem for m in models)
    print(f"# ensemble: {len(models)} models, {E} elements, built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    bt = DesBatch(models)
    bt.run(16, abi.DES_MIXED)          # the autotuner (every group's first 8 steps), before the warm-up
    bt.run(args.warmup, abi.DES_MIXED)
    bt.sync()
    t0 = time.perf_counter()
    bt.run(args.steps, abi.DES_MIXED)
    bt.sync()
    secs = time.perf_counter() - t0
    value = E * args.steps / secs
    print(f"# batch: {1e3 * secs / args.steps:.4f} ms per step of all models", file=sys.stderr, flush=True)
    from dynearthsol_amd.lib import TUNE_MEANING
    tuning = {f"{nd}d": {p: TUNE_MEANING[p][v] for p, (v, _) in bt.tuning_all(nd, abi.DES_MIXED).items()}
              for nd in (2, 3)}
    # the same batch in one persistent cooperative launch per run (both
    # dimension groups in one kernel, grid barriers between the passes)
    bt.set_launch_mode(1)
    psteps = max(1, min(args.steps, 50))
    bt.run(psteps, abi.DES_MIXED)
    bt.sync()
    t0 = time.perf_counter()
    bt.run(psteps, abi.DES_MIXED)
    bt.sync()
    psecs = time.perf_counter() - t0
    print(f"# batch, persistent launch: {1e3 * psecs / psteps:.4f} ms per step", file=sys.stderr, flush=True)
    bt.close()
    # one at a time: every model, seq_steps timed steps each after a short warmup
    # (--seq-steps 0 leaves it out)
    seq_secs, seq_work = 0.0, 0
    for m in models if args.seq_steps > 0 else ():
        g = DesModel(m)
        g.run(20, abi.DES_MIXED)
        dt = timed(g, args.seq_steps, abi.DES_MIXED)
        seq_secs += dt
        seq_work += m.nelem * args.seq_steps
        g.close()
    seq_value = seq_work / seq_secs if seq_secs else None
    print(f"# one at a time: {seq_value} elem-steps/s", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": value, "unit": "elem-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 deviator/rheology + f64 pressure/geometry/force accumulation", "data": "synthetic",
        "config": {"workload": WORKLOADS["ensemble"][0], "models": len(models), "models_2d": n2,
                   "nelem_total": E, "precision": "mixed", "schedule": "batched fused six-pass",
                   "tuning": tuning},
        "one_at_a_time": {"value": seq_value, "steps_per_model": args.seq_steps,
                          "batch_speedup": value / seq_value if seq_value else None},
        "persistent_launch": {"value": E * psteps / psecs, "ms_per_step": 1e3 * psecs / psteps, "steps": psteps,
                              "vs_per_pass_launches": (E * psteps / psecs) / value},
    }
    print(json.dumps(out), flush=True)


def adjoint_timing(nsteps, window, nelem, case="maxwell"):
    """The differentiated run of the adjoint configuration (visco-elastic or
    plastic-enabled): plain forward, checkpointed (window > 0) or taped
    (window 0) forward and reverse sweep, ms per step."""
    from dynearthsol_amd import model as M
    from dynearthsol_amd.lib import DesModel
    m = M.config_adjoint(case, nelem_target=nelem)
    g = DesModel(m)
    g.run(20)          # first launches of every kernel, outside the timing
    g.reset()
    t_plain = timed(g, nsteps, 0)
    g.reset()
    g.sync()
    t0 = time.perf_counter()
    g.adj_forward(nsteps, window)
    g.sync()
    t_fwd = time.perf_counter() - t0
    top = (m.bcflag & 32) != 0
    cb = np.zeros((m.nnode, 3))
    cb[top, 2] = 1.0
    t0 = time.perf_counter()
    g.adj_backward(cb)
    t_bwd = time.perf_counter() - t0
    g.close()
    rheology = "maxwell" if case == "maxwell" else "elasto-visco-plastic"
    return {"workload": f"adjoint box 50x50x10 km, {m.nelem} tets, {rheology}, log10 viscosity U[22,24]",
            "steps": nsteps, "window": window,
            "forward_ms_per_step": 1e3 * t_plain / nsteps,
            "checkpointed_forward_ms_per_step": 1e3 * t_fwd / nsteps,
            "backward_ms_per_step": 1e3 * t_bwd / nsteps,
            "gradient_over_forward": (t_fwd + t_bwd) / t_plain}


def bench_adjoint(args):
    """The differentiated run (visco-elastic adjoint configuration, 200 k tets,
    2,000 steps, checkpoint every 100): gradient cost over forward cost."""
    adj = adjoint_timing(args.adjoint_steps, 100, args.adjoint_nelem)
    print(f"# visco-elastic case: {adj['gradient_over_forward']:.3f}x", file=sys.stderr, flush=True)
    plastic = adjoint_timing(args.adjoint_steps, 100, args.adjoint_nelem, case="plastic")
    taped = adjoint_timing(args.adjoint_steps, 0, args.adjoint_nelem)
    print(f"# taped visco-elastic case: {taped['gradient_over_forward']:.3f}x", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": adj["gradient_over_forward"], "unit": "x forward time",
        "n_gpus": 1, "steps": args.adjoint_steps, "warmup": 0,
        "ms_per_step": adj["checkpointed_forward_ms_per_step"] + adj["backward_ms_per_step"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": {"workload": adj["workload"], "window": 100}, "adjoint": adj,
        "adjoint_plastic": plastic, "adjoint_taped": taped,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="3d-8m")
    ap.add_argument("--nelem", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps per CPU-baseline run (default: about 10 s of CPU work per run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-persistent", action="store_true",
                    help="leave out the persistent-launch comparison (cooperative launches)")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-nelem", type=int, default=200_000)
    ap.add_argument("--ensemble-scale", type=float, default=1.0)
    ap.add_argument("--seq-steps", type=int, default=100)
    ap.add_argument("--sweep-order", type=int, choices=(0, 1), default=1,
                    help="1: the node passes walk each XCD's blocks downwards (default); 0: upwards")
    ap.add_argument("--mesh", choices=("structured", "delaunay-joggled"), default="structured",
                    help="box mesh of the 3D workloads: jittered Kuhn tetrahedra (default) or the Delaunay "
                         "mesh of a jittered lattice (irregular node degrees, as a mesher's)")
    ap.add_argument("--force-layout", type=int, choices=(0, 1, 2), default=0,
                    help="FP64 force contributions: 0 node-major slots found from list start + one-byte "
                         "position (default), 1 element-major records, 2 node-major slots found by full positions")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is "
              "not part of this build", file=sys.stderr)
        sys.exit(2)
    from dynearthsol_amd import lib as L
    L.set_sweep_order(args.sweep_order)
    L.set_force_layout(args.force_layout)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)

    from dynearthsol_amd import abi, model as M
    from dynearthsol_amd.lib import DesModel

    desc, ne_default = WORKLOADS[args.workload]
    nelem = args.nelem or ne_default
    primary = abi.DES_MIXED if args.workload == "3d-mixed" else abi.DES_FP64
    t0 = time.perf_counter()
    m = M.config_benchmark(MODEL_OF[args.workload], nelem_target=nelem, mesh=args.mesh)
    E, N = m.nelem, m.nnode
    print(f"# {args.workload}: {E} elements, {N} nodes built in {time.perf_counter() - t0:.1f} s; "
          f"warmup {args.warmup}", file=sys.stderr, flush=True)
    has_params = m.elem_visc is not None

    def pass_table(precision):
        table = dict(PASS_BYTES_3D_MIXED if precision == abi.DES_MIXED else PASS_BYTES_3D)
        if has_params:
            be, bn = table["pass_c"]
            table["pass_c"] = (be + ELEM_PARAM_BYTES["mixed" if precision == abi.DES_MIXED else "fp64"], bn)
        return table

    # The autotuner runs before any warm-up, on a throwaway model of the same
    # mesh (its first 8 steps time every variant of passes A, C, D, E, F);
    # the timed models get its choices.
    from dynearthsol_amd.lib import TUNE_MEANING
    tuning = {}

    def tuned(precision):
        if precision not in tuning:
            t = DesModel(m)
            t.run(16, precision)
            t.sync()
            tuning[precision] = t.tuning_all(precision)
            t.close()
            assert all(v >= 0 for v, _ in tuning[precision].values()), tuning
        return tuning[precision]

    # Two identical models per precision.  g's warm-up and the twin's warm-up
    # and timed steps run under per-pass events that the dispatches timestamp
    # (kernel execution time); g's timed steps run plain (the timing) and see
    # the same states as the twin's (the step is deterministic).  So every
    # launch of a kernel in this process is either event-timed or mirrored by
    # an event-timed twin launch, and the per-kernel averages below are the
    # ones a kernel trace of this command reports.
    kernel_totals = {}

    def book(precision, passes, times=1):
        for p, (ms, c) in passes.items():
            k = kernel_name(precision, p)
            t = kernel_totals.setdefault(k, [0.0, 0])
            t[0] += times * ms
            t[1] += times * c

    def timed_pair(precision):
        choice = tuned(precision)
        a, b = DesModel(m), DesModel(m)
        a.set_tuning_all(precision, choice)
        b.set_tuning_all(precision, choice)
        wa = a.run_timed(args.warmup, precision) if args.warmup else {}
        wb = b.run_timed(args.warmup, precision) if args.warmup else {}
        a.sync()
        b.sync()
        secs = timed(a, args.steps, precision)
        passes = b.run_timed(args.steps, precision)
        b.sync()
        book(precision, wa)
        book(precision, wb)
        book(precision, passes, times=2)
        b.close()
        return a, secs, passes

    g, secs, passes = timed_pair(primary)
    ms_per_step = 1e3 * secs / args.steps
    value = E * args.steps / secs
    print(f"# timed {args.steps} steps: {ms_per_step:.4f} ms/step", file=sys.stderr, flush=True)
    avg = {k: ms / max(c, 1) for k, (ms, c) in passes.items()}
    table = pass_table(primary)
    dom = max((k for k in avg if k in table), key=lambda k: avg[k])
    be, bn = table[dom]
    step_bytes = sum(be_ * E + bn_ * N for be_, bn_ in table.values())
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9
    kernel = kernel_name(primary, dom)

    # the other precision from the same initial state, over the same steps
    compare = None
    if not args.no_mixed:
        other = abi.DES_FP64 if primary == abi.DES_MIXED else abi.DES_MIXED
        g2, t2, passes2 = timed_pair(other)
        names = ["vel", "stress", "coord", "plstrain", "temperature"]
        a, b = g2.download(names), g.download(names)
        t64, t32 = (secs, t2) if primary == abi.DES_FP64 else (t2, secs)
        if primary != abi.DES_FP64:
            a, b = b, a
        mixed_passes = passes2 if primary == abi.DES_FP64 else passes
        mt = pass_table(abi.DES_MIXED)
        mixed_bytes = sum(be_ * E + bn_ * N for be_, bn_ in mt.values())
        compare = {"steps": args.warmup + args.steps,
                   "fp64_ms_per_step": 1e3 * t64 / args.steps, "mixed_ms_per_step": 1e3 * t32 / args.steps,
                   "mixed_speedup": t64 / t32,
                   "mixed_speedup_bound_by_bytes": step_bytes / mixed_bytes if primary == abi.DES_FP64 else None,
                   "mixed_pass_ms": {k: round(ms / max(c, 1), 5) for k, (ms, c) in mixed_passes.items()},
                   "mixed_rel_err_vs_fp64": {k: rel(a[k], b[k]) for k in names}}
        g2.close()
    g.close()
    # the same model in one persistent cooperative launch per run (grid
    # barriers between the passes), for comparison with the per-pass launches
    persistent = None
    if not args.no_persistent:
        gp = DesModel(m)
        gp.set_launch_mode(1)
        gp.run(10, primary)
        psteps = max(1, min(args.steps, 50))
        tp = timed(gp, psteps, primary)
        gp.close()
        persistent = {"steps": psteps, "ms_per_step": 1e3 * tp / psteps,
                      "vs_per_pass_launches": (secs / args.steps) / (tp / psteps)}
        print(f"# persistent launch: {1e3 * tp / psteps:.4f} ms/step", file=sys.stderr, flush=True)
    tot_ms, tot_n = kernel_totals[kernel]
    proc_avg = tot_ms / tot_n
    achieved = (be * E + bn * N) / (proc_avg * 1e-3) / 1e9

    adjoint = None
    if not args.no_adjoint and args.workload == "3d-8m":
        adj = adjoint_timing(200, 100, args.adjoint_nelem)
        adjoint = {"gradient_over_forward": adj["gradient_over_forward"], **adj}
        print(f"# adjoint/forward {adj['gradient_over_forward']:.3f}x", file=sys.stderr, flush=True)

    cpu = port = None
    if not args.no_cpu:
        # about 7 s of CPU work per reference run (measured 3.7 M elem-steps/s on 16 cores)
        # and about 4 s per run of the restatement (39 M elem-steps/s)
        steps = args.cpu_steps or max(1, int(2.5e7 / E))
        port_steps = args.cpu_steps or max(1, int(1.6e8 / E))
        cpu_model_ = M.config_benchmark(MODEL_OF[args.workload], nelem_target=nelem, mesh=args.mesh)
        cpu, port = cpu_baselines(cpu_model_, steps, port_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elem-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if primary == abi.DES_FP64 else "f32 deviator/rheology + f64 pressure/geometry/force accumulation",
        "data": "synthetic",
        "config": {"workload": desc if args.mesh == "structured" else f"{desc}, {args.mesh} mesh",
                   "nelem": E, "nnode": N, "schedule": "fused six-pass",
                   "sweep_order": args.sweep_order, "force_layout": args.force_layout,
                   "precision": "fp64" if primary == abi.DES_FP64 else "mixed",
                   "tuning": {("fp64" if p == abi.DES_FP64 else "mixed"):
                              {k: {"variant": TUNE_MEANING[k][v], "ms": [round(x, 5) for x in ms if x > 0]}
                               for k, (v, ms) in t.items()} for p, t in tuning.items()}},
        "roofline": {"bound": "hbm", "kernel": kernel, "pass": dom, "avg_ms": proc_avg,
                     "launches": tot_n, "timed_window_avg_ms": avg[dom],
                     "bytes_per_launch": be * E + bn * N,
                     "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (TRAFFIC.get(args.workload, {}).get(kernel)
                                 if nelem == ne_default and args.mesh == "structured" else None)},
        "step_hbm": {"algorithmic_bytes": step_bytes, "achieved_GBs": step_gbs,
                     "frac": step_gbs / HBM_PEAK_GBS},
        "pass_ms": {k: round(v, 5) for k, v in avg.items()},
        "fp64_vs_mixed": compare,
        "persistent_launch": persistent,
        "adjoint_over_forward": adjoint["gradient_over_forward"] if adjoint else None,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
        "cpu_port": port,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
