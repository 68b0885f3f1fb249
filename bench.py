#!/usr/bin/env python3
"""Benchmark of the per-step vortex-element time step on one MI355X.

Workload: circular cylinder, radius 0.5, 1000 equal panels, U = (1, 0),
Re = 1000 (nu = 1e-3), dt = 0.005, blob radius = panel length, with a
synthetic wake of N = 2^20 vortices (uniform in [-0.6, 10] x [-2, 2] minus the
disk, circulations +-U(0, 2 * eps) alternating in sign, seed 1) resident in
HBM.  One step = one full time step: birth from the panel system, Biot-Savart
+ nearest-3, vortex and wall diffusive velocities, move / absorb / delete,
force sums, compaction and two collapse kinds (up to 2 passes each).

Prints one JSON line: steps/s of the mixed-precision path (value), its
pair-interactions/s, the plain FP64 GPU path's steps/s on the same state, the
drag-coefficient difference of one step from the identical state, the
roofline of the dominant kernel and the FP64 CPU oracle timed on a bounded
sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32
PEAK_FP64 = 78.6   # TFLOP/s, MI355X vector FP64
BS_FLOP_PER_PAIR = 12
DIFF_FLOP_PER_PAIR = 17


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TREE_MIN_N = 65536   # direct sums below 64k vortices, Morton tree above
ORDER, THETA = 16, 0.5


def make_case(prec, n, P=1000, dt=0.005, tree_min_n=TREE_MIN_N):
    from vortex_amd import geometry as geo
    from vortex_amd.solver import Case, make_params
    from vortex_amd.state import uniform_wake
    pan = geo.circle(P, 0.5)
    eps = float(pan["length"][0])
    prm = make_params(pan, dt=dt, ve_size=eps, nu=1e-3, rho=1.0, vinf=(1.0, 0.0),
                      box=(-2.0, 10.0, -10.0, 10.0), r_col_diff_sign=0.3, r_col_same_sign=0.3,
                      max_ve_g=2 * eps, n_col=2, precision=prec, tree_min_n=tree_min_n,
                      order=ORDER, theta=THETA)
    x, y, g = uniform_wake(n, seed=1, max_ve_g=2 * eps)
    c = Case(prm, pan)
    c.reserve(n + 8 * P + 64)
    return c, (x, y, g), prm


def run_path(prec, n, steps, warmup, M=None, tree_min_n=TREE_MIN_N, calibrate=False, prof_steps=3):
    c, (x, y, g), prm = make_case(prec, n, tree_min_n=tree_min_n)
    if M is None:
        c.build_matrix()
        M = c.matrix()
    else:
        c.set_matrix(M)
    c.set_state(x, y, g)
    if calibrate:
        # a kernel of known traffic for HBM counter calibration: the mixed
        # right-hand side packs the n sources to FP32 (k_pack_f32 reads
        # 3 x 8 B and writes 3 x 4 B per vortex, all coalesced)
        log(f"[bench] calibration: k_pack_f32 over {n} vortices ({24 * n} B read, {12 * n} B written)")
        c.right_part(n, (1.0, 0.0))
    vinf = (1.0, 0.0)
    first = None
    for _ in range(warmup):
        f = c.time_step(vinf)
        first = f if first is None else first
    c.sync()
    n_start = c.n
    # the timed steps run uninstrumented: stage timing synchronises the host
    # at every stage boundary
    t = time.perf_counter()
    forces = []
    marks = [t]
    for _ in range(steps):
        forces.append(c.time_step(vinf))  # ends with the step's one read-back
        marks.append(time.perf_counter())
    c.sync()
    wall = time.perf_counter() - t
    n_end = c.n
    # stage breakdown and kernel durations (HIP events on the case stream)
    # from further steps with timing on
    c.enable_timing(True)
    t0 = c.timings()
    for _ in range(prof_steps):
        c.time_step(vinf)
    c.sync()
    t1 = c.timings()
    c.enable_timing(False)
    tree = c.tree_stats() if tree_min_n else None
    per = lambda k: (t1[k] - t0[k]) / max(1, prof_steps)
    res = {
        "tree": tree, "bs_kernel_ms": per("bs_kernel_ms"), "tree_build_ms": per("tree_build_ms"),
        "diff_kernel_ms": per("diff_kernel_ms"),
        "wall_ms": per("wall_ms"),
        "wall": wall, "n_start": n_start, "n_end": n_end, "prof_steps": prof_steps,
        "step_wall_ms": [1e3 * (b - a) for a, b in zip(marks, marks[1:])],
        "bs_ms": (t1["bs_ms"] - t0["bs_ms"]) / max(1, t1["bs_launches"] - t0["bs_launches"]),
        "diff_ms": per("diff_ms"),
        "speed_ms": per("speed_ms"),
        "creation_ms": per("creation_ms"),
        "step_ms": per("step_ms"),
        "first_forces": first, "forces": forces,
    }
    c.close()
    return res, M


def one_step_cd(prec, n, M, tree_min_n=TREE_MIN_N):
    """Drag of one full step from the synthetic state (Cd = Fx / (0.5 rho U^2 D), D = 1)."""
    c, (x, y, g), prm = make_case(prec, n, tree_min_n=tree_min_n)
    c.set_matrix(M)
    c.set_state(x, y, g)
    f = c.time_step((1.0, 0.0))
    c.close()
    return f[1] / (0.5 * 1.0 * 1.0 * 1.0)


def cpu_baseline(n, sample_targets, threads):
    """FP64 oracle (multithreaded C) on the host: Biot-Savart + diffusion for a
    sample of targets against all n sources, scaled to a full step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import vx_ref
    from vortex_amd.state import uniform_wake
    eps = math.pi / 1000
    x, y, g = uniform_wake(n, seed=1, max_ve_g=2 * eps)
    vx_ref.lib().vxr_set_num_threads(threads)
    m = sample_targets
    t = time.perf_counter()
    u, v, d = vx_ref.biot_savart(x, y, g, eps * eps, (1.0, 0.0), n_tgt=m)
    t_bs = time.perf_counter() - t
    t = time.perf_counter()
    dfull = np.full(n, 0.005)
    dfull[:m] = d
    vx_ref.diffusion(x, y, g, dfull, 1e-3, eps, u, v, n_tgt=m)
    t_diff = time.perf_counter() - t
    scale = n / m
    step_s = (t_bs + t_diff) * scale
    return {
        "value": 1.0 / step_s, "unit": "steps/s", "cores": threads, "kind": "port",
        "sample": (f"FP64 oracle Biot-Savart + nearest-3 and vortex diffusion for the first {m} "
                   f"targets against all {n} sources ({t_bs + t_diff:.1f} s), scaled by N/{m} to one "
                   f"step; panel, wall and O(N) stages omitted (they add under 1% at this N)"),
        "pairs_per_s": m * (n - 1) / t_bs,
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "tree_step_mixed_1M_counters.txt")


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same
    counter's reading on a kernel of known traffic), or None.  Lines read
    `traffic_bytes_per_launch <kernel> <bytes>`."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        tok = line.split()
        if len(tok) == 3 and tok[0] == "traffic_bytes_per_launch" and tok[1] == kernel:
            try:
                return float(tok[2])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    # the synthetic random wake sheds its closest pairs in the first steps
    # (collapse); three untimed steps pass that transient
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    # default: as many as the mixed path, so both time the same steps
    ap.add_argument("--f64-steps", type=int, default=None)
    ap.add_argument("--cpu-targets", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--direct", action="store_true",
                    help="direct O(N^2) sums at every N (the reference's algorithm) instead of the tree")
    ap.add_argument("--calibrate", action="store_true",
                    help="run one kernel of known HBM traffic before the steps (for counter runs)")
    args = ap.parse_args()
    if args.f64_steps is None:
        args.f64_steps = args.steps
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the step would shard by targets, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    from vortex_amd.solver import VX_PREC_F64, VX_PREC_MIXED
    n = args.n
    log(f"[bench] mixed path: N={n}, warmup={args.warmup}, steps={args.steps}")
    tmin = 0 if args.direct else TREE_MIN_N
    mixed, M = run_path(VX_PREC_MIXED, n, args.steps, args.warmup, tree_min_n=tmin, calibrate=args.calibrate)
    sps = args.steps / mixed["wall"]
    n_avg = 0.5 * (mixed["n_start"] + mixed["n_end"])
    s_bs = math.ceil(n_avg / 32) * 32
    pairs_eff = s_bs * (s_bs - 1)
    bs_stage_s = mixed["bs_ms"] / 1e3
    log(f"[bench] mixed: {sps:.4f} steps/s; speed {mixed['speed_ms']:.2f} ms (tree build "
        f"{mixed['tree_build_ms']:.2f}, BS kernel {mixed['bs_kernel_ms']:.2f}, diffusion {mixed['diff_ms']:.2f} "
        f"(kernel {mixed['diff_kernel_ms']:.2f}), "
        f"wall {mixed['wall_ms']:.2f}), creation {mixed['creation_ms']:.2f} ms, step {mixed['step_ms']:.2f} ms")
    tr = mixed["tree"] or {}
    log(f"[bench] tree: {tr}")
    # dominant kernel of the step and its algorithmic FLOPs per launch
    # mixed far field: one M2L per (bucket, far node), 4p^2 + 24p FLOP (the
    # p x p binomial product, powers and scaling), and one L2P per target,
    # 8(p - 1) FLOP; the near field 12 FLOP per pair
    bs_flop = (BS_FLOP_PER_PAIR * tr.get("bs_pairs", 0) + (4 * ORDER * ORDER + 24 * ORDER) * tr.get("m2l", 0)
               + 8 * (ORDER - 1) * tr.get("m", 0))
    diff_flop = DIFF_FLOP_PER_PAIR * tr.get("diff_pairs", 0)
    if args.direct:
        # every target against every source of the padded count s
        bs_flop = diff_flop = None
        kname, kms = "k_biot_savart", mixed["bs_ms"]
        kflop = BS_FLOP_PER_PAIR * s_bs * s_bs
        if mixed["diff_ms"] > kms:
            kname, kms, kflop = "k_diffusion", mixed["diff_ms"], DIFF_FLOP_PER_PAIR * s_bs * s_bs
    elif mixed["bs_kernel_ms"] >= mixed["diff_kernel_ms"]:
        kname, kms, kflop = "k_tree_bs", mixed["bs_kernel_ms"], bs_flop
    else:
        kname, kms, kflop = "k_tree_diffusion", mixed["diff_kernel_ms"], diff_flop
    achieved = kflop / (kms / 1e3) / 1e12 if kms > 0 else 0.0

    f64 = None
    dcd = None
    if not args.no_f64 and args.f64_steps > 0:
        log(f"[bench] FP64 path: steps={args.f64_steps}")
        f64, _ = run_path(VX_PREC_F64, n, args.f64_steps, args.warmup, M=M, tree_min_n=tmin)
        log(f"[bench] f64: {args.f64_steps / f64['wall']:.4f} steps/s; speed {f64['speed_ms']:.2f} ms, "
            f"BS kernel {f64['bs_kernel_ms']:.2f} ms, diffusion {f64['diff_ms']:.2f} ms")
        cd_m = one_step_cd(VX_PREC_MIXED, n, M, tmin)
        cd_f = one_step_cd(VX_PREC_F64, n, M, tmin)
        dcd = abs(cd_m - cd_f) / max(abs(cd_f), 1e-300)
        log(f"[bench] one-step Cd mixed {cd_m:.9g} f64 {cd_f:.9g} rel diff {dcd:.3g}")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] CPU oracle sample: {args.cpu_targets} targets x {n} sources, {threads} threads")
        cpu = cpu_baseline(n, args.cpu_targets, threads)
        log(f"[bench] cpu: {cpu['value']:.3g} steps/s ({cpu['pairs_per_s']:.3g} BS pairs/s)")

    out = {
        "metric": ("time steps/s at N=1M (mixed precision; direct sums)" if args.direct else
                   "time steps/s at N=1M (mixed precision; Morton tree above 64k vortices)"),
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (uniform wake, seed 1)",
        "config": {"workload": "cylinder P=1000, Re=1000, dt=0.005, eps=pi/1000, N=2^20 wake; " +
                               ("direct O(N^2) sums" if args.direct else "tree leaf 64, order 16, theta 0.5"),
                   "n_vortices": int(n_avg), "panels": 1000, "parallelism": "single GPU"},
        "pair_interactions_per_s": pairs_eff / bs_stage_s,
        "pair_interactions_kind": ("N(N-1) / Biot-Savart stage time" if args.direct else
                                   "effective N(N-1) / Biot-Savart stage time (tree build included)"),
        "evaluated": ({"bs_pairs": s_bs * s_bs, "diff_pairs": s_bs * s_bs} if args.direct else
                      {"bs_pairs": tr.get("bs_pairs"), "far_interactions": tr.get("m2p_evals"),
                       "m2l_translations": tr.get("m2l"),
                       "diff_pairs": tr.get("diff_pairs"), "nearest3_fallbacks": tr.get("fallback")}),
        "f64_steps_per_s": (args.f64_steps / f64["wall"]) if f64 else None,
        "mixed_over_f64": (sps / (args.f64_steps / f64["wall"])) if f64 else None,
        "delta_cd_rel_one_step": dcd,
        "stage_ms": {k: mixed[k] for k in ("creation_ms", "speed_ms", "tree_build_ms", "bs_kernel_ms", "diff_ms",
                                           "diff_kernel_ms", "wall_ms", "step_ms")},
        "stage_steps": mixed["prof_steps"],  # stage_ms: further steps with stage timing on
        "step_wall_ms": mixed["step_wall_ms"],
        # compute-bound FP32 work on the vector ALUs, whose peak equals the
        # dense FP32 matrix peak (157.3 TFLOP/s); "mfma" is the contract's
        # name for the compute roof
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": traffic_from_profiles(kname),
                     "kernel": kname + {"k_tree_bs": f"<float,{ORDER}>", "k_tree_diffusion": "<float>"}.get(
                         kname, "<float,4>"),
                     "pipe": "VALU FP32", "flop_per_launch": kflop,
                     "avg_launch_ms": kms},
        "cpu_baseline": cpu and {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
