#!/usr/bin/env python3
"""Benchmark of the per-step vortex-element time step on one MI355X.

Workload: circular cylinder, radius 0.5, 1000 equal panels, U = (1, 0),
Re = 1000 (nu = 1e-3), dt = 0.005, blob radius = panel length (pi/1000),
with a synthetic wake of N = 2^20 + 16384 vortices resident in HBM: one
vortex per cell of a square grid over [-0.6, 10] x [-2, 2] minus the disk,
jittered so that no two lie closer than 0.003 (above the collapse radii), in
two shear layers: circulation -U(1e-5/3, 1e-5) above y = 0 and +U(...) below
(seed 1).  Births replace
what leaves the box and what collapses, so the count stays near 2^20 through
the warm-up and timed steps; the line reports the smallest and largest N
timed and labels the metric N >= 2^20 only when the smallest is.

One step = one full time step of the reference loop: birth from the panel
system, the 500 probe velocities (velocity_control), Biot-Savart + nearest-3,
vortex and wall diffusive velocities, move / absorb / delete, force sums,
compaction and two collapse kinds (up to 2 passes each).

Prints one JSON line: steps/s of the mixed-precision path (value), its
pair-interactions/s, the plain FP64 GPU path's steps/s on the same state (at
its own and at the mixed path's diffusion cut-off), the roofline of the
dominant kernel, and the CPU baseline: the FP64 oracle (multithreaded C) timed
on the host cores on a bounded sample of the same step, plus full C0 steps
(200-panel cylinder) whose drag and lift histories are the reference for the
mixed and FP64 GPU paths' Cd/Cl errors.
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
BS_FLOP_PER_PAIR = 12
DIFF_FLOP_PER_PAIR = 17

TREE_MIN_N = 65536   # direct sums below 64k vortices, Morton tree above
ORDER, THETA = 16, 0.5
N_DEFAULT = (1 << 20) + 16384
NPROBE = 500


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def probe_points():
    """The reference's 500 control points (src/main.c:521-526)."""
    k = np.arange(NPROBE)
    return np.full(NPROBE, -0.15), 0.01 + 0.002 * k


def wake_state(n):
    """Two shear layers (negative circulation above y = 0, positive below) of
    U(1e-5 / 3, 1e-5) per vortex, about 5 per layer in all, as a wake of
    length 10 behind a unit body at U = 1 carries; 0.003 minimum spacing, above
    the collapse radii.  Births replace what leaves the box at x = 10."""
    from vortex_amd.state import jittered_wake
    return jittered_wake(n, seed=1, max_ve_g=1e-5, min_sep=0.003, layers=True)


def make_case(prec, P=1000, dt=0.005, tree_min_n=TREE_MIN_N, diff_cut=0.0, cap=0):
    from vortex_amd import geometry as geo
    from vortex_amd.solver import Case, make_params
    pan = geo.circle(P, 0.5)
    eps = float(pan["length"][0])
    prm = make_params(pan, dt=dt, ve_size=eps, nu=1e-3, rho=1.0, vinf=(1.0, 0.0),
                      box=(-2.0, 10.0, -10.0, 10.0), r_col_diff_sign=0.3, r_col_same_sign=0.3,
                      max_ve_g=2 * eps, n_col=2, precision=prec, tree_min_n=tree_min_n,
                      order=ORDER, theta=THETA, diff_cut=diff_cut)
    c = Case(prm, pan)
    if cap:
        c.reserve(cap)
    return c, prm


def diff_cut_default(prec, diff_cut):
    from vortex_amd.solver import VX_PREC_MIXED
    return diff_cut if diff_cut > 0 else (18.0 if prec == VX_PREC_MIXED else 30.0)


def run_path(prec, state, steps, warmup, M=None, tree_min_n=TREE_MIN_N, calibrate=False, prof_steps=3,
             diff_cut=0.0):
    x, y, g = state
    c, prm = make_case(prec, tree_min_n=tree_min_n, diff_cut=diff_cut, cap=x.size + 8 * 1000 + 64)
    if M is None:
        c.build_matrix()
        M = c.matrix()
    else:
        c.set_matrix(M)
    c.set_state(x, y, g)
    c.set_probes(*probe_points())
    if calibrate:
        # a kernel of known traffic for HBM counter calibration: the mixed
        # right-hand side packs the n sources to FP32 (k_pack_f32 reads
        # 3 x 8 B and writes 3 x 4 B per vortex, all coalesced)
        log(f"[bench] calibration: k_pack_f32 over {x.size} vortices ({24 * x.size} B read, {12 * x.size} B written)")
        c.right_part(x.size, (1.0, 0.0))
    vinf = (1.0, 0.0)
    for _ in range(warmup):
        c.time_step(vinf)
    c.sync()
    # the timed steps run uninstrumented: stage timing synchronises the host
    # at every stage boundary; each step ends with its one read-back (forces
    # and the vortex count), so the count of every timed step is known
    counts = []
    t = time.perf_counter()
    marks = [t]
    for _ in range(steps):
        counts.append(c.n)                 # vortices entering the step
        c.time_step(vinf)
        marks.append(time.perf_counter())
    c.sync()
    wall = time.perf_counter() - t
    pu, pv = c.probes()
    # stage breakdown and kernel durations (HIP events on the case stream)
    # from further steps with timing on
    c.enable_timing(True)
    t0 = c.timings()
    for _ in range(prof_steps):
        c.time_step(vinf)
    c.sync()
    t1 = c.timings()
    c.enable_timing(False)
    tree = None
    if tree_min_n:
        # one more step with the diffusion's census on: the pairs it evaluates
        # that lie within their target's cut-off radius
        c.tree_census(True)
        c.time_step(vinf)
        c.sync()
        tree = c.tree_stats()
        c.tree_census(False)
    per = lambda k: (t1[k] - t0[k]) / max(1, prof_steps)
    res = {
        "tree": tree, "bs_kernel_ms": per("bs_kernel_ms"), "tree_build_ms": per("tree_build_ms"),
        "diff_kernel_ms": per("diff_kernel_ms"), "wall_ms": per("wall_ms"),
        "wall": wall, "counts": counts, "prof_steps": prof_steps,
        "step_wall_ms": [1e3 * (b - a) for a, b in zip(marks, marks[1:])],
        "bs_ms": (t1["bs_ms"] - t0["bs_ms"]) / max(1, t1["bs_launches"] - t0["bs_launches"]),
        "diff_ms": per("diff_ms"), "speed_ms": per("speed_ms"), "creation_ms": per("creation_ms"),
        "step_ms": per("step_ms"), "probe_u_range": (float(pu.min()), float(pu.max())),
        "diff_cut": diff_cut_default(prec, diff_cut),
    }
    c.close()
    return res, M


def one_step_cd(prec, state, M, tree_min_n=TREE_MIN_N):
    """Drag and lift of one full step from the synthetic state (D = 1, rho = U = 1)."""
    x, y, g = state
    c, prm = make_case(prec, tree_min_n=tree_min_n, cap=x.size + 2048)
    c.set_matrix(M)
    c.set_state(x, y, g)
    f = c.time_step((1.0, 0.0))
    c.close()
    return f[1] / 0.5, f[2] / 0.5


# ------------------------------------------------------------------ CPU baseline
def c0_case_params():
    """Configuration C0: cylinder radius 0.5, 200 panels, U = (1, 0), Re = 1000,
    dt = 0.01, eps = panel length, births 0.5 panel length off the surface,
    same-sign collapse within 0.3 eps (the reference's conf values for the
    rest: opposite-sign radius 1 eps, max_ve_g 0.00628162, 2 passes)."""
    from vortex_amd import geometry as geo
    from vortex_amd.solver import make_params
    pan = geo.circle(200, 0.5)
    return pan, dict(dt=0.01, nu=1e-3, rho=1.0, vinf=(1.0, 0.0), box=(-2.0, 10.0, -10.0, 10.0),
                     r_col_diff_sign=1.0, r_col_same_sign=0.3, max_ve_g=0.00628162, n_col=2,
                     birth_offset_rel=0.5, tree_min_n=0)


GRAPH_CHUNK = 25  # steps per graph call in the C0 graph leg


def c0_step_rates(steps=500):
    """Configuration C0 run as specified (500 steps from the empty start,
    direct FP64 sums, and the same in mixed precision) on the GPU, once as
    stream steps (one host wait per step, as the reference loop) and once as
    replays of the captured step graph, GRAPH_CHUNK steps per call (one host
    wait per call).  A graph call sizes every launch for the most vortices
    its steps can reach (no deaths), so the case reserves no more than that:
    one 500-step call would size the launches for 1e5 slots while about
    3e3 are live.  The difference per step is the launch and host-wait
    overhead the graph removes."""
    from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
    pan, kw = c0_case_params()
    out = {}
    for name, prec in (("f64", VX_PREC_F64), ("mixed", VX_PREC_MIXED)):
        for mode in ("stream", "graph"):
            c = Case(make_params(pan, precision=prec, **kw), pan)
            c.build_matrix()
            if mode == "stream":
                c.reserve(steps * pan.size + 1024)
            c.sync()
            t = time.perf_counter()
            if mode == "stream":
                for _ in range(steps):
                    c.time_step((1.0, 0.0))
            else:
                for _ in range(steps // GRAPH_CHUNK):
                    c.run_steps_graph(GRAPH_CHUNK, (1.0, 0.0))
                c.run_steps_graph(steps % GRAPH_CHUNK, (1.0, 0.0))
            c.sync()
            dt = time.perf_counter() - t
            out[f"{name}_{mode}_steps_per_s"] = steps / dt
            out[f"{name}_n_end"] = c.n
            c.close()
        out[f"{name}_overhead_us_per_step"] = 1e6 * (1.0 / out[f"{name}_stream_steps_per_s"] -
                                                     1.0 / out[f"{name}_graph_steps_per_s"])
    log(f"[bench] C0 GPU step rates ({steps} steps): " +
        ", ".join(f"{k} {v:.4g}" for k, v in out.items()))
    return out


def c3_batch_rates(ncases=64, steps=20, seed=3):
    """Configuration C3's batch: flat plates with panel counts drawn from
    {200, 1000, 2040, 3320}, angle of attack uniform in [-10, 20] degrees, Re
    log-uniform in [500, 5000], dt = 0.005, mixed precision with the FP32 LU
    and FP64 refinement, each case from rest for `steps` steps (direct sums:
    N <= steps P).  Timed twice on fresh cases: advanced together (vx_batch:
    one graph launch for the steps of all cases) and one case after another
    with stream steps.  Reported as case-steps per second."""
    from vortex_amd import geometry as geo
    from vortex_amd.solver import Batch, Case, make_params, VX_PREC_MIXED, VX_FLAG_LU_REFINE
    rng = np.random.default_rng(seed)
    spec = []
    for _ in range(ncases):
        P = int(rng.choice([200, 1000, 2040, 3320]))
        alpha = float(rng.uniform(-10.0, 20.0))
        re = float(np.exp(rng.uniform(np.log(500.0), np.log(5000.0))))
        spec.append((geo.flat_plate(P, chord=1.0, thickness=0.02, alpha_deg=alpha), 1.0 / re))

    def make():
        cs = []
        for pan, nu in spec:
            c = Case(make_params(pan, dt=0.005, nu=nu, r_col_diff_sign=1.0, r_col_same_sign=0.3,
                                 precision=VX_PREC_MIXED, flags=VX_FLAG_LU_REFINE), pan)
            c.build_matrix()
            c.reserve(steps * pan.size + 1024)
            cs.append(c)
        for c in cs:
            c.sync()
        return cs

    out = {"cases": ncases, "steps": steps, "panels": [int(pan.size) for pan, _ in spec]}
    cs = make()
    t = time.perf_counter()
    for c in cs:
        c.run_steps(steps, (1.0, 0.0))
    for c in cs:
        c.sync()
    t_one = time.perf_counter() - t
    n_one = [c.n for c in cs]
    for c in cs:
        c.close()
    cs = make()
    b = Batch(cs)
    t = time.perf_counter()
    b.run_steps(steps, [(1.0, 0.0)] * ncases)
    t_batch = time.perf_counter() - t
    n_batch = [c.n for c in cs]
    b.close()
    for c in cs:
        c.close()
    out.update({"one_by_one_case_steps_per_s": ncases * steps / t_one,
                "batch_case_steps_per_s": ncases * steps / t_batch,
                "batch_over_one_by_one": t_one / t_batch,
                "n_end_min": min(n_batch), "n_end_max": max(n_batch),
                "same_counts": n_one == n_batch})
    log(f"[bench] C3 batch ({ncases} cases x {steps} steps, N per case up to {max(n_batch)}): batch "
        f"{out['batch_case_steps_per_s']:.4g} case-steps/s, one by one {out['one_by_one_case_steps_per_s']:.4g}, "
        f"x{out['batch_over_one_by_one']:.2f}")
    return out


def c0_accuracy_and_cpu(vx_ref, steps=500, samples=(50, 100, 200, 300, 400, 500), window=200):
    """C0 full steps on the FP64 oracle (timed: the CPU baseline's complete
    steps) and on both GPU paths from the same start; Cd/Cl errors of the GPU
    paths against the oracle three ways:
    Every error is relative to the oracle's RMS Cd over the run.
      one_step   -- one step from the oracle's state at the sampled steps;
      horizon    -- per-step, from the common start, up to the step before the
                    vortex counts first differ (an integer decision flips);
      window     -- mean Cd and Cl over the last `window` steps."""
    from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
    pan, kw = c0_case_params()
    prm64 = make_params(pan, precision=VX_PREC_F64, **kw)
    rp = vx_ref.Params(**{k: getattr(prm64, k) for k, _ in vx_ref.Params._fields_})
    M = vx_ref.invert(vx_ref.panel_matrix(pan), 1)
    vinf = (1.0, 0.0)
    q = 0.5  # 0.5 rho U^2 D
    # oracle run, timed, keeping the states at the sampled steps
    st = vx_ref.State(pan, rp, M, capacity=8192)
    ref_f, ref_n, snaps = [], [], {}
    t = time.perf_counter()
    for k in range(1, steps + 1):
        f = st.step(vinf)
        ref_f.append((f[1] / q, f[2] / q))
        ref_n.append(st.n)
        if k in samples:
            snaps[k] = (st.x.copy(), st.y.copy(), st.g.copy(), st.n, st.d_g)
    t_cpu = time.perf_counter() - t

    def gpu_case(prec):
        c = Case(make_params(pan, precision=prec, **kw), pan)
        c.set_matrix(M)
        return c

    out = {"config": "C0: cylinder P=200, Re=1000, dt=0.01, eps=panel length, birth offset 0.5 l, direct sums",
           "steps": steps, "n_end": ref_n[-1]}
    cd_ref = np.array([a for a, _ in ref_f]); cl_ref = np.array([b for _, b in ref_f])
    # The impulse forces of a single step swing through zero (the oracle's
    # C0 Cd has a standard deviation several times its mean), so errors are
    # relative to the oracle's RMS Cd over the run, not to the step's own Cd.
    cd_rms = float(np.sqrt(np.mean(cd_ref ** 2)))
    out["cd_rms_ref"] = cd_rms
    out["cd_mean_ref"] = float(cd_ref.mean())
    for name, prec in (("mixed", VX_PREC_MIXED), ("f64", VX_PREC_F64)):
        c = gpu_case(prec)
        fs, ns = [], []
        for k in range(steps):
            f = c.time_step(vinf)
            fs.append((f[1] / q, f[2] / q))
            ns.append(c.n)
        c.close()
        cd = np.array([a for a, _ in fs]); cl = np.array([b for _, b in fs])
        same = [i for i in range(steps) if ns[i] != ref_n[i]]
        h = same[0] if same else steps
        hor_cd = float(np.max(np.abs(cd[:h] - cd_ref[:h])) / cd_rms) if h else None
        hor_cl = float(np.max(np.abs(cl[:h] - cl_ref[:h])) / cd_rms) if h else None
        w = slice(steps - window, steps)
        win_cd = abs(cd[w].mean() - cd_ref[w].mean()) / cd_rms
        win_cl = abs(cl[w].mean() - cl_ref[w].mean()) / cd_rms
        one = []
        for k in samples:
            x, y, g, n, dg = snaps[k]
            s = min(x.size, (n + pan.size + 31) // 32 * 32)
            c = gpu_case(prec)
            c.set_state(x[:s], y[:s], g[:s])
            c.n = n
            c.d_g = dg
            f = c.time_step(vinf)
            c_n = c.n
            c.close()
            st1 = vx_ref.State(pan, rp, M, capacity=x.size)
            st1.x[:], st1.y[:], st1.g[:] = x, y, g
            st1.n, st1.d_g = n, dg
            fr = st1.step(vinf)
            one.append({"step": k, "n": n, "cd_ref": fr[1] / q, "dcd_rel": abs(f[1] - fr[1]) / q / cd_rms,
                        "dcl_rel": abs(f[2] - fr[2]) / q / cd_rms, "n_match": c_n == st1.n})
        out[name] = {
            "one_step_max_dcd_rel": max(o["dcd_rel"] for o in one),
            "one_step_max_dcl_rel": max(o["dcl_rel"] for o in one),
            "one_step": one,
            "horizon_steps": h, "horizon_max_dcd_rel": hor_cd, "horizon_max_dcl_rel": hor_cl,
            "window_steps": window, "window_dcd_rel": float(win_cd), "window_dcl_rel": float(win_cl),
        }
        log(f"[bench] C0 {name}: one-step dCd {out[name]['one_step_max_dcd_rel']:.3g}, horizon {h} steps "
            f"dCd {hor_cd if hor_cd is None else f'{hor_cd:.3g}'}, window dCd {win_cd:.3g}")
    return out, t_cpu, steps, ref_n


def cpu_baseline(state, sample_targets, threads, vx_ref):
    """FP64 oracle on the host, on the 1M wake: Biot-Savart + nearest-3 and
    the vortex diffusion for a sample of targets against all sources (scaled
    by N / sample to one step), the wall diffusion and the birth right-hand
    side over the full state."""
    from vortex_amd import geometry as geo
    x, y, g = state
    n = x.size
    eps = math.pi / 1000
    pan = geo.circle(1000, 0.5)
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
    t = time.perf_counter()
    vx_ref.diffusion_wall(x, y, dfull, pan, 1e-3, 20)
    t_wall = time.perf_counter() - t
    t = time.perf_counter()
    vx_ref.right_part(x, y, g, pan, (1.0, 0.0))
    t_rhs = time.perf_counter() - t
    scale = n / m
    step_s = (t_bs + t_diff) * scale + t_wall + t_rhs
    return {
        "value": 1.0 / step_s, "unit": "steps/s", "cores": threads, "kind": "port",
        "sample": (f"FP64 oracle on {threads} threads: Biot-Savart + nearest-3 and vortex diffusion for the "
                   f"first {m} of {n} targets against all sources ({t_bs + t_diff:.1f} s, scaled by N/{m}), "
                   f"wall diffusion and birth right-hand side over all {n} vortices ({t_wall + t_rhs:.1f} s); "
                   f"O(N) move/compaction and the collapse search omitted"),
        "pairs_per_s": m * (n - 1) / t_bs,
        "stage_s": {"bs_sampled": t_bs, "diffusion_sampled": t_diff, "wall": t_wall, "rhs": t_rhs},
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


def host_threads():
    """Host cores this process may use: its CPU affinity, bounded by the
    OpenMP thread count the host sets for one GPU's share (if any)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(aff, omp) if omp > 0 else aff), aff


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    # the first steps of a case run slower (3.2-3.3 ms against 3.05-3.1 at
    # N = 2^20 from about the eighth step on): 10 untimed steps reach the
    # steady state that long runs see
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    # default: as many as the mixed path, so both time the same steps
    ap.add_argument("--f64-steps", type=int, default=None)
    ap.add_argument("--cpu-targets", type=int, default=8192)
    ap.add_argument("--c0-steps", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c0", action="store_true", help="skip the C0 GPU step-rate runs")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 batch runs")
    ap.add_argument("--c3-cases", type=int, default=64)
    ap.add_argument("--c3-steps", type=int, default=20)
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
    state = wake_state(args.n)
    log(f"[bench] mixed path: N0={args.n}, warmup={args.warmup}, steps={args.steps}")
    tmin = 0 if args.direct else TREE_MIN_N
    mixed, M = run_path(VX_PREC_MIXED, state, args.steps, args.warmup, tree_min_n=tmin, calibrate=args.calibrate)
    sps = args.steps / mixed["wall"]
    nmin, nmax = min(mixed["counts"]), max(mixed["counts"])
    n_avg = float(np.mean(mixed["counts"]))
    s_bs = math.ceil(n_avg / 32) * 32
    pairs_eff = s_bs * (s_bs - 1)
    bs_stage_s = mixed["bs_ms"] / 1e3
    log(f"[bench] mixed: {sps:.4f} steps/s; N {nmin}..{nmax}; speed {mixed['speed_ms']:.2f} ms (tree build "
        f"{mixed['tree_build_ms']:.2f}, BS kernel {mixed['bs_kernel_ms']:.2f}, diffusion {mixed['diff_ms']:.2f} "
        f"(kernel {mixed['diff_kernel_ms']:.2f}, cut {mixed['diff_cut']} d), wall {mixed['wall_ms']:.2f}), "
        f"creation {mixed['creation_ms']:.2f} ms, step {mixed['step_ms']:.2f} ms")
    tr = mixed["tree"] or {}
    log(f"[bench] tree: {tr}")
    diff_waste = (tr["diff_pairs"] / tr["diff_in_radius"]) if tr.get("diff_in_radius", -1) > 0 else None
    if diff_waste:
        log(f"[bench] diffusion: evaluated / in-radius pairs {diff_waste:.3f}")
    # dominant kernel of the step and its algorithmic FLOPs per launch
    # mixed far field: one M2L per (bucket, far node), 4p^2 + 24p FLOP (the
    # p x p binomial product, powers and scaling), and one L2P per target,
    # 8(p - 1) FLOP; the near field 12 FLOP per pair
    bs_flop = (BS_FLOP_PER_PAIR * tr.get("bs_pairs", 0) + (4 * ORDER * ORDER + 24 * ORDER) * tr.get("m2l", 0)
               + 8 * (ORDER - 1) * tr.get("m", 0))
    diff_flop = DIFF_FLOP_PER_PAIR * tr.get("diff_pairs", 0)
    if args.direct:
        kname, kms = "k_biot_savart", mixed["bs_ms"]
        kflop = BS_FLOP_PER_PAIR * s_bs * s_bs
        if mixed["diff_ms"] > kms:
            kname, kms, kflop = "k_diffusion", mixed["diff_ms"], DIFF_FLOP_PER_PAIR * s_bs * s_bs
    elif mixed["bs_kernel_ms"] >= mixed["diff_kernel_ms"]:
        kname, kms, kflop = "k_tree_bs", mixed["bs_kernel_ms"], bs_flop
    else:
        kname, kms, kflop = "k_tree_diffusion", mixed["diff_kernel_ms"], diff_flop
    achieved = kflop / (kms / 1e3) / 1e12 if kms > 0 else 0.0

    f64 = f64s = None
    dcd = dcl = None
    if not args.no_f64 and args.f64_steps > 0:
        log(f"[bench] FP64 path: steps={args.f64_steps}")
        f64, _ = run_path(VX_PREC_F64, state, args.f64_steps, args.warmup, M=M, tree_min_n=tmin)
        log(f"[bench] f64 (cut {f64['diff_cut']} d): {args.f64_steps / f64['wall']:.4f} steps/s; speed "
            f"{f64['speed_ms']:.2f} ms, BS kernel {f64['bs_kernel_ms']:.2f} ms, diffusion {f64['diff_ms']:.2f} ms")
        if not args.direct:
            f64s, _ = run_path(VX_PREC_F64, state, args.f64_steps, args.warmup, M=M, tree_min_n=tmin,
                               diff_cut=mixed["diff_cut"])
            log(f"[bench] f64 (cut {f64s['diff_cut']} d): {args.f64_steps / f64s['wall']:.4f} steps/s; speed "
                f"{f64s['speed_ms']:.2f} ms, diffusion {f64s['diff_ms']:.2f} ms")
        cd_m, cl_m = one_step_cd(VX_PREC_MIXED, state, M, tmin)
        cd_f, cl_f = one_step_cd(VX_PREC_F64, state, M, tmin)
        dcd = abs(cd_m - cd_f) / max(abs(cd_f), 1e-300)
        dcl = abs(cl_m - cl_f) / max(abs(cd_f), 1e-300)
        log(f"[bench] one-step at N={args.n}: Cd mixed {cd_m:.9g} f64 {cd_f:.9g} rel diff {dcd:.3g}; "
            f"Cl diff / Cd {dcl:.3g}")

    c0_rates = None if args.no_c0 else c0_step_rates(args.c0_steps)
    c3 = None if args.no_c3 else c3_batch_rates(args.c3_cases, args.c3_steps)

    cpu = None
    c0 = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import vx_ref
        path = vx_ref.select_build(True)
        threads, aff = host_threads()
        log(f"[bench] CPU oracle ({os.path.basename(path)}), {threads} threads of {aff} available "
            f"({os.cpu_count()} logical CPUs on the host)")
        vx_ref.lib().vxr_set_num_threads(threads)
        c0, t_c0, c0_steps, c0_n = c0_accuracy_and_cpu(vx_ref, steps=args.c0_steps)
        log(f"[bench] CPU C0: {c0_steps} full steps in {t_c0:.2f} s (N up to {max(c0_n)})")
        cpu = cpu_baseline(state, args.cpu_targets, threads, vx_ref)
        cpu["c0_full_steps_per_s"] = c0_steps / t_c0
        cpu["c0_steps"] = c0_steps
        log(f"[bench] cpu: {cpu['value']:.3g} steps/s ({cpu['pairs_per_s']:.3g} BS pairs/s); {cpu['stage_s']}")

    label = "N>=2^20" if nmin >= (1 << 20) else f"N={nmin}..{nmax}"
    out = {
        "metric": (f"time steps/s at {label} (mixed precision; direct sums)" if args.direct else
                   f"time steps/s at {label} (mixed precision; Morton tree above 64k vortices)"),
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
        "data": "synthetic (jittered two-layer wake, seed 1)",
        "config": {"workload": "cylinder P=1000, Re=1000, dt=0.005, eps=pi/1000, N=2^20+16384 jittered wake, "
                               "500 probes per step; " +
                               ("direct O(N^2) sums" if args.direct else "tree leaf 64, order 16, theta 0.5"),
                   "n_vortices_min": nmin, "n_vortices_max": nmax, "n_vortices_mean": n_avg,
                   "panels": 1000, "parallelism": "single GPU"},
        "pair_interactions_per_s": pairs_eff / bs_stage_s,
        "pair_interactions_kind": ("N(N-1) / Biot-Savart stage time" if args.direct else
                                   "effective N(N-1) / Biot-Savart stage time (tree build included)"),
        "evaluated": ({"bs_pairs": s_bs * s_bs, "diff_pairs": s_bs * s_bs} if args.direct else
                      {"bs_pairs": tr.get("bs_pairs"), "far_interactions": tr.get("m2p_evals"),
                       "m2l_translations": tr.get("m2l"),
                       "diff_pairs": tr.get("diff_pairs"), "diff_pairs_in_radius": tr.get("diff_in_radius"),
                       "diff_evaluated_over_in_radius": diff_waste, "nearest3_fallbacks": tr.get("fallback")}),
        "diff_cut": {"mixed": mixed["diff_cut"], "f64": f64["diff_cut"] if f64 else None},
        "f64_steps_per_s": (args.f64_steps / f64["wall"]) if f64 else None,
        "mixed_over_f64": (sps / (args.f64_steps / f64["wall"])) if f64 else None,
        "f64_same_cut_steps_per_s": (args.f64_steps / f64s["wall"]) if f64s else None,
        "mixed_over_f64_same_cut": (sps / (args.f64_steps / f64s["wall"])) if f64s else None,
        "delta_cd_rel_one_step_1M": dcd,
        "delta_cl_rel_to_cd_one_step_1M": dcl,
        "accuracy_c0_vs_fp64_oracle": c0,
        "c0_gpu_step_rates": c0_rates,
        "c3_batch": c3,
        "stage_ms": {k: mixed[k] for k in ("creation_ms", "speed_ms", "tree_build_ms", "bs_kernel_ms", "diff_ms",
                                           "diff_kernel_ms", "wall_ms", "step_ms")},
        "stage_steps": mixed["prof_steps"],  # stage_ms: further steps with stage timing on
        "step_wall_ms": mixed["step_wall_ms"],
        # compute-bound FP32 work on the vector ALUs (MI355X FP32 vector peak)
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": traffic_from_profiles(kname),
                     "kernel": kname + {"k_tree_bs": f"<float,{ORDER}>", "k_tree_diffusion": "<float>"}.get(
                         kname, "<float,4>"),
                     "pipe": "VALU FP32", "flop_per_launch": kflop,
                     "avg_launch_ms": kms},
        "cpu_baseline": cpu and {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                     "c0_full_steps_per_s", "c0_steps")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
