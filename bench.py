#!/usr/bin/env python3
"""Contact-step benchmark on one MI355X.

Workload (default): the 10 000-box pile -- per-axis edges U(0.05, 0.15) m, density 1000, 5 layers
on a 0.2 m grid with jitter and random yaw, mu U(0.3, 0.8) per body -- stepped by the ColorCache
solver with 16 colour classes, 2 substeps, 16 velocity and 10 position iterations, dt = 1/240,
float32, warm-started. A "step" is one ContactEngine.step(dt, substeps=2) over the whole pile; the
warm-up steps let the pile fall into contact first (warm-up + timed = 600 steps by default).

Prints ONE JSON line: body-steps/s (whole job), ms/step, the roofline of the dominant kernel
(algorithmic bytes per launch / HIP-event launch time, against 8 TB/s) and the CPU oracle timed
on this host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "body-steps/s at 10k boxes & 256-scene batch; adjoint ms/step; HBM % of peak"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(pairs, invM, solver, V, P, S, N, C):
    """Compulsory SoA traffic, each field once per pass: velocity chunk per pair
    20 + 100*n_dyn + 64*k B; position chunk and warm apply 20 + 100*n_dyn + 52*k B."""
    a, b, k = pairs["pbi"], pairs["pbj"], pairs["pcount"].astype(np.int64)
    ndyn = (invM[a] > 0).astype(np.int64) + ((b >= 0) & (invM[np.maximum(b, 0)] > 0)).astype(np.int64)
    vel = int(np.sum(20 + 100 * ndyn + 64 * k))
    pos = int(np.sum(20 + 100 * ndyn + 52 * k))
    warm = pos if solver == "color_cache" else 0
    pre = 116 * N + 108 * C + 24 * N
    integ = 108 * N
    return dict(vel_sweep=vel, pos_sweep=pos, warm=warm,
                step=S * (pre + warm + V * vel + P * pos + integ))


def cpu_baseline(x, q, v, w, bodies_set, solver, V, P, S, n_colors, budget_s, omp=False):
    from oracle import oracle as orc
    bs = bodies_set
    bs.x[:], bs.q[:], bs.v[:], bs.w[:] = x, q, v, w
    o = orc.OracleWorld(bs, solver, vit=V, pit=P, n_colors=n_colors, omp=omp)
    steps, secs = 0, 0.0
    while secs < budget_s and steps < 400:
        secs += o.time_steps(1 / 240, S, 2)
        steps += 2
    return steps, secs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed steps before the warm-up, so that the timed window sees the settled pile "
                         "(default 400 for the piles and the batch)")
    ap.add_argument("--bodies", type=int, default=10000)
    ap.add_argument("--solver", default="color_cache", choices=["color_cache", "pair_chunks", "jacobi"])
    ap.add_argument("--colors", type=int, default=16)
    ap.add_argument("--substeps", type=int, default=2)
    ap.add_argument("--vit", type=int, default=16)
    ap.add_argument("--pit", type=int, default=10)
    ap.add_argument("--tiled", type=int, default=1)
    ap.add_argument("--tile-bodies", type=int, default=0,
                    help="bodies per tile of the tiled solve (0: 128; 224 for the 256-scene batch, the size the "
                         "autotuner's tile search picks for it; 256 for the adjoint's 64-box piles, four per tile, "
                         "so that the reverse tile kernel's 256 tiles run in one round)")
    ap.add_argument("--graph", type=int, default=0, help="replay each step as one captured hipGraph (tiled path)")
    ap.add_argument("--fast-math", type=int, default=1,
                    help="pile workloads under ColorCache: also time the fast-arithmetic engine from the exact "
                         "engine's state at the run's end (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="pile", choices=["pile", "batched", "large", "adjoint", "ncp", "pgs", "lattice",
                                                                 "pile_tuned"])
    ap.add_argument("--rollouts", type=int, default=3, help="adjoint workload: timed forward+backward rollouts")
    ap.add_argument("--extras", type=int, default=1,
                    help="pile workload: also measure the 256-scene batch, the 100 000-box pile, the config-4 "
                         "adjoint, the NCP sensitivity, the reference's lattice fixture and the exact-cone PGS, "
                         "and attach their results to the JSON line")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--lattice-runs",
                    default="color_cache:1,color_cache:0,pair_chunks:1,pair_chunks:0,jacobi:0,colored:1,colored:0",
                    help="lattice workload: engines to run, solver:tiled pairs")
    ap.add_argument("--seq-rounds", type=int, default=3)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--tune-steps", type=int, default=40, help="autotuner: steps per trial")
    ap.add_argument("--trace-summary", default=None, metavar="CSV",
                    help="print a per-kernel table of a rocprofv3 kernel_trace.csv over the substeps given by "
                         "--trace-window, and exit (no GPU work)")
    ap.add_argument("--trace-window", default=None, metavar="FIRST:LAST",
                    help="substep range [FIRST, LAST) of --trace-summary, counted by k_body_pre dispatches "
                         "(two substeps per step)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the run, one plain device-to-device copy of MIB MiB (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    a = ap.parse_args()
    if a.trace_summary:
        trace_summary(a.trace_summary, a.trace_window)
        return
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              f"(independent scenes and islands shard without communication; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    if a.workload == "batched":
        print(json.dumps(bench_batched(a)), flush=True)
        return
    if a.workload == "adjoint":
        print(json.dumps(bench_adjoint(a)), flush=True)
        return
    if a.workload == "ncp":
        print(json.dumps(bench_ncp(a)), flush=True)
        return
    if a.workload == "pgs":
        print(json.dumps(bench_pgs(a)), flush=True)
        return
    if a.workload == "lattice":
        print(json.dumps(bench_lattice(a)), flush=True)
        return
    if a.workload == "pile_tuned":
        print(json.dumps(bench_pile_tuned(a)), flush=True)
        return
    if a.workload == "large":
        print(json.dumps(bench_pile(a, large=True)), flush=True)
        return
    out = bench_pile(a)
    if a.extras and a.bodies == 10000:
        import torch
        torch.cuda.empty_cache()
        out["batched_256"] = _extra(bench_batched, a, warmup=20, steps=100, settle=400, seq_rounds=1, scenes=256,
                                    autotune=True, tune_steps=24, tile_bodies=0,
                                    keys=("value", "ms_per_step", "speedup_vs_sequential_tiled",
                                          "speedup_vs_sequential_plain", "sequential_tiled", "sequential_plain_launches",
                                          "graph_replay", "autotune", "roofline", "config"))
        torch.cuda.empty_cache()
        out["large_100k"] = _extra(bench_pile, a, large=True, settle=200, warmup=5, steps=95, cpu_seconds=0.0,
                                   keys=("value", "ms_per_step", "roofline", "run_300", "config", "kernel_ms"))
        torch.cuda.empty_cache()
        out["adjoint"] = _extra(bench_adjoint, a, rollouts=2,
                                keys=("value", "ms_per_step", "adjoint", "config"))
        out["ncp_sensitivity"] = _extra(bench_ncp, a, keys=("value", "unit", "ms_per_call", "vs_reference",
                                                            "config", "cpu_baseline"))
        out["reference_lattice_10000"] = _extra(bench_lattice, a, keys=("value", "unit", "vs_reference",
                                                                        "reference", "runs"))
        out["exact_cone_pgs"] = _extra(bench_pgs, a, keys=("value", "unit", "vs_reference", "solves", "kernel",
                                                           "lattice_10000", "cpu_baseline"))
    print(json.dumps(out), flush=True)
    if a.calibrate_copy:
        _copy_kernel(a.calibrate_copy)


def trace_summary(path, window):
    """Per-kernel calls, total and average duration of a rocprofv3 kernel trace (CSV), restricted to
    the dispatches that start within substeps [first, last); substep k starts at the k-th dispatch of
    k_body_pre, the first kernel of every substep."""
    import csv
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    starts = [t0 for t0, _, n in rows if "k_body_pre" in n]
    first, last = (int(x) for x in window.split(":")) if window else (0, len(starts))
    lo = starts[first]
    hi = starts[last] if last < len(starts) else float("inf")
    agg = {}
    for t0, t1, n in rows:
        if lo <= t0 < hi:
            short = n.split("(")[0].replace("void ", "")
            c, tot = agg.get(short, (0, 0))
            agg[short] = (c + 1, tot + (t1 - t0))
    span = (max(t1 for t0, t1, n in rows if lo <= t0 < hi) - lo) / 1e6
    print(f"substeps [{first}, {last}) of {len(starts)}: {span:.3f} ms from the first dispatch to the last end\n")
    print("| kernel | calls | total ms | avg us |")
    print("|---|---|---|---|")
    for n, (c, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{n}` | {c} | {tot / 1e6:.3f} | {tot / c / 1e3:.2f} |")


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


# kernels timed with HIP events on the stream each is launched on; the tiled solve's region
# (K_TILED) is not a kernel and never competes for "dominant"
def _kernel_names(L):
    return {L.K_SWEEP_VEL: "k_sweep_vel", L.K_SWEEP_POS: "k_sweep_pos", L.K_NARROW: "k_narrow",
            L.K_JAC_VEL: "k_jac_vel", L.K_TILE_BIG: "k_t_big", L.K_TILE_SMALL: "k_t_small",
            L.K_TILE_GIANT: "k_t_giant", L.K_ISLANDS: "k_t_islands_lds", L.K_COLOR: "k_t_color"}


def read_tile_work(e, L):
    """The tile kernels' device work counters since me_profile(e, 1): read right after the window,
    since they keep counting while the events are off."""
    return [int(v) for v in e.buffer(L.BUF_TILE_WORK, np.int64, (12,))]


def kernel_roofline(e, L, solver, V, P, ncol_pairs=None, work=None):
    """Per-kernel HIP-event times over the profiled window and the roofline of the dominant kernel
    (largest summed time). Algorithmic bytes of the tile kernels come from the work they counted on
    the device over the same window (ME_BUF_TILE_WORK): per launch warm + V * vel + P * pos with
    vel = 20*pairs + 100*dynamic pair ends + 64*contacts, pos = 20*pairs + 100*ends + 52*contacts,
    warm = pos under ColorCache. Kernels without a byte model report achieved = null."""
    import ctypes as C
    Lib = L.lib()
    prof = {}
    for kid, nm in _kernel_names(L).items():
        tot, n = C.c_double(), C.c_int64()
        L.check(Lib.me_profile_read(e._h, kid, C.byref(tot), C.byref(n)), "me_profile_read")
        if n.value:
            prof[nm] = (tot.value, n.value)
    region = None
    tot, n = C.c_double(), C.c_int64()
    L.check(Lib.me_profile_read(e._h, L.K_TILED, C.byref(tot), C.byref(n)), "me_profile_read")
    if n.value:
        region = (tot.value, n.value)
    if work is None:
        work = read_tile_work(e, L)
    per_launch, unit_work = {}, {}
    for k, nm in enumerate(("k_t_big", "k_t_small", "k_t_giant")):
        if nm not in prof:
            continue
        tiles, bp, bd, bc = work[4 * k: 4 * k + 4]
        nl = max(prof[nm][1], 1)
        vel = 20 * bp + 100 * bd + 64 * bc
        pos = 20 * bp + 100 * bd + 52 * bc
        warm = pos if solver == "color_cache" else 0
        per_launch[nm] = (warm + V * vel + P * pos) / nl
        unit_work[nm] = {"tiles": tiles / nl, "pairs": bp / nl, "dynamic_pair_ends": bd / nl, "contacts": bc / nl}
    if ncol_pairs is not None:
        vel_b, pos_b, ncol = ncol_pairs
        per_launch["k_sweep_vel"] = vel_b / ncol
        per_launch["k_sweep_pos"] = pos_b / ncol
    roof = None
    if prof:
        dom = max(prof, key=lambda k: prof[k][0])
        avg_s = prof[dom][0] / prof[dom][1] / 1e3
        b = per_launch.get(dom)
        ach = None if b is None else b / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": None if ach is None else round(ach, 3),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None if ach is None else round(ach / HBM_PEAK_GBS, 6),
                "traffic": None, "bytes_per_launch": None if b is None else int(b),
                "avg_launch_us": round(avg_s * 1e6, 3), "launches_timed": prof[dom][1],
                "work_per_launch": unit_work.get(dom)}
    kms = {k: {"ms_total": round(v[0], 4), "launches": v[1], "avg_us": round(1e3 * v[0] / v[1], 3)}
           for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    if region:
        kms["(tiled solve region on the step stream)"] = {"ms_total": round(region[0], 4), "launches": region[1],
                                                          "avg_us": round(1e3 * region[0] / region[1], 3)}
    return roof, kms


SETTLE_PILE = 400


def bench_pile(a, large=False):
    """Config 2 (10 000-box pile) or, with large=True, config 5 (100 000 boxes, mass ratio
    log-U[1, 100]). The pile first falls and settles for `settle` untimed steps (400 by default:
    the 10 k pile carries about 32 000 contacts by then), then `warmup` untimed steps, then `steps`
    timed steps with every kernel of interest timed by HIP events; the run then continues untimed
    to the config's full length (600 steps; 300 for the large pile) so that the whole run's
    average is reported beside the settled window."""
    import torch
    pkg = importlib.import_module("3fold-motion-engine_amd")
    L = importlib.import_module("3fold-motion-engine_amd._lib")
    large = large or bool(getattr(a, "large", False))
    nb = 100000 if large and a.bodies == 10000 else a.bodies
    full = 300 if large else 600
    settle = a.settle if a.settle is not None else SETTLE_PILE
    bodies = (pkg.scenes.box_pile(nb, seed=a.seed, mass_ratio=100.0) if large
              else pkg.scenes.box_pile(nb, seed=a.seed))
    cls = {"color_cache": pkg.ColorCacheContactEngine, "pair_chunks": pkg.GraphColoredContactEngine,
           "jacobi": pkg.RelaxedJacobiContactEngine}[a.solver]
    kw = dict(vit=a.vit, pit=a.pit, check_every_step=False)
    if a.solver == "pair_chunks":
        kw.update(pair_chunks=True, manifold_reduce=True)
    if a.solver != "jacobi":
        kw.update(n_colors=a.colors, tiled=bool(a.tiled or large), tile_bodies=a.tile_bodies or 128,
                  graph_capture=bool(a.graph))
    e = cls(**kw)
    e.add_boxes(bodies)
    dt, S = 1 / 240, a.substeps
    Lib = L.lib()
    seg = {}

    def run(n, name):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(n):
            e.step(dt, S)
            if (i + 1) % 100 == 0:
                torch.cuda.synchronize()
                print(f"[bench] {name} {i + 1}/{n} {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        seg[name] = time.perf_counter() - t

    run(settle, "settle")
    run(a.warmup, "warmup")
    start = e.refresh_counters()
    print(f"[bench] at timed start (step {settle + a.warmup}): {start}", file=sys.stderr, flush=True)
    x0, q0, v0, w0 = e.state_numpy()
    fast = None
    if a.fast_math and a.solver == "color_cache" and (a.tiled or large):
        fast = fast_math_engines(e, bodies, a)  # copies of the state at the window's start
    # kernel timing needs per-launch events, so steps taken while profiling run eagerly; with --graph
    # the timed region replays graphs and a second window of the same length is profiled eagerly
    if not a.graph:
        L.check(Lib.me_profile(e._h, 1), "me_profile")
    run(a.steps, "timed")
    elapsed = seg["timed"]
    eager_ms = None
    if a.graph:
        L.check(Lib.me_profile(e._h, 1), "me_profile")
        run(a.steps, "eager_profiled")
        eager_ms = 1e3 * seg["eager_profiled"] / a.steps
    L.check(Lib.me_profile(e._h, 0), "me_profile")
    cnt = e.refresh_counters()
    work = read_tile_work(e, L)
    print(f"[bench] at timed end: {cnt}", file=sys.stderr, flush=True)
    pairs = e.pairs()
    done = settle + a.warmup + a.steps * (2 if a.graph else 1)
    if done < full:
        run(full - done, "rest")
    ms_per_step = 1e3 * elapsed / a.steps
    value = nb * a.steps / elapsed
    ab = algorithmic_bytes(pairs, bodies.invM, a.solver, a.vit, a.pit, S, nb, cnt["n_contacts"])
    roofline, kms = kernel_roofline(e, L, a.solver, a.vit, a.pit,
                                    (ab["vel_sweep"], ab["pos_sweep"], max(cnt["n_colors"], 1)), work=work)
    dbg = e.buffer(L.BUF_DEBUG, np.int64, (16,))
    if dbg[6] > dbg[0] > 0:
        ph = np.diff(dbg[:7]).astype(np.float64)
        fr = ph / ph.sum()
        print(f"[bench] largest big tile ({dbg[7]} bodies, {dbg[8]} pairs), last substep, share of its time: "
              f"load {fr[0]:.3f} count {fr[1]:.3f} records {fr[2]:.3f} vel {fr[3]:.3f} pos {fr[4]:.3f} "
              f"writeback {fr[5]:.3f} ({int(ph.sum())} wall-clock ticks)", file=sys.stderr, flush=True)
    whole = sum(seg[k] for k in ("settle", "warmup", "timed", "rest") if k in seg)
    n_whole = settle + a.warmup + a.steps + (full - done if done < full else 0)
    cpu = cpu_all = None
    if not a.no_cpu and a.cpu_seconds > 0:
        model = cpu_model()
        steps_c, secs_c = cpu_baseline(x0, q0, v0, w0, bodies, a.solver, a.vit, a.pit, S,
                                       a.colors if a.solver != "jacobi" else 0, a.cpu_seconds)
        cpu = {"value": round(nb * steps_c / secs_c, 1), "unit": "body-steps/s", "cores": 1, "kind": "port",
               "sample": f"CPU oracle (C, -O3, 1 thread) from the GPU state at the start "
                         f"of the timed window (step {settle + a.warmup}): {steps_c} steps of the same {nb}-box "
                         f"pile in {secs_c:.1f} s on {model} ({os.cpu_count()} visible cores)"}
        if a.solver != "jacobi":
            threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
            os.environ["OMP_NUM_THREADS"] = str(threads)
            steps_m, secs_m = cpu_baseline(x0, q0, v0, w0, bodies, a.solver, a.vit, a.pit, S, a.colors,
                                           a.cpu_seconds, omp=True)
            cpu_all = {"value": round(nb * steps_m / secs_m, 1), "unit": "body-steps/s", "cores": threads,
                       "kind": "port",
                       "sample": f"the same oracle built -O3 with OpenMP over the pairs of a colour ({threads} "
                                 f"threads of {model}; contact generation, colouring and integration stay "
                                 f"serial): {steps_m} steps in {secs_m:.1f} s"}
    tag = (f"{nb}-box pile" + (", mass ratio log-U[1,100]" if large else "") +
           f", {a.solver}, {a.colors} colours, S{S} V{a.vit} P{a.pit}, dt 1/240" +
           (", tiled" if (a.tiled or large) else ", plain launches"))
    out = {"metric": METRIC, "value": round(value, 1), "unit": "body-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded box pile)",
           "config": {"workload": tag, "bodies": nb, "solver": a.solver, "colors": a.colors, "substeps": S,
                      "vit": a.vit, "pit": a.pit, "tiled": bool(a.tiled or large), "graph": bool(a.graph),
                      "settle_steps_untimed": settle, "timed_window_steps": [settle + a.warmup,
                                                                            settle + a.warmup + a.steps],
                      "contacts_at_timed_start": start["n_contacts"], "pairs_at_timed_start": start["n_pairs"],
                      "contacts": cnt["n_contacts"], "pairs": cnt["n_pairs"], "colors_used": cnt["n_colors"],
                      "islands": cnt["n_islands"], "tiles": cnt["n_tiles"]},
           "roofline": roofline, "cpu_baseline": cpu, "cpu_baseline_all_cores": cpu_all,
           f"run_{full}": {"steps": n_whole, "ms_per_step": round(1e3 * whole / n_whole, 4),
                           "note": "every step from the initial state, timed window included (kernel events on)"},
           "kernel_ms": kms}
    if eager_ms is not None:
        out["kernel_timing_window"] = {"mode": "eager, right after the timed graph replays", "steps": a.steps,
                                       "ms_per_step": round(eager_ms, 4)}
    if fast is not None:
        out["fast_math"] = fast_math_leg(fast, a, L, S, settle + a.warmup, round(ms_per_step, 4), nb)
    del e
    return out


def copy_engine_state(src, dst):
    """dst (built, same configuration but the arithmetic flag) takes src's complete state: the
    leading workspace bytes me_state_size names (warm-start store, colour cache, status) and the bodies."""
    import ctypes as C
    L = importlib.import_module("3fold-motion-engine_amd._lib")
    sb = C.c_size_t()
    L.check(L.lib().me_state_size(src._h, C.byref(sb)), "me_state_size")
    n = int(sb.value)
    dst._ws[dst._ws_off:dst._ws_off + n].copy_(src._ws[src._ws_off:src._ws_off + n])
    for k in ("x", "q", "v", "w"):
        getattr(dst, k).copy_(getattr(src, k))


def fast_math_engines(e, bodies, a):
    """A fast-arithmetic engine and a second exact one, both holding e's complete state (taken at the
    start of the timed window)."""
    pkg = importlib.import_module("3fold-motion-engine_amd")
    kw = dict(vit=a.vit, pit=a.pit, check_every_step=False, n_colors=a.colors, tiled=True,
              tile_bodies=a.tile_bodies or 128)
    out = []
    for fm in (True, False):
        g = pkg.ColorCacheContactEngine(fast_math=fm, **kw)
        g.add_boxes(bodies)
        g._build()
        copy_engine_state(e, g)
        out.append(g)
    return out


def fast_math_leg(engines, a, L, S, at_step, exact_ms, nb):
    """The pile under fast arithmetic (ColorCacheContactEngine(fast_math=True)) from the exact engine's
    state at the timed window's start (step `at_step`): the fast engine and an exact copy take one step
    (the fast step's deviation from the exact one: centres in m, orientations, impulses relative in the
    2-norm and relative to the largest of their kind), then the fast engine steps `steps` more along its
    own trajectory, timed with kernel events as the exact window was."""
    import torch
    f, g = engines
    dt = 1 / 240
    L.check(L.lib().me_profile(f._h, 1), "me_profile")
    torch.cuda.synchronize()
    t = time.perf_counter()
    f.step(dt, S)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t
    g.step(dt, S)
    f.refresh_counters()
    g.refresh_counters()
    xe, qe, _, _ = g.state_numpy()
    xf, qf, _, _ = f.state_numpy()
    ce, cf = g.contacts(), f.contacts()
    dev = {"x_m": float(np.max(np.abs(xf.astype(np.float64) - xe))),
           "q": float(np.max(np.abs(qf.astype(np.float64) - qe))), "same_contacts": bool(np.array_equal(ce["key"], cf["key"]))}
    if dev["same_contacts"] and len(ce["key"]):
        for k in ("jn", "jt1", "jt2", "jp"):
            d = cf[k].astype(np.float64) - ce[k]
            dev[k + "_rel_norm"] = float(np.linalg.norm(d)) / max(float(np.linalg.norm(ce[k])), 1e-30)
            dev[k + "_max_over_max"] = float(np.max(np.abs(d))) / max(float(np.max(np.abs(ce[k]))), 1e-30)
    del g
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps - 1):
        f.step(dt, S)
    torch.cuda.synchronize()
    el = time.perf_counter() - t + t_first
    L.check(L.lib().me_profile(f._h, 0), "me_profile")
    _, kms = kernel_roofline(f, L, a.solver, a.vit, a.pit)
    cnt = f.refresh_counters()
    print(f"[bench] fast-math leg: {1e3 * el / a.steps:.3f} ms/step", file=sys.stderr, flush=True)
    return {"note": "ColorCacheContactEngine(fast_math=True) from the exact engine's complete state at the "
                    f"timed window's start (step {at_step}), {a.steps} steps along its own trajectory, kernel "
                    "events on as in the exact window; deviation: its first step against an exact copy's",
            "window_steps": [at_step, at_step + a.steps], "ms_per_step": round(1e3 * el / a.steps, 4),
            "value": round(nb * a.steps / el, 1), "exact_ms_per_step": exact_ms,
            "kernel_ms": {k: v for k, v in list(kms.items())[:6]},
            "one_step_deviation_from_exact": dev, "contacts_at_end": cnt["n_contacts"],
            "islands_at_end": cnt["n_islands"], "max_tile_bodies_at_end": cnt["max_tile_bodies"]}


def _extra(fn, a, keys, **over):
    """Run another workload with its own settings and keep the named fields of its result."""
    b = argparse.Namespace(**vars(a))
    for k, v in over.items():
        setattr(b, k, v)
    print(f"[bench] extra workload {fn.__name__}", file=sys.stderr, flush=True)
    r = fn(b)
    return {k: r[k] for k in keys if k in r}


def _copy_kernel(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] calibration copy: {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)


def bench_adjoint(a):
    """Differentiable rollout: 1024 scenes of 64-box piles (v0, w0 ~ N(0, 0.1), targets U(-0.5, 0.5)^3),
    32 steps of the ColorCache step (S2 V16 P10, 16 colours), checkpoints every k = 8 steps; loss =
    sum of squared distances of the final centres to the targets; gradient w.r.t. v0, w0 and the
    per-body friction through the hand-written adjoint behind torch.autograd.Function. One warm-up
    rollout, then --rollouts timed forward + backward rollouts."""
    import torch
    pkg = importlib.import_module("3fold-motion-engine_amd")
    n_sc, n, T, k = 1024, 64, 32, 8
    b, target = pkg.scenes.fixed_piles(n_sc, n, seed=a.seed)
    N = b.n
    dt, S = 1 / 240, a.substeps
    ro = pkg.adjoint.Rollout(b, steps=T, dt=dt, substeps=S, vit=a.vit, pit=a.pit, n_colors=a.colors, k=k,
                             tile_bodies=a.tile_bodies or 256)
    fn = pkg.adjoint.contact_rollout(ro)
    dev = ro.eng.x.device
    f = lambda x: torch.tensor(np.asarray(x), dtype=torch.float32, device=dev)
    tgt = f(target)
    times = []
    gnorm = None
    for r in range(1 + a.rollouts):
        v0 = f(b.v).requires_grad_(True)
        w0 = f(b.w).requires_grad_(True)
        mu = f(b.mu).requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xT = fn(v0, w0, mu)
        loss = torch.sum((xT - tgt) ** 2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        print(f"[bench] rollout {r}: forward {1e3 * (t1 - t0):.1f} ms, backward {1e3 * (t2 - t1):.1f} ms, "
              f"loss {loss.item():.4f}", file=sys.stderr, flush=True)
        if r > 0:
            times.append((t1 - t0, t2 - t1))
        gnorm = [float(torch.linalg.vector_norm(g).item()) for g in (v0.grad, w0.grad, mu.grad)]
    fw = sum(t[0] for t in times) / len(times)
    bw = sum(t[1] for t in times) / len(times)
    ms_per_step = 1e3 * (fw + bw) / T
    out = {"metric": METRIC, "value": round(N * T / (fw + bw), 1), "unit": "body-steps/s", "n_gpus": 1,
           "steps": T * a.rollouts, "warmup": T, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded box piles)",
           "config": {"workload": f"{n_sc} x {n}-box piles, {T} steps, checkpoint k={k}, ColorCache S{S} V{a.vit} "
                                  f"P{a.pit}, forward + hand-written adjoint (grad of v0, w0, mu)",
                      "bodies": N, "scenes": n_sc, "steps_per_rollout": T, "checkpoint_k": k,
                      "tile_bodies": a.tile_bodies or 256},
           "adjoint": {"forward_ms_per_step": round(1e3 * fw / T, 4), "backward_ms_per_step": round(1e3 * bw / T, 4),
                       "backward_over_forward": round(bw / fw, 2),
                       "adjoint_ms_per_step": round(ms_per_step, 4), "rollouts_timed": len(times),
                       "grad_norms_v0_w0_mu": gnorm},
           "roofline": None, "cpu_baseline": None}
    return out


def bench_ncp(a):
    """NCP-level sensitivity, the reference's batched adjoint workload: 4096 stance problems of 4 point
    contacts on 18 generalised coordinates (synthetic G, b, mu with exactly known solutions, every
    contact drawn open / stick / slip), dlam/db along 2 directions per problem plus dlam/dmu of every
    contact, float64, inputs resident on the GPU; one me_ncp_sensitivity launch per call. The
    reference's number is 119.6 ms per 2 directions at B = 4096 (68 488 env-grads/s, RTX 5070)."""
    import ctypes as C
    import torch
    ncp = importlib.import_module("3fold-motion-engine_amd.ncp")
    L = importlib.import_module("3fold-motion-engine_amd._lib")
    E, n_c, m = 4096, 4, 2
    G, b, mu, lam, labels = ncp.synthetic_problems(E, n_c=n_c, seed=a.seed)
    D = np.random.default_rng(a.seed + 1).standard_normal((E, 3 * n_c, m))
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(np.asarray(x), dtype=torch.float64, device=dev).contiguous()
    Gd, bd, md, ld, Dd = t(G), t(b), t(mu), t(lam), t(D)
    dmu = torch.empty((E, 3 * n_c, n_c), dtype=torch.float64, device=dev)
    dlb = torch.empty((E, 3 * n_c, m), dtype=torch.float64, device=dev)
    lab = torch.empty((E, n_c), dtype=torch.int32, device=dev)
    st = torch.empty(E, dtype=torch.int32, device=dev)
    lib = L.lib()
    nb = C.c_size_t()
    L.check(lib.me_ncp_workspace_size(E, n_c, m, C.byref(nb)), "me_ncp_workspace_size")
    ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    p = lambda x: C.c_void_p(x.data_ptr())
    args = (E, n_c, m, p(Gd), p(bd), p(md), p(ld), p(Dd), None, 1e-10, 1e-6, 1e-12, p(dmu), p(dlb), p(lab),
            p(st), p(ws), int(nb.value), C.c_void_p(stream))
    for _ in range(20):
        L.check(lib.me_ncp_sensitivity(*args), "me_ncp_sensitivity")
    torch.cuda.synchronize()
    reps = 200
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        L.check(lib.me_ncp_sensitivity(*args), "me_ncp_sensitivity")
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    assert np.array_equal(lab.cpu().numpy(), labels) and int(st.abs().sum()) == 0
    value = E * m / (ms * 1e-3)
    cpu = None
    if not a.no_cpu:
        from oracle import ncp_oracle as R
        n_s, t0 = 256, time.perf_counter()
        for e in range(n_s):
            R.sensitivity(lam[e], G[e], b[e], mu[e], [R.LABELS[k] for k in labels[e]], db_dparam=D[e])
        dt_cpu = time.perf_counter() - t0
        cpu = {"value": round(n_s * m / dt_cpu, 1), "unit": "env-grads/s", "cores": 1, "kind": "port",
               "sample": f"{n_s} of the {E} problems through the numpy oracle (SVD rank guard + LAPACK solve)"}
    return {"metric": "NCP sensitivity env-grads/s (B=4096, n_c=4, 2 directions)", "value": round(value, 1),
            "unit": "env-grads/s", "ms_per_call": round(ms, 5), "higher_is_better": True,
            "vs_reference": round(value / 68488.0, 1), "dtype": "f64", "data": "synthetic (exactly solved problems)",
            "config": {"workload": f"{E} stance problems x {n_c} contacts, nv 18, dlam/db along {m} directions + "
                                   f"dlam/dmu, one launch", "envs": E, "contacts": n_c, "directions": m},
            "cpu_baseline": cpu}


def pgs_algorithmic_bytes(bs, real_bytes):
    """Compulsory traffic of one sweep over every contact: per contact its order slot and body pair
    (12 B), rho, mu, bias and lambda read + write (9 reals); per present side its 3x6 block and the
    body's velocity (24 reals); per dynamic side the inverse mass, v_free, the written velocity
    (48 reals) and the int64 total read + written (96 B)."""
    R = real_bytes
    pab = np.asarray(bs.body_pairs)
    dyn = np.abs(bs.Minv).reshape(bs.n_b, 36).max(axis=1) > 0
    side = (pab >= 0)
    dside = side & dyn[np.maximum(pab, 0)]
    return int(bs.n_c * (12 + 9 * R) + side.sum() * 24 * R + dside.sum() * (48 * R + 96))


def bench_pgs(a):
    """The exact-cone PGS (NCPSolverGPU's selftest workload, ncp_gpu.py:1872-1880): a full solve() of
    synthetic_scene(n_c) over 512 bodies with 50 iterations, colour-cache hit, inputs uploaded from the
    host each call as in the reference, at n_c = 4, 32, 256, 4096; the device-resident 50-iteration
    launch at n_c = 4096 timed with HIP events (its roofline); and the N = 10 000 two-layer lattice
    step (host-side scene build + solve, 120 iterations, float32). The reference's numbers are
    0.73 / 1.44 / 4.32 / 47.5 ms per solve and 49.8 ms per lattice step (gpu_scene_reference.json)."""
    import torch
    m = importlib.import_module("3fold-motion-engine_amd.pgs")
    ref_ms = {4: 0.73, 32: 1.44, 256: 4.32, 4096: 47.5}
    solves = {}
    for n in (4, 32, 256, 4096):
        sc = m.synthetic_scene(n, seed=a.seed)
        s = m.NCPSolver(sc.n_b, n)
        for _ in range(3):
            s.solve(sc, iters=50, measure_residual=False)
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            r = s.solve(sc, iters=50, measure_residual=False)
        ms = (time.perf_counter() - t0) / reps * 1e3
        assert r.cache_hit
        solves[n] = {"ms": round(ms, 4), "reference_ms": ref_ms[n], "speedup": round(ref_ms[n] / ms, 1),
                     "colours": s.n_colors}
    # device-resident launch of 50 sweeps at n_c = 4096 (the dominant kernel, k_pgs_sweeps)
    sc = m.synthetic_scene(4096, seed=a.seed)
    s = m.NCPSolver(sc.n_b, sc.n_c)
    s.solve(sc, iters=1, measure_residual=False)
    for _ in range(3):
        s._iterate(50, 1)
    torch.cuda.synchronize()
    reps = 20
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        s._iterate(50, 1)
    ev1.record()
    torch.cuda.synchronize()
    k_ms = ev0.elapsed_time(ev1) / reps
    ab = 50 * pgs_algorithmic_bytes(sc, 4)
    gbs = ab / (k_ms * 1e-3) / 1e9
    # the lattice step
    w = m.LatticeWorld(n=10000, iters=120)
    f = m.lattice_pushed_forces(w)
    for _ in range(4):
        w.step(f_ext=f)
    torch.cuda.synchronize()
    ts, tsolve = [], []
    for _ in range(10):
        t0 = time.perf_counter()
        r = w.step(f_ext=f)
        ts.append((time.perf_counter() - t0) * 1e3)
        tsolve.append(sum(r.ms.values()))
    lat_ms = float(np.median(ts))
    cpu = None
    if not a.no_cpu:
        from oracle import pgs_oracle as PO
        t0 = time.perf_counter()
        nrep = 0
        while time.perf_counter() - t0 < 10.0:
            PO.solve(sc, iters=50, dtype="float32")
            nrep += 1
        cpu_ms = (time.perf_counter() - t0) / nrep * 1e3
        cpu = {"value": round(cpu_ms, 3), "unit": "ms per 50-iteration solve (n_c = 4096)", "cores": 1,
               "kind": "port", "sample": f"{nrep} solves of the n_c = 4096 scene through the C oracle"}
    return {"metric": "exact-cone PGS ms per 50-iteration solve (n_c = 4096, 512 bodies, colour-cache hit)",
            "value": solves[4096]["ms"], "unit": "ms", "higher_is_better": False,
            "vs_reference": round(ref_ms[4096] / solves[4096]["ms"], 1), "dtype": "f32",
            "data": "synthetic (the reference's synthetic_scene)", "solves": {str(k): v for k, v in solves.items()},
            "kernel": {"name": "k_pgs_sweeps", "ms_per_50_iterations": round(k_ms, 4),
                       "contact_updates_per_s": round(4096 * 50 / (k_ms * 1e-3), 1),
                       "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None,
                                    "algorithmic_bytes_per_launch": ab}},
            "lattice_10000": {"ms_per_step": round(lat_ms, 3), "solver_ms": round(float(np.median(tsolve)), 3),
                              "reference_ms": 49.8, "contacts": w.n_contacts, "colours": w.n_colors,
                              "iterations": 120},
            "config": {"workload": "synthetic_scene(n_c) over 512 bodies, 50 iterations; two-layer lattice N=10000",
                       "iters": 50},
            "cpu_baseline": cpu}


# published wall-clock ms/step of the reference engines on this fixture (all RTX 5070): README.md:37
# (ColorCache), docs/RUNNING.md:738 (pair chunks and Jacobi, first of two runs), docs/RUNNING.md:639
# (the default coloured engine: per-contact colouring of the unreduced contacts, uncaptured)
LATTICE_REF_MS = {"color_cache": 3.7, "pair_chunks": 4.931, "jacobi": 5.591, "colored": 30.629}


def bench_lattice(a):
    """The reference's own headline fixture (scripts/colored_vs_jacobi.py:104-135): N = 10 000 bead
    boxes 0.3 x 0.3 x 0.2 m in a two-layer lattice 0.35 m apart, dropped onto the ground; each engine
    built with its default arguments (vit 40, pit 20, mu 0.5) and stepped with step(1/240, substeps=1),
    10 warm-up steps then 30 timed, wall clock per step. The engines read their counters back after
    every step, as the reference's do. The coloured engines run their default island-tiled solve
    and, for comparison, the per-colour launches (tiled=False)."""
    import torch
    pkg = importlib.import_module("3fold-motion-engine_amd")
    cen = pkg.scenes.lattice(10000)
    cls = {"color_cache": pkg.ColorCacheContactEngine, "pair_chunks": pkg.GraphColoredContactEngine,
           "jacobi": pkg.RelaxedJacobiContactEngine, "colored": pkg.GraphColoredContactEngine}
    runs = []
    runs_wanted = [(r.split(":")[0], r.split(":")[1] == "1") for r in a.lattice_runs.split(",")]
    for solver, tiled in runs_wanted:
        ekw = {} if solver == "jacobi" else {"tiled": tiled}
        if solver == "pair_chunks":
            ekw.update(pair_chunks=True, manifold_reduce=True)
        e = cls[solver](**ekw)
        for c in cen:
            e.add_body(c)
        for _ in range(10):
            e.step(1 / 240, substeps=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(30):
            e.step(1 / 240, substeps=1)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / 30 * 1e3
        cnt = e.refresh_counters()
        r = {"solver": solver, "tiled": tiled, "ms_per_step": round(ms, 4), "reference_ms": LATTICE_REF_MS[solver],
             "speedup": round(LATTICE_REF_MS[solver] / ms, 2), "contacts": cnt["n_contacts"],
             "contacts_solved": cnt["n_solved"], "pairs": cnt["n_pairs"], "colours": cnt["n_colors"]}
        print(f"[bench] lattice {r}", file=sys.stderr, flush=True)
        runs.append(r)
        del e
        torch.cuda.empty_cache()
    best = min((r for r in runs if r["solver"] == "color_cache"), key=lambda r: r["ms_per_step"], default=runs[0])
    return {"metric": f"ms/step, reference lattice fixture N=10000 ({best['solver']}, vit 40, pit 20, S1)",
            "value": best["ms_per_step"], "unit": "ms/step", "higher_is_better": False,
            "vs_reference": best["speedup"], "reference": "RTX 5070, README.md:37 / docs/RUNNING.md:738",
            "dtype": "f32", "data": "the reference's lattice(10000) fixture", "runs": runs,
            "config": {"workload": "two-layer bead lattice, N=10000, dt 1/240, substeps 1, 10 warm-up + 30 timed "
                                   "steps", "bodies": 10000}}


def bench_pile_tuned(a):
    """The 10 000-box pile under the autotuner's choice of colour classes, substeps, iterations and
    tile size at a maximum penetration of 1e-3 m (the batch's penetration target). The trials start
    from the pile as settled by the default configuration; the chosen configuration then steps the
    pile from its initial state through the same warm-up, is timed over the same steps, and its
    maximum penetration is read over the next 100 steps (and the default configuration's over 100
    steps after its own warm-up, for comparison)."""
    import torch
    pkg = importlib.import_module("3fold-motion-engine_amd")
    at = importlib.import_module("3fold-motion-engine_amd.autotune")
    bodies = pkg.scenes.box_pile(a.bodies, seed=a.seed)
    dt = 1 / 240
    mk = lambda **k: pkg.ColorCacheContactEngine(tiled=True, check_every_step=False, **k)

    def max_pen_over(e, S, n):
        worst = 0.0
        for _ in range(n):
            e.step(dt, S)
            e.refresh_counters()
            worst = max(worst, at._max_penetration(e))
        return worst

    warm = mk(n_colors=a.colors, vit=a.vit, pit=a.pit, tile_bodies=a.tile_bodies or 128)
    warm.add_boxes(bodies)
    for _ in range(a.warmup):
        warm.step(dt, a.substeps)
    st = warm.state_numpy()
    pen_default = max_pen_over(warm, a.substeps, 100)
    del warm
    torch.cuda.empty_cache()
    res = at.autotune(mk, bodies, warm_state=st, target_pen=1e-3,
                      log=lambda m: print(m, file=sys.stderr, flush=True))
    if res.best is None:
        return {"error": "no configuration of the search met the penetration target", "trials": len(res.trials)}
    b = res.best
    e = mk(**res.as_kwargs())
    e.add_boxes(bodies)
    for _ in range(a.warmup):
        e.step(dt, b.substeps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        e.step(dt, b.substeps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    pen = max_pen_over(e, b.substeps, 100)
    return {"value": round(a.bodies * a.steps / el, 1), "unit": "body-steps/s", "ms_per_step": round(1e3 * el / a.steps, 4),
            "max_penetration_m": pen, "target_m": 1e-3, "default_config_max_penetration_m": pen_default,
            "config": {"colors": b.colours, "substeps": b.substeps, "vit": b.vit, "pit": b.pit,
                       "tile_bodies": b.tile_bodies},
            "trials": [{"colors": t.colours, "substeps": t.substeps, "vit": t.vit, "tile_bodies": t.tile_bodies,
                        "max_pen": t.max_pen, "ms_per_step": round(t.ms_per_step, 3)} for t in res.trials]}


def bench_batched(a):
    """Config 3: 256 independent piles (N_s log-uniform in [8, 2048]) in one tiled engine -- one
    launch sequence per substep serves every scene, tiles and big-tile lists being the work units --
    after `settle` untimed steps (400 by default) and `warmup` more, `steps` timed steps with the
    kernels timed by HIP events. With --autotune the autotuner first picks colours, substeps,
    iterations and tile size under the 1e-3 m penetration target (trials start from the settled
    state); when no configuration meets it, the line says so and the default configuration runs.
    The same window is then replayed as one captured hipGraph per step, and the same scenes, from
    the same settled state, are stepped one engine after another, once with the same tiled solve
    and once with plain per-colour launches (the sequential baselines)."""
    import torch
    pkg = importlib.import_module("3fold-motion-engine_amd")
    L = importlib.import_module("3fold-motion-engine_amd._lib")
    bodies, sizes = pkg.scenes.batched_piles(a.scenes, seed=a.seed)
    N = bodies.n
    dt = 1 / 240
    tb = a.tile_bodies or 224
    settle = a.settle if a.settle is not None else SETTLE_PILE
    kw = dict(n_colors=a.colors, vit=a.vit, pit=a.pit, tile_bodies=tb)
    S = a.substeps
    mk = lambda **k: pkg.ColorCacheContactEngine(tiled=True, check_every_step=False, **k)

    def steps(e, n, S_):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            e.step(dt, S_)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    base = mk(**kw)
    base.add_boxes(bodies)
    t_settle = steps(base, settle, S)
    print(f"[bench] batch settled: {settle} steps in {t_settle:.1f}s", file=sys.stderr, flush=True)
    st = base.state_numpy()
    tune = None
    tune_out = None
    if a.autotune:
        at = importlib.import_module("3fold-motion-engine_amd.autotune")
        res = at.autotune(mk, bodies, warm_state=st, steps=a.tune_steps, tile_bodies=(tb,),
                          log=lambda m: print(m, file=sys.stderr, flush=True))
        best_seen = min(res.trials, key=lambda t: t.max_pen) if res.trials else None
        tune_out = {"target_m": 1e-3, "trials": len(res.trials), "trial_steps": a.tune_steps,
                    "met_target": res.best is not None}
        if res.best is not None:
            tune = res.best
            tune_out.update(chosen={"colours": tune.colours, "substeps": tune.substeps, "vit": tune.vit,
                                    "pit": tune.pit, "tile_bodies": tune.tile_bodies},
                            max_penetration_m=tune.max_pen)
        else:
            tune_out.update(note="no configuration of the search met the 1e-3 m target; the default "
                                 "configuration runs", smallest_max_penetration_m=best_seen.max_pen if best_seen else None,
                            at={"colours": best_seen.colours, "substeps": best_seen.substeps,
                                "vit": best_seen.vit} if best_seen else None)
    if tune is not None:
        kw = res.as_kwargs()
        S = res.substeps
        del base
        base = mk(**kw)
        base.add_boxes(bodies)
        steps(base, settle, S)
    e = base
    steps(e, a.warmup, S)
    start = e.refresh_counters()
    x0, q0, v0, w0 = e.state_numpy()
    Lib = L.lib()
    L.check(Lib.me_profile(e._h, 1), "me_profile")
    t_b = steps(e, a.steps, S) / a.steps
    L.check(Lib.me_profile(e._h, 0), "me_profile")
    cnt = e.refresh_counters()
    roofline, kms = kernel_roofline(e, L, "color_cache", kw["vit"], kw["pit"], work=read_tile_work(e, L))
    pen_now = float(e.buffer_tensor(L.BUF_PEN, np.float32, cnt["n_solved"]).max().item()) if cnt["n_solved"] else 0.0

    def from_state(eng):
        eng.add_boxes(bodies)
        eng._build()
        with torch.no_grad():
            for name, arr in zip(("x", "q", "v", "w"), (x0, q0, v0, w0)):
                getattr(eng, name).copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(eng.x.device))
        return eng

    # the same window replayed as captured graphs (state restored, warm impulses rebuilt by a few steps)
    g = from_state(pkg.ColorCacheContactEngine(tiled=True, check_every_step=False, graph_capture=True, **kw))
    steps(g, 3, S)
    t_g = steps(g, a.steps, S) / a.steps
    del g
    # sequential baselines: one engine per scene, from the settled state
    offs = np.concatenate([[0], np.cumsum(sizes)])
    seq = {}
    for label, tiled in (("tiled", True), ("plain", False)):
        if a.seq_rounds <= 0:
            break
        engines = []
        for k in range(a.scenes):
            sl = slice(offs[k], offs[k + 1])
            part = pkg.bodies.BodySet(x=x0[sl].copy(), q=q0[sl].copy(), v=v0[sl].copy(), w=w0[sl].copy(),
                                      invM=bodies.invM[sl], IbInv=bodies.IbInv[sl], kin=bodies.kin[sl],
                                      half=bodies.half[sl], mu=bodies.mu[sl],
                                      scene=np.zeros(offs[k + 1] - offs[k], np.int32), M=bodies.M[sl])
            ek = pkg.ColorCacheContactEngine(tiled=tiled, check_every_step=False, **kw)
            ek.add_boxes(part)
            ek.step(dt, S)
            engines.append(ek)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(a.seq_rounds):
            for ek in engines:
                ek.step(dt, S)
        torch.cuda.synchronize()
        seq[label] = (time.perf_counter() - t0) / a.seq_rounds
        print(f"[bench] sequential {label}: {1e3 * seq[label]:.1f} ms per round of {a.scenes} scenes",
              file=sys.stderr, flush=True)
        del engines
        torch.cuda.empty_cache()
    sq = lambda k: ({"value": round(N / seq[k], 1), "unit": "body-steps/s", "ms_per_round": round(1e3 * seq[k], 3)}
                    if k in seq else None)
    out = {"metric": METRIC, "value": round(N / t_b, 1), "unit": "body-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(1e3 * t_b, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded box piles)",
           "config": {"workload": f"{a.scenes} independent box piles, N_s log-uniform [8, 2048], one tiled launch "
                                  f"sequence per substep", "bodies": int(N), "scenes": a.scenes,
                      "colors": kw.get("n_colors"), "substeps": S, "vit": kw.get("vit"), "pit": kw.get("pit"),
                      "autotuned": tune is not None, "settle_steps_untimed": settle,
                      "contacts_at_timed_start": start["n_contacts"], "pairs": cnt["n_pairs"],
                      "islands": cnt["n_islands"], "tile_bodies": kw.get("tile_bodies"), "tiles": cnt["n_tiles"],
                      "max_tile_bodies": cnt["max_tile_bodies"], "max_penetration_m_at_window_end": pen_now},
           "roofline": roofline, "kernel_ms": kms,
           "graph_replay": {"ms_per_step": round(1e3 * t_g, 4), "value": round(N / t_g, 1),
                            "note": "the same window, each step one captured hipGraph"},
           "autotune": tune_out,
           "sequential_tiled": sq("tiled"), "sequential_plain_launches": sq("plain"),
           "speedup_vs_sequential_tiled": round(seq["tiled"] / t_b, 2) if "tiled" in seq else None,
           "speedup_vs_sequential_plain": round(seq["plain"] / t_b, 2) if "plain" in seq else None}
    return out


if __name__ == "__main__":
    main()
