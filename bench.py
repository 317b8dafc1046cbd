#!/usr/bin/env python3
"""Benchmark: CFR+ iterations/s on the heads-up turn+river subgame, batched
BF16 hi+lo split showdown path vs the plain FP32 path, on one MI355X.

Workload (seed 0): a seeded flop; the root deals the turn (49 cards), turn
betting, river chance (48), river betting, showdown.  Pot 6bb, stacks 40bb,
per street 33%/75% pot, one 2.5x raise and all-in; Dirichlet(0.5) ranges
over the 1326 combos (926,542 nodes).

A second batched leg computes the river showdown values by exact prefix
scans in strength order instead of the GEMM (same game, same qualifier
measurement; `batched_river_scan`; skip with --no-scan).

Each path: `--warmup` untimed iterations, `--steps` timed iterations
(`value` = the batched path's timed iterations/s, all state resident in
HBM), then untimed iterations up to `--solve-iters` (default 1250, what
CFR+ needs on this game to bring the FP64 best-response exploitability
under 0.5% of the pot) with the exploitability measured on the GPU every
50 iterations, so that every run reports the iteration count at which the
qualifier is met, the exploitability at the configuration's 500 iterations
and at the end, for both paths, and the strategy L-inf between them.

Prints ONE JSON line, with the per-kernel roofline of the dominant kernel
(HIP events on the solver's stream) and the CPU oracle timed on a bounded
sample on this host at one thread and at all threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "CFR+ iter/s at ≤0.5% pot exploitability: batched BF16-MFMA vs plain FP32"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# INT32 VALU: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane operations per
# second; the count builder spends 2 (packed 16-bit) operations per
# pair-runout comparison, which counts both W and T of an unordered pair
VALU_LANE_OPS = 256 * 64 * 2.4e9
PAIR_RUNOUTS_PER_FLOP = 635628 * 990          # unordered unblocked pairs x runouts avoiding both hands


def load(name, path):
    import importlib.util
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def river_sweep_mode():
    """The library's river showdown mode (environment, as the library reads it)."""
    return not os.environ.get("MTT_RIVER_GEMM") and os.environ.get("MTT_RIVER_SD") == "sweep"


def algorithmic_work(S, game):
    """Per-iteration algorithmic work of the river kernels of the batched path
    (n = 1081 live combos per river board, 49x48 boards).

    River showdown GEMM (default): 2 flops per (hero, villain, column) over
    the live combos of each board: 2 * n^2 * (2 * showdowns); the hi+lo
    split's second product is not counted.
    River showdown, exact strength-order sweep (MTT_RIVER_SD=sweep): each
    column's reach read and its values written (4 + 4 B per live combo) and
    the column's 52 card totals read from both forward halves (832 B).
    River showdown by prefix scans (MTT_RIVER_SD=scan, the bench's second
    batched leg): each column read and written (4 + 4 B per live combo) and
    the board's scan tables read once.
    River backward: regrets read and written (8 B per action slot and
    combo), showdown values read (4 B per column entry), card sums of each
    terminal read (2 players x 53 x 8 B).
    River forward: regrets read (4 B), strategy sums read-modify-write (8 B),
    reach columns out (4 B per column entry: FP32, or BF16 hi + lo for the
    GEMM), card sums out."""
    nodes, ni, nm = S.tree(game)
    river = nodes["river_idx"] >= 64
    river_dec = int(((nodes["type"] == 0) & river).sum())
    river_sd = int(((nodes["type"] == 2) & river).sum())
    river_terms = int((((nodes["type"] == 1) | (nodes["type"] == 2)) & river).sum())
    act = nodes["num_actions"].astype(np.int64)
    river_actions = int(act[(nodes["type"] == 0) & river].sum())
    boards = 49 * 48
    n = 1081
    sums = river_terms * 2 * 53 * 8
    gemm_flops = 2.0 * n * n * (2 * river_sd)
    showdown_bytes = 2 * river_sd * (n * 8 + 2 * 52 * 8)
    backward_bytes = river_actions * n * 8 + 2 * river_sd * n * 4 + sums
    forward_bytes = river_actions * n * (4 + 8) + 2 * river_sd * n * 4 + sums
    return dict(boards=boards, river_decisions=river_dec, river_showdowns=river_sd, river_terminals=river_terms,
                river_actions=river_actions, gemm_flops=gemm_flops, showdown_bytes=showdown_bytes,
                backward_bytes=backward_bytes,
                forward_bytes=forward_bytes)


def hbm_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (profiles/hbm_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        return k["fetch_bytes"] + k["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_info():
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
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, os.cpu_count(), avail


def cpu_baseline(S, O, game, r0, r1, seconds=15.0):
    """Oracle (FP32 host restatement of the same recurrence) on the same
    game: whole CFR+ iterations of the full game, as many as fit in
    `seconds` after the first, at all OpenMP threads and at one thread."""
    nodes, ni, _ = S.tree(game)
    board = list(game.board)[:3]
    gi, _ = S.alive_index(board)
    all_threads = O.max_threads()

    h = O.HU(nodes, ni, board, r0[gi], r1[gi], 1)
    t0 = time.time()
    h.iterate(1)                        # first touch of the regret / strategy-sum pages
    first = time.time() - t0

    def sample(threads, budget, cap):
        O.set_threads(threads)
        n, t1 = 0, time.time()
        while n < cap and (n == 0 or time.time() - t1 < budget):
            h.iterate(1)
            n += 1
        return n / (time.time() - t1), n

    v_all, n_all = sample(all_threads, seconds, 50)
    v_one, n_one = sample(1, seconds, 1)
    O.set_threads(all_threads)
    h.close()
    model, ncpu, avail = cpu_info()
    return {"value": v_all, "unit": "iter/s", "cores": all_threads, "kind": "port",
            "single_thread_value": v_one, "cpu_model": model, "nproc": ncpu, "cpus_available": avail,
            "sample": f"full CFR+ iterations of the same turn+river game on the FP32 oracle "
                      f"(OpenMP over turn cards), one handle: an untimed first iteration ({first:.1f}s, first "
                      f"touch of the state), then {n_all} timed at {all_threads} threads and {n_one} at 1 thread"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solve_checkpoints(solver, total, every=50, name=""):
    """Untimed iterations up to `total` with the FP64 best-response
    exploitability (% of the starting pot) every `every` iterations and at
    500: returns [(iterations, exploitability)]."""
    marks = sorted(set([c for c in range(every, total + 1, every)] + [500, total]))
    out = []
    for c in marks:
        if c < solver.iterations or c > total:
            continue
        if c > solver.iterations:
            solver.run(c - solver.iterations)
        out.append((solver.iterations, float(solver.exploitability()[5])))
        log(f"{name}: {out[-1][0]} iterations, exploitability {out[-1][1]:.4f}% pot")
    return out


def expl_summary(curve):
    first = next((n for n, e in curve if e <= 0.5), None)
    at = {n: e for n, e in curve}
    return {"iterations_to_0p5pct": first, "exploitability_at_500": at.get(500),
            "exploitability_pct_pot": curve[-1][1] if curve else None,
            "iterations": curve[-1][0] if curve else None,
            "meets_0p5pct": bool(curve and curve[-1][1] <= 0.5),
            "curve": [[n, round(e, 5)] for n, e in curve]}


def strategy_linf(a, b, S_ref, ni, n_alive, chunk=8192):
    """Average-strategy differences over (infoset, combo) entries with
    strategy mass (sum over actions of the reference path's strategy sums
    >= 1e-5, resp. >= 0.1, of the infoset's largest): L-inf over each set and
    the fraction of the first set that differs by more than 1e-2."""
    a = a.reshape(ni, 4, -1)
    b = b.reshape(ni, 4, -1)
    w = S_ref.reshape(ni, 4, -1)
    live, heavy, n_live, n_off = 0.0, 0.0, 0, 0
    for lo in range(0, ni, chunk):
        hi = min(ni, lo + chunk)
        d = np.abs(a[lo:hi, :, :n_alive] - b[lo:hi, :, :n_alive]).max(1)
        m = w[lo:hi, :, :n_alive].sum(1)
        mm = m.max(1, keepdims=True)
        keep = (m >= 1e-5 * mm) & (m > 0)
        big = (m >= 0.1 * mm) & (m > 0)
        if keep.any():
            live = max(live, float(d[keep].max()))
            n_live += int(keep.sum())
            n_off += int((d[keep] > 1e-2).sum())
        if big.any():
            heavy = max(heavy, float(d[big].max()))
    return {"linf_mass_1e-5": live, "linf_mass_0.1": heavy,
            "frac_mass_1e-5_off_by_1e-2": n_off / max(n_live, 1)}


def run_flop_counts(args):
    """Exhaustive win/tie counts of all 1,755 canonical flops (global 1326
    layout, int32 W and T), `--warmup` untimed passes then `--steps` timed
    passes; a step is one pass over all flops in batches of 128."""
    import ctypes
    import torch
    mtt = load("mtt", os.path.join(ROOT, "mtt-gpu-solver_amd", "mtt.py"))
    canon = np.zeros((1755, 4), np.uint8)
    assert mtt.lib().mtt_canonical_flops(canon.ctypes.data) == 1755
    flops = np.ascontiguousarray(canon[:, :3])
    B = 128
    W = torch.empty((B, 1326, 1326), dtype=torch.int32, device="cuda:0")
    T = torch.empty_like(W)

    def one_pass():
        for f0 in range(0, 1755, B):
            n = min(B, 1755 - f0)
            mtt.call("mtt_flop_counts", flops[f0:f0 + n].ctypes.data, n, W.data_ptr(), T.data_ptr(),
                     mtt.LAYOUT_GLOBAL1326, None)

    for _ in range(args.warmup):
        one_pass()
    torch.cuda.synchronize()
    mtt.call("mtt_flop_counts_timing", 1)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    t = time.time()
    for _ in range(args.steps):
        one_pass()
    ev1.record()
    torch.cuda.synchronize()
    dt = time.time() - t
    ktimes = {}
    k = 0
    while True:
        name = ctypes.create_string_buffer(32)
        ms, n = ctypes.c_double(), ctypes.c_int64()
        if mtt.lib().mtt_flop_counts_kernel_times(k, name, 32, ctypes.byref(ms), ctypes.byref(n)) != 0:
            break
        ktimes[name.value.decode()] = (ms.value, n.value)
        k += 1
    mtt.call("mtt_flop_counts_timing", 0)
    launches = ktimes["fc_pairs"][1]
    avg_ms = ktimes["fc_pairs"][0] / launches
    flops_per_launch = 1755 * args.steps / launches
    peak = VALU_LANE_OPS / 2 / 1e12
    achieved = PAIR_RUNOUTS_PER_FLOP * flops_per_launch / (avg_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        try:
            O = load("oracle_ctypes", os.path.join(ROOT, "oracle", "oracle_ctypes.py"))
            O.flop_counts(flops[0])
            n_s, t0 = 0, time.time()
            while n_s < 40 and (n_s < 2 or time.time() - t0 < 15.0):
                O.flop_counts(flops[(n_s * 97) % 1755])
                n_s += 1
            model, ncpu, avail = cpu_info()
            cpu = {"value": n_s / (time.time() - t0), "unit": "flops/s", "cores": O.max_threads(), "kind": "port",
                   "cpu_model": model, "nproc": ncpu,
                   "sample": f"{n_s} canonical flops' exhaustive counts on the C oracle (OpenMP over runouts and rows)"}
        except Exception as ex:
            cpu = {"error": str(ex)}
    out = {
        "metric": "exhaustive flop W/T count builds per second (all 1,755 canonical flops, bit-exact)",
        "value": 1755 * args.steps / dt, "unit": "flops/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (the 1,755 canonical flops)",
        "config": {"workload": "flop_counts_1755", "layout": "global1326", "batch_flops": B,
                   "pair_runouts_per_flop": PAIR_RUNOUTS_PER_FLOP},
        "event_ms_per_step": ev0.elapsed_time(ev1) / args.steps,
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in ktimes.items()},
        "roofline": {"kernel": "k_fc_pairs", "bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "Tcompare/s", "frac": achieved / peak, "traffic": None, "avg_ms": avg_ms,
                     "note": "algorithmic pair-runout comparisons (unordered unblocked pairs x 990 runouts) per "
                             "second; peak = INT32 VALU lane rate / 2 operations per comparison"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_preflop(args):
    """Preflop DCFR co-solver (preflop_grid_generator.cu:426-614) on the
    reference's 100-flop subset (its data/mtt_subset.bin, a committed
    fixture): the 3000-sample all-in matrix and the subset's class matrices
    once (timed, untimed warmup first), then `--warmup` untimed and `--steps`
    timed calls of the 300-iteration DCFR loop (the README's figure: 45.4 ms
    for 300 iterations on a T4)."""
    import ctypes
    import torch
    mtt = load("mtt", os.path.join(ROOT, "mtt-gpu-solver_amd", "mtt.py"))
    pf = load("mtt_preflop", os.path.join(ROOT, "mtt-gpu-solver_amd", "preflop.py"))
    meta = json.load(open(os.path.join(ROOT, "tests", "golden", "ev_dump_meta.json")))
    cards, w = pf.read_subset(bytes.fromhex(meta["subset_bin_hex"]))
    n = len(cards)
    iters = 300

    def timed(fn):
        torch.cuda.synchronize()
        t = time.time()
        r = fn()
        torch.cuda.synchronize()
        return r, time.time() - t

    pf.allin_matrix(3000)
    A, t_allin = timed(lambda: pf.allin_matrix(3000))
    pf.class_matrices(cards)
    mtt.call("mtt_preflop_timing", 1)
    M, t_class = timed(lambda: pf.class_matrices(cards))
    out = torch.empty((4, 3, 169), dtype=torch.float32, device="cuda:0")

    def solve():
        mtt.call("mtt_preflop_solve", M.data_ptr(), w.ctypes.data, n, A.data_ptr(), iters, 20.0, 1.35, 2.5,
                 out.data_ptr(), None)

    # the timed loop runs without per-kernel events; one more call measures the kernels
    ct = {}
    k = 0
    while True:
        name = ctypes.create_string_buffer(32)
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        if mtt.lib().mtt_preflop_kernel_times(k, name, 32, ctypes.byref(ms), ctypes.byref(cnt)) != 0:
            break
        ct[name.value.decode()] = ms.value
        k += 1
    mtt.call("mtt_preflop_timing", 0)
    for _ in range(args.warmup):
        solve()
    _, dt = timed(lambda: [solve() for _ in range(args.steps)])
    mtt.call("mtt_preflop_timing", 1)
    solve()
    ktimes = {}
    k = 0
    while True:
        name = ctypes.create_string_buffer(32)
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        if mtt.lib().mtt_preflop_kernel_times(k, name, 32, ctypes.byref(ms), ctypes.byref(cnt)) != 0:
            break
        ktimes[name.value.decode()] = (ms.value, cnt.value)
        k += 1
    mtt.call("mtt_preflop_timing", 0)
    ms_solve = 1e3 * dt / args.steps
    # class-equity kernel: villain-runout comparisons (hero combos off the flop x 1,326 villain lanes x 192
    # sampled runouts) at 3 packed 16-bit VALU operations per comparison
    heroes = sum(sum(1 for a in range(52) for b in range(a + 1, 52) if a not in fl and b not in fl)
                 for fl in (set(int(x) for x in c) for c in cards))
    compares = heroes * 1326 * 192
    ce_ms, ce_n = ct.get("pf_class_eq", 0.0), 2
    ce_peak = VALU_LANE_OPS / 3 / 1e12
    ce_ach = compares / (ce_ms * 1e-3) / 1e12 if ce_ms else None
    # the DCFR loop's dominant kernel: one workgroup per iteration
    tr_ms, tr_n = ktimes.get("pf_tree", (0.0, 1))
    tree_flops = 169 * (n * 2 + 5 * 169 * 2 + 60)           # flop sum, five 169-long FMA chains, tree + DCFR
    cpu = None
    if not args.no_cpu:
        try:
            O = load("oracle_ctypes", os.path.join(ROOT, "oracle", "oracle_ctypes.py"))
            Mh, Ah = M.cpu().numpy(), A.cpu().numpy().reshape(169, 169)
            O.preflop_solve(Mh, w, Ah, iters=5)
            t0 = time.time()
            O.preflop_solve(Mh, w, Ah, iters=iters)
            t_loop = time.time() - t0
            model, ncpu, avail = cpu_info()
            cpu = {"value": iters / t_loop, "unit": "DCFR iter/s", "cores": 1, "kind": "port",
                   "cpu_model": model, "nproc": ncpu, "ms_per_300_iters": 1e3 * t_loop,
                   "sample": f"the C oracle's {iters}-iteration loop on the same 100 class matrices, one thread"}
        except Exception as ex:
            cpu = {"error": str(ex)}
    value = iters / (ms_solve * 1e-3)
    baseline = 300 / 45.4e-3
    print(json.dumps({
        "metric": "preflop DCFR co-solver iterations/s (100 subset flops, 300-iteration loop)",
        "value": value, "unit": "DCFR iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_solve, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": value / baseline, "baseline": "README.md:8 (45.4 ms for 300 iterations, Tesla T4)",
        "dtype": "fp32", "data": "the reference's 100-flop subset (fixture); synthetic Monte-Carlo all-in matrix",
        "config": {"workload": "preflop_cosolve_100", "flops": n, "iterations": iters, "stack_bb": 20.0,
                   "bubble_factor": 1.35, "dead_pot": 2.5, "allin_samples": 3000},
        "ms_allin_matrix": 1e3 * t_allin, "ms_class_matrices": 1e3 * t_class,
        "kernel_ms_class_matrices": ct,
        "kernel_ms_per_step": {k: v[0] for k, v in ktimes.items()},
        "roofline": {"kernel": "k_pf_tree", "bound": "latency (one workgroup, dependent FP32 chains)",
                     "achieved": tree_flops / (tr_ms / tr_n * 1e-3) / 1e12 if tr_ms else None,
                     "peak": 157.3, "unit": "TFLOP/s",
                     "frac": tree_flops / (tr_ms / tr_n * 1e-3) / 157.3e12 if tr_ms else None, "traffic": None,
                     "avg_ms": tr_ms / tr_n if tr_ms else None},
        "roofline_class_eq": {"kernel": "k_pf_class_eq", "bound": "valu", "achieved": ce_ach, "peak": ce_peak,
                              "unit": "Tcompare/s", "frac": ce_ach / ce_peak if ce_ach else None,
                              "total_ms": ce_ms, "launches": ce_n, "compares": compares},
        "cpu_baseline": cpu,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="timed plain FP32 iterations (default: the same as --steps)")
    ap.add_argument("--solve-iters", type=int, default=1250,
                    help="untimed iterations after the timed window, up to this total, with the FP64 "
                         "exploitability every 50 iterations (0: stop after the timed window)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-scan", action="store_true",
                    help="skip the batched leg whose river showdown is computed by exact prefix scans")
    ap.add_argument("--workload", default="hu_turn_river", choices=["hu_turn_river", "flop_counts", "preflop"],
                    help="hu_turn_river: the metric's CFR+ game (default); flop_counts: exhaustive W/T counts of "
                         "all 1,755 canonical flops; preflop: the preflop DCFR co-solver on the 100-flop subset")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (3x): the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 30
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": n, "bytes_written": n, "launches": 3}}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; sharding across GPUs is designed but not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "flop_counts":
        run_flop_counts(args)
        return
    if args.workload == "preflop":
        run_preflop(args)
        return
    S = load("mtt_solver", os.path.join(ROOT, "mtt-gpu-solver_amd", "solver.py"))
    game, r0, r1 = S.config3(args.seed)
    total = args.warmup + args.steps

    solve_total = max(args.solve_iters, total) if args.solve_iters > 0 else total

    # ------------------------------------------------ batched BF16 split ---
    t = time.time()
    sb = S.Solver(game, r0, r1, S.MODE_BF16_SPLIT)
    setup_b = time.time() - t
    sb.run(args.warmup)
    sb.timing(True)
    t = time.time()
    sb.run(args.steps)
    dt_b = time.time() - t
    ktimes = sb.kernel_times()
    sb.timing(False)
    ips_b = args.steps / dt_b
    t = time.time()
    log(f"batched: setup {setup_b:.1f}s, {args.steps} timed iterations {ips_b:.2f} it/s")
    curve_b = solve_checkpoints(sb, solve_total, name="batched")
    solve_b_s = time.time() - t
    info = sb.info.tolist()
    st_b = sb.strategy() if not args.no_plain else None
    sb.close()
    del sb

    # ------------------------------- batched, river showdown by prefix scans ---
    scan = None
    if not args.no_scan and not river_sweep_mode():
        prev = os.environ.get("MTT_RIVER_SD")
        os.environ["MTT_RIVER_SD"] = "scan"
        try:
            t = time.time()
            ss = S.Solver(game, r0, r1, S.MODE_BF16_SPLIT)
            setup_s = time.time() - t
        finally:
            if prev is None:
                del os.environ["MTT_RIVER_SD"]
            else:
                os.environ["MTT_RIVER_SD"] = prev
        ss.run(args.warmup)
        ss.timing(True)
        t = time.time()
        ss.run(args.steps)
        dt_s = time.time() - t
        kts = ss.kernel_times()
        ss.timing(False)
        log(f"batched, river showdown by prefix scans: {args.steps} timed iterations {args.steps / dt_s:.2f} it/s")
        t = time.time()
        curve_s = solve_checkpoints(ss, solve_total, name="batched scan")
        scan = {"iter_per_s": args.steps / dt_s, "timed_steps": args.steps, "setup_s": setup_s,
                "kernel_ms_per_iter": {k: v[0] / max(v[1], 1) for k, v in kts.items()},
                "untimed_solve_s": time.time() - t, **expl_summary(curve_s)}
        ss.close()
        del ss

    # ----------------------------------------------------- plain FP32 ---
    plain = None
    linf = None
    if not args.no_plain:
        sp = S.Solver(game, r0, r1, S.MODE_PLAIN_FP32)
        pw = args.warmup
        psteps = args.plain_steps or args.steps
        sp.run(pw)
        sp.timing(True)
        t = time.time()
        sp.run(psteps)
        dt_p = time.time() - t
        ktp = sp.kernel_times()
        sp.timing(False)
        t = time.time()
        log(f"plain: {psteps} timed iterations {psteps / dt_p:.2f} it/s")
        curve_p = solve_checkpoints(sp, solve_total, name="plain")
        solve_p_s = time.time() - t
        plain = {"iter_per_s": psteps / dt_p, "timed_steps": psteps, "warmup": pw,
                 "kernel_ms_per_iter": {k: v[0] / max(psteps, 1) for k, v in ktp.items()},
                 "untimed_solve_s": solve_p_s, **expl_summary(curve_p)}
        if st_b is not None and sp.iterations == solve_total:
            st_p = sp.strategy()
            _, S_p = sp.state()
            linf = {"iterations": solve_total, **strategy_linf(st_b, st_p, S_p, int(info[1]), int(info[2])),
                    "note": "average strategies of the batched BF16-split and plain FP32 paths over (infoset, "
                            "combo) entries whose plain strategy sums are >= 1e-5 (>= 0.1) of the infoset's largest"}
            del st_p, S_p
        sp.close()
        del sp
    del st_b

    # ------------------------------------------------------- roofline ---
    work = algorithmic_work(S, game)
    per_iter = {k: v[0] / max(v[1], 1) for k, v in ktimes.items()}
    sd_ms = per_iter.get("river_showdown", float("nan"))
    sweep = river_sweep_mode()
    dominant = max(per_iter, key=per_iter.get) if per_iter else None

    def hbm_roof(kernel, byt, ms):
        ach = byt / (ms * 1e-3) / 1e9
        return {"kernel": kernel, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": hbm_traffic(kernel), "avg_ms": ms, "algorithmic_bytes": byt}

    def mfma_roof(kernel, ms):
        ach = work["gemm_flops"] / (ms * 1e-3) / 1e12
        return {"kernel": kernel, "bound": "mfma", "achieved": ach, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / BF16_PEAK_TFLOPS, "traffic": hbm_traffic(kernel), "avg_ms": ms}

    byts = {"river_backward": work["backward_bytes"], "river_forward": work["forward_bytes"],
            "river_showdown": work["showdown_bytes"]}
    if dominant == "river_showdown" and not sweep:
        roof = mfma_roof(dominant, sd_ms)
    else:
        roof = hbm_roof(dominant, byts.get(dominant, float("nan")), per_iter.get(dominant, float("nan")))
    sd_roof = hbm_roof("river_showdown", work["showdown_bytes"], sd_ms) if sweep else mfma_roof("river_showdown", sd_ms)
    sd_roof["form"] = "strength-order sweep" if sweep else "sign x blocker GEMM"
    if scan:
        # the scan reads each column and writes its values (4 + 4 B per live combo) and reads its board's
        # tables (3 x 1088 + 23 x 64 words) once
        scan_bytes = 2 * work["river_showdowns"] * 1081 * 8 + work["boards"] * (3 * 1088 + 23 * 64) * 4
        scan["roofline_showdown"] = hbm_roof("river_showdown", scan_bytes,
                                             scan["kernel_ms_per_iter"].get("river_showdown", float("nan")))
        scan["roofline_showdown"]["form"] = "prefix scans (exact integer sums)"
        scan["roofline_showdown"]["traffic"] = None

    cpu = None
    if not args.no_cpu:
        try:
            O = load("oracle_ctypes", os.path.join(ROOT, "oracle", "oracle_ctypes.py"))
            log("cpu baseline: FP32 oracle on the host")
            cpu = cpu_baseline(S, O, game, r0, r1)
        except Exception as ex:  # the baseline is reported, never required
            cpu = {"error": str(ex)}

    out = {
        "metric": METRIC,
        "value": ips_b,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_b / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if sweep else "bf16",
        "data": "synthetic (seeded board and Dirichlet(0.5) ranges)",
        "config": {"workload": "hu_turn_river_cfr_plus", "seed": args.seed, "board": [int(c) for c in game.board[:3]],
                   "pot_bb": 6, "stack_bb": 40, "bets": "33%/75% + 2.5x raise + all-in per street",
                   "timed_iterations": args.steps, "solve_iterations": solve_total,
                   "nodes": info[0], "infosets": info[1]},
        **expl_summary(curve_b),
        "untimed_solve_s": solve_b_s,
        "setup_s": setup_b,
        "kernel_ms_per_iter": per_iter,
        "plain_fp32": plain,
        "speedup_vs_plain_fp32": (ips_b / plain["iter_per_s"]) if plain else None,
        "batched_river_scan": scan,
        "scan_speedup_vs_plain_fp32": (scan["iter_per_s"] / plain["iter_per_s"]) if (scan and plain) else None,
        "strategy_linf_batched_vs_plain": linf,
        "roofline": roof,
        "roofline_showdown": sd_roof,
        "work_per_iter": work,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
