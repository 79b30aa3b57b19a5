#!/usr/bin/env python3
"""Throughput of the sparse level-set engine on one MI355X.

A step is one complete segmentation run of the workload: restore the initial
state from an HBM-resident checkpoint, then run the workload's iterations
(update, marking, duplicate removal and compaction) through the C ABI.  Inputs
are resident in HBM before the timed region starts.

The headline workload is the 512^3 uint16 vessel tree (2000 iterations, less
than 1% of the volume active): the configuration the HBM-roofline target is
quoted on.  The 181x217x181 float32 phantom (1000 iterations) is reported
beside it under "phantom", and the batched and differentiable forms under
"batched" and "adjoint".

Printed (one JSON line): active-voxel updates/s (sum of n_k over the run's
iterations divided by the wall time), the roofline of the dominant kernel (its
algorithmic bytes per launch, from the per-iteration counters, over its average
launch time measured with HIP events on the engine's stream, both over the
launches of iterations with a non-empty list), the whole-path fraction of HBM
peak, and the CPU baseline (the oracle: the reference's update equations
compiled as host C with -march=native, timed on this machine's cores).

Algorithmic bytes per iteration k, FP32 phi (4 bytes), source S bytes/voxel:
  update   (38 + S) n_k   coordinate, 7 phi taps, source, phi write, flag, mask byte
  mark     33 n_k         coordinate, flag, 7 phi taps
  keep     5 c_k          candidate read-back, tag probe
  compact  5 c_k + 5 m_k  candidate write, tag set; list write, clear
with n_k active voxels, c_k candidate emissions and m_k new active voxels (the
gather model's 10 c_k + 5 m_k for duplicate removal and compaction, split over
the two kernels that do that work).  Each iteration is three launches: the
update, which from the second iteration on runs in k_update_fused_defer
together with the previous iteration's compaction and rebuilds the traversal
the keep step left as kept directions, k_mark and k_keep_defer; the run ends
with the last iteration's compaction (k_compact) and a traversal rebuild
without updates (k_settle_traversal, not a profiled launch).  3 L + 1 profiled
launches for L launched iterations (launching stops once the list is empty).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "active-voxel updates/s + % HBM peak; batched/adjoint vs plain at Dice≥0.9999"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_sync():
    """hipDeviceSynchronize of the HIP runtime the engine library is linked
    against (resolved through the engine's own dependencies, so that with torch
    loaded it is torch's runtime, the one the engine uses)."""
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import lib
    rc = lib().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def host_cores() -> int:
    try:
        return max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "64"))))
    except Exception:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def algorithmic_bytes(n, c, m, S, P=4):
    """The gather model with P bytes per phi value: update n (4 + 7P + S + P + 2),
    mark n (5 + 7P), compaction 10 c + 5 m."""
    n = np.asarray(n, np.float64)
    c = np.asarray(c, np.float64)
    m = np.asarray(m, np.float64)
    return {"update": (6 + 8 * P + S) * n, "mark": (5 + 7 * P) * n, "keep": 5.0 * c, "compact": 5.0 * c + 5.0 * m}


# ---------------------------------------------------------------------------
# launches of a per-launch profiled run, by iteration
# ---------------------------------------------------------------------------
# The launch order of a run of L launched iterations: k_update (iteration 0),
# k_mark, k_keep, then (k_update_fused, k_mark, k_keep) for iterations 1..L-1,
# then the last compaction.  A launch belongs to a non-empty iteration when
# that iteration's list has n_k > 0; only those enter the averages, so that the
# iterations the early stop launches over an empty list do not dilute them.
# By default the keep step leaves the next traversal to the update launch
# (k_keep_defer, k_update_fused_defer); LSSEG_DEFER_KEEP=0 selects the placing
# keep step (k_keep, k_update_fused).  The launch order is the same.
DEFER_KEEP = os.environ.get("LSSEG_DEFER_KEEP", "1") != "0"
KERNEL_OF_PART = {"update_fused": "k_update_fused_defer" if DEFER_KEEP else "k_update_fused", "mark": "k_mark",
                  "keep": "k_keep_defer" if DEFER_KEEP else "k_keep"}
SUMMARY_KEY_OF_PART = {"update_fused": "k_update_fused", "mark": "k_mark", "keep": "k_keep"}


def launch_parts(per_launch, n, c, m, S, P):
    """Per-launch times (ms) and algorithmic bytes of the non-empty iterations'
    launches, by part."""
    per_launch = np.asarray(per_launch, np.float64)
    L = (per_launch.size - 1) // 3
    assert per_launch.size == 3 * L + 1 and L >= 2, f"unexpected launch count {per_launch.size}"
    n, c, m = (np.asarray(a)[:L] for a in (n, c, m))
    B = algorithmic_bytes(n, c, m, S, P)
    head = per_launch[0:3 * L:3]
    live = n > 0
    parts = {
        # launch k >= 1: the update of iteration k and the compaction of k - 1
        "update_fused": (head[1:], B["update"][1:] + B["compact"][:-1], n[1:], live[1:]),
        "mark": (per_launch[1:3 * L:3], B["mark"], n, live),
        "keep": (per_launch[2:3 * L:3], B["keep"], n, live),
    }
    out = {}
    for k, (ms, by, nk, sel) in parts.items():
        ms, by, nk = ms[sel], by[sel], nk[sel]
        launches = int(ms.size)
        tot_ms = float(ms.sum())
        avg_us = 1e3 * tot_ms / max(launches, 1)
        bpl = float(by.sum()) / max(launches, 1)
        big = nk >= np.percentile(nk, 90) if launches else nk > 0
        out[k] = {"kernel": KERNEL_OF_PART[k], "launches": launches, "avg_launch_us": avg_us,
                  "bytes_per_launch": bpl, "achieved_GBs": bpl / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0,
                  "total_ms": tot_ms,
                  "busiest_decile_GBs": (float(by[big].sum()) / (float(ms[big].sum()) * 1e-3) / 1e9
                                         if big.any() and ms[big].sum() > 0 else 0.0)}
    return out, L


def _kernel_key(name):
    name = name.split("(")[0]
    name = name[5:] if name.startswith("void ") else name
    if name.startswith("lsseg::k_update_fused<") or name.startswith("lsseg::k_update_fused_defer<"):
        return "k_update_fused"
    if name.startswith("lsseg::k_update<"):
        return "k_update"
    if name.startswith("lsseg::k_mark<"):
        return "k_mark"
    if name.startswith("lsseg::k_keep<") or name.startswith("lsseg::k_keep_defer<"):
        return "k_keep"
    if name == "lsseg::k_compact<2>":
        return "k_compact"
    return None


def iteration_dispatches(rows, n):
    """Rows (dicts with Kernel_Name, in dispatch order) of a process whose last
    iteration run is the profiled one -> {kernel: [(iteration, row), ...]} for
    that run's iteration launches (its k_update, the last in the process,
    starts them; earlier runs, such as the autotuner's, are skipped)."""
    out = {"k_update_fused": [], "k_mark": [], "k_keep": []}
    keys = [_kernel_key(r["Kernel_Name"]) for r in rows]
    first = max((i for i, k in enumerate(keys) if k == "k_update"), default=len(rows))
    it_mark = it_keep = 0
    it_fused = 1
    for r, k in zip(rows[first + 1:], keys[first + 1:]):
        if k is None or k == "k_compact":
            continue
        if k == "k_update_fused":
            out[k].append((it_fused, r))
            it_fused += 1
        elif k == "k_mark":
            out[k].append((it_mark, r))
            it_mark += 1
        elif k == "k_keep":
            out[k].append((it_keep, r))
            it_keep += 1
    for k in out:
        out[k] = [(i, r) for i, r in out[k] if i < len(n) and n[i] > 0]
    return out


def _csv_rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def _dispatch_order(rows):
    """Rows in dispatch order (by Dispatch_Id where the CSV has it, else as written)."""
    key = next((k for k in ("Dispatch_Id", "Correlation_Id") if rows and k in rows[0]), None)
    return sorted(rows, key=lambda r: int(r[key])) if key else rows


def _find_csv(root, d, suffix):
    hits = glob.glob(os.path.join(root, d, "**", f"*{suffix}"), recursive=True)
    if not hits:
        raise SystemExit(f"no *{suffix} under {os.path.join(root, d)}")
    return hits[0]


def _load_counts(root):
    with open(os.path.join(root, "counts.json")) as f:
        d = json.load(f)
    return d


def summarize_pmc(root, workload):
    """HBM traffic per launch from two rocprofv3 --pmc runs of `bench.py
    --profile-only` (FETCH_SIZE, then WRITE_SIZE) and the same two counters over
    a plain copy of known size (--calibrate-copy), which give each counter's
    scale.  Averaged over the launches of non-empty iterations, the launches the
    roofline's algorithmic bytes cover."""
    copy_bytes = None
    with open(os.path.join(root, "copy_fetch.json")) as f:
        for line in f:
            if line.startswith("{"):
                copy_bytes = json.loads(line)["calibrate_copy_bytes"]

    def means(d):
        acc = {}
        for r in _csv_rows(_find_csv(root, d, "counter_collection.csv")):
            acc.setdefault(r["Kernel_Name"].split("(")[0], []).append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}

    cf, cw = means("copy_fetch"), means("copy_write")
    copy_name = max(cf, key=lambda k: cf[k][0])
    f_scale = copy_bytes / (cf[copy_name][0] * 1024.0)
    w_scale = copy_bytes / (cw[copy_name][0] * 1024.0)
    counts = _load_counts(root)
    n = np.asarray(counts["n"])
    S, P = counts["S"], counts["P"]
    B = algorithmic_bytes(n, counts["c"], counts["m"], S, P)
    alg = {"k_update_fused": lambda i: B["update"][i] + B["compact"][i - 1], "k_mark": lambda i: B["mark"][i],
           "k_keep": lambda i: B["keep"][i]}
    per = {}
    for d, key in (("fetch", "fetch"), ("write", "write")):
        rows = _dispatch_order(_csv_rows(_find_csv(root, d, "counter_collection.csv")))
        for k, lst in iteration_dispatches(rows, n).items():
            per.setdefault(k, {})[key] = [float(r["Counter_Value"]) for _, r in lst]
            per[k]["iters"] = [i for i, _ in lst]
    out = {"workload": workload, "copy_kernel": copy_name, "copy_bytes": copy_bytes, "fetch_scale": f_scale,
           "write_scale": w_scale, "launches": "non-empty iterations of one per-launch profiled run",
           "kernels": {}}
    for k, v in per.items():
        fe, wr = np.asarray(v["fetch"]), np.asarray(v["write"])
        assert fe.size == wr.size == len(v["iters"]), k
        rd, wt = fe.mean() * 1024.0 * f_scale, wr.mean() * 1024.0 * w_scale
        ab = float(np.mean([alg[k](i) for i in v["iters"]]))
        out["kernels"][k] = {"dispatches": int(fe.size), "fetch_KiB_raw": float(fe.mean()),
                             "write_KiB_raw": float(wr.mean()), "read_bytes": rd, "write_bytes": wt,
                             "bytes_per_launch": rd + wt, "algorithmic_bytes_per_launch": ab,
                             "physical_over_algorithmic": (rd + wt) / ab if ab else None}
    return out


def summarize_trace(root):
    """Kernel durations from a rocprofv3 --kernel-trace run of `bench.py
    --profile-only`, over the launches of non-empty iterations (the same set as
    the bench line's per-launch figures), and over every dispatch."""
    counts = _load_counts(root)
    n = np.asarray(counts["n"])
    rows = _dispatch_order(_csv_rows(_find_csv(root, "kt", "kernel_trace.csv")))
    out = {"workload": counts["workload"], "kernels": {}}
    for k, lst in iteration_dispatches(rows, n).items():
        us = np.array([(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for _, r in lst])
        out["kernels"][k] = {"launches": int(us.size), "avg_us": float(us.mean()), "min_us": float(us.min()),
                             "max_us": float(us.max()), "total_ms": float(us.sum() / 1e3)}
    allk = {}
    for r in rows:
        allk.setdefault(r["Kernel_Name"].split("(")[0], []).append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out["all_dispatches"] = {k: {"calls": len(v), "avg_us": float(np.mean(v)), "total_ms": float(np.sum(v) / 1e3)}
                             for k, v in sorted(allk.items(), key=lambda kv: -sum(kv[1]))}
    return out


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        return json.load(f), os.path.relpath(path, ROOT)


def make_workload(name, iterations):
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
    if name == "ellipsoids":
        w = synth.ellipsoids()
    elif name == "phantom":
        w = synth.phantom()
    elif name == "vessels":
        w = synth.vessels()
    else:
        raise SystemExit(f"unknown workload {name}")
    if iterations:
        w.iterations = iterations
    return w


def describe(w, X, Y, Z):
    z, y, x = w.volume.shape
    return (f"{w.name} {x}x{y}x{z} {w.volume.dtype.name} (padded {X}x{Y}x{Z}), {w.iterations} iterations, "
            f"T={w.T:g} eps={w.eps:g} alpha={w.alpha:g} dt={w.dt:g}, phi0 = sphere SDF clamped +-4, "
            f"active band |phi0|<{w.band:g}")


# ---------------------------------------------------------------------------
# CPU baseline: the oracle compiled for this host (-march=native)
# ---------------------------------------------------------------------------
def _oracle_native():
    import oracle as O
    try:
        O.build_native()
        O.lib(native=True)
        return True, "-O3 -march=native"
    except Exception as e:  # noqa: BLE001 -- fall back to the portable build, and say so
        log(f"native oracle build failed ({e}); timing the portable x86-64-v3 build")
        return False, "-O3 -march=x86-64-v3"


def cpu_baseline(w, gpu_mask, seconds_1core, phi_storage=0, window_at=None, window_len=20):
    """The oracle on this host.  All cores: the full workload (which also gives
    the mask for the Dice check).  With window_at (the 512^3 workload): the
    first 200 iterations and a steady-state window of window_len iterations
    from iteration window_at, on all cores and on one core; otherwise one core
    over the workload's first iterations.  Single-core legs are bounded by
    seconds_1core each."""
    import oracle as O
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
    native, flags = _oracle_native()
    cores = host_cores()

    def new_state():
        s = O.State(w.volume, phi_storage=phi_storage, native=native)
        s.set_phi(w.phi0, w.band)
        return s

    ref = new_state()
    legs = {}
    t0 = time.perf_counter()
    n_tot = 0
    done = 0
    saved = None
    while done < w.iterations:
        k = min(50, w.iterations - done)
        if window_at is not None and done < window_at < done + k:
            k = window_at - done
        if window_at is not None and done == 200 and "all_first_200" not in legs:
            legs["all_first_200"] = {"iterations": 200, "updates": n_tot, "s": time.perf_counter() - t0}
        if window_at is not None and done == window_at and saved is None:
            t_pause = time.perf_counter()
            saved = ref.state(), ref.active()
            t0 += time.perf_counter() - t_pause  # the copy is not oracle work
            tw = time.perf_counter()
            n, _, _ = ref.iterate(w.T, w.eps, w.alpha, w.dt, window_len, threads=cores)
            legs["all_window"] = {"iterations": window_len, "from": window_at, "updates": int(n.sum()),
                                  "s": time.perf_counter() - tw}
            n_tot += int(n.sum())
            done += window_len
            continue
        n, _, _ = ref.iterate(w.T, w.eps, w.alpha, w.dt, k, threads=cores)
        n_tot += int(n.sum())
        done += k
    t_all = time.perf_counter() - t0
    mask = ref.mask()
    ref.close()
    dice = synth.dice(gpu_mask, mask)
    identical = bool(np.array_equal(gpu_mask, mask))

    def one_core(state, iters_max, label, start):
        t = time.perf_counter()
        upd = it = 0
        while it < iters_max and time.perf_counter() - t < seconds_1core:
            k = min(5, iters_max - it)
            n, _, _ = state.iterate(w.T, w.eps, w.alpha, w.dt, k, threads=1)
            upd += int(n.sum())
            it += k
        dt = time.perf_counter() - t
        return {"iterations": it, "from": start, "updates": upd, "s": dt, "value": upd / dt if dt > 0 else 0.0,
                "what": label}

    one = new_state()
    if window_at is None:
        single = one_core(one, w.iterations, "the workload's first iterations", 0)
        one.close()
        single_all = [single]
    else:
        first = one_core(one, 200, "first 200 iterations", 0)
        one.close()
        (r, wr, e), lst = saved
        win = O.State(w.volume, phi_storage=phi_storage, native=native)
        win.set_state(r, wr, e, lst)
        window = one_core(win, window_len, f"steady state: iterations {window_at}..{window_at + window_len - 1} "
                                           f"from the all-core run's state", window_at)
        win.close()
        single_all = [first, window]
        for k in ("all_first_200", "all_window"):
            if k in legs:
                legs[k]["value"] = legs[k]["updates"] / legs[k]["s"] if legs[k]["s"] > 0 else 0.0
    out = {
        "value": n_tot / t_all, "unit": "active-voxel updates/s", "cores": cores, "kind": "port",
        "sample": f"full workload ({w.iterations} iterations, {n_tot} updates) in {t_all:.2f} s on {cores} "
                  f"threads; oracle/lsoracle.c {flags} -ffp-contract=off, OpenMP over the active list and "
                  f"prefix-sum compactions in the reference's order; {cpu_model()}",
        "single_core": [dict(s, cores=1) for s in single_all],
    }
    if legs:
        out["all_cores_legs"] = legs
    return out, dice, identical


# ---------------------------------------------------------------------------
# the plain form on one workload
# ---------------------------------------------------------------------------
def plain_leg(name, args, with_cpu, window_at=None):
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import PHI_F32, PHI_I8, Context
    t = time.perf_counter()
    w = make_workload(name, args.iterations if name == args.workload else 0)
    log(f"{name}: workload generated in {time.perf_counter() - t:.1f} s")
    phi_storage = PHI_I8 if args.phi == "i8" else PHI_F32
    ctx = Context(w.volume, phi_storage=phi_storage)
    ctx.set_phi(w.phi0, w.band)
    ctx.checkpoint_save(0)
    S = w.volume.dtype.itemsize
    P = 1 if args.phi == "i8" else 4
    log(f"{name}: context ready: {describe(w, ctx.X, ctx.Y, ctx.Z)}; n0={ctx.active_count}")
    # launch shapes and widths timed on this workload (outside the timed region)
    t = ctx.autotune(w.T, w.eps, w.alpha, w.dt, w.iterations)
    tuning = {"update_blocks": t.update_blocks, "compact_blocks": t.compact_blocks, "keep_items": t.keep_items,
              "compact_tile": t.compact_tile, "tunings_timed": t.candidates, "cached": bool(t.cached)}
    log(f"{name}: autotuned: {t.update_blocks} update/mark workgroups, {t.compact_blocks} keep/compaction "
        f"workgroups, keep step {t.keep_items} positions per thread, compaction tiles of {t.compact_tile} "
        f"positions below the small-list bound ({t.candidates} tunings timed)")

    for _ in range(args.warmup):
        ctx.checkpoint_restore(0)
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    device_sync()
    t0 = time.perf_counter()
    last = None
    for _ in range(args.steps):
        ctx.checkpoint_restore(0)
        last = ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    device_sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    updates_per_step = int(last.sum_active) if last else 0
    value = updates_per_step * args.steps / elapsed if elapsed > 0 else 0.0
    gpu_mask = ctx.mask()
    counts = ctx.iteration_counts()
    log(f"{name}: timed {args.steps} steps: {ms_per_step:.2f} ms/step, {updates_per_step} updates/step, "
        f"few/many iterations {last.few_path_iters}/{last.many_path_iters}, final n={last.active_final}, "
        f"{last.empty_skipped} trailing iterations over an empty list not launched")

    # the same run from and to host arrays: context creation (upload, pad and
    # tile the source), phi0 upload, the iterations and the mask download
    device_sync()
    t0 = time.perf_counter()
    hctx = Context(w.volume, phi_storage=phi_storage)
    hctx.set_phi(w.phi0, w.band)
    hst = hctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    hmask = hctx.mask()
    host_ms = 1e3 * (time.perf_counter() - t0)
    assert np.array_equal(hmask, gpu_mask), "host-inclusive run diverged from the timed run"
    hctx.close()
    host_inclusive = {"ms": host_ms, "value": int(hst.sum_active) / (host_ms * 1e-3), "unit": "active-voxel updates/s",
                      "what": "lsseg_create (source upload, pad, tile) + lsseg_set_phi + lsseg_iterate + "
                              "lsseg_get_mask, host arrays in and out, one run"}
    log(f"{name}: host-inclusive run: {host_ms:.2f} ms ({host_inclusive['value']:.3e} updates/s)")

    # per-launch device times (HIP events on the engine's stream), in issue order
    ctx.checkpoint_restore(0)
    pst, kt, per_launch = ctx.iterate_profiled(w.T, w.eps, w.alpha, w.dt, w.iterations)
    pc = ctx.iteration_counts()
    assert np.array_equal(pc.n, counts.n), "profiled run diverged from the timed run"
    kernels, L = launch_parts(per_launch, pc.n, pc.c, pc.m, S, P)
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    # the committed counter summaries are of FP32 phi runs
    traffic, traffic_src = load_traffic(name) if P == 4 else (None, None)
    traffic_bytes = None
    tkey = SUMMARY_KEY_OF_PART[dom]  # the counter summaries key launches by part, whatever kernel ran it
    if traffic and traffic.get("kernels", {}).get(tkey):
        traffic_bytes = traffic["kernels"][tkey]["bytes_per_launch"]
    B = algorithmic_bytes(pc.n, pc.c, pc.m, S, P)
    path_bytes = float(sum(v.sum() for v in B.values()))
    path_gbs = path_bytes / (ms_per_step * 1e-3) / 1e9
    live = int((np.asarray(pc.n) > 0).sum())
    log(f"{name}: per launch over the {live} non-empty of {L} launched iterations: "
        + "; ".join(f"{k} {v['avg_launch_us']:.2f} us, {v['bytes_per_launch'] / 1e6:.2f} MB, "
                    f"{v['achieved_GBs']:.0f} GB/s (busiest decile {v['busiest_decile_GBs']:.0f})"
                    for k, v in kernels.items())
        + f"; kernel time sum {kt.total_ms:.2f} ms vs {ms_per_step:.2f} ms/step")

    floor = random_access_floor(ctx, int(round(float(np.mean(pc.n[pc.n > 0])))), P * ctx.X * ctx.Y * ctx.Z)
    log(f"{name}: random-access floor at {floor['threads']} threads over {floor['buffer_bytes'] / 2**20:.0f} MiB: "
        f"one load {floor['load_us']:.2f} us, seven loads {floor['load7_us']:.2f} us, one store {floor['store_us']:.2f} us")

    cpu = dice = identical = None
    if with_cpu:
        log(f"{name}: running the CPU oracle")
        cpu, dice, identical = cpu_baseline(w, gpu_mask, args.cpu_seconds, 1 if args.phi == "i8" else 0,
                                            window_at=window_at)
        log(f"{name}: cpu baseline {cpu['value']:.3e} updates/s on {cpu['cores']} cores; Dice vs oracle {dice:.6f}")
    out = {
        "value": value, "unit": "active-voxel updates/s", "ms_per_step": ms_per_step,
        "config": {
            "workload": describe(w, ctx.X, ctx.Y, ctx.Z),
            "volume_xyz": list(w.shape_xyz), "iterations": w.iterations,
            "active_voxel_updates_per_step": updates_per_step,
            "non_empty_iterations": live,
            "parallelism": "single GPU",
            "phi_storage": "fixed-point int8 (LEVEL_SET_CHAR)" if args.phi == "i8" else "FP32",
            "tuning": tuning,
        },
        "roofline": {
            "bound": "hbm", "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d["achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic_bytes, "traffic_source": traffic_src,
            "kernel": d["kernel"], "bytes_per_launch": d["bytes_per_launch"], "avg_launch_us": d["avg_launch_us"],
            "launches": f"the {d['launches']} launches of non-empty iterations in one per-launch profiled run",
        },
        "path_roofline": {"achieved": path_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": path_gbs / HBM_PEAK_GBS, "bytes_per_step": path_bytes},
        "kernels": kernels,
        "random_access_floor": floor,
        "cpu_baseline": cpu,
        "host_inclusive": host_inclusive,
        "dice_vs_oracle": dice,
        "mask_identical_to_oracle": identical,
    }
    ctx.close()
    return out


def random_access_floor(ctx, threads, nbytes):
    """Launch times of the random-access floor (lsseg_probe_random_access) at
    the workload's mean list length, over the context's phi write buffer
    (nbytes; its contents are spent: the store mode zeroes words of it): what
    one launch of `threads` threads costs when each makes one (or seven) loads,
    or one store, in 32-byte sectors spread over a buffer of that size."""
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import engine
    _, buf, stream = ctx.device_buffers()
    out = {"threads": threads, "buffer_bytes": nbytes}
    for key, mode in (("load_us", engine.PROBE_LOAD), ("load7_us", engine.PROBE_LOAD7),
                      ("store_us", engine.PROBE_STORE)):
        out[key] = engine.probe_random_access(buf, nbytes, threads, mode, stream=stream)
    out["what"] = ("median of 21 launches; each of `threads` threads (the mean list length of the non-empty "
                   "iterations) makes one pseudo-random 4-byte load, seven such loads, or one store, each in its own "
                   "32-byte sector of a buffer the size of one phi buffer")
    return out


def profile_only(args):
    """One per-launch profiled run of the workload and nothing else (for
    rocprofv3 kernel-trace and counter passes); writes the run's per-iteration
    counts to --counts-out so that the summaries can select the launches of
    non-empty iterations."""
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Context
    w = make_workload(args.workload, args.iterations)
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    # the bench line's tuning; under a profiler, run this once beforehand with
    # LSSEG_TUNING_CACHE set so that the profiled runs take it from the cache
    t = ctx.autotune(w.T, w.eps, w.alpha, w.dt, w.iterations)
    st, kt, per = ctx.iterate_profiled(w.T, w.eps, w.alpha, w.dt, w.iterations)
    pc = ctx.iteration_counts()
    if os.environ.get("LSSEG_FUSE", "1") == "0":  # four launches per iteration: no per-part summary
        kernels, L = None, int(np.count_nonzero(pc.n))
    else:
        kernels, L = launch_parts(per, pc.n, pc.c, pc.m, w.volume.dtype.itemsize, 4)
    out = {"workload": describe(w, ctx.X, ctx.Y, ctx.Z), "S": w.volume.dtype.itemsize, "P": 4,
           "n": [int(x) for x in pc.n], "c": [int(x) for x in pc.c], "m": [int(x) for x in pc.m],
           "launched_iterations": L, "kernels": kernels,
           "tuning": {"update_blocks": t.update_blocks, "compact_blocks": t.compact_blocks,
                      "keep_items": t.keep_items, "compact_tile": t.compact_tile, "cached": bool(t.cached)}}
    if args.counts_out:
        os.makedirs(os.path.dirname(os.path.abspath(args.counts_out)), exist_ok=True)
        with open(args.counts_out, "w") as f:
            json.dump(out, f)
    print(json.dumps({k: v for k, v in out.items() if k not in ("n", "c", "m")}), flush=True)
    ctx.close()


def batch_vs_sequential(steps, iterations):
    """The batched form on the 64-volume workload (uneven uint8 volumes, per-volume
    parameters) against 64 sequential plain runs of the same volumes on the same
    GPU; masks must be bit-identical."""
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Batch, Context
    t = time.perf_counter()
    ws = synth.batch(iterations=iterations) if iterations else synth.batch()
    iters = ws[0].iterations
    log(f"batch workload: {len(ws)} volumes, {sum(w.volume.size for w in ws) / 1e6:.1f} M voxels, "
        f"{iters} iterations, generated in {time.perf_counter() - t:.1f} s")
    params = [(w.T, w.eps, w.alpha, w.dt, 0) for w in ws]
    b = Batch([w.volume for w in ws])
    for i, w in enumerate(ws):
        b.volume(i).set_phi(w.phi0, w.band)
    b.checkpoint_save(0)
    tune = b.autotune(params, min(iters, 100))
    log(f"autotuned: {tune.update_blocks} update/mark workgroups, {tune.compact_blocks} compaction workgroups, "
        f"keep step {tune.keep_items} positions per thread, compaction tiles of {tune.compact_tile} positions "
        f"({tune.candidates} tunings timed{', cached' if tune.cached else ''})")
    b.checkpoint_restore(0)
    b.iterate(params, iters)  # warm-up (graph capture)
    device_sync()
    t0 = time.perf_counter()
    upd = 0
    for _ in range(steps):
        b.checkpoint_restore(0)
        st = b.iterate(params, iters)
        upd = sum(int(x.sum_active) for x in st)
    device_sync()
    t_batch = (time.perf_counter() - t0) / steps
    batch_masks = [b.volume(i).mask() for i in range(len(ws))]
    b.close()
    # sequential plain runs: each volume warmed up once, then timed
    t_seq = 0.0
    upd_seq = 0
    same = True
    for i, w in enumerate(ws):
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        ctx.checkpoint_save(0)
        ctx.iterate(*params[i][:4], iters)
        ctx.checkpoint_restore(0)
        device_sync()
        t0 = time.perf_counter()
        st = ctx.iterate(*params[i][:4], iters)
        device_sync()
        t_seq += time.perf_counter() - t0
        upd_seq += int(st.sum_active)
        same &= bool(np.array_equal(ctx.mask(), batch_masks[i]))
        ctx.close()
    assert upd_seq == upd, "batched and sequential runs did different work"
    out = {
        "workload": f"{len(ws)} uint8 volumes, edges log-uniform in [64,256], per-volume T/eps/alpha, "
                    f"dt=0.5, {iters} iterations each",
        "volumes": len(ws), "iterations": iters, "updates_per_step": upd,
        "ms_per_step": 1e3 * t_batch, "value": upd / t_batch, "unit": "active-voxel updates/s",
        "sequential_ms": 1e3 * t_seq, "sequential_value": upd / t_seq,
        "speedup_vs_sequential": t_seq / t_batch, "masks_identical_to_sequential": same,
        "tuning": {"update_blocks": tune.update_blocks, "compact_blocks": tune.compact_blocks,
                   "keep_items": tune.keep_items, "compact_tile": tune.compact_tile,
                   "tunings_timed": tune.candidates},
    }
    log(f"batched: {1e3 * t_batch:.2f} ms/step ({upd / t_batch:.3e} updates/s); sequential plain runs: "
        f"{1e3 * t_seq:.2f} ms ({upd / t_seq:.3e} updates/s); speedup {t_seq / t_batch:.2f}x; "
        f"masks identical: {same}")
    return out


def _soft_dice_grad(phi, mask, temp=0.5):
    """1 - softDice(sigmoid(-phi / temp), mask) and its gradient in phi (float64, host)."""
    phi = phi.astype(np.float64)
    m = mask.astype(np.float64)
    s = 1.0 / (1.0 + np.exp(phi / temp))
    a = (s * m).sum()
    b = s.sum() + m.sum()
    return 1.0 - 2.0 * a / b, (-2.0 * (m * b - a) / (b * b)) * (-s * (1.0 - s) / temp)


REPS = 3  # timed calls per volume and side in the differentiable-form comparison


def adjoint_vs_plain(iterations, fd_h=1e-3, fd64_h=1e-6):
    """The differentiable form on the 16-volume 96^3 float workload.  Timed:
    the plain forward (lsseg_iterate) against forward + loss + backward through
    the torch.autograd.Function (adjoint.segment, soft-Dice loss on GPU
    tensors, loss.backward()), volume after volume.  Checked, on every volume,
    the GPU adjoint against central finite differences of the recorded map:
      - of the FP64 replay (the oracle) at h = 1e-6, three-point: the
        derivative of the map itself;
      - of the GPU's FP32 replay at h = 1e-3, five-point, and its Richardson
        extrapolation from h and h/2.  Over 200 iterations the map's higher
        derivatives are large, and the h = 1e-3 differences carry ~1e-2
        truncation error even in FP64 (the FP64 differences at h = 1e-3 are
        reported beside them); FP32 rounding rules out smaller steps there;
    and against the FP64 host adjoint on one 64^3 instance."""
    import torch
    import oracle as O
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import adjoint, synth
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Context
    ws = synth.adjoint(iterations=iterations) if iterations else synth.adjoint()
    K = ws[0].iterations
    t_plain = t_adj = 0.0
    upd = 0
    worst = worst_dir = worst_r = worst_dir_r = worst_torch = 0.0
    w64 = w64_dir = w64_h3 = 0.0
    for i, w in enumerate(ws):
        T, eps, alpha = w.T, w.eps, w.alpha
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        ctx.checkpoint_save(0)
        ctx.iterate(T, eps, alpha, w.dt, K)  # warm-up
        # each side is timed REPS times per volume and its median counted, so
        # that one slow call (host jitter, a first touch of fresh device
        # memory) does not decide the ratio
        reps = []
        for _ in range(REPS):
            ctx.checkpoint_restore(0)
            device_sync()
            t0 = time.perf_counter()
            st = ctx.iterate(T, eps, alpha, w.dt, K)
            device_sync()
            reps.append(time.perf_counter() - t0)
        t_plain += float(np.median(reps))
        upd += int(st.sum_active)
        # the torch path, timed: phi0 and the mask already on the GPU
        phi0_t = torch.from_numpy(np.ascontiguousarray(w.phi0)).cuda().requires_grad_()
        mask_t = torch.from_numpy(np.ascontiguousarray(w.truth.astype(np.float32))).cuda()
        theta_t = torch.tensor([T, eps, alpha], dtype=torch.float64, requires_grad=True)
        if i == 0:  # torch's own first-call set-up stays out of the timing, as the plain warm-up does
            adjoint.soft_dice_loss(adjoint.segment(ctx, phi0_t, theta_t, dt=w.dt, iters=K, band=w.band),
                                   mask_t).backward()
            phi0_t.grad = None
            theta_t.grad = None
        reps = []
        for _ in range(REPS):
            phi0_t.grad = None
            theta_t.grad = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            phi_t = adjoint.segment(ctx, phi0_t, theta_t, dt=w.dt, iters=K, band=w.band)
            adjoint.soft_dice_loss(phi_t, mask_t).backward()
            torch.cuda.synchronize()
            reps.append(time.perf_counter() - t0)
            del phi_t
        t_adj += float(np.median(reps))
        # the same gradients through the host path, for the checks
        ctx.checkpoint_restore(0)
        rec = ctx.fwd_record(T, eps, alpha, w.dt, K)
        loss, dphi = _soft_dice_grad(ctx.phi(0), w.truth)
        g0, th = rec.bwd(dphi.astype(np.float32))
        torch_vs_host = max(abs(float(theta_t.grad[k]) - th[k]) / max(abs(th[k]), 1e-12) for k in range(3))
        worst_torch = max(worst_torch, torch_vs_host)

        def L(p, t, e, a):
            return _soft_dice_grad(rec.replay(p, t, e, a, w.dt), w.truth)[0]

        def central(f, h):
            return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)

        def rel(a, b):
            return abs(a - b) / max(abs(b), 1e-12)

        # the FP64 replay of the same recorded map (oracle)
        n, lists, dec = rec.lists()
        nz, ny, nx = w.volume.shape
        X, Y, Z = O.padded_dims(nx, ny, nz)
        src = O.padded_tiled_source(w.volume)

        def L64(p, t, e, a):
            out, _ = O.replay_f64(X, Y, Z, src, t, e, a, w.dt, n, lists, dec, O.pad_tile_phi(p, 1.0))
            return _soft_dice_grad(O.unpad_untile(out, nx, ny, nz), w.truth)[0]

        base = [T, eps, alpha]
        p64 = w.phi0.astype(np.float64)
        for k in range(3):
            def f(step, k=k):
                v = list(base)
                v[k] += step
                return L(w.phi0, *v)

            def f64(step, k=k):
                v = list(base)
                v[k] += step
                return L64(p64, *v)
            d1, d2 = central(f, fd_h), central(f, fd_h / 2)
            worst = max(worst, rel(th[k], d1))
            worst_r = max(worst_r, rel(th[k], (16 * d2 - d1) / 15))
            w64 = max(w64, rel(th[k], (f64(fd64_h) - f64(-fd64_h)) / (2 * fd64_h)))
            w64_h3 = max(w64_h3, rel(th[k], central(f64, fd_h)))
        d = (g0 / max(np.linalg.norm(g0), 1e-30)).astype(np.float32)
        adj_d = float((g0.astype(np.float64) * d).sum())
        fdir = (lambda step: L(w.phi0 + step * d, *base))
        d1, d2 = central(fdir, fd_h), central(fdir, fd_h / 2)
        worst_dir = max(worst_dir, rel(adj_d, d1))
        worst_dir_r = max(worst_dir_r, rel(adj_d, (16 * d2 - d1) / 15))
        d64 = d.astype(np.float64)
        w64_dir = max(w64_dir, rel(adj_d, (L64(p64 + fd64_h * d64, *base) - L64(p64 - fd64_h * d64, *base))
                                   / (2 * fd64_h)))
        rec.close()
        ctx.close()
    # FP64 host adjoint of the same recorded map on one 64^3 instance
    w = synth.adjoint(count=1, edge=64, iterations=K)[0]
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    rec = ctx.fwd_record(w.T, w.eps, w.alpha, w.dt, K)
    _, dphi = _soft_dice_grad(ctx.phi(0), w.truth)
    g0, th = rec.bwd(dphi.astype(np.float32))
    n, lists, dec = rec.lists()
    nz, ny, nx = w.volume.shape
    X, Y, Z = O.padded_dims(nx, ny, nz)
    src = O.padded_tiled_source(w.volume)
    out64, hist = O.replay_f64(X, Y, Z, src, w.T, w.eps, w.alpha, w.dt, n, lists, dec, O.pad_tile_phi(w.phi0, 1.0),
                               keep_history=True)
    _, dphi64 = _soft_dice_grad(O.unpad_untile(out64, nx, ny, nz), w.truth)
    g64, th64 = O.adjoint_f64(X, Y, Z, src, w.T, w.eps, w.alpha, w.dt, n, lists, dec, hist,
                              O.pad_tile_phi(dphi64, 0.0))
    g64 = O.unpad_untile(g64, nx, ny, nz)
    fp64_theta = float(max(abs(th[k] - th64[k]) / max(abs(th64[k]), 1e-12) for k in range(3)))
    fp64_phi0 = float(np.linalg.norm(g0 - g64) / max(np.linalg.norm(g64), 1e-30))
    rec.close()
    ctx.close()
    del hist
    out = {
        "workload": f"{len(ws)} float32 volumes 96^3 (ellipsoids / 255), K={K} iterations, dt=0.5, T=0.6 eps=0.15 "
                    f"alpha=0.1; loss 1 - softDice(sigmoid(-phi/0.5), mask)",
        "plain_forward_ms": 1e3 * t_plain, "forward_record_plus_adjoint_ms": 1e3 * t_adj,
        "timing": f"per volume the median of {REPS} timed calls of each side, summed over the volumes",
        "timed_path": "adjoint.segment (torch.autograd.Function) + soft-Dice loss + backward, GPU tensors",
        "torch_vs_host_max_rel_err_theta": worst_torch,
        "adjoint_over_plain": t_adj / t_plain if t_plain > 0 else None,
        "updates": upd, "value": 2 * upd / t_adj if t_adj > 0 else 0.0,
        "unit": "active-voxel updates/s (forward + adjoint, each counted once)",
        "fd_fp64_max_rel_err_theta": w64, "fd_fp64_max_rel_err_phi0_direction": w64_dir,
        "fd_fp64": f"central three-point differences at h = {fd64_h:g} of the FP64 replay (oracle) of each "
                   f"volume's recorded map, all volumes, against the GPU adjoint; phi0 along the normalised gradient",
        "fd_fp64_h1e-3_five_point_max_rel_err_theta": w64_h3,
        "fd_max_rel_err_theta": worst, "fd_max_rel_err_phi0_direction": worst_dir,
        "fd_richardson_max_rel_err_theta": worst_r, "fd_richardson_max_rel_err_phi0_direction": worst_dir_r,
        "fd": "central five-point stencil at h = 1e-3 of the GPU's FP32 replay, and (16 D(h/2) - D(h)) / 15, "
              "all volumes; the same stencil on the FP64 replay gives the truncation error alone "
              "(fd_fp64_h1e-3_five_point_max_rel_err_theta)",
        "fp64_64cubed_max_rel_err_theta": fp64_theta, "fp64_64cubed_rel_l2_err_phi0": fp64_phi0,
    }
    log(f"adjoint: plain forward {1e3 * t_plain:.1f} ms, forward record + loss + adjoint (torch) {1e3 * t_adj:.1f} ms "
        f"({out['adjoint_over_plain']:.2f}x; torch vs host theta {worst_torch:.1e}); FD of the FP64 replay "
        f"(h={fd64_h:g}): theta {w64:.2e}, phi0 dir {w64_dir:.2e}; FD (h=1e-3) of the FP32 replay: theta "
        f"{worst:.2e}, phi0 dir {worst_dir:.2e} (FP64 replay at h=1e-3: theta {w64_h3:.2e}); "
        f"FP64 adjoint 64^3: theta {fp64_theta:.2e}, phi0 {fp64_phi0:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="vessels", choices=["ellipsoids", "phantom", "vessels"])
    ap.add_argument("--iterations", type=int, default=0, help="override the workload's iteration count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--phi", default="f32", choices=["f32", "i8"],
                    help="phi storage: FP32 (default) or the reference's fixed-point signed char")
    ap.add_argument("--no-phantom", action="store_true", help="skip the phantom workload beside the headline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-sequential comparison")
    ap.add_argument("--batch-iterations", type=int, default=0, help="override the batch workload's iterations")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-form comparison")
    ap.add_argument("--adjoint-iterations", type=int, default=0, help="override the adjoint workload's iterations")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="bound of each single-core baseline sample")
    ap.add_argument("--window-at", type=int, default=450,
                    help="first iteration of the 512^3 workload's steady-state CPU window")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (HBM counter calibration)")
    ap.add_argument("--profile-only", action="store_true",
                    help="only one per-launch profiled run of --workload (for rocprofv3 passes)")
    ap.add_argument("--counts-out", default="", help="with --profile-only: where to write the run's counts (JSON)")
    ap.add_argument("--summarize-pmc", default="", metavar="DIR",
                    help="only summarize rocprofv3 --pmc runs under DIR (fetch/, write/, copy_fetch/, copy_write/, "
                         "copy_fetch.json, counts.json) into profiles/hbm_traffic_<workload>.json")
    ap.add_argument("--summarize-trace", default="", metavar="DIR",
                    help="only summarize a rocprofv3 --kernel-trace run under DIR (kt/, counts.json)")
    args = ap.parse_args()

    if args.summarize_pmc:
        out = summarize_pmc(args.summarize_pmc, args.workload)
        path = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.workload}.json")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out, indent=1))
        return
    if args.summarize_trace:
        print(json.dumps(summarize_trace(args.summarize_trace), indent=1))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "reads_per_launch_KiB": 4 * n / 1024,
                          "writes_per_launch_KiB": 4 * n / 1024}), flush=True)
        return
    if args.profile_only:
        profile_only(args)
        return
    if not args.no_adjoint:
        # the differentiable form runs through torch, whose HIP runtime must be
        # the process's: imported before the engine library is loaded
        import torch  # noqa: F401

    window_at = args.window_at if args.workload == "vessels" and not args.iterations else None
    head = plain_leg(args.workload, args, not args.no_cpu_baseline, window_at=window_at)
    phantom = None
    if args.workload != "phantom" and not args.no_phantom:
        phantom = plain_leg("phantom", args, not args.no_cpu_baseline)
    batched = None
    if not args.no_batch:
        batched = batch_vs_sequential(max(1, min(args.steps, 3)), args.batch_iterations)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_vs_plain(args.adjoint_iterations)

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "active-voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.phi == "f32" else "i8",
        "data": "synthetic",
        "config": head["config"],
        "roofline": head["roofline"],
        "path_roofline": head["path_roofline"],
        "kernels": head["kernels"],
        "random_access_floor": head["random_access_floor"],
        "cpu_baseline": head["cpu_baseline"],
        "host_inclusive": head["host_inclusive"],
        "dice_vs_oracle": head["dice_vs_oracle"],
        "mask_identical_to_oracle": head["mask_identical_to_oracle"],
        "phantom": phantom,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
