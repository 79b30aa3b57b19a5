#!/usr/bin/env python3
"""Throughput of the sparse level-set engine on one MI355X.

A step is one complete segmentation run of the workload: restore the initial
state from an HBM-resident checkpoint, then run the workload's iterations
(update, marking, duplicate removal and compaction) through the C ABI.  Inputs
are resident in HBM before the timed region starts.

Printed (one JSON line): active-voxel updates/s (sum of n_k over the run's
iterations divided by the wall time), the roofline of the dominant kernel (its
algorithmic bytes per launch, from the per-iteration counters, over its average
launch time measured with HIP events on the engine's stream), the whole-path
fraction of HBM peak, and the CPU baseline (the oracle: the reference's update
equations compiled as host C, timed on this machine's cores).

Algorithmic bytes per iteration k, FP32 phi (4 bytes), source S bytes/voxel:
  update   (38 + S) n_k   coordinate, 7 phi taps, source, phi write, flag, mask byte
  mark     33 n_k         coordinate, flag, 7 phi taps
  keep     5 c_k          candidate read-back, tag probe
  compact  5 c_k + 5 m_k  candidate write, tag set; list write, clear
with n_k active voxels, c_k candidate emissions and m_k new active voxels (the
gather model's 10 c_k + 5 m_k for duplicate removal and compaction, split over
the two kernels that do that work).  Each iteration is three launches: the
update, which from the second iteration on runs in k_update_fused together with
the previous iteration's compaction, k_mark and k_keep; the run ends with the
last iteration's compaction (k_compact).  3 K + 1 launches for K iterations.
"""
from __future__ import annotations

import argparse
import ctypes
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
    n = n.astype(np.float64)
    c = c.astype(np.float64)
    m = m.astype(np.float64)
    return {"update": (6 + 8 * P + S) * n, "mark": (5 + 7 * P) * n, "keep": 5.0 * c, "compact": 5.0 * c + 5.0 * m}


KERNELS = ("k_update_fused", "k_mark", "k_keep", "k_update", "k_compact")


def _pmc_means(path):
    """Mean counter value (KiB) per dispatch for each kernel in a rocprofv3
    counter_collection.csv."""
    import csv
    acc = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"].split("(")[0]
            name = name[5:] if name.startswith("void ") else name
            acc.setdefault(name, []).append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def summarize_pmc(root, workload):
    """HBM traffic per launch from two rocprofv3 --pmc runs of this benchmark
    (FETCH_SIZE, then WRITE_SIZE) and the same two counters over a plain copy of
    known size (--calibrate-copy), which give the scale factor of each counter."""
    def csv_in(d):
        import glob
        hits = glob.glob(os.path.join(root, d, "**", "*counter_collection.csv"), recursive=True)
        if not hits:
            raise SystemExit(f"no counter_collection.csv under {os.path.join(root, d)}")
        return hits[0]
    copy_bytes = None
    with open(os.path.join(root, "copy_fetch.json")) as f:
        for line in f:
            if line.startswith("{"):
                copy_bytes = json.loads(line)["calibrate_copy_bytes"]
    cf = _pmc_means(csv_in("copy_fetch"))
    cw = _pmc_means(csv_in("copy_write"))
    copy_name = max(cf, key=lambda k: cf[k][0])
    f_scale = copy_bytes / (cf[copy_name][0] * 1024.0)
    w_scale = copy_bytes / (cw[copy_name][0] * 1024.0)
    fe = _pmc_means(csv_in("fetch"))
    wr = _pmc_means(csv_in("write"))
    out = {"workload": workload, "copy_kernel": copy_name, "copy_bytes": copy_bytes,
           "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for k in KERNELS:
        fk = [v for n, v in fe.items() if n == "lsseg::" + k or n.startswith("lsseg::" + k + "<")]
        wk = [v for n, v in wr.items() if n == "lsseg::" + k or n.startswith("lsseg::" + k + "<")]
        if not fk or not wk:
            continue
        fk = max(fk, key=lambda v: v[1])
        wk = max(wk, key=lambda v: v[1])
        out["kernels"][k] = {"dispatches": fk[1], "fetch_KiB_raw": fk[0], "write_KiB_raw": wk[0],
                             "read_bytes": fk[0] * 1024.0 * f_scale, "write_bytes": wk[0] * 1024.0 * w_scale,
                             "bytes_per_launch": fk[0] * 1024.0 * f_scale + wk[0] * 1024.0 * w_scale}
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


def cpu_baseline(w, gpu_mask, seconds_1core, phi_storage=0):
    """The oracle on this host: full workload on all cores (also gives the mask
    for the Dice check), plus a bounded single-core sample."""
    import oracle as O
    cores = host_cores()
    ref = O.State(w.volume, phi_storage=phi_storage)
    ref.set_phi(w.phi0, w.band)
    t0 = time.perf_counter()
    n_tot = 0
    done = 0
    while done < w.iterations:
        k = min(50, w.iterations - done)
        n, _, _ = ref.iterate(w.T, w.eps, w.alpha, w.dt, k, threads=cores)
        n_tot += int(n.sum())
        done += k
    t_all = time.perf_counter() - t0
    mask = ref.mask()
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
    dice = synth.dice(gpu_mask, mask)
    identical = bool(np.array_equal(gpu_mask, mask))
    # single core: the first iterations of the same workload, bounded in time
    one = O.State(w.volume, phi_storage=phi_storage)
    one.set_phi(w.phi0, w.band)
    t0 = time.perf_counter()
    n1 = 0
    it1 = 0
    while it1 < w.iterations and time.perf_counter() - t0 < seconds_1core:
        n, _, _ = one.iterate(w.T, w.eps, w.alpha, w.dt, 10, threads=1)
        n1 += int(n.sum())
        it1 += 10
    t1 = time.perf_counter() - t0
    return {
        "value": n_tot / t_all, "unit": "active-voxel updates/s", "cores": cores, "kind": "port",
        "sample": f"full workload ({w.iterations} iterations, {n_tot} updates) in {t_all:.2f} s, "
                  f"oracle/lsoracle.c -O3 OpenMP over the active list, {cpu_model()}",
        "single_core": {"value": n1 / t1 if t1 > 0 else 0.0, "cores": 1,
                        "sample": f"first {it1} iterations ({n1} updates) in {t1:.2f} s"},
    }, dice, identical


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
    log(f"autotuned launch shape: {tune.update_blocks} update/mark workgroups, {tune.compact_blocks} compaction "
        f"workgroups ({tune.candidates} shapes)")
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
        "tuning": {"update_blocks": tune.update_blocks, "compact_blocks": tune.compact_blocks},
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


def adjoint_vs_plain(iterations, fd_h=1e-3):
    """The differentiable form on the 16-volume 96^3 float workload.  Timed:
    the plain forward (lsseg_iterate) against forward + loss + backward through
    the torch.autograd.Function (adjoint.segment, soft-Dice loss on GPU
    tensors, loss.backward()), volume after volume.  Checked: the gradients
    against central finite differences of the recorded map on every volume
    (five-point stencil at h = 1e-3, and its Richardson extrapolation from h
    and h/2), and against the FP64 host adjoint on one 64^3 instance.  Over
    200 iterations the map is nonlinear enough that the h = 1e-3 difference
    carries ~1e-2 truncation error even in exact arithmetic; the extrapolation
    removes most of it."""
    import torch
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd import adjoint, synth
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Context
    ws = synth.adjoint(iterations=iterations) if iterations else synth.adjoint()
    K = ws[0].iterations
    t_plain = t_adj = 0.0
    upd = 0
    worst = worst_dir = worst_r = worst_dir_r = worst_torch = 0.0
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

        base = [T, eps, alpha]
        for k in range(3):
            def f(step, k=k):
                v = list(base)
                v[k] += step
                return L(w.phi0, *v)
            d1, d2 = central(f, fd_h), central(f, fd_h / 2)
            worst = max(worst, rel(th[k], d1))
            worst_r = max(worst_r, rel(th[k], (16 * d2 - d1) / 15))
        d = (g0 / max(np.linalg.norm(g0), 1e-30)).astype(np.float32)
        adj_d = float((g0.astype(np.float64) * d).sum())
        fdir = (lambda step: L(w.phi0 + step * d, *base))
        d1, d2 = central(fdir, fd_h), central(fdir, fd_h / 2)
        worst_dir = max(worst_dir, rel(adj_d, d1))
        worst_dir_r = max(worst_dir_r, rel(adj_d, (16 * d2 - d1) / 15))
        rec.close()
        ctx.close()
    # FP64 host adjoint of the same recorded map on one 64^3 instance
    import oracle as O
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
        "fd_max_rel_err_theta": worst, "fd_max_rel_err_phi0_direction": worst_dir,
        "fd_richardson_max_rel_err_theta": worst_r, "fd_richardson_max_rel_err_phi0_direction": worst_dir_r,
        "fd": "central five-point stencil at h = 1e-3, and (16 D(h/2) - D(h)) / 15, on the recorded map, "
              "all volumes; phi0 along the normalised gradient",
        "fp64_64cubed_max_rel_err_theta": fp64_theta, "fp64_64cubed_rel_l2_err_phi0": fp64_phi0,
    }
    log(f"adjoint: plain forward {1e3 * t_plain:.1f} ms, forward record + loss + adjoint (torch) {1e3 * t_adj:.1f} ms "
        f"({out['adjoint_over_plain']:.2f}x; torch vs host theta {worst_torch:.1e}); FD (h=1e-3) max rel err theta {worst:.2e}, phi0 dir {worst_dir:.2e}; "
        f"extrapolated: theta {worst_r:.2e}, phi0 dir {worst_dir_r:.2e}; "
        f"FP64 64^3: theta {fp64_theta:.2e}, phi0 {fp64_phi0:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="phantom", choices=["ellipsoids", "phantom", "vessels"])
    ap.add_argument("--iterations", type=int, default=0, help="override the workload's iteration count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--phi", default="f32", choices=["f32", "i8"],
                    help="phi storage: FP32 (default) or the reference's fixed-point signed char")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-sequential comparison")
    ap.add_argument("--batch-iterations", type=int, default=0, help="override the batch workload's iterations")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-form comparison")
    ap.add_argument("--adjoint-iterations", type=int, default=0, help="override the adjoint workload's iterations")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="single-core baseline sample budget")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (HBM counter calibration)")
    ap.add_argument("--skip-profiled", action="store_true",
                    help="stop after the timed steps (for hardware-counter passes over the timed dispatches only)")
    ap.add_argument("--summarize-pmc", default="", metavar="DIR",
                    help="only summarize rocprofv3 --pmc runs under DIR (fetch/, write/, copy_fetch/, copy_write/, "
                         "copy_fetch.json) into profiles/hbm_traffic_<workload>.json")
    args = ap.parse_args()
    if not (args.no_adjoint or args.summarize_pmc or args.calibrate_copy):
        # the differentiable form runs through torch, whose HIP runtime must be
        # the process's: imported before the engine library is loaded
        import torch  # noqa: F401

    if args.summarize_pmc:
        out = summarize_pmc(args.summarize_pmc, args.workload)
        path = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.workload}.json")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out, indent=1))
        return

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

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import PHI_F32, PHI_I8, Context

    t = time.perf_counter()
    w = make_workload(args.workload, args.iterations)
    log(f"workload generated in {time.perf_counter() - t:.1f} s")
    phi_storage = PHI_I8 if args.phi == "i8" else PHI_F32
    ctx = Context(w.volume, phi_storage=phi_storage)
    ctx.set_phi(w.phi0, w.band)
    ctx.checkpoint_save(0)
    S = w.volume.dtype.itemsize
    P = 1 if args.phi == "i8" else 4
    log(f"context ready: {describe(w, ctx.X, ctx.Y, ctx.Z)}; n0={ctx.active_count}")

    for _ in range(args.warmup):
        ctx.checkpoint_restore(0)
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    log("warmup done")

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
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {updates_per_step} updates/step, "
        f"few/many iterations {last.few_path_iters}/{last.many_path_iters}, final n={last.active_final}")
    if args.skip_profiled:
        ctx.close()
        return

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
    log(f"host-inclusive run: {host_ms:.2f} ms ({host_inclusive['value']:.3e} updates/s)")

    # per-launch device times (HIP events on the engine's stream), in issue
    # order: update, marking and keep of each iteration (the update fused with
    # the previous compaction from the second iteration on), then the last
    # compaction.
    ctx.checkpoint_restore(0)
    pst, kt, per_launch = ctx.iterate_profiled(w.T, w.eps, w.alpha, w.dt, w.iterations)
    pc = ctx.iteration_counts()
    assert np.array_equal(pc.n, counts.n), "profiled run diverged from the timed run"
    B = algorithmic_bytes(pc.n, pc.c, pc.m, S, P)
    it = w.iterations
    assert it >= 2 and per_launch.size == 3 * it + 1, f"expected {3 * it + 1} launches, got {per_launch.size}"
    per_launch = per_launch.astype(np.float64)
    head = per_launch[0:3 * it:3]
    parts = {
        # launch k >= 1: the update of iteration k and the compaction of k - 1
        "update_fused": {"kernel": "k_update_fused", "bytes": B["update"][1:] + B["compact"][:-1], "ms": head[1:]},
        "mark": {"kernel": "k_mark", "bytes": B["mark"], "ms": per_launch[1:3 * it:3]},
        "keep": {"kernel": "k_keep", "bytes": B["keep"], "ms": per_launch[2:3 * it:3]},
        # the first update and the last compaction run on their own
        "ends": {"kernel": "k_update + k_compact", "bytes": np.array([B["update"][0], B["compact"][-1]]),
                 "ms": np.array([head[0], per_launch[-1]])},
    }
    kernels = {}
    for k, v in parts.items():
        tot_ms = float(v["ms"].sum())
        launches = v["ms"].size
        avg_us = 1e3 * tot_ms / launches
        bpl = float(v["bytes"].sum()) / launches
        kernels[k] = {"kernel": v["kernel"], "avg_launch_us": avg_us, "bytes_per_launch": bpl,
                      "achieved_GBs": bpl / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0, "total_ms": tot_ms}
        kernels[k]["launches"] = launches
        nk = pc.n[1:] if k == "update_fused" else (pc.n if v["ms"].size == it else pc.n[[0, -1]])
        big = nk >= np.percentile(nk, 90)
        kernels[k]["busiest_decile_GBs"] = (float(v["bytes"][big].sum()) / (float(v["ms"][big].sum()) * 1e-3) / 1e9
                                            if big.any() and v["ms"][big].sum() > 0 else 0.0)
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    traffic, traffic_src = load_traffic(args.workload)
    traffic_bytes = None
    if traffic and traffic.get("kernels", {}).get(d["kernel"]):
        traffic_bytes = traffic["kernels"][d["kernel"]]["bytes_per_launch"]
    avg_us, bytes_per_launch, achieved = d["avg_launch_us"], d["bytes_per_launch"], d["achieved_GBs"]
    path_bytes = float(sum(v.sum() for v in B.values()))
    path_gbs = path_bytes / (ms_per_step * 1e-3) / 1e9
    log("per launch: " + "; ".join(f"{k} {v['avg_launch_us']:.2f} us, {v['bytes_per_launch'] / 1e6:.2f} MB, "
                                   f"{v['achieved_GBs']:.0f} GB/s (busiest decile {v['busiest_decile_GBs']:.0f})"
                                   for k, v in kernels.items())
        + f"; kernel time sum {kt.total_ms:.2f} ms vs {ms_per_step:.2f} ms/step")

    batched = None
    if not args.no_batch:
        batched = batch_vs_sequential(max(1, min(args.steps, 3)), args.batch_iterations)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_vs_plain(args.adjoint_iterations)

    cpu = None
    dice = None
    identical = None
    if not args.no_cpu_baseline:
        log("running the CPU oracle (full workload on all cores, then a single-core sample)")
        cpu, dice, identical = cpu_baseline(w, gpu_mask, args.cpu_seconds, 1 if args.phi == "i8" else 0)
        log(f"cpu baseline {cpu['value']:.3e} updates/s on {cpu['cores']} cores; Dice vs oracle {dice:.6f}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "active-voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": describe(w, ctx.X, ctx.Y, ctx.Z),
            "volume_xyz": list(w.shape_xyz),
            "iterations": w.iterations,
            "active_voxel_updates_per_step": updates_per_step,
            "parallelism": "single GPU",
            "phi_storage": "fixed-point int8 (LEVEL_SET_CHAR)" if args.phi == "i8" else "FP32",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes,
            "traffic_source": traffic_src,
            "kernel": d["kernel"], "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_us,
        },
        "path_roofline": {"achieved": path_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": path_gbs / HBM_PEAK_GBS, "bytes_per_step": path_bytes},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "host_inclusive": host_inclusive,
        "batched": batched,
        "adjoint": adjoint,
        "dice_vs_oracle": dice,
        "mask_identical_to_oracle": identical,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
