"""Benchmark of the per-electron basis path on one MI355X.

A step is one pass of the path over the whole batch: the point coordinates
are re-stamped (qmckl_context_touch_device) and en_distance, ao_vgl and
mo_vgl are provided again through the C ABI, with the inputs already
resident in HBM.  Default workload: the 64-water cluster at 128 walkers
(81,920 electron-points, 192 nuclei, 4,160 AOs, 320 MOs).

Prints one JSON line (rank 0, one GPU).  The value is electron-points/s;
`roofline` is computed for the kernel that takes the most time, from its
algorithmic work and its average duration measured with HIP events on the
context's stream over the timed steps; `cpu_baseline` times the FP64 CPU
oracle on a bounded sample of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "f16": 2500.0}
METRIC = "AO+MO VGL electron-points/s: mixed/batched vs FP64 plain, MO rel err ≤1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(name, walkers):
    from qmckl_sycl_amd import systems
    if name == "water64":
        return systems.water64(walk_num=walkers or 128)
    if name == "chbrclf_sized":
        return systems.chbrclf_sized(walk_num=walkers or 256)
    if name == "n2_like":
        return systems.n2_like(walk_num=walkers or 1)
    raise SystemExit(f"unknown config {name}")


def kernel_work(name, s):
    """(kind, amount per launch, unit) of algorithmic work."""
    P, N, A, M = s.point_num, s.nucl_num, s.ao_num, s.mo_num
    if name == "mo_gemm_f64":
        return "mfma", 10.0 * P * A * M, "f64"
    if name == "ao_vgl":
        return "hbm", P * (24.0 + 40.0 * A), None
    if name == "en_distance":
        return "hbm", P * (24.0 + 8.0 * N), None
    return None, 0.0, None


def run_plain(s, steps, warmup):
    from qmckl_sycl_amd.context import Context, lib
    L = lib()
    ctx = Context(0)
    ctx.set_system(s)
    h = ctx.handle

    def step():
        ctx.check(L.qmckl_context_touch_device(h))
        ctx.check(L.qmckl_provide_en_distance_device(h))
        ctx.check(L.qmckl_provide_mo_basis_mo_vgl_device(h))

    for _ in range(warmup):
        step()
    import torch
    torch.cuda.synchronize()
    ctx.check(L.qmckl_reset_kernel_timing_device(h))
    ctx.check(L.qmckl_set_kernel_timing_device(h, True))
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.check(L.qmckl_set_kernel_timing_device(h, False))
    import ctypes
    kernels = {}
    for name in ("en_distance", "ao_vgl", "mo_gemm_f64"):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        L.qmckl_get_kernel_timing_device(h, name.encode(), ctypes.byref(n), ctypes.byref(ms))
        if n.value:
            kernels[name] = {"launches": int(n.value), "avg_ms": ms.value / n.value}
    # parity spot check of the timed output against the oracle
    mo = ctx.mo_vgl(s.point_num)
    ctx.destroy()
    return (t1 - t0) / steps, kernels, mo


def parity_sample(s, mo, n=64, seed=0):
    from oracle import oracle
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(s.point_num, size=min(n, s.point_num), replace=False))
    coord = np.ascontiguousarray(s.coord_T()[:, idx])
    ao, am = oracle.ao_vgl(s, coord, with_mag=True)
    ref, mm = oracle.mo_vgl_from_ao(s.mo_coef, ao, am)
    got = mo[idx]
    d = np.abs(got - ref)
    rel_mag = (d / np.where(mm > 0, mm, 1.0)).max(axis=(0, 2))
    rowmax = np.abs(ref).max(axis=2, keepdims=True)
    rel_row = (d / np.where(rowmax > 0, rowmax, 1.0)).max(axis=(0, 2))
    return {"points_checked": int(idx.size),
            "max_err_over_term_sum": [float(x) for x in rel_mag],
            "max_err_over_row_max": [float(x) for x in rel_row]}


def cpu_baseline(s, target_s=10.0):
    """FP64 oracle (OpenMP over points) on a walker subset of the workload."""
    from oracle import oracle
    cores = os.cpu_count() or 1
    oracle.set_num_threads(cores)
    E = s.elec_num

    def run(w):
        coord = np.ascontiguousarray(s.coord_T()[:, :w * E])
        t = time.perf_counter()
        ao = oracle.ao_vgl(s, coord)
        oracle.mo_vgl_from_ao(s.mo_coef, ao)
        return time.perf_counter() - t

    run(1)                       # warm-up
    t1 = run(1)
    w = int(max(1, min(s.walk_num, 32, round(target_s / max(t1, 1e-6)))))
    t = run(w) if w > 1 else t1
    pts = w * E
    return {"value": pts / t, "unit": "electron-points/s", "cores": cores, "kind": "port",
            "sample": f"{w} of {s.walk_num} walkers ({pts} points), FP64 oracle "
                      f"(ao_vgl then per-point mo_vgl loop), {t:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="water64")
    ap.add_argument("--walkers", type=int, default=0)
    ap.add_argument("--form", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    s = make_system(args.config, args.walkers)
    log("workload", s.summary())
    sec, kernels, mo = run_plain(s, args.steps, args.warmup)
    value = s.point_num / sec
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    kind, work, prec = kernel_work(dom, s)
    avg_s = kernels[dom]["avg_ms"] / 1e3
    if kind == "mfma":
        ach = work / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                "frac": ach / PEAK_TFLOPS[prec], "traffic": None, "kernel": dom}
    else:
        ach = work / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": dom}
    for k, v in kernels.items():
        kk, w, p = kernel_work(k, s)
        if kk == "mfma":
            v["achieved_tflops"] = w / (v["avg_ms"] / 1e3) / 1e12
        elif kk == "hbm":
            v["achieved_gbs"] = w / (v["avg_ms"] / 1e3) / 1e9
    parity = parity_sample(s, mo)
    out = {
        "metric": METRIC, "value": value, "unit": "electron-points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{s.name}: {s.nucl_num} nuclei, {s.ao_num} AOs, {s.mo_num} MOs, "
                               f"{s.walk_num} walkers x {s.elec_num} electrons",
                   "points": s.point_num, "form": args.form},
        "roofline": roof, "kernels": kernels, "parity": parity,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
