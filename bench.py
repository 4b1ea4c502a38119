"""Benchmark of the per-electron basis path on one MI355X.

A step is one pass of the path over the whole batch: the point coordinates
are re-stamped (qmckl_context_touch_device) and en_distance, ao_vgl and
mo_vgl are provided again through the C ABI, with the inputs already
resident in HBM.  Default workload: the 64-water cluster at 128 walkers
(81,920 electron-points, 192 nuclei, 4,160 AOs, 320 MOs).

The headline form (--form, default fp16x2: FP64 distances and AO assembly,
split-fp16 MO contraction) is timed next to the FP64 plain form on the same
context; both are checked against the FP64 oracle on sampled points.
Prints one JSON line (one GPU).  The value is electron-points/s of the
headline form; `roofline` is computed for the kernel that takes the most
time, from its algorithmic work and its average duration measured with HIP
events on the context's stream over the timed steps; `cpu_baseline` times
the FP64 CPU oracle on a bounded sample of the same workload on the host
cores.
"""
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
METRIC = "AO+MO VGL electron-points/s: mixed/batched vs FP64 plain, MO rel err ≤1e-6"
FORM_IDS = {"plain": 0, "fp32": 1, "fp16x2": 2, "bf16x2": 3, "bf16x3": 4, "int8x5": 5}
GEMM_KERNEL = {"plain": "mo_gemm_f64", "fp32": "mo_gemm_fp32", "fp16x2": "mo_gemm_fp16x2",
               "bf16x2": "mo_gemm_bf16x2", "bf16x3": "mo_gemm_bf16x3", "int8x5": "mo_gemm_int8x5"}
# published dense peaks (TFLOP/s) of the matrix-core type each GEMM runs on
GEMM_PEAK = {"mo_gemm_f64": ("f64", 78.6), "mo_gemm_fp32": ("f32", 157.3),
             "mo_gemm_fp16x2": ("f16", 2500.0), "mo_gemm_bf16x2": ("bf16", 2500.0),
             "mo_gemm_bf16x3": ("bf16", 2500.0), "mo_gemm_int8x5": ("i8", 5000.0)}
PRODUCTS = {"mo_gemm_f64": 1, "mo_gemm_fp32": 1, "mo_gemm_fp16x2": 3, "mo_gemm_bf16x2": 3,
            "mo_gemm_bf16x3": 6, "mo_gemm_int8x5": 15}


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
    """(bound, algorithmic work per launch) of a kernel: bytes for HBM-bound
    kernels, flops (10 P A M, logical FMAs counted once) for the MO GEMMs."""
    P, N, A, M = s.point_num, s.nucl_num, s.ao_num, s.mo_num
    if name in GEMM_PEAK:
        return "mfma", 10.0 * P * A * M
    if name == "ao_vgl":
        return "hbm", P * (24.0 + 40.0 * A)
    if name == "ao_vgl_f32":
        return "hbm", P * (24.0 + 20.0 * A)
    if name == "ao_vgl_f32_lo":  # value and residual, the INT8X5 form's operand
        return "hbm", P * (24.0 + 40.0 * A)
    if name == "en_distance":
        return "hbm", P * (24.0 + 8.0 * N)
    return None, 0.0


def run_form(ctx, s, form, steps, warmup):
    import torch
    from qmckl_sycl_amd.context import lib
    L = lib()
    h = ctx.handle
    ctx.check(L.qmckl_set_mo_basis_precision_device(h, FORM_IDS[form]))

    def step():
        ctx.check(L.qmckl_context_touch_device(h))
        ctx.check(L.qmckl_provide_en_distance_device(h))
        ctx.check(L.qmckl_provide_mo_basis_mo_vgl_device(h))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.check(L.qmckl_reset_kernel_timing_device(h))
    ctx.check(L.qmckl_set_kernel_timing_device(h, True))
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.check(L.qmckl_set_kernel_timing_device(h, False))
    kernels = {}
    for name in ("en_distance", "point_order", "ao_vgl", "ao_vgl_f32", "ao_vgl_f32_lo", GEMM_KERNEL[form]):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        L.qmckl_get_kernel_timing_device(h, name.encode(), ctypes.byref(n), ctypes.byref(ms))
        if n.value:
            k = {"launches": int(n.value), "avg_ms": ms.value / n.value}
            kind, w = kernel_work(name, s)
            if kind == "mfma":
                k["achieved_tflops"] = w / (k["avg_ms"] / 1e3) / 1e12
            elif kind == "hbm":
                k["achieved_gbs"] = w / (k["avg_ms"] / 1e3) / 1e9
            kernels[name] = k
    mo = ctx.mo_vgl(s.point_num)
    return (t1 - t0) / steps, kernels, mo


def bench_batch(forms, steps, warmup, scale=1.0, eager_only=False):
    """Batching form: the 48-molecule ragged batch (2-64 nuclei, 30-1,000 AOs,
    16-512 walkers) through qmckl_get_mo_basis_mo_vgl_batch_device, one
    context per molecule, timed per form with the default tile choices;
    next to it the same molecules one inplace getter call after another
    (each call synchronises), the batch replayed as a HIP graph, and for the
    mixed forms both again after qmckl_mo_basis_autotune_device on every
    context (the default tiles are the fixed-tile baseline)."""
    import torch
    from qmckl_sycl_amd import systems
    from qmckl_sycl_amd.context import BatchGraph, Context, DeviceArray, batch_mo_vgl, lib
    L = lib()
    mols = systems.ragged_batch(48, scale=scale)
    points = int(sum(m.point_num for m in mols))
    ctxs, outs = [], []
    for m in mols:
        c = Context(0)
        c.set_system(m)
        ctxs.append(c)
        outs.append(DeviceArray(c, (m.point_num, 5, m.mo_num)))
    res = {"molecules": len(mols), "points": points,
           "ao_range": [int(min(m.ao_num for m in mols)), int(max(m.ao_num for m in mols))],
           "points_range": [int(min(m.point_num for m in mols)), int(max(m.point_num for m in mols))]}

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    try:
        for form in forms:
            for c in ctxs:
                c.check(L.qmckl_set_mo_basis_precision_device(c.handle, FORM_IDS[form]))
            sec = timed(lambda: batch_mo_vgl(ctxs, outs))
            res[form] = {"ms_per_step": sec * 1e3, "value": points / sec}
            if eager_only:
                log("batch", form, json.dumps(res[form]))
                continue

            def one_by_one():
                for c, o in zip(ctxs, outs):
                    c.check(L.qmckl_get_mo_basis_mo_vgl_inplace_device(c.handle, o.ptr, o.size))
            sec = timed(one_by_one)
            res[form]["one_call_per_molecule"] = {"ms_per_step": sec * 1e3, "value": points / sec}

            def graph_rate():
                g = BatchGraph(ctxs, outs)
                try:
                    def launch():
                        rc = g.launch()
                        if rc != 0:
                            raise RuntimeError(f"graph launch returned {rc}")
                    return timed(launch)
                finally:
                    g.destroy()
            sec = graph_rate()
            res[form]["graph"] = {"ms_per_step": sec * 1e3, "value": points / sec}
            if form != "plain":
                for c in ctxs:
                    c.check(L.qmckl_mo_basis_autotune_device(c.handle))
                sec = timed(lambda: batch_mo_vgl(ctxs, outs))
                res[form]["autotuned"] = {"ms_per_step": sec * 1e3, "value": points / sec}
                sec = graph_rate()
                res[form]["autotuned_graph"] = {"ms_per_step": sec * 1e3, "value": points / sec}
                for c in ctxs:  # back to the fixed tiles for any later form
                    c.check(L.qmckl_set_mo_basis_tile_config_device(c.handle, -1, 0))
            res[form]["mo_error_vs_fp64_oracle"] = batch_errors(mols, outs, form)
            log("batch", form, json.dumps(res[form]))
    finally:
        for c in ctxs:
            c.destroy()
    return res


def batch_errors(mols, outs, form, n_pts=64):
    """The batch's outputs against the FP64 oracle on 64 sampled points of
    eight molecules (the two largest by P A M and six spread over the
    batch); the worst per-component error over them."""
    by_work = sorted(range(len(mols)), key=lambda i: mols[i].point_num * mols[i].ao_num * mols[i].mo_num)
    picks = sorted(set(by_work[-2:]) | set(np.linspace(0, len(mols) - 1, 6).astype(int).tolist()))
    worst = {"value": 0.0, "gradient": 0.0, "laplacian": 0.0}
    for i in picks:
        m = mols[i]
        idx = sample_points(m, n_pts, seed=i)
        e = errors_vs_oracle(m, {form: outs[i].numpy()[idx]}, idx)[form]
        for k in worst:
            worst[k] = max(worst[k], e[k])
    return {"molecules_checked": [int(i) for i in picks], "points_per_molecule": n_pts, **worst}


# device kernel name (as the profiler lists it) of each timed kernel
# (the FP64 ao_vgl kernel depends on the basis: lane-per-point for 1024 AOs
# or more in whole 64-byte rows, lane-per-AO otherwise)
KERNEL_SYMBOL = {"en_distance": "en_distance_kernel", "ao_vgl": ("ao_vgl_f64_kernel", "ao_kernel<5, false, double>"),
                 "ao_vgl_f32": "ao_rows_kernel<false>", "ao_vgl_f32_lo": "ao_rows_kernel<true>",
                 "mo_gemm_f64": "gemm_f64_frag_kernel<",
                 "mo_gemm_fp32": "gemm_f32_kernel<",
                 "mo_gemm_fp16x2": ("gemm_split_kernel<2,", "gemm_split_live_kernel<2,"),
                 "mo_gemm_bf16x2": ("gemm_split_kernel<3,", "gemm_split_live_kernel<3,"),
                 "mo_gemm_bf16x3": ("gemm_split_kernel<4,", "gemm_split_live_kernel<4,"),
                 "mo_gemm_int8x5": "gemm_i8_kernel<"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_{workload}.json")


def measured(s, kernel, key="bytes_per_launch"):
    """A per-launch figure of a kernel on this workload from the committed
    rocprofv3 counter measurement (bench.py --traffic): HBM bytes
    (FETCH_SIZE + WRITE_SIZE, calibrated) or issued MFMA flops; None when
    there is none."""
    path = TRAFFIC_FILE.format(workload=f"{s.name}_{s.walk_num}w")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get(key)


def measured_adjoint(kernel):
    """The adjoint kernels' record in the committed counter measurement."""
    path = TRAFFIC_FILE.format(workload=ADJOINT_WORKLOAD)
    try:
        with open(path) as f:
            return json.load(f).get("kernels", {}).get(kernel)
    except (OSError, ValueError):
        return None


def bench_value_only(steps, warmup):
    """The value-only path (mo_value; src/qmckl_mo.cpp:185-453) at the C3
    size, FP64: per step the points are touched and mo_value provided, which
    computes ao_value and the value-only contraction (mo_vgl is not
    computed)."""
    import torch
    from qmckl_sycl_amd.context import Context, lib
    L = lib()
    s = make_system("water64", 0)
    ctx = Context(0)
    try:
        ctx.set_system(s)
        h = ctx.handle

        def step():
            ctx.check(L.qmckl_context_touch_device(h))
            ctx.check(L.qmckl_provide_mo_basis_mo_value_device(h))
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        res = {"workload": f"water64 at {s.walk_num} walkers ({s.point_num} points), FP64 mo_value "
                           "(ao_value, then the value-only contraction)",
               "ms_per_step": sec * 1e3, "value": s.point_num / sec, "unit": "electron-points/s"}
        ctx.check(L.qmckl_reset_kernel_timing_device(h))
        ctx.check(L.qmckl_set_kernel_timing_device(h, True))
        for _ in range(2):
            step()
        ctx.check(L.qmckl_set_kernel_timing_device(h, False))
        kernels = {}
        for name in ("point_order", "ao_value", "mo_value_gemm_f64"):
            n, ms = ctypes.c_int64(0), ctypes.c_double(0.0)
            L.qmckl_get_kernel_timing_device(h, name.encode(), ctypes.byref(n), ctypes.byref(ms))
            if n.value:
                kernels[name] = {"launches": int(n.value), "avg_ms": ms.value / n.value}
        if "ao_value" in kernels:  # algorithmic bytes: 8 per (point, AO) written, 24 per point read
            kernels["ao_value"]["achieved_gbs"] = (s.point_num * (24.0 + 8.0 * s.ao_num) /
                                                   (kernels["ao_value"]["avg_ms"] / 1e3) / 1e9)
        res["kernels"] = kernels
        log("value_only", json.dumps(res))
        return res
    finally:
        ctx.destroy()


ADJOINT_WORKLOAD = "water64_diff_32w"
VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wavefront VALU instructions/s: 1024 SIMDs, 4 cycles each at 2.4 GHz


def bench_adjoint(steps, warmup, forward=True):
    """Differentiation form at C5 (water64, 32 walkers, weights w ~ N(0,1)
    shaped like mo_vgl): the forward FP64 mo_vgl then the adjoint (dL/dR,
    dL/dalpha, dL/dC of L = sum w mo_vgl), timed per step.  forward=False
    runs the adjoint alone (the counter runs use it to separate the
    adjoint's G GEMM from the forward contraction, which is the same
    kernel)."""
    import torch
    from qmckl_sycl_amd import systems
    from qmckl_sycl_amd.context import Context, DeviceArray, lib
    L = lib()
    s, w = systems.water64_diff(walk_num=32)
    ctx = Context(0)
    try:
        ctx.set_system(s)
        wd = DeviceArray.from_numpy(ctx, w)
        dR = DeviceArray(ctx, (s.nucl_num, 3))
        dA = DeviceArray(ctx, (s.prim_num,))
        dC = DeviceArray(ctx, (s.mo_num, s.ao_num))
        mo = DeviceArray(ctx, (s.point_num, 5, s.mo_num))
        h = ctx.handle

        def step():
            if forward:
                ctx.check(L.qmckl_get_mo_basis_mo_vgl_inplace_device(h, mo.ptr, mo.size))
            ctx.check(L.qmckl_get_mo_basis_mo_vgl_adjoint_device(h, wd.ptr, wd.size, dR.ptr, dA.ptr, dC.ptr))
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        res = {"workload": f"water64 at 32 walkers ({s.point_num} points), FP64 "
                           f"{'forward + ' if forward else ''}adjoint",
               "ms_per_step": sec * 1e3, "value": s.point_num / sec, "unit": "electron-points/s"}
        # per-kernel breakdown from a separate run with the HIP-event timers on
        ctx.check(L.qmckl_reset_kernel_timing_device(h))
        ctx.check(L.qmckl_set_kernel_timing_device(h, True))
        for _ in range(2):
            step()
        ctx.check(L.qmckl_set_kernel_timing_device(h, False))
        kernels = {}
        for name in ("ao_vgl", "mo_gemm_f64", "adjoint_g_gemm_f64", "ao_adjoint", "adjoint_dc_gemm_f64"):
            n, ms = ctypes.c_int64(0), ctypes.c_double(0.0)
            L.qmckl_get_kernel_timing_device(h, name.encode(), ctypes.byref(n), ctypes.byref(ms))
            if n.value:
                kernels[name] = {"launches": int(n.value), "avg_ms": ms.value / n.value}
        # Per-kernel rooflines.  The FP64 GEMMs (G = w C over [5P][M] x [M][A],
        # dL/dC = AO^T w, the forward contraction) on the MFMA flops they
        # issued, from the committed counter run of the adjoint alone
        # (they skip zero coefficient blocks and screened AO blocks, so the
        # dense 10 P A M, also given, can read above the peak); ao_adjoint,
        # which recomputes the primitives and the product rule per point, on
        # its VALU issue rate (wavefront VALU instructions per second
        # against one per 4 cycles on each of the 1,024 SIMDs).  The forward
        # kernels (ao_vgl, mo_gemm_f64) are the plain form's, whose
        # rooflines the plain_fp64 record gives.
        P, A, M = s.point_num, s.ao_num, s.mo_num
        for name, k in kernels.items():
            sec_k = k["avg_ms"] / 1e3
            prof = measured_adjoint(name)
            if name in ("ao_vgl", "mo_gemm_f64"):
                continue
            if name.endswith("gemm_f64"):
                dense = 10.0 * P * A * M / sec_k / 1e12
                rec = {"bound": "mfma", "achieved": dense, "peak": 78.6, "unit": "TFLOP/s",
                       "frac": dense / 78.6, "basis": "dense 10 P A M"}
                issued = (prof or {}).get("mfma_flops_per_launch")
                if issued:
                    rec.update(achieved=issued / sec_k / 1e12, frac=issued / sec_k / 1e12 / 78.6,
                               basis="issued MFMA flops (live blocks)",
                               dense={"achieved": dense, "frac": dense / 78.6},
                               dense_fraction=issued / (10.0 * P * A * M))
                k["roofline"] = rec
            elif name == "ao_adjoint" and prof and prof.get("valu_insts_per_launch"):
                ach = prof["valu_insts_per_launch"] / sec_k
                k["roofline"] = {"bound": "valu", "achieved": ach, "peak": VALU_PEAK,
                                 "unit": "wavefront VALU instructions/s", "frac": ach / VALU_PEAK,
                                 "hbm_bytes_per_launch": prof.get("bytes_per_launch")}
        res["kernels"] = kernels
        log("adjoint", json.dumps(res))
        return res
    finally:
        ctx.destroy()


def roofline(kernels, s, basis="dense"):
    """Roofline record of the kernel that takes the most time.  For a GEMM,
    basis "dense" prices the algorithmic work 10 P A M (logical FMAs counted
    once) against the pipe's peak; basis "issued" prices the MFMA flops the
    kernel issued (SQ_INSTS_VALU_MFMA_MOPS_* x 512 from the committed
    counter run, split products included), its utilisation of the matrix
    cores.  The kernels skip exact-zero blocks (screened AOs, zero
    coefficient blocks), so dense work can exceed what the matrix cores run:
    the FP64 contraction issues 41% of the dense products, and its dense
    figure would read above 1."""
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    kind, work = kernel_work(dom, s)
    avg_s = kernels[dom]["avg_ms"] / 1e3
    traffic = measured(s, dom)
    if kind == "mfma":
        prec, peak = GEMM_PEAK[dom]
        dense = work / avg_s / 1e12
        issued = measured(s, dom, "mfma_flops_per_launch")
        rec = {"bound": "mfma", "achieved": dense, "peak": peak, "unit": "TFLOP/s",
               "frac": dense / peak, "traffic": traffic, "kernel": dom, "peak_dtype": prec,
               "mfma_products_per_fma": PRODUCTS[dom], "basis": "dense 10 P A M"}
        if issued:
            rec["issued"] = {"flops_per_launch": issued, "achieved": issued / avg_s / 1e12,
                             "frac": issued / avg_s / 1e12 / peak,
                             "dense_fraction": issued / (work * PRODUCTS[dom])}
            if basis == "issued":
                rec["dense"] = {"flops_per_launch": work, "achieved": dense, "frac": dense / peak}
                rec.update(achieved=issued / avg_s / 1e12, frac=issued / avg_s / 1e12 / peak,
                           basis="issued MFMA flops (live blocks x split products)")
        return rec
    ach = work / avg_s / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom}


def _pmc_per_kernel(counters, argv):
    """Runs rocprofv3 --pmc <counters> (one pass) over a child python
    process running this file with argv; returns {counter: {kernel name:
    mean value per dispatch}}."""
    import csv
    import subprocess
    import tempfile
    counters = counters.split()
    out = tempfile.mkdtemp(prefix="qmckl_pmc_")
    cmd = ["rocprofv3", "--pmc"] + counters + ["--output-format", "csv", "-d", out, "-o", "run", "--",
                                               sys.executable, os.path.abspath(__file__)] + argv
    log("running", " ".join(cmd))
    subprocess.run(cmd, check=True, timeout=300, stdout=sys.stderr, stderr=sys.stderr)
    vals = {c: {} for c in counters}
    for dirpath, _, files in os.walk(out):
        for fn in files:
            if fn.endswith("counter_collection.csv"):
                with open(os.path.join(dirpath, fn)) as f:
                    for r in csv.DictReader(f):
                        if r["Counter_Name"] in vals:
                            # one dispatch can appear once per counter
                            d = vals[r["Counter_Name"]].setdefault(r["Kernel_Name"], {})
                            key = r.get("Dispatch_Id", len(d))
                            d[key] = d.get(key, 0.0) + float(r["Counter_Value"])
    return {c: {k: sum(v.values()) / len(v) for k, v in per.items()} for c, per in vals.items()}


def traffic_profile(args):
    """Measures HBM bytes per launch of each timed kernel of the workload:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) in two profiler runs, each
    scaled by what the same counter reads for a plain 1 GiB device copy.
    Writes profiles/hbm_traffic_<workload>.json and prints it."""
    s = make_system(args.config, args.walkers)
    run = ["--steps", "2", "--warmup", "0", "--no-cpu", "--no-batch", "--form", args.form,
           "--config", args.config] + (["--walkers", str(args.walkers)] if args.walkers else [])
    cal = ["--copy-calibration", "1024"]
    copy_bytes = 1024 * (1 << 20)
    factors, per = {}, {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        c = _pmc_per_kernel(counter, cal)[counter]
        copy_kib = max(v for k, v in c.items() if "copy" in k.lower() or "elementwise" in k)
        factors[counter] = copy_bytes / (copy_kib * 1024.0)
        per.update(_pmc_per_kernel(counter, run))
    mops = ("SQ_INSTS_VALU_MFMA_MOPS_F64", "SQ_INSTS_VALU_MFMA_MOPS_F16", "SQ_INSTS_VALU_MFMA_MOPS_BF16",
            "SQ_INSTS_VALU_MFMA_MOPS_F32", "SQ_INSTS_VALU_MFMA_MOPS_I8")
    per.update(_pmc_per_kernel(" ".join(mops), run))

    def mean_of(counter, sym):
        syms = (sym,) if isinstance(sym, str) else sym
        hits = [v for k, v in per[counter].items()
                if any(x in k.replace("(anonymous namespace)::", "") for x in syms)]
        return sum(hits) / len(hits) if hits else None

    kernels = {}
    for timer, sym in KERNEL_SYMBOL.items():
        rec = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            v = mean_of(counter, sym)
            if v is not None:
                rec[counter.lower() + "_bytes"] = v * 1024.0 * factors[counter]
        if len(rec) == 2:
            rec["bytes_per_launch"] = rec["fetch_size_bytes"] + rec["write_size_bytes"]
            kind, work = kernel_work(timer, s)
            if kind == "hbm":
                rec["algorithmic_bytes_per_launch"] = work
            else:
                rec["mfma_flops_per_launch"] = 512.0 * sum(mean_of(c, sym) or 0.0 for c in mops)
                rec["algorithmic_flops_per_launch"] = work
            kernels[timer] = rec
    out = {"workload": f"{s.name}_{s.walk_num}w", "form": args.form,
           "counter_scale": factors,
           "mfma_counters": list(mops),
           "command": "python bench.py --traffic " + " ".join(run[6:]),
           "kernels": kernels}
    path = TRAFFIC_FILE.format(workload=out["workload"])
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)

    # the adjoint alone (its G GEMM is the forward's kernel, so the forward
    # stays out of these runs)
    arun = ["--adjoint-only", "--steps", "2", "--warmup", "0"]
    aper = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        aper.update(_pmc_per_kernel(counter, arun))
    aper.update(_pmc_per_kernel(" ".join(mops) + " SQ_INSTS_VALU", arun))
    per = aper
    adj = {}
    for timer, sym in (("adjoint_g_gemm_f64", "gemm_f64_frag_kernel<5, false>"),
                       ("adjoint_dc_gemm_f64", "gemm_f64_frag_kernel<5, true>"),
                       ("ao_adjoint", "ao_adjoint_kernel")):
        rec = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            v = mean_of(counter, sym)
            if v is not None:
                rec[counter.lower() + "_bytes"] = v * 1024.0 * factors[counter]
        if len(rec) == 2:
            rec["bytes_per_launch"] = rec["fetch_size_bytes"] + rec["write_size_bytes"]
        f = 512.0 * sum(mean_of(c, sym) or 0.0 for c in mops)
        if f:
            rec["mfma_flops_per_launch"] = f
        v = mean_of("SQ_INSTS_VALU", sym)
        if v is not None:
            rec["valu_insts_per_launch"] = v
        adj[timer] = rec
    aout = {"workload": ADJOINT_WORKLOAD, "counter_scale": factors, "mfma_counters": list(mops),
            "command": "python bench.py --traffic (adjoint runs: bench.py " + " ".join(arun) + ")",
            "kernels": adj}
    with open(TRAFFIC_FILE.format(workload=ADJOINT_WORKLOAD), "w") as f:
        json.dump(aout, f, indent=1)
    print(json.dumps(aout), flush=True)


def sample_points(s, n=256, seed=0):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(s.point_num, size=min(n, s.point_num), replace=False))


def errors_vs_oracle(s, mos, idx):
    """Per-form error on the sampled points idx (mos[form] holds mo_vgl at
    those points): for each (point, component) row max_i |d| / max_i |ref|,
    maximised over rows, for value / gradient / Laplacian."""
    from oracle import oracle
    coord = np.ascontiguousarray(s.coord_T()[:, idx])
    ref = oracle.mo_vgl_from_ao(s.mo_coef, oracle.ao_vgl(s, coord))
    rm = np.abs(ref).max(axis=2, keepdims=True)
    rm = np.where(rm > 0, rm, 1.0)
    out = {"points_checked": int(idx.size),
           "measure": "max over (point, component) rows of max_i|d|/max_i|ref|"}
    for form, mo in mos.items():
        e = (np.abs(mo - ref) / rm).max(axis=(0, 2))
        out[form] = {"value": float(e[0]), "gradient": float(e[1:4].max()),
                     "laplacian": float(e[4])}
    return out


def errors_all_points(s, mo, chunk=8192):
    """The same measure as errors_vs_oracle over every point of the
    workload (mo: mo_vgl of all points), the oracle run in chunks."""
    from oracle import oracle
    coord = s.coord_T()
    worst = np.zeros(5)
    for p0 in range(0, s.point_num, chunk):
        sl = slice(p0, min(s.point_num, p0 + chunk))
        ref = oracle.mo_vgl_from_ao(s.mo_coef, oracle.ao_vgl(s, np.ascontiguousarray(coord[:, sl])))
        rm = np.abs(ref).max(axis=2, keepdims=True)
        e = (np.abs(mo[sl] - ref) / np.where(rm > 0, rm, 1.0)).max(axis=(0, 2))
        worst = np.maximum(worst, e)
    return {"points_checked": int(s.point_num), "value": float(worst[0]), "gradient": float(worst[1:4].max()),
            "laplacian": float(worst[4])}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(s, target_s=10.0):
    """FP64 oracle (OpenMP over points) on a walker subset of the workload,
    built for this host (-O3 -march=native) into a temporary directory; the
    portable x86-64-v3 build if that compile fails."""
    import subprocess
    import tempfile
    from oracle import oracle
    build = "gcc -O3 -march=native -fopenmp"
    native = os.path.join(tempfile.mkdtemp(prefix="qmckl_oracle_"), "libqmckl_oracle_native.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_OUT={native}"],
                       check=True, timeout=180, stdout=sys.stderr, stderr=sys.stderr)
        oracle.use_library(native)
    except (OSError, subprocess.SubprocessError) as e:
        log("native oracle build failed, timing the x86-64-v3 build:", e)
        build = "gcc -O3 -march=x86-64-v3 -fopenmp"
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
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
    # a walker subset of about a tenth of target_s per evaluation (at most
    # all walkers), evaluated repeatedly until target_s has passed; the rate
    # is the median evaluation's
    w = int(max(1, min(s.walk_num, round(0.1 * target_s / max(t1, 1e-6)))))
    times = []
    while not times or (sum(times) < target_s and len(times) < 50):
        times.append(run(w))
    t = float(np.median(times))
    pts = w * E
    return {"value": pts / t, "unit": "electron-points/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "build": build, "threads": oracle.num_threads(),
            "sample": f"{w} of {s.walk_num} walkers ({pts} points), FP64 oracle "
                      f"(ao_vgl then per-point mo_vgl loop), median of {len(times)} evaluations "
                      f"({t:.2f} s each, {sum(times):.1f} s in all)"}


def copy_calibration(mib):
    """A plain copy kernel of known size: reads and writes mib MiB each."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the kernels reach their steady rate after a few launches (on MI355X the
    # FP16X2 contraction's first launches of a run take 2.7, 2.5, 2.4 ms,
    # then 2.31-2.33 ms), so the untimed warm-up is ten steps
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="water64")
    ap.add_argument("--walkers", type=int, default=0)
    ap.add_argument("--form", default="fp16x2", choices=list(FORM_IDS))
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64 plain comparison run")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the other mixed precision forms at the same size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 48-molecule ragged batch measurement")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the 48-molecule ragged batch measurement")
    ap.add_argument("--batch-eager", action="store_true",
                    help="with --batch-only: only the eager batch calls (no per-molecule calls, graph or "
                         "autotuner), for counter runs")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the adjoint (no forward pass) at the differentiation workload")
    ap.add_argument("--traffic", action="store_true",
                    help="measure HBM bytes per launch of each kernel with rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE runs (calibrated on a plain copy) and write them under profiles/")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference "
                         "kernel for scaling HBM traffic counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.traffic:
        traffic_profile(args)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint_c5": bench_adjoint(args.steps, args.warmup, forward=False)}), flush=True)
        return
    if args.batch_only:
        forms = [args.form] if args.no_plain else ["plain", args.form]
        print(json.dumps({"batch_48_ragged": bench_batch(forms, args.steps, args.warmup,
                                                         eager_only=args.batch_eager)}), flush=True)
        return

    from qmckl_sycl_amd.context import Context
    s = make_system(args.config, args.walkers)
    log("workload", s.summary())
    ctx = Context(0)
    ctx.set_system(s)
    results, mos = {}, {}
    head_mo = None  # the headline form's mo_vgl at every point
    idx = sample_points(s)
    forms = [args.form] if (args.no_plain or args.form == "plain") else ["plain", args.form]
    extra = [] if args.no_forms else [f for f in FORM_IDS if f not in forms]
    for form in forms + extra:
        steps = args.steps if form in forms else max(3, args.steps // 2)
        sec, kernels, mo = run_form(ctx, s, form, steps, args.warmup)
        results[form] = {"ms_per_step": sec * 1e3, "value": s.point_num / sec, "kernels": kernels,
                         "steps": steps}
        mos[form] = mo[idx]
        if form == args.form:
            head_mo = mo
        del mo
        log(form, json.dumps(results[form]))
    ctx.destroy()
    errs = errors_vs_oracle(s, mos, idx)
    head = results[args.form]
    out = {
        "metric": METRIC, "value": head["value"], "unit": "electron-points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.form == "plain" else args.form,
        "data": "synthetic",
        "config": {"workload": f"{s.name}: {s.nucl_num} nuclei, {s.ao_num} AOs, {s.mo_num} MOs, "
                               f"{s.walk_num} walkers x {s.elec_num} electrons",
                   "points": s.point_num, "form": args.form},
        "roofline": roofline(head["kernels"], s),
        "kernels": head["kernels"],
        "mo_error_vs_fp64_oracle": {k: v for k, v in errs.items() if k not in extra},
        "mo_error_all_points": errors_all_points(s, head_mo),
    }
    del head_mo
    if "plain" in results and args.form != "plain":
        out["plain_fp64"] = {"value": results["plain"]["value"],
                             "ms_per_step": results["plain"]["ms_per_step"],
                             "kernels": results["plain"]["kernels"],
                             "roofline": roofline(results["plain"]["kernels"], s, basis="issued")}
        out["speedup_vs_plain_fp64"] = head["value"] / results["plain"]["value"]
    # every precision form at the same size, each checked on the same points
    out["forms"] = {}
    for form in [f for f in forms + extra if f != "plain"]:
        r = results[form]
        rec = {"value": r["value"], "ms_per_step": r["ms_per_step"], "steps": r["steps"],
               "mo_error_vs_fp64_oracle": errs[form], "roofline": roofline(r["kernels"], s),
               "kernels": r["kernels"]}
        if "plain" in results:
            rec["speedup_vs_plain_fp64"] = r["value"] / results["plain"]["value"]
        out["forms"][form] = rec
    if not args.no_batch:
        out["batch_48_ragged"] = bench_batch(["plain", args.form], max(5, args.steps), args.warmup)
        out["adjoint_c5"] = bench_adjoint(args.steps, args.warmup)
        out["value_only_fp64"] = bench_value_only(args.steps, args.warmup)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
