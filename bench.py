#!/usr/bin/env python3
"""Benchmark of the masked 3D diffusion on one MI355X.

A step is one full diffusion of the workload (all repetitions) with the mask
and both state buffers already resident in HBM.  Default workload: 1024^3
mask (cylinder r=300 + 2000 spheres r~U[5,40]), sigma = 10 (R = 40, 81 taps),
40 repetitions; the result is written as uint16.

Prints ONE JSON line.  `value` is Gvoxel*iterations/s over the K timed steps,
which run with no per-launch instrumentation.  The default form is fp32_split,
the FP32-class form (FP16 hi + lo state, three FP16 MFMA products per tap,
FP32 accumulation: within the FP32 bar of the reference path); the run also
times the plain FP32 form (`fp32`) and the BF16 form (`bf16`), with
`bf16_over_fp32` against the plain form.  `roofline` describes the
dominant kernel: its average launch duration comes from a second timed region
of K steps in which HIP events are recorded on the launch stream around every
pass launch.  `error_vs_fp64` gives each form's error against the FP64
oracle on a crop of the workload (and the FP32 forms' against the FP32
reference path).  `cpu_baseline` times the
reference's own CPU module (built from its sources into oracle/_ref) on a
bounded crop of the same workload.
"""
import argparse
import dataclasses
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_MFMA = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0}  # dense TFLOP/s
PEAK_VALU_FP32 = 78.6  # non-packed FP32 FMA, TFLOP/s
PRECISION_NOTE = {
    "fp32_split": "FP32-class: state held as FP16 hi + lo (u = hi + lo), each tap product as "
                  "hi*k_hi + hi*k_lo + lo*k_hi on FP16 MFMAs with FP32 accumulation; z-pass plus "
                  "fused x->y pass; error against the FP32 reference path and FP64 in error_vs_fp64",
    "fp32": "plain FP32: FP32 state, FP32 MFMA chain in tap order (equal to one FMA per tap); "
            "z, y, x passes",
    "bf16": "BF16 state, BF16 MFMA passes (z-pass plus fused x->y), last repetition in the FP32-class split form",
}


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="c2", choices=["c1", "c2", "c4", "batch", "fields", "edt"])
    p.add_argument("--precision", default="fp32_split",
                   choices=["fp32", "bf16", "fp16", "fp32_fma", "fp32_exact", "fp32_split"])
    p.add_argument("--reps", type=int, default=None, help="override the workload's repetitions")
    p.add_argument("--sigma", type=float, default=None,
                   help="override the workload's sigma (same mask; R = round(4 sigma))")
    p.add_argument("--shape", type=int, nargs=3, default=None, metavar=("NZ", "NY", "NX"),
                   help="override the workload's volume shape (same mask construction)")
    p.add_argument("--no-e2e", action="store_true",
                   help="skip the end-to-end timing of the drop-in call on host buffers")
    p.add_argument("--no-companion", action="store_true",
                   help="FP32 runs: skip the BF16 companion timing and the FP64 error report")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="target CPU-baseline duration (0 disables it)")
    p.add_argument("--chunked", type=int, default=0, metavar="Z",
                   help="time the host-resident chunked driver (mxb_diffusion_chunked) with z-chunks of Z "
                        "layers instead of the in-HBM path (PCIe-inclusive; a separate record)")
    p.add_argument("--temporal", type=int, default=1,
                   help="repetitions per chunk visit of --chunked (halo = temporal x radius)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run only a plain elementwise copy of 1 GiB (read) + 1 GiB (write), "
                        "the reference kernel for scaling HBM byte counters")
    p.add_argument("--ceilings", action="store_true",
                   help="measure the achievable ceilings instead (HBM read / write / copy of 4 GiB and "
                        "FP32 and BF16 MFMA throughput, with the library's probe kernels) as one JSON line")
    return p.parse_args()


def ceilings():
    """The achievable rates the kernels' rooflines are read against: HBM pure
    read, pure write and copy (plain and non-temporal) of 4 GiB buffers, and
    dense MFMA throughput, all with the library's probe kernels."""
    import ctypes

    from maxibone_amd import capi, hip
    L = capi.lib()
    nbytes = 4 << 30
    src, dst = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
    src.fill_bytes(0)
    dst.fill_bytes(0)
    stream = hip.Stream()
    rec = {"ceilings": "MI355X achievable rates", "buffer_bytes": nbytes, "hbm_TBps": {}, "mfma_TFLOPs": {}}
    ms = ctypes.c_float()
    for name, kind, moved in (("read", 0, 1), ("write", 1, 1), ("copy", 2, 2), ("copy_nontemporal", 3, 2)):
        capi.check(L.mxb_probe_hbm(kind, src.ptr, dst.ptr, nbytes, 10, stream.handle, ctypes.byref(ms)))
        rec["hbm_TBps"][name] = round(moved * nbytes / (ms.value * 1e-3) / 1e12, 3)
    flops = ctypes.c_double()
    for name, kind, iters in (("fp32_16x16x4", 0, 4096), ("bf16_16x16x32", 1, 4096)):
        capi.check(L.mxb_probe_mfma(kind, iters, dst.ptr, 5, stream.handle, ctypes.byref(ms), ctypes.byref(flops)))
        rec["mfma_TFLOPs"][name] = round(flops.value / (ms.value * 1e-3) / 1e12, 1)
    src.free()
    dst.free()
    rec["peaks"] = {"hbm_TBps": 8.0, "fp32_mfma_TFLOPs": 157.3, "bf16_mfma_TFLOPs": 2500.0}
    print(json.dumps(rec), flush=True)


def calibrate_copy():
    """One elementwise kernel that reads 1 GiB and writes 1 GiB (float32)."""
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "elementwise copy", "read_bytes": 4 * n, "write_bytes": 4 * n}))


def algorithmic(workload, precision):
    """Bytes and useful flops per voxel for one launch of each kernel class:
    every pass reads and writes the state once; the illuminating pass (x, or
    the fused x -> y pass) also reads the mask, packed to one bit per voxel
    once per diffusion."""
    sb = 2 if precision in ("bf16", "fp16") else 4
    K = 2 * workload.radius + 1
    f = 2 * K * (3 if precision == "fp32_split" else 1)  # the split form: three FP16 products per tap
    # the fused x -> y pass moves one pass's bytes for two passes' products
    return {"zpass": (2 * sb, f), "ypass": (2 * sb, f), "xpass": (2 * sb + 0.125, f),
            "xypass": (2 * sb + 0.125, 2 * f)}


def cpu_baseline(workload, target_s):
    """Time the reference CPU module on a central crop of the workload mask."""
    import oracle
    ref = oracle.ref_module()
    kind = "reference"
    if ref is None:
        kind = "port"
    mask = workload.mask()
    c = [min(s, 256) for s in mask.shape]
    o = [(s - cc) // 2 for s, cc in zip(mask.shape, c)]
    crop = np.ascontiguousarray(mask[o[0]:o[0] + c[0], o[1]:o[1] + c[1], o[2]:o[2] + c[2]])
    k = workload.kernel().astype(np.float32)
    out = np.empty(crop.shape, np.uint16)

    def run(reps):
        t = time.perf_counter()
        if ref is not None:
            ref.diffusion(crop, k, out, reps)
        else:
            oracle.diffusion_f32(crop, k, reps)
        return time.perf_counter() - t

    t1 = run(1)
    reps = int(max(1, min(workload.reps, round(target_s / max(t1, 1e-3)))))
    t = run(reps)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": crop.size * reps / t / 1e9, "unit": "Gvoxel*iter/s", "cores": cores,
            "kind": kind,
            "sample": f"central {c[0]}x{c[1]}x{c[2]} crop of the workload mask, "
                      f"{len(k)} taps, {reps} repetitions, {t:.1f} s"}


def bench_batch(args):
    """64 sub-volumes (dims U[64,384], 5-50 spheres, sigma~U[2,10]), 20 repetitions:
    the captured persistent graph vs per-volume launches, same GPU, same buffers."""
    from maxibone_amd import capi, hip
    from maxibone_amd.synth import batched_volumes

    reps = 20 if args.reps is None else args.reps
    vols = batched_volumes()
    dm = [hip.DeviceBuffer.from_array(v.mask) for v in vols]
    oq = [hip.DeviceBuffer(2 * v.mask.size) for v in vols]
    kernels = [v.kernel().astype(np.float32) for v in vols]
    hip.synchronize()
    t0 = time.perf_counter()
    plan = capi.BatchPlan([b.ptr for b in dm], [v.mask.shape for v in vols], kernels, reps,
                          outs_u16=[b.ptr for b in oq], precision=args.precision)
    plan.sync()
    t_plan = time.perf_counter() - t0
    info = plan.info()
    vox_iter = float(sum(v.mask.size for v in vols)) * reps

    def timed(mode):
        for _ in range(args.warmup):
            plan.run(mode)
        plan.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            plan.run(mode)
        plan.sync()
        return (time.perf_counter() - t) / args.steps

    t_graph = timed(0) if info["persistent"] else None
    t_single = timed(1)
    t_main = t_graph if t_graph is not None else t_single
    rec = {
        "metric": "Gvoxel*iter/s (masked 3D diffusion, batched)",
        "value": round(vox_iter / t_main / 1e9, 3),
        "unit": "Gvoxel*iter/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_main * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision.startswith("fp32") else args.precision,
        "data": "synthetic",
        "config": {"workload": "64 sub-volumes, dims U[64,384], 5-50 spheres r~U[3,20], "
                               "sigma~U[2,10], radius round(4 sigma)",
                   "voxels": int(sum(v.mask.size for v in vols)), "repetitions": reps,
                   "precision": args.precision, "output": "uint16"},
        "persistent_graph_ms": None if t_graph is None else round(t_graph * 1e3, 3),
        "per_volume_launch_ms": round(t_single * 1e3, 3),
        "speedup_vs_per_volume": None if t_graph is None else round(t_single / t_graph, 3),
        "autotune": info, "plan_seconds": round(t_plan, 3),
    }
    plan.close()
    print(json.dumps(rec), flush=True)


def synth_ct(shape, seed):
    """uint16 CT-like volumes for the field consumers, generated slice by slice:
    voxels = material classes by radius (implant, bone, soft tissue, air) with
    noise; gauss = a smooth field decaying with the distance from the axis;
    edt = the distance from the implant surface."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(ny) - ny / 2, np.arange(nx) - nx / 2, indexing="ij")
    r = np.hypot(yy, xx) / (min(ny, nx) / 2)
    base = np.select([r < 0.15, r < 0.45, r < 0.9], [52000.0, 24000.0, 9000.0], 1500.0)
    gauss2 = np.clip(65535.0 * np.exp(-np.maximum(r - 0.15, 0) * 6), 0, 65535)
    edt2 = np.clip(np.abs(r - 0.15) * 40000.0, 0, 65535)
    vox = np.empty(shape, np.uint16)
    gauss = np.empty(shape, np.uint16)
    edt = np.empty(shape, np.uint16)
    for z in range(nz):
        t = 1.0 + 0.1 * np.sin(z / 37.0)
        vox[z] = np.clip(base + rng.normal(0.0, 700.0, (ny, nx)), 0, 65535).astype(np.uint16)
        gauss[z] = np.clip(gauss2 * t, 0, 65535).astype(np.uint16)
        edt[z] = edt2.astype(np.uint16)
    return vox, gauss, edt


def bench_fields(args):
    """The consumers of the diffusion field on 1024^3 uint16 volumes resident in
    HBM: the 0975 combination (gauss, edt -> uint16), the field histogram of the
    voxels against the half-resolution Gaussian field (4096 x 2048 bins) and
    the axis histograms (4096 bins).  One step = the three operations once."""
    from maxibone_amd import capi, hip
    from maxibone_amd.helpers import generate_cylinder_mask

    L = capi.lib()
    shape = (1024, 1024, 1024)
    nz, ny, nx = shape
    n = nz * ny * nx
    vox, gauss, edt = synth_ct(shape, 2024)
    fld = np.ascontiguousarray(gauss[::2, ::2, ::2])
    cyl = generate_cylinder_mask(ny, nx).astype(np.uint8)
    vb, fbins, vrange, frange = 4096, 2048, (1.0, 40000.0), (1.0, 65535.0)
    d_vox, d_g, d_e, d_f, d_cyl = (hip.DeviceBuffer.from_array(a) for a in (vox, gauss, edt, fld, cyl))
    d_out = hip.DeviceBuffer(2 * n)
    d_bins = hip.DeviceBuffer(8 * fbins * vb)
    d_ab = [hip.DeviceBuffer(8 * m * vb) for m in (nx, ny, nz, 725)]
    ws_c = hip.DeviceBuffer(L.mxb_combine_workspace_bytes())
    ws_h = hip.DeviceBuffer(L.mxb_histogram_workspace_bytes(*shape))
    for b in [d_bins] + d_ab:
        b.fill_bytes(0)
    st = hip.Stream()
    ops = {
        "combine": lambda: L.mxb_device_combine_fields(d_g.ptr, d_e.ptr, *shape, d_cyl.ptr, d_out.ptr, ws_c.ptr,
                                                       st.handle),
        "field_histogram": lambda: L.mxb_device_field_histogram(d_vox.ptr, *shape, d_f.ptr, *fld.shape,
                                                                d_bins.ptr, fbins, vb, *vrange, *frange,
                                                                ws_h.ptr, st.handle),
        "axis_histograms": lambda: L.mxb_device_axis_histograms(d_vox.ptr, *shape, 0, *(b.ptr for b in d_ab), nz,
                                                                725, vb, ny // 2, nx // 2, *vrange, ws_h.ptr,
                                                                st.handle),
    }
    # algorithmic HBM bytes per voxel: combine reads two uint16 fields and
    # writes one; the field histogram reads the voxel and 1/8 of a field value
    # per voxel; the axis histograms read the voxel
    alg = {"combine": 6.0, "field_histogram": 2.25, "axis_histograms": 2.0}
    for name, f in ops.items():
        for _ in range(args.warmup):
            capi.check(f())
    st.synchronize()
    per = {}
    for name, f in ops.items():
        e0, e1 = hip.Event(), hip.Event()
        e0.record(st)
        for _ in range(args.steps):
            capi.check(f())
        e1.record(st)
        ms = e0.elapsed_ms(e1) / args.steps
        per[name] = {"ms": round(ms, 4), "Gvoxel_per_s": round(n / ms / 1e6, 2),
                     "algorithmic_GBps": round(alg[name] * n / ms / 1e6, 1)}
    total_ms = sum(v["ms"] for v in per.values())
    dom = max(per, key=lambda k: per[k]["ms"])
    rec = {
        "metric": "Gvoxel/s (field consumers: combine + field histogram + axis histograms)",
        "value": round(n / total_ms / 1e6, 3),
        "unit": "Gvoxel/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": "1024^3 uint16 voxels, Gaussian and EDT fields; field histogram against the "
                               "512^3 field, 4096 voxel x 2048 field bins; axis histograms 4096 bins",
                   "shape": list(shape)},
        "per_op": per,
        "roofline": {"bound": "hbm", "achieved": per[dom]["algorithmic_GBps"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(per[dom]["algorithmic_GBps"] / PEAK_HBM_GBS, 4),
                     "traffic": None, "kernel": dom},
    }
    for b in (d_vox, d_g, d_e, d_f, d_cyl, d_out, d_bins, ws_c, ws_h, *d_ab):
        b.free()
    if args.cpu_seconds > 0:
        rec["cpu_baseline"] = fields_cpu_baseline(vox, gauss, edt, fld, cyl, vb, fbins, vrange, frange)
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


def fields_cpu_baseline(vox, gauss, edt, fld, cyl, vb, fbins, vrange, frange):
    """The reference's CPU forms on a 64-layer slab of the same volumes: the
    histograms module compiled from its sources (OpenMP, all host cores) and
    the numpy combination of 0975 with the compiled to_int."""
    import oracle
    H = oracle.ref_module("histograms")
    G = oracle.ref_module("general")
    if H is None or G is None:
        return None
    slab = 64
    v, g, e, f = vox[:slab], gauss[:slab], edt[:slab], fld[:slab // 2]
    n = v.size
    t = {}
    t0 = time.perf_counter()
    c = g.astype(np.float32) / g.max()
    c -= e.astype(np.float32) / e.max()
    c += np.abs(c.min())
    c /= c.max()
    c *= cyl.astype(bool)
    q = np.empty(c.shape, np.uint16)
    G.normalized_convert(c, q)
    q *= cyl.astype(bool)
    t["combine"] = time.perf_counter() - t0
    bins = np.zeros((fbins, vb), np.uint64)
    t0 = time.perf_counter()
    H.field_histogram(v, f, (0, 0, 0), bins, vrange, frange, 0)
    t["field_histogram"] = time.perf_counter() - t0
    ab = [np.zeros((m, vb), np.uint64) for m in (v.shape[2], v.shape[1], v.shape[0], 725)]
    t0 = time.perf_counter()
    H.axis_histograms(v, (0, 0, 0), *ab, (v.shape[1] // 2, v.shape[2] // 2), vrange, 0)
    t["axis_histograms"] = time.perf_counter() - t0
    total = sum(t.values())
    return {"value": round(n / total / 1e9, 4), "unit": "Gvoxel/s", "cores": os.cpu_count(),
            "kind": "reference",
            "sample": f"64 x 1024 x 1024 slab: combine {t['combine']:.2f} s (numpy + compiled to_int), "
                      f"field histogram {t['field_histogram']:.2f} s, axis histograms "
                      f"{t['axis_histograms']:.2f} s (compiled CPU histograms module)"}


def synth_implant(shape):
    """uint8 implant_solid mask: a threaded rod along z (radius 0.12 of the
    slice, modulated by a helical thread) from a quarter of the depth on, and a
    sparse scatter of bright specks, generated slice by slice."""
    nz, ny, nx = shape
    rng = np.random.default_rng(77)
    yy, xx = np.meshgrid(np.arange(ny) - ny / 2, np.arange(nx) - nx / 2, indexing="ij")
    r = np.hypot(yy, xx) / (min(ny, nx) / 2)
    phi = np.arctan2(yy, xx)
    m = np.zeros(shape, np.uint8)
    for z in range(nz // 4, nz):
        thread = 0.015 * np.sin(phi + 2 * np.pi * z / 48.0)
        m[z] = r < 0.12 + thread
    flat = m.reshape(-1)
    flat[rng.integers(0, flat.size, flat.size // 200000)] = 1
    return m


def bench_edt(args):
    """The EDT field of 0950 (EDT of the negated implant mask, cylinder mask,
    uint16 conversion) on a 1024^3 implant mask resident in HBM; one step =
    one mxb_device_edt_field call."""
    from maxibone_amd import capi, hip
    from maxibone_amd.helpers import generate_cylinder_mask

    L = capi.lib()
    shape = (1024, 1024, 1024)
    nz, ny, nx = shape
    n = nz * ny * nx
    implant = synth_implant(shape)
    cyl = generate_cylinder_mask(ny, nx).astype(np.uint8)
    d_m, d_c = hip.DeviceBuffer.from_array(implant), hip.DeviceBuffer.from_array(cyl)
    d_out = hip.DeviceBuffer(2 * n)
    ws = hip.DeviceBuffer(L.mxb_edt_workspace_bytes(*shape))
    st = hip.Stream()

    def run():
        capi.check(L.mxb_device_edt_field(d_m.ptr, *shape, d_c.ptr, d_out.ptr, ws.ptr, st.handle))

    for _ in range(args.warmup):
        run()
    st.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(st)
    for _ in range(args.steps):
        run()
    e1.record(st)
    ms = e0.elapsed_ms(e1) / args.steps
    # algorithmic HBM bytes per voxel: the mask (1), three passes of uint32
    # distances (pass x writes 4, y reads and writes 8, z reads 4 and writes the
    # float field 4), the uint16 conversion (reads 4, writes 2)
    alg_bytes = 27.0
    gbps = alg_bytes * n / ms / 1e6
    rec = {
        "metric": "Gvoxel/s (EDT field: distance transform + cylinder mask + uint16)",
        "value": round(n / ms / 1e6, 3),
        "unit": "Gvoxel/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "1024^3 implant mask (threaded rod along z from z=256, radius 0.12 of "
                               "the slice, scattered specks), 0950 EDT field",
                   "shape": list(shape), "implant_voxels": int(implant.sum(dtype=np.int64))},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbps / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "k_edt_pass (3 passes) + k_edt_to_u16"},
    }
    for b in (d_m, d_c, d_out, ws):
        b.free()
    rec["cpu_baseline"] = edt_cpu_baseline() if args.cpu_seconds > 0 else None
    print(json.dumps(rec), flush=True)


def edt_cpu_baseline():
    """The oracle's exact EDT (one host thread: the 1D passes of the `edt`
    package's algorithm in C) on a 256^3 mask of the same construction."""
    import oracle.edt as OE
    m = synth_implant((256, 256, 256))
    t0 = time.perf_counter()
    OE.edt_sq(m == 0)
    t = time.perf_counter() - t0
    return {"value": round(m.size / t / 1e9, 4), "unit": "Gvoxel/s", "cores": 1, "kind": "port",
            "sample": f"256^3 implant mask of the same construction, squared EDT in {t:.2f} s"}


def bench_chunked(args, w, reps):
    """The host-resident chunked driver (the reference's on-disk loop, with the
    state in host memory): each repetition block streams every z-chunk, its
    halo and the mask over PCIe, diffuses it on the GPU and copies the
    interior back.  One step = one full diffusion of `reps` repetitions."""
    from maxibone_amd import capi
    mask = w.mask()
    k = w.kernel().astype(np.float32)
    n = mask.size
    for _ in range(args.warmup):
        capi.diffusion_chunked(mask, k, reps, args.chunked, args.temporal, precision=args.precision)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        capi.diffusion_chunked(mask, k, reps, args.chunked, args.temporal, precision=args.precision)
    t = (time.perf_counter() - t0) / args.steps
    rec = {
        "metric": "Gvoxel*iter/s (masked 3D diffusion, host-resident chunked driver)",
        "value": round(float(n) * reps / t / 1e9, 3), "unit": "Gvoxel*iter/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision.startswith("fp32") else args.precision, "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "shape": list(w.shape), "sigma": w.sigma,
                   "radius": w.radius, "repetitions": reps, "path": "chunked, host-resident state",
                   "chunk_z": args.chunked, "temporal": args.temporal,
                   "halo": args.temporal * w.radius, "precision": args.precision, "output": "uint16"},
        "note": "PCIe-inclusive: mask, state and outputs live in host memory; the chunked driver keeps "
                + ("16-bit" if args.precision in ("bf16", "fp16") else "FP32") + " state between chunk visits",
    }
    print(json.dumps(rec), flush=True)


def load_traffic(workload_name, precision, kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path)).get(f"{workload_name}/{precision}/{kernel}")
        return None if d is None else d["bytes_per_launch"]
    except (ValueError, OSError):
        return None


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.ceilings:
        ceilings()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "fields":
        bench_fields(args)
        return
    if args.workload == "edt":
        bench_edt(args)
        return
    from maxibone_amd.synth import WORKLOADS

    w = WORKLOADS[args.workload]
    if args.sigma is not None:
        w = dataclasses.replace(w, sigma=args.sigma,
                                description=f"{w.description}; sigma overridden to {args.sigma:g} "
                                            f"(R={round(4 * args.sigma)})")
    if args.shape is not None:
        w = dataclasses.replace(w, shape=tuple(args.shape),
                                description=f"{w.description}; shape overridden to "
                                            f"{'x'.join(str(v) for v in args.shape)}")
    reps = w.reps if args.reps is None else args.reps
    if args.chunked > 0:
        bench_chunked(args, w, reps)
        return
    mask_h = w.mask()
    k = w.kernel().astype(np.float32)
    main_run = device_run(args, w, mask_h, k, reps, args.precision)
    rec = {
        "metric": "Gvoxel*iter/s (masked 3D diffusion)",
        "value": main_run["value"],
        "unit": "Gvoxel*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision.startswith("fp32") else args.precision,
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "shape": list(w.shape),
                   "sigma": w.sigma, "radius": w.radius, "repetitions": reps,
                   "precision": args.precision, "output": "uint16"},
        "precision_note": PRECISION_NOTE.get(args.precision, ""),
        "hbm_fraction_fused_accounting": main_run["hbm_fraction_fused_accounting"],
        "per_kernel": main_run["per_kernel"],
        "roofline": main_run["roofline"],
        "stream_ms_per_step": main_run["stream_ms_per_step"],
        "instrumented_ms_per_step": main_run["instrumented_ms_per_step"],
    }
    if args.precision in ("fp32", "fp32_split") and not args.no_companion:
        # the other FP32 form and the BF16 precision form on the same workload
        keys = ("value", "ms_per_step", "hbm_fraction_fused_accounting", "per_kernel", "roofline")
        other = "fp32" if args.precision == "fp32_split" else "fp32_split"
        oth = device_run(args, w, mask_h, k, reps, other)
        rec[other] = {x: oth[x] for x in keys}
        lo = device_run(args, w, mask_h, k, reps, "bf16")
        rec["bf16"] = {x: lo[x] for x in keys}
        plain = main_run if args.precision == "fp32" else oth
        # the BF16 target is set against the plain FP32 form (FP32 MFMA chain)
        rec["bf16_over_fp32"] = round(lo["value"] / plain["value"], 3)
        rec["bf16_over_fp32_split"] = round(lo["value"] / (oth if args.precision == "fp32" else main_run)["value"], 3)
    if not args.no_e2e:
        rec["end_to_end"] = end_to_end(w, mask_h, k, reps, args.precision)
    if args.cpu_seconds > 0:
        rec["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
        if not args.no_companion:
            rec["error_vs_fp64"] = error_report(w, mask_h, k, reps)
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


def device_run(args, w, mask_h, k, reps, precision):
    """Time `steps` full diffusions of the workload resident in HBM (after
    `warmup` untimed ones), then, in a separate run of the same steps, the
    per-kernel durations from HIP events on the launch stream."""
    from maxibone_amd import capi, hip
    L = capi.lib()
    nz, ny, nx = mask_h.shape
    n = mask_h.size
    prec = capi.PRECISION[precision]
    sbytes = L.mxb_state_bytes(prec, nz, ny, nx)
    mask_d = hip.DeviceBuffer.from_array(mask_h)
    s0 = hip.DeviceBuffer(sbytes)
    s1 = hip.DeviceBuffer(sbytes)
    out = hip.DeviceBuffer(2 * n)
    stream = hip.Stream()

    def step():
        capi.check(L.mxb_device_diffusion(mask_d.ptr, nz, ny, nx, k.ctypes.data, len(k), reps,
                                          prec, s0.ptr, s1.ptr, out.ptr, None, stream.handle))

    for _ in range(args.warmup):
        step()
    hip.synchronize()
    # the headline: no per-launch instrumentation inside the timed region
    ev0, ev1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    hip.synchronize()
    elapsed = time.perf_counter() - t0
    stream_ms = ev0.elapsed_ms(ev1)
    # per-kernel durations: one instrumented step fills the event pool, then
    # the same number of steps is timed launch by launch
    capi.set_kernel_timing(True)
    step()
    hip.synchronize()
    capi.set_kernel_timing(True)  # resets the totals, keeps the events
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    hip.synchronize()
    instrumented_s = time.perf_counter() - t0
    times = {name: capi.kernel_times(c) for name, c in
             (("zpass", capi.K_ZPASS), ("ypass", capi.K_YPASS), ("xpass", capi.K_XPASS),
              ("xypass", capi.K_XYPASS), ("other", capi.K_OTHER))}
    capi.set_kernel_timing(False)
    for b in (mask_d, s0, s1, out):
        b.free()

    vox_iter = float(n) * reps * args.steps
    value = vox_iter / elapsed / 1e9
    alg = algorithmic(w, precision)
    per_kernel = {}
    for name in ("zpass", "ypass", "xpass", "xypass"):
        ms, cnt = times[name]
        if cnt == 0:
            continue
        avg_s = ms / cnt / 1e3
        b, f = alg[name]
        per_kernel[name] = {"avg_ms": ms / cnt, "launches": cnt,
                            "GBps": b * n / avg_s / 1e9, "TFLOPs": f * n / avg_s / 1e12}
    # the dominant kernel: the most time over the step
    dom = max(per_kernel, key=lambda kk: per_kernel[kk]["avg_ms"] * per_kernel[kk]["launches"])
    d = per_kernel[dom]
    mfma_peak = PEAK_MFMA["fp16"] if precision == "fp32_split" else PEAK_MFMA.get(precision.split("_")[0],
                                                                                    PEAK_MFMA["fp32"])
    compute_peak = PEAK_VALU_FP32 if precision in ("fp32_fma", "fp32_exact") else mfma_peak
    frac_hbm = d["GBps"] / PEAK_HBM_GBS
    frac_cmp = d["TFLOPs"] / compute_peak
    # the committed counter readings are for the workload's own shape and sigma
    traffic = None if args.shape or args.sigma else load_traffic(args.workload, precision, dom)
    if frac_cmp >= frac_hbm:
        roof = {"bound": "mfma", "achieved": round(d["TFLOPs"], 3), "peak": compute_peak,
                "unit": "TFLOP/s", "frac": round(frac_cmp, 4), "traffic": traffic}
    else:
        roof = {"bound": "hbm", "achieved": round(d["GBps"], 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(frac_hbm, 4), "traffic": traffic}
    roof["kernel"] = dom
    roof["avg_launch_ms"] = round(d["avg_ms"], 4)
    fused_bytes = (2 if precision in ("bf16", "fp16") else 4) * 2 + 0.125
    return {
        "value": round(value, 3),
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "stream_ms_per_step": round(stream_ms / args.steps, 3),
        "instrumented_ms_per_step": round(instrumented_s / args.steps * 1e3, 3),
        "hbm_fraction_fused_accounting": round(value * fused_bytes / PEAK_HBM_GBS, 4),
        "per_kernel": {kk: {x: round(v, 4) if isinstance(v, float) else v for x, v in vv.items()}
                       for kk, vv in per_kernel.items()},
        "roofline": roof,
    }


def end_to_end(w, mask_h, k, reps, precision):
    """One drop-in call on host buffers: allocation, H2D of the mask, the
    diffusion and D2H of the uint16 result (PCIe-inclusive, not `value`)."""
    from maxibone_amd import capi
    L = capi.lib()
    nz, ny, nx = mask_h.shape
    q = np.empty(mask_h.shape, np.uint16)
    t0 = time.perf_counter()
    capi.check(L.mxb_diffusion_ex(mask_h.ctypes.data, nz, ny, nx, k.ctypes.data, len(k), reps,
                                  capi.PRECISION[precision], q.ctypes.data, None, 0))
    t = time.perf_counter() - t0
    return {"value": round(float(mask_h.size) * reps / t / 1e9, 3), "unit": "Gvoxel*iter/s",
            "seconds": round(t, 3),
            "what": "one drop-in call on host buffers incl. device allocation, "
                    "H2D of the mask and D2H of the uint16 output"}


def error_report(w, mask_h, k, reps, edge=112):
    """Max-abs and RMS error of the FP32 and BF16 forms against the FP64
    oracle, all repetitions, on a crop of the workload mask centred on the
    cylinder surface (the oracle at full size would take hours)."""
    import oracle
    from maxibone_amd import capi
    nz, ny, nx = mask_h.shape
    c = min(edge, nz, ny, nx)
    oz, oy = (nz - c) // 2, (ny - c) // 2
    ox = min(max(nx // 2 + w.cyl_r - c // 2, 0), nx - c)
    m = np.ascontiguousarray(mask_h[oz:oz + c, oy:oy + c, ox:ox + c])
    t0 = time.perf_counter()
    u64 = oracle.diffusion_f64(m, k.astype(np.float64), reps)
    u32, _ = oracle.diffusion_f32(m, k, reps)
    t = time.perf_counter() - t0
    rep = {"crop": f"{c}^3 at (z, y, x) = ({oz}, {oy}, {ox}), centred on the cylinder surface, "
                   f"{reps} repetitions, {len(k)} taps", "oracle_seconds": round(t, 2)}
    for prec in ("fp32", "fp32_split", "bf16"):
        u = capi.diffusion(m, k, reps, precision=prec, u16=False, f32=True)[1]
        d = u.astype(np.float64) - u64
        rep[prec] = {"max_abs": float(f"{np.abs(d).max():.3e}"),
                     "rms": float(f"{np.sqrt(np.mean(d * d)):.3e}")}
        if prec.startswith("fp32"):
            rep[prec]["max_abs_vs_fp32_oracle"] = float(f"{np.abs(u - u32).max():.3e}")
    return rep


if __name__ == "__main__":
    main()
