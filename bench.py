#!/usr/bin/env python3
"""MTTKRP throughput on one MI355X.

Workload (a "step" = one MTTKRP of every mode over the resident CSF):
  3-mode tensor 200k x 150k x 50k, 50M distinct nonzeros with Zipf(1.2)
  indices in every mode (compacted), lognormal(0,1) values, rank 32.
  Primary line: FP32 values/factors/accumulation on the all-mode CSF
  (every mode uses the root kernel). FP64 and the one-CSF layout (root,
  internal and leaf kernels) are reported beside it.

value = algorithmic bytes of the K timed steps / wall time of those steps,
with the CSF and factors already in HBM. Algorithmic bytes per MTTKRP:
  nnz*(s_v+4) + sum_internal nfibs*4 + sum (nfibs+1)*4 + sum_m dims[m]*R*s_v
(4-byte indices, s_v = value size; every factor read once, output written
once). roofline.achieved divides the same bytes by the walker kernel's
average duration from HIP events recorded around each launch in the timed
region. The CPU baseline runs the oracle (a restatement of the reference's
OpenMP CSF MTTKRP) on a random sample of the same nonzeros.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "splatt-stream_amd"))
import pysplatt as ps  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def dev_factors(mats, dtype):
    bufs = [ps.DevBuf.from_host(m.astype(dtype)) for m in mats]
    arr = (C.c_void_p * len(bufs))(*[b.ptr for b in bufs])
    return bufs, arr


def run_sweeps(L, csf, opts, arr, outs, prec, R, nmodes, count):
    for _ in range(count):
        for m in range(nmodes):
            rc = L.splatt_amd_mttkrp_dev(m, R, csf.ptr, arr, outs[m].ptr, prec, opts.ptr, None)
            if rc != ps.SUCCESS:
                raise RuntimeError("mttkrp failed: " + ps.last_error())


def measure(L, csf, opts, mats, dims, R, prec, steps, warmup):
    dtype = np.float32 if prec == ps.F32 else np.float64
    bufs, arr = dev_factors(mats, dtype)
    outs = [ps.DevBuf(d * R * np.dtype(dtype).itemsize) for d in dims]
    N = len(dims)
    bytes_m, flops_m = [], []
    for m in range(N):
        b, f = C.c_double(), C.c_double()
        if L.splatt_amd_mttkrp_cost(m, R, csf.ptr, prec, opts.ptr, C.byref(b), C.byref(f)) != ps.SUCCESS:
            raise RuntimeError(ps.last_error())
        bytes_m.append(b.value)
        flops_m.append(f.value)
    run_sweeps(L, csf, opts, arr, outs, prec, R, N, warmup)
    L.splatt_amd_sync()
    L.splatt_amd_mttkrp_timing(1)
    L.splatt_amd_sync()
    t0 = time.perf_counter()
    run_sweeps(L, csf, opts, arr, outs, prec, R, N, steps)
    L.splatt_amd_sync()
    t1 = time.perf_counter()
    ktimes = np.zeros(steps * N)
    got = L.splatt_amd_mttkrp_kernel_times(ktimes.ctypes.data_as(ps.f64p), steps * N)
    L.splatt_amd_mttkrp_timing(0)
    ktimes = ktimes[:got].reshape(steps, N)
    wall = t1 - t0
    per_mode_ms = ktimes.mean(axis=0)
    res = dict(
        wall_s=wall,
        ms_per_step=1e3 * wall / steps,
        gbps=sum(bytes_m) * steps / wall / 1e9,
        kernel_gbps=sum(bytes_m) / (per_mode_ms.sum() * 1e-3) / 1e9,
        bytes_per_step=sum(bytes_m),
        flops_per_step=sum(flops_m),
        modes=[dict(mode=m, kernel_ms=float(per_mode_ms[m]), bytes=bytes_m[m],
                    gbps=bytes_m[m] / (per_mode_ms[m] * 1e-3) / 1e9) for m in range(N)],
    )
    outs_host = [o.to_host((d, R), dtype) for o, d in zip(outs, dims)]
    for b in bufs + outs:
        b.free()
    return res, outs_host


def cpu_baseline(inds, vals, dims, mats, R, sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    sel = slice(0, sample)   # the nonzeros are stored in shuffled order
    coo = orc.Coo.from_arrays([x[sel] for x in inds], vals[sel], dims)
    N = len(dims)
    csfs = [orc.Csf(coo, orc.mode_order(dims, 3, m)) for m in range(N)]
    m32 = [m.astype(np.float32) for m in mats]
    # one warm-up pass, then passes over all modes until about 12 s of CPU
    # work have been timed; the median pass is reported
    for m in range(N):
        csfs[m].mttkrp(m32, m, nthreads=threads, f32=True)
    times = []
    t_start = time.perf_counter()
    while len(times) < 10 or time.perf_counter() - t_start < 12.0:
        t0 = time.perf_counter()
        for m in range(N):
            csfs[m].mttkrp(m32, m, nthreads=threads, f32=True)
        times.append(time.perf_counter() - t0)
        if len(times) >= 1000:
            break
    best = float(np.median(times))
    total = float(np.sum(times))
    # algorithmic bytes of the sample, same model as the GPU figure
    nbytes = 0.0
    for m in range(N):
        a = csfs[m].arrays()["tiles"][0]
        nb = sample * 8.0
        nb += sum(a["nfibs"][l] * 4.0 for l in range(1, N - 1))
        nb += sum((a["nfibs"][l] + 1) * 4.0 for l in range(N - 1))
        nb += sum(d * R * 4.0 for d in dims)
        nbytes += nb
    return dict(value=nbytes / best / 1e9, unit="GB/s", cores=threads, kind="port",
                sample="%d of the workload's nonzeros (its first %d in shuffled order), "
                       "all-mode CSF, FP32, all %d modes; median of %d passes "
                       "(%.3f s each, %.1f s timed in total)"
                       % (sample, sample, N, len(times), best, total))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nnz", type=int, default=50_000_000)
    ap.add_argument("--rank", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--primary-only", action="store_true",
                    help="only the FP32 all-mode measurement (for profiling)")
    ap.add_argument("--cpu-sample", type=int, default=20_000_000)
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a streaming kernel of known traffic (FP64->FP32 "
                         "conversion of 2^26 values: 512 MiB read, 256 MiB written) "
                         "so hardware byte counters can be scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    L = ps.lib()
    R = args.rank
    if args.calibrate:
        n = 1 << 26
        src = ps.DevBuf(n * 8)
        dst = ps.DevBuf(n * 4)
        L.spd_f64_to_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.spd_f64_to_f32(src.ptr, dst.ptr, n, None)
        L.splatt_amd_sync()
        src.free()
        dst.free()
        log("calibration kernel done (k_f64_to_f32, %d values)" % n)
    t0 = time.perf_counter()
    inds, vals, dims = ps.synth_coord([200_000, 150_000, 50_000], args.nnz,
                                      zipf=[1.2, 1.2, 1.2], vals="lognormal",
                                      seed=args.seed)
    log("generated %d nnz dims %s in %.1fs" % (len(vals), dims, time.perf_counter() - t0))
    mats = [ps.synth_matrix(d, R, args.seed + m) for m, d in enumerate(dims)]

    t0 = time.perf_counter()
    opts_all = ps.Opts(csf_alloc=ps.CSF_ALLMODE)
    csf_all = ps.CsfSet.convert(inds, vals, opts_all)
    build_all = time.perf_counter() - t0
    log("ALLMODE CSF built (GPU sort + compression + host mirror) in %.2fs" % build_all)

    primary, out32 = measure(L, csf_all, opts_all, mats, dims, R, ps.F32, args.steps, args.warmup)
    log("ALLMODE fp32: %.1f GB/s (kernel %.1f GB/s), %.3f ms/step"
        % (primary["gbps"], primary["kernel_gbps"], primary["ms_per_step"]))
    extra = {}
    parity = {}
    if not args.primary_only:
        f64, out64 = measure(L, csf_all, opts_all, mats, dims, R, ps.F64, args.steps, args.warmup)
        log("ALLMODE fp64: %.1f GB/s" % f64["gbps"])
        extra["allmode_fp64"] = f64
        # size-independent checks at full size: <X,[[A,B,C]]> is the same
        # through every mode's MTTKRP; FP32 stays within 1e-5 of FP64
        inner = [float(np.sum(out64[m] * mats[m])) for m in range(3)]
        parity["inner_product_fp64_spread_rel"] = (max(inner) - min(inner)) / abs(np.mean(inner))
        parity["fp32_vs_fp64_rel"] = max(
            float(np.linalg.norm(out32[m] - out64[m]) / np.linalg.norm(out64[m])) for m in range(3))
        del csf_all
        t0 = time.perf_counter()
        opts_one = ps.Opts(csf_alloc=ps.CSF_ONEMODE)
        csf_one = ps.CsfSet.convert(inds, vals, opts_one)
        extra["onemode_build_s"] = time.perf_counter() - t0
        o32, _ = measure(L, csf_one, opts_one, mats, dims, R, ps.F32, args.steps, args.warmup)
        o64, o64out = measure(L, csf_one, opts_one, mats, dims, R, ps.F64, args.steps, args.warmup)
        log("ONEMODE fp32: %.1f GB/s, fp64: %.1f GB/s" % (o32["gbps"], o64["gbps"]))
        extra["onemode_fp32"] = o32
        extra["onemode_fp64"] = o64
        parity["onemode_vs_allmode_fp64_rel"] = max(
            float(np.linalg.norm(o64out[m] - out64[m]) / np.linalg.norm(out64[m])) for m in range(3))
        del csf_one

    cpu = None
    if not args.skip_cpu:
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        cpu = cpu_baseline(inds, vals, dims, mats, R, min(args.cpu_sample, len(vals)), threads)
        log("cpu baseline %.2f GB/s on %d threads (%.1fs)" % (cpu["value"], threads,
                                                              time.perf_counter() - t0))

    line = {
        "metric": "MTTKRP HBM GB/s",
        "value": round(primary["gbps"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(primary["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "MTTKRP of all 3 modes, 200000x150000x50000 tensor, %d nnz, "
                        "Zipf(1.2) indices, lognormal values, rank %d, FP32, all-mode CSF"
                        % (len(vals), R),
            "dims": dims, "nnz": len(vals), "rank": R, "csf": "ALLMODE",
            "step": "one MTTKRP per mode",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(primary["kernel_gbps"], 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(primary["kernel_gbps"] / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu,
        "modes": primary["modes"],
        "csf_build_s": round(build_all, 3),
        "extra": {k: (v if not isinstance(v, dict) else
                      {kk: vv for kk, vv in v.items() if kk != "modes"} | {"modes": v["modes"]})
                  for k, v in extra.items()},
        "parity": parity,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
