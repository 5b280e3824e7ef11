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
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "splatt-stream_amd"))
import pysplatt as ps  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per root-kernel launch from the FETCH_SIZE / WRITE_SIZE counter
# runs of this workload (scaled by a copy of known size), kept with the
# rocprofv3 summary they come from
PMC_SUMMARY = os.path.join(ROOT, "profiles", "mttkrp_root3_f32_zipf50m_pmc.md")


def measured_traffic():
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# Host threads of the CPU baselines: the box's CPU share for one GPU (16),
# whatever larger count the container reports.
CPU_SHARE = 16


def host_cpu():
    """CPU model and counts of the host the CPU baselines run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    visible = os.cpu_count() or 1
    return dict(model=model, cpu_count=visible, affinity=affinity,
                threads_used=min(CPU_SHARE, visible),
                note="threads capped at the host CPU share of one GPU (%d)" % CPU_SHARE)


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


def time_dropin(csf, mats, dims, reps=3):
    """The drop-in splatt_mttkrp with host factors and output (the caller's
    arrays: upload of the input factors, the kernel, download of the
    output), FP64, per mode: median wall time of reps calls after one
    warm-up call."""
    per = []
    for m in range(len(dims)):
        csf.mttkrp(m, mats)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            csf.mttkrp(m, mats)
            ts.append(time.perf_counter() - t0)
        per.append(1e3 * float(np.median(ts)))
    host_bytes = [8.0 * sum(d * mats[0].shape[1] for d in dims)] * len(dims)
    return dict(ms_per_mode=per, precision="f64",
                host_bytes_moved_per_call=host_bytes,
                note="includes host-to-device copies of the other modes' factors and "
                     "the device-to-host copy of the output")


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
                host=host_cpu(),
                sample="%d of the workload's nonzeros (its first %d in shuffled order), "
                       "all-mode CSF, FP32, all %d modes; median of %d passes "
                       "(%.3f s each, %.1f s timed in total)"
                       % (sample, sample, N, len(times), best, total))


def bench_cpd_c1(threads, seed=20261203):
    """CPD-ALS, rank 16, 20 iterations, FP64, 1000x800x600 with 1M uniform
    nonzeros (values U(0,1)), two-CSF allocation as splatt_cpd uses it. The
    GPU run goes through splatt_cpd; the oracle runs the same iterations
    from the same start on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    R, its = 16, 20
    inds, vals, dims = ps.synth_coord([1000, 800, 600], 1_000_000, vals="uniform", seed=seed)
    opts = ps.Opts(csf_alloc=ps.CSF_TWOMODE)
    cs = ps.CsfSet.convert(inds, vals, opts)
    ps.run_cpd(cs, R, max_its=1, tol=0.0, seed=1)      # warm-up (plans, scratch)
    t0 = time.perf_counter()
    res = ps.run_cpd(cs, R, max_its=its, tol=0.0, seed=1)
    gpu_s = time.perf_counter() - t0
    if res["rc"] != ps.SUCCESS:
        raise RuntimeError("splatt_cpd failed: " + ps.last_error())
    coo = orc.Coo.from_arrays(inds, vals, dims)
    perm0 = cs.arrays(0)["perm"]
    ocsfs = [orc.Csf(coo, perm0), orc.Csf(coo, cs.arrays(1)["perm"])]
    mode_csf = [1 if perm0[-1] == m else 0 for m in range(3)]
    t0 = time.perf_counter()
    facs, lam, fits = orc.cpd_als_con(ocsfs, mode_csf, res["init"], its, 0.0,
                                      [(0, 0, 0.0)] * 3, True, nthreads=threads)
    cpu_s = time.perf_counter() - t0
    cs.free()
    return dict(workload="CPD-ALS rank 16, 20 iterations, FP64, 1000x800x600, %d nnz "
                         "uniform, two-CSF allocation" % len(vals),
                gpu_s=gpu_s, gpu_ms_per_iteration=1e3 * gpu_s / its,
                cpu_s=cpu_s, cpu_ms_per_iteration=1e3 * cpu_s / its, cpu_threads=threads,
                cpu_kind="port", fit_gpu=res["fit"], fit_cpu=float(fits[-1]),
                fit_abs_diff=abs(res["fit"] - float(fits[-1])))


def bench_cpd_c3(nnz, R=64, its=25, seed=20261208):
    """CPD-ALS rank 64, 25 iterations, 4-mode 500000x300000x100000x2000 with
    nnz Zipf(1.1) nonzeros in the first three modes and uniform in the
    fourth, values U(-1, 1), two-CSF allocation. MTTKRP in FP32 (Grams,
    solves and the fit in FP64), then the same iterations from the same start
    with FP64 MTTKRP: the fit and every factor (scaled by lambda) of the FP32
    run are compared with the FP64 run."""
    inds, vals, dims = ps.synth_coord([500_000, 300_000, 100_000, 2_000], nnz,
                                      zipf=[1.1, 1.1, 1.1, 0.0], vals="symmetric",
                                      compact=True, seed=seed)
    opts = ps.Opts(csf_alloc=ps.CSF_TWOMODE)
    t0 = time.perf_counter()
    cs = ps.CsfSet.convert(inds, vals, opts)
    build_s = time.perf_counter() - t0
    del inds, vals
    L = ps.lib()
    out = {}
    try:
        for name, prec in (("f32", 1), ("f64", 0)):
            L.splatt_amd_set_mttkrp_precision(prec)
            one = ps.run_cpd(cs, R, max_its=1, tol=0.0, seed=1)     # warm-up
            res = ps.run_cpd(cs, R, max_its=its, tol=0.0, seed=1)
            if res["rc"] != ps.SUCCESS or one["rc"] != ps.SUCCESS:
                raise RuntimeError("splatt_cpd failed: " + ps.last_error())
            # per iteration: the splatt_cpd call with its iterations less the
            # call with one (factor upload/download and set-up cancel)
            out[name] = (res, res["cpd_s"], (res["cpd_s"] - one["cpd_s"]) / (its - 1))
    finally:
        L.splatt_amd_set_mttkrp_precision(0)
        cs.free()
    r32, s32, it32 = out["f32"]
    r64, s64, it64 = out["f64"]
    ferr = []
    for m in range(4):
        a = r32["factors"][m] * r32["lam"]
        b = r64["factors"][m] * r64["lam"]
        ferr.append(float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    return dict(workload="CPD-ALS rank %d, %d iterations, 500000x300000x100000x2000 "
                         "(compacted to %s), %d nnz Zipf(1.1) x3 + uniform, values U(-1,1), "
                         "two-CSF allocation; FP32 MTTKRP vs an all-FP64 run from the same "
                         "start" % (R, its, "x".join(str(d) for d in dims), nnz),
                csf_build_s=build_s, f32_cpd_call_s=s32, f32_ms_per_iteration=1e3 * it32,
                f64_cpd_call_s=s64, f64_ms_per_iteration=1e3 * it64,
                fit_f32=r32["fit"], fit_f64=r64["fit"],
                fit_abs_diff=abs(r32["fit"] - r64["fit"]),
                factor_rel_err_vs_f64=ferr)


def stream_c4_slices(nslices, seed=20261206):
    """The first nslices time slices of the streaming workload: 500000 x
    200000 non-stream dims, per-slice nnz lognormal (median 2e5, sigma 1)
    clipped to [1e4, 2e6], Zipf(1.2) coordinates distinct within a slice,
    values U(0,1)."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    nn = np.clip(np.round(np.exp(np.log(2e5) + rng.standard_normal(365))), 1e4, 2e6)
    nn = nn.astype(np.int64)[:nslices]

    def one(t):
        return ps.synth_coord([500_000, 200_000], int(nn[t]), zipf=[1.2, 1.2],
                              vals="uniform", compact=False, seed=1000 + t, shuffle=False)

    # slices are independent (each its own seed): generated on 8 host threads
    with ThreadPoolExecutor(8) as ex:
        parts = list(ex.map(one, range(len(nn))))
    I, J, V, K = [], [], [], []
    for t, (ij, v, _) in enumerate(parts):
        I.append(ij[0]); J.append(ij[1]); V.append(v)
        K.append(np.full(len(v), t, dtype=np.uint64))
    inds = [np.concatenate(I), np.concatenate(J), np.concatenate(K)]
    vals = np.concatenate(V)
    return inds, vals, [500_000, 200_000, nslices]


def bench_stream_c4(nslices, cpu_slices, threads=1):
    """Streaming CPD, rank 32, frob 1e-2 on the time mode, 5 outer
    iterations per slice, FP64, over the first nslices slices. The oracle's
    restatement of the reference runs the first cpu_slices slices (host
    threads on its row-parallel triangular solves and history product) and
    must report the same per-slice errors."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    R = 32
    inds, vals, dims = stream_c4_slices(nslices)
    t0 = time.perf_counter()
    res = ps.run_stream(inds, vals, dims, R, 2, forget=0.95, max_its=5, seed=4, frob=1e-2)
    wall = time.perf_counter() - t0
    if res is None:
        raise RuntimeError("stream failed: " + ps.last_error())
    bs = res["batch_seconds"]
    out = dict(workload="streaming CPD rank 32, frob 1e-2 on time, 5 outer its, FP64, "
                        "first %d slices of 500000x200000x365 (%d nnz)" % (nslices, len(vals)),
               slices=int(res["nbatches"]), slices_per_s=float(len(bs) / bs.sum()),
               p50_slice_s=float(np.percentile(bs, 50)), p99_slice_s=float(np.percentile(bs, 99)),
               wall_s=wall, final_err=res["final_err"],
               mean_local_err=float(np.mean(res["local_err"])))
    if cpu_slices > 0:
        sel = inds[2] < cpu_slices
        cinds = [i[sel] for i in inds]
        coo = orc.Coo.from_arrays(cinds, vals[sel], [dims[0], dims[1], cpu_slices])
        orc.stream_prepare(coo, 2)
        ps.srand(4)
        orc.set_threads(threads)
        t0 = time.perf_counter()
        try:
            ref = orc.stream_cpd(coo, 2, R, 0.95, 5, [(2, 1, 0.0), (2, 1, 0.0), (1, 0, 1e-2)])
        finally:
            orc.set_threads(1)
        cpu_s = time.perf_counter() - t0
        out.update(cpu_slices=cpu_slices, cpu_s=cpu_s, cpu_slices_per_s=cpu_slices / cpu_s,
                   cpu_threads=threads, cpu_kind="port", cpu_host=host_cpu(),
                   gpu_slices_per_s_same_slices=float(cpu_slices / bs[:cpu_slices].sum()),
                   local_err_max_abs_diff=float(np.max(np.abs(
                       res["local_err"][:cpu_slices] - ref["local_err"]))))
    return out


def serve_c5_streams(nstreams, per_stream, seed=20261207):
    """The serving workload: independent 3-mode streams (time last) with
    non-stream dims from U[10k, 1M] and rank from {8, 16, 32, 64}; per-slice
    nnz lognormal (median 1e5, sigma 1) clipped to [5e3, 2e6], Zipf(1.2)
    coordinates distinct within a slice, values U(0,1)."""
    rng = np.random.default_rng(seed)
    streams = []
    for k in range(nstreams):
        dims = [int(rng.integers(10_000, 1_000_001)), int(rng.integers(10_000, 1_000_001))]
        R = int(rng.choice([8, 16, 32, 64]))
        nn = np.clip(np.round(np.exp(np.log(1e5) + rng.standard_normal(per_stream))), 5e3, 2e6)
        sl = []
        for t in range(per_stream):
            ij, v, _ = ps.synth_coord(dims, int(nn[t]), zipf=[1.2, 1.2], vals="uniform",
                                      compact=False, seed=seed + 1000 * k + t, shuffle=True)
            sl.append(([ij[0], ij[1], np.full(len(v), t, dtype=np.uint64)], v))
        streams.append(dict(dims=dims, R=R, slices=sl))
    return streams


def run_server(streams, arrivals, max_inflight):
    """Submit every slice at its arrival time (seconds from start, stream
    order kept), wait, return per-slice latencies and the span."""
    sv = ps.Server(budget=0, max_inflight=max_inflight)
    ids = [sv.open(s["dims"] + [1], s["R"], 2, forget=0.95, max_its=5, frob=1e-2, seed=1 + k)
           for k, s in enumerate(streams)]
    order = sorted(arrivals)          # (time, stream, slice)
    tickets = []
    t0 = time.perf_counter()
    for at, k, j in order:
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        tickets.append(sv.submit(ids[k], *streams[k]["slices"][j]))
    for sid in ids:
        sv.close(sid)
    sv.drain()
    span = time.perf_counter() - t0
    info = [sv.slice(t) for t in tickets]
    sv.free()
    lat = np.array([i["latency"] for i in info])
    return dict(slices=len(lat), span_s=span, slices_per_s=len(lat) / span,
                p50_latency_s=float(np.percentile(lat, 50)),
                p99_latency_s=float(np.percentile(lat, 99)),
                mean_gpu_s=float(np.mean([i["gpu"] for i in info])))


def bench_serve_c5(nstreams, per_stream, rate, inflight, steady_slices=0):
    """Serving: nstreams streams, Poisson arrivals at `rate` slices/s in
    aggregate, served with `inflight` slices in flight against the
    one-slice-at-a-time loop (the same server with one slice in flight)."""
    streams = serve_c5_streams(nstreams, per_stream)
    rng = np.random.default_rng(7)
    arrivals = []
    t = 0.0
    nxt = [0] * nstreams
    total = nstreams * per_stream
    while len(arrivals) < total:
        t += rng.exponential(1.0 / rate)
        k = int(rng.integers(0, nstreams))
        while nxt[k] >= per_stream:
            k = (k + 1) % nstreams
        arrivals.append((t, k, nxt[k]))
        nxt[k] += 1
    batched = run_server(streams, arrivals, inflight)
    serial = run_server(streams, arrivals, 1)
    nnz = sum(len(v) for s in streams for _, v in s["slices"])
    out = dict(workload="serving: %d streams x %d slices (%d nnz), Poisson arrivals %.0f "
                        "slices/s, dims U[10k,1M], rank in {8,16,32,64}, frob 1e-2, 5 its, "
                        "FP64" % (nstreams, per_stream, nnz, rate),
               inflight=inflight, batched=batched, one_at_a_time=serial,
               throughput_ratio=batched["slices_per_s"] / serial["slices_per_s"])
    if steady_slices > 0:
        out["steady"] = bench_serve_steady(nstreams, steady_slices,
                                           0.4 * batched["slices_per_s"], inflight)
    return out


def bench_serve_steady(nstreams, per_stream, rate, inflight):
    """Steady-state serving: the same kind of streams with per_stream slices
    each, Poisson arrivals at `rate` (below the concurrent service rate, so
    latencies are service times, not queue growth)."""
    streams = serve_c5_streams(nstreams, per_stream, seed=20261210)
    rng = np.random.default_rng(11)
    arrivals = []
    t = 0.0
    nxt = [0] * nstreams
    while len(arrivals) < nstreams * per_stream:
        t += rng.exponential(1.0 / rate)
        k = int(rng.integers(0, nstreams))
        while nxt[k] >= per_stream:
            k = (k + 1) % nstreams
        arrivals.append((t, k, nxt[k]))
        nxt[k] += 1
    r = run_server(streams, arrivals, inflight)
    r.update(streams=nstreams, slices_per_stream=per_stream, arrival_rate=rate,
             inflight=inflight)
    return r


def _coord_equal(p, q):
    a, b = p[0], q[0]
    if a.nmodes != b.nmodes or a.nnz != b.nnz:
        return False
    n = int(a.nnz)
    for m in range(int(a.nmodes)):
        if a.dims[m] != b.dims[m]:
            return False
        if not np.array_equal(np.ctypeslib.as_array(a.ind[m], shape=(n,)),
                              np.ctypeslib.as_array(b.ind[m], shape=(n,))):
            return False
    return np.array_equal(np.ctypeslib.as_array(a.vals, shape=(n,)).view(np.uint64),
                          np.ctypeslib.as_array(b.vals, shape=(n,)).view(np.uint64))


def bench_ingest(nnz, threads, seed=20261209):
    """Text input (src/io.c:273-348): a .tns file of the primary workload's
    shape read by the host reader (all host threads) and by the GPU reader,
    and splatt_csf_load (file to all-mode CSFs) against the same steps on
    the host reader's tensor. The file is read once beforehand, so both
    readers take it from the page cache."""
    inds, vals, dims = ps.synth_coord([200_000, 150_000, 50_000], nnz, zipf=[1.2] * 3,
                                      vals="lognormal", seed=seed)
    tmp = tempfile.mkdtemp(prefix="splatt_ingest_", dir="/tmp")
    try:
        path = os.path.join(tmp, "ingest.tns")
        t0 = time.perf_counter()
        ps.write_tns(path, inds, vals)
        size = os.path.getsize(path)
        log("ingest: wrote %.2f GB of text (%d lines) in %.1fs"
            % (size / 1e9, nnz, time.perf_counter() - t0))
        with open(path, "rb") as f:
            text = f.read()
        L = ps.lib()
        L.spd_tns_parse.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.c_void_p,
                                    C.POINTER(C.c_void_p), ps.u64p]
        L.spd_tns_parse.restype = C.c_int
        # the GPU part alone: text in host memory to a device COO
        parse_s = []
        for _ in range(3):
            h = C.c_void_p()
            nh = C.c_uint64(0)
            L.splatt_amd_sync()
            t0 = time.perf_counter()
            rc = L.spd_tns_parse(text, len(text), 3, None, C.byref(h), C.byref(nh))
            parse_s.append(time.perf_counter() - t0)
            if rc != 0:
                raise RuntimeError("spd_tns_parse: " + ps.last_error())
            L.spd_coo_free(h)
        del text
        t0 = time.perf_counter()
        hp = L.splatt_tt_read(path.encode())
        host_s = time.perf_counter() - t0
        dp = L.splatt_alloc_coord(1, 1)
        nh = C.c_uint64(0)
        t0 = time.perf_counter()
        rc = L.splatt_amd_coord_load_dev(path.encode(), dp, C.byref(nh))
        dev_s = time.perf_counter() - t0
        if rc != ps.SUCCESS:
            raise RuntimeError("splatt_amd_coord_load_dev: " + ps.last_error())
        same = _coord_equal(hp, dp)
        L.splatt_free_coord(dp)
        L.splatt_tt_free(hp)
        # splatt_csf_load against read + compact + build on the host path
        opts = ps.Opts(csf_alloc=ps.CSF_ALLMODE)
        out = C.POINTER(ps.SplattCsf)()
        nm = C.c_uint64(0)
        t0 = time.perf_counter()
        rc = L.splatt_csf_load(path.encode(), C.byref(nm), C.byref(out), opts.ptr)
        load_s = time.perf_counter() - t0
        if rc != ps.SUCCESS:
            raise RuntimeError("splatt_csf_load: " + ps.last_error())
        L.splatt_free_csf(out, opts.ptr)
        t0 = time.perf_counter()
        hp = L.splatt_tt_read(path.encode())
        L.splatt_tt_remove_empty(hp)
        cs = L.splatt_csf_alloc(hp, opts.ptr)
        hostpath_s = time.perf_counter() - t0
        L.splatt_free_csf(cs, opts.ptr)
        L.splatt_tt_free(hp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    gpu = min(parse_s)
    return {
        "workload": "read a %.2f GB .tns text (%d lines, 3 modes, %%.17g values) into a COO"
                    % (size / 1e9, nnz),
        "text_bytes": size,
        "host_reader_s": host_s, "host_reader_threads": threads,
        "host_reader_gbps": size / host_s / 1e9,
        "gpu_reader_s": dev_s, "gpu_reader_gbps": size / dev_s / 1e9,
        "gpu_parse_s": gpu, "gpu_parse_gbps": size / gpu / 1e9,
        "gpu_reader_host_lines": int(nh.value),
        "gpu_equals_host_reader": bool(same),
        "csf_load_allmode_s": load_s,
        "host_read_compact_build_allmode_s": hostpath_s,
    }


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
    ap.add_argument("--skip-extras", action="store_true",
                    help="skip the CPD (1000x800x600) and streaming measurements")
    ap.add_argument("--stream-slices", type=int, default=365)
    ap.add_argument("--stream-cpu-slices", type=int, default=6)
    ap.add_argument("--serve-streams", type=int, default=64)
    ap.add_argument("--serve-slices", type=int, default=3)
    ap.add_argument("--serve-rate", type=float, default=500.0)
    ap.add_argument("--serve-inflight", type=int, default=16)
    ap.add_argument("--serve-steady-slices", type=int, default=20,
                    help="slices per stream of the steady-state serving run (0: skip)")
    ap.add_argument("--c3-nnz", type=int, default=100_000_000)
    ap.add_argument("--ingest-nnz", type=int, default=20_000_000)
    ap.add_argument("--only", choices=["cpd", "c3", "stream", "serve", "ingest"], default=None,
                    help="run one secondary workload alone and print its JSON (profiling)")
    ap.add_argument("--root-variant", type=int, default=0, choices=[0, 1, 2, 3],
                    help="3-mode root MTTKRP kernel: 0 per-nonzero k_root3 (default), "
                         "1 persistent with hot leaf rows in LDS, 2 persistent, "
                         "3 k_root3 gathering A once per fiber")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a streaming kernel of known traffic (FP64->FP32 "
                         "conversion of 2^26 values: 512 MiB read, 256 MiB written) "
                         "so hardware byte counters can be scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.only is not None:
        threads = min(CPU_SHARE, os.cpu_count() or 1)
        if args.only == "cpd":
            out = bench_cpd_c1(threads)
        elif args.only == "c3":
            out = bench_cpd_c3(args.c3_nnz)
        elif args.only == "ingest":
            out = bench_ingest(args.ingest_nnz, threads)
        elif args.only == "stream":
            out = bench_stream_c4(args.stream_slices,
                                  0 if args.skip_cpu else args.stream_cpu_slices, threads)
        else:
            out = bench_serve_c5(args.serve_streams, args.serve_slices, args.serve_rate,
                                 args.serve_inflight, args.serve_steady_slices)
        print(json.dumps(out))
        return

    L = ps.lib()
    R = args.rank
    L.spd_mttkrp_set_root_variant(args.root_variant)
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
        extra["dropin_fp64"] = time_dropin(csf_all, mats, dims)
        log("drop-in splatt_mttkrp (host arrays, FP64): %s ms per mode"
            % ", ".join("%.1f" % x for x in extra["dropin_fp64"]["ms_per_mode"]))
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
        threads = min(CPU_SHARE, os.cpu_count() or 1)
        t0 = time.perf_counter()
        cpu = cpu_baseline(inds, vals, dims, mats, R, min(args.cpu_sample, len(vals)), threads)
        log("cpu baseline %.2f GB/s on %d threads (%.1fs)" % (cpu["value"], threads,
                                                              time.perf_counter() - t0))

    if not args.primary_only and not args.skip_extras:
        threads = min(CPU_SHARE, os.cpu_count() or 1)
        c1 = bench_cpd_c1(threads)
        log("CPD c1: GPU %.2f ms/it, CPU %.2f ms/it (%d threads), fit diff %.2e"
            % (c1["gpu_ms_per_iteration"], c1["cpu_ms_per_iteration"], threads,
               c1["fit_abs_diff"]))
        extra["cpd_c1"] = c1
        c3 = bench_cpd_c3(args.c3_nnz)
        log("CPD c3: FP32 %.1f ms/it, FP64 %.1f ms/it, fit diff %.2e, factor errors %s"
            % (c3["f32_ms_per_iteration"], c3["f64_ms_per_iteration"], c3["fit_abs_diff"],
               ", ".join("%.1e" % e for e in c3["factor_rel_err_vs_f64"])))
        extra["cpd_c3"] = c3
        c4 = bench_stream_c4(args.stream_slices, 0 if args.skip_cpu else args.stream_cpu_slices,
                             threads)
        log("stream c4: %.2f slices/s, p50 %.4fs p99 %.4fs" % (
            c4["slices_per_s"], c4["p50_slice_s"], c4["p99_slice_s"]))
        extra["stream_c4"] = c4
        c5 = bench_serve_c5(args.serve_streams, args.serve_slices, args.serve_rate,
                            args.serve_inflight, args.serve_steady_slices)
        log("serve c5: %.2f slices/s (p50 %.3fs p99 %.3fs) vs one-at-a-time %.2f slices/s"
            % (c5["batched"]["slices_per_s"], c5["batched"]["p50_latency_s"],
               c5["batched"]["p99_latency_s"], c5["one_at_a_time"]["slices_per_s"]))
        if "steady" in c5:
            log("serve c5 steady state: %.2f slices/s at arrivals %.1f/s, p50 %.3fs p99 %.3fs"
                % (c5["steady"]["slices_per_s"], c5["steady"]["arrival_rate"],
                   c5["steady"]["p50_latency_s"], c5["steady"]["p99_latency_s"]))
        extra["serve_c5"] = c5
        ing = bench_ingest(args.ingest_nnz, threads)
        log("ingest: host reader %.2f GB/s (%d threads), GPU reader %.2f GB/s (parse alone "
            "%.2f GB/s), same result %s; csf_load %.2fs vs host path %.2fs"
            % (ing["host_reader_gbps"], threads, ing["gpu_reader_gbps"], ing["gpu_parse_gbps"],
               ing["gpu_equals_host_reader"], ing["csf_load_allmode_s"],
               ing["host_read_compact_build_allmode_s"]))
        extra["ingest"] = ing

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
            "traffic": measured_traffic(),
        },
        "cpu_baseline": cpu,
        "modes": primary["modes"],
        "csf_build_s": round(build_all, 3),
        "extra": extra,
        "parity": parity,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
