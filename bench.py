#!/usr/bin/env python3
"""Benchmark of the packed FFT hot path on one MI355X.

Workload: batched FP32 C2C FFT, N = 2^16, batch 4096 (2^28 complex points, 2 GiB split re/im
input, pack = N), synthetic standard-normal re/im generated in HBM from a seed.  A step is one
forward transform of the whole batch in natural order, out of place, inputs resident in HBM.
The inverse and the bit-reversed order are timed beside it (extra fields).

Prints one JSON line: metric/value (GFLOP/s at 5*N*log2(N) per transform), roofline of the FFT
against the 8 TB/s HBM peak at 16 algorithmic bytes per point, and the host-core baseline (the
oracle's FP32 split-complex radix-4 port, OpenMP over the batch) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_LOG2 = 16
BATCH = 4096
HBM_PEAK_GBS = 8000.0
SEED = 0x5EED0002


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-core baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--quick", action="store_true", help="time the forward step only (for profiler runs)")
    ap.add_argument("--log2n", type=int, default=N_LOG2)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x3", "fp16x2"],
                    help="arithmetic of the default workload and of --config 3 (bf16x3 / fp16x2: "
                         "split-BF16 / split-FP16 matrix-core butterflies)")
    ap.add_argument("--dtype", default="float32", choices=["float32", "float64"],
                    help="element type of the default workload (float64: the double-precision path)")
    ap.add_argument("--copy", action="store_true",
                    help="time a plain device copy of the same bytes instead (calibrates the HBM counters)")
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                    help="2: the headline batched FFT (default); 1: N=1024 latency; 3: fxcorr; "
                         "4: uneven batch; 5: split-precision MFMA")
    return ap.parse_args()


def profile_figure(name: str, key: str):
    """A figure (`key = value` line) from a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key):
            return int(line.split("=")[1])
    return None


def measured_traffic(log2n: int, batch: int, dtype: str = "float32"):
    """HBM bytes per step from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (counters scaled by a plain copy of the same bytes); None for other shapes."""
    tag = "" if dtype == "float32" else "f64_"
    return profile_figure(f"fft_fourstep_{tag}n{1 << log2n}_b{batch}_hbm_traffic.md", "traffic_bytes_per_step")


def time_steps(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(st)
    for _ in range(steps):
        fn()
    t1.record(st)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def cpu_share() -> int:
    """Host cores this process may use: its CPU affinity, bounded by OMP_NUM_THREADS when set (a
    GPU job's host share)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    return max(1, cores)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _port_rate(n: int, seconds: float, threads: int):
    """GFLOP/s of the oracle's FP32 port over a batch sized to about `seconds` of work, best of 3
    after a warm-up, and the batch size."""
    import numpy as np

    from oracle import oracle as O

    plan = O.PortPlan(n)
    cal = threads * 2
    re = np.empty(cal * n, np.float32)
    im = np.empty(cal * n, np.float32)
    O.lib().orc_gen_normal(SEED, 1.0, O._f32(re), O._f32(im), cal * n)
    plan.batch(re, im, n, cal, threads=threads)  # warm-up
    t = time.perf_counter()
    plan.batch(re, im, n, cal, threads=threads)
    per = (time.perf_counter() - t) / cal
    count = int(max(cal, min(4096, seconds / 4 / max(per, 1e-9))))
    count = (count // threads) * threads or threads
    re = np.empty(count * n, np.float32)
    im = np.empty(count * n, np.float32)
    O.lib().orc_gen_normal(SEED, 1.0, O._f32(re), O._f32(im), count * n)
    plan.batch(re, im, n, count, threads=threads)  # warm-up
    best = float("inf")
    for _ in range(3):  # best of 3
        t = time.perf_counter()
        plan.batch(re, im, n, count, threads=threads)
        best = min(best, time.perf_counter() - t)
    return 5.0 * n * math.log2(n) * count / best / 1e9, count


CALIBRATION_NOTE = ("kind=port: the oracle's FP32 split-complex radix-4 restatement (C, OpenMP over the batch), "
                    "not the reference's AVX2 code.  The reference cannot be compiled in this image (fft.hpp:15 "
                    "includes <bits/utility.h>, which libstdc++ 11 does not ship), so the port is not calibrated "
                    "against it.")


def cpu_baseline(n: int, seconds: float):
    """Times the oracle's FP32 port on host cores: bounded samples of the same workload on all of
    this process's cores and on one."""
    threads = cpu_share()
    value, count = _port_rate(n, seconds, threads)
    value1, count1 = _port_rate(n, seconds / 2, 1)
    return {
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{count} forward FP32 C2C FFTs of N=2^{int(math.log2(n))}, split re/im, normal order, "
                  f"best of 3, {threads} OpenMP threads; 1 thread: {count1} FFTs",
        "cores_1": 1,
        "value_1": round(value1, 2),
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "note": CALIBRATION_NOTE + f" cores = this process's CPU share (affinity {cpu_share()} of nproc "
                                   f"{os.cpu_count()}, bounded by OMP_NUM_THREADS).",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.config != 2:
        print(json.dumps(CONFIGS[args.config](args)))
        return
    if args.copy:
        print(json.dumps(run_copy(args)))
        return
    import numpy as np
    import torch

    import packed_amd as pa
    from packed_amd._lib import check, lib

    n = 1 << args.log2n
    batch = args.batch
    pts = n * batch
    f64 = args.dtype == "float64"
    if f64 and args.precision != "fp32":
        raise SystemExit("bench.py: --precision applies to float32 data only")
    esize = 8 if f64 else 4
    lib()  # load the HIP library (fails loudly if it is not built)
    tdt = torch.float64 if f64 else torch.float32
    x = torch.empty(2 * pts, dtype=tdt, device="cuda")
    y = torch.empty(2 * pts, dtype=tdt, device="cuda")
    z = torch.empty(2 * pts, dtype=tdt, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    # standard-normal re/im; split layout: transform b = [re N values][im N values]
    gen = lib().pcx_generate_f64 if f64 else lib().pcx_generate
    check(gen(x.data_ptr(), n, pts, SEED, 1, 1.0, stream))
    torch.cuda.synchronize()

    fwd = pa.fft_unit(n, order="normal", precision=args.precision, dtype=args.dtype)
    fwd_br = pa.fft_unit(n, order="bit_reversed", dtype=args.dtype)

    def step():
        fwd.batch(y.data_ptr(), x.data_ptr(), batch, n, 2 * n, n, 2 * n)

    def step_inv():
        fwd.batch(z.data_ptr(), y.data_ptr(), batch, n, 2 * n, n, 2 * n, inverse=True)

    def step_br():
        fwd_br.batch(z.data_ptr(), x.data_ptr(), batch, n, 2 * n, n, 2 * n)

    ms = time_steps(step, args.steps, args.warmup)
    if args.quick:
        print(json.dumps({"ms_per_step": round(ms, 4)}))
        return
    ms_inv = time_steps(step_inv, max(3, args.steps // 2), 1)
    ms_br = time_steps(step_br, max(3, args.steps // 2), 1)  # FP32 (the matrix-core form is natural order only)
    from oracle import oracle as O

    zin = z[: 2 * n].cpu().numpy()  # written by step_br (bit-reversed forward of x)
    xin = x[: 2 * n].cpu().numpy()
    want_br = np.fft.fft(xin[:n].astype(np.float64) + 1j * xin[n:])[O.slot_freq("bit_reversed", n)]
    err_br = O.rel_l2(zin[:n] + 1j * zin[n:], want_br)
    # the split-precision matrix-core forms on the same batch (the metric's "MFMA-split vs FP32")
    split = {}
    if not f64 and args.precision == "fp32":
        for prec in ("bf16x3", "fp16x2"):
            u = pa.fft_unit(n, precision=prec)
            ms_p = time_steps(lambda: u.batch(z.data_ptr(), x.data_ptr(), batch, n, 2 * n, n, 2 * n),
                              max(3, args.steps // 2), 1)
            errs_p = []
            for b in (0, batch - 1):
                xin = x[b * 2 * n:(b + 1) * 2 * n].cpu().numpy()
                want = np.fft.fft(xin[:n].astype(np.float64) + 1j * xin[n:].astype(np.float64))
                zo = z[b * 2 * n:(b + 1) * 2 * n].cpu().numpy()
                errs_p.append(O.rel_l2(zo[:n] + 1j * zo[n:], want))
            split[prec] = {"ms": round(ms_p, 4), "gflops": round(5.0 * n * math.log2(n) * batch / (ms_p * 1e-3) / 1e9, 1),
                           "speedup_vs_fp32": round(ms / ms_p, 3), "rel_l2_vs_fp64": float(f"{max(errs_p):.3e}")}

    flops = 5.0 * n * math.log2(n) * batch
    gflops = flops / (ms * 1e-3) / 1e9
    algo_bytes = 4.0 * esize * pts  # read and write one complex value per point
    achieved = algo_bytes / (ms * 1e-3) / 1e9

    # parity spot check on the timed outputs: transform 0 and the last one vs the FP64 oracle
    errs = []
    for b in (0, batch - 1):
        xin = x[b * 2 * n:(b + 1) * 2 * n].cpu().numpy()
        xc = xin[:n].astype(np.float64) + 1j * xin[n:].astype(np.float64)
        want = np.fft.fft(xc)
        yo = y[b * 2 * n:(b + 1) * 2 * n].cpu().numpy()
        errs.append(O.rel_l2(yo[:n] + 1j * yo[n:], want))
    traffic = measured_traffic(args.log2n, batch, args.dtype) if args.precision == "fp32" else None
    out = {
        "metric": ("C2C FFT GFLOP/s (5N\u00b7log2N) & %HBM peak, N=2^16 batch; MFMA-split vs FP32" if not f64
                   else "C2C FFT GFLOP/s (5N*log2N), N=2^16 batch 4096, FP64 split re/im"),
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if f64 else ("f32" if args.precision == "fp32" else
                                    f"f32 ({args.precision} split operands, f32 accumulation)"),
        "data": "synthetic (standard-normal re/im from splitmix64 seed 0x5EED0002, generated in HBM)",
        "config": {"workload": f"batched C2C FFT forward, N=2^{args.log2n}, batch {batch}, split re/im "
                               f"(pack=N), natural order, out of place",
                   "fft_size": n, "batch": batch, "order": "normal",
                   "precision": "fp64" if f64 else args.precision, "element_type": args.dtype},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": (f"profiles/fft_fourstep_{'f64_' if f64 else ''}n{n}_b{batch}_hbm_traffic.md: "
                               "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this workload, scaled by a plain copy"
                               if traffic else None),
            "kernel": "four-step column pass k_col + row pass k_row (one step = one launch of each)",
        },
        "inverse": {"ms": round(ms_inv, 4), "gflops": round(flops / (ms_inv * 1e-3) / 1e9, 1)},
        "bit_reversed": {"ms": round(ms_br, 4), "gflops": round(flops / (ms_br * 1e-3) / 1e9, 1),
                         "rel_l2_vs_fp64": float(f"{err_br:.3e}")},
        "rel_l2_vs_fp64": float(f"{max(errs):.3e}"),
    }
    if split:
        out["split"] = split
    if not args.no_cpu and not f64:  # the host-core port is FP32
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    print(json.dumps(out))


# ---------------------------------------------------------------------------- other workloads


def run_copy(args):
    """Plain device-to-device copy of the headline's bytes (2^28 complex points: read 2 GiB, write
    2 GiB).  Its counter readings give the scale factors from FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch

    pts = (1 << args.log2n) * args.batch
    x = torch.ones(2 * pts, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    ms = time_steps(lambda: y.copy_(x), args.steps, args.warmup)
    return {"workload": f"device copy of {2 * pts * 4} bytes", "ms_per_copy": round(ms, 4),
            "gb_per_s": round(2 * 2 * pts * 4 / (ms * 1e-3) / 1e9, 1), "bytes_read": 2 * pts * 4,
            "bytes_written": 2 * pts * 4}


def run_config1(args):
    """N=1024 single FP32 FFT, pack 8, forward then inverse: latency, order exactness, round trip."""
    import numpy as np
    import torch

    import packed_amd as pa
    from oracle import oracle as O

    n = 1024
    x = O.gen_uniform(0x5EED0001, n)
    res = {"metric": "N=1024 single FFT forward+inverse latency (pack 8)", "unit": "us", "higher_is_better": False,
           "config": {"workload": "single C2C FFT N=1024, pack 8, forward then inverse"}, "orders": {}}
    for order in ("normal", "bit_reversed", "unordered"):
        unit = pa.fft_unit(n, order=order)
        v = pa.vector.from_complex(x, 8)
        w = pa.vector(n, 8)
        unit(w, v)
        torch.cuda.synchronize()
        _, want = O.fft_unit(order, False, True, n, O.to_packed(x, 8), 8, 8)
        err = O.rel_l2(w.to_complex(), want)
        unit.ifft(v, w)
        torch.cuda.synchronize()
        rt = O.rel_l2(v.to_complex(), x)

        def pair():
            unit(w, v)
            unit.ifft(v, w)

        us = time_steps(pair, 2000, 50) * 1e3
        # device time per pair without host launch overhead: 100 pairs captured in one graph
        # (the path makes no allocations and no memsets), replayed; every replay is checked
        # against the eager result
        v.data.copy_(torch.from_numpy(O.to_packed(x, 8)).cuda())
        torch.cuda.synchronize()
        for _ in range(100):
            pair()
        torch.cuda.synchronize()
        eager = v.data.clone()
        v.data.copy_(torch.from_numpy(O.to_packed(x, 8)).cuda())
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            g.capture_begin()
            for _ in range(100):
                pair()
            g.capture_end()
        torch.cuda.synchronize()
        replays_ok = True
        for _ in range(3):
            v.data.copy_(torch.from_numpy(O.to_packed(x, 8)).cuda())
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            replays_ok &= bool(torch.equal(v.data, eager))
        g_us = time_steps(g.replay, 20, 2) * 1e3 / 100
        res["orders"][order] = {"fwd_inv_us": round(us, 2), "fwd_inv_us_graph": round(g_us, 2),
                                "graph_replays_equal_eager": replays_ok, "rel_l2_vs_fp64": float(f"{err:.3e}"),
                                "round_trip_rel_l2": float(f"{rt:.3e}")}
    res["value"] = res["orders"]["normal"]["fwd_inv_us"]
    return res


def run_config3(args):
    """Overlap-save cross-correlation: 2^24 complex samples, 16 copies of a 4096-tap unit-modulus
    kernel at SNR 0 dB, FFT block 16384 (step 12288, 1366 blocks), in place and out of place."""
    import numpy as np
    import torch

    import packed_amd as pa
    from oracle import oracle as O

    L, M, N = 1 << 24, 4096, 16384
    f = O.gen_normal(0x5EED0003, L, math.sqrt(0.5))
    ph = O.gen_uniform(0x5EED0013, M)
    g = np.exp(1j * np.pi * (np.real(ph).astype(np.float64) + 1.0)).astype(np.complex64)
    rng = np.random.default_rng(0x5EED0003)
    offs = np.sort(rng.choice(np.arange(0, L - M, 2 * M), 16, replace=False))
    for o in offs:
        f[o:o + M] += g
    gv = pa.vector.from_complex(g, 8)
    xc = pa.fxcorr_unit(gv, N, precision=args.precision)
    src = pa.vector.from_complex(f, 8)
    work = pa.vector(L, 8)
    out = pa.vector(L, 8)
    step = xc.step(8)
    blocks = -(-L // step)

    def inplace():
        work.data.copy_(src.data)
        xc(work)

    def copy_only():
        work.data.copy_(src.data)

    def oop():
        xc(src, out)

    ms_oop = time_steps(oop, max(5, args.steps), 2)
    ms_in = time_steps(inplace, max(5, args.steps), 2) - time_steps(copy_only, max(5, args.steps), 2)
    res_out = out.to_complex()
    top = np.sort(np.argsort(np.abs(res_out))[-16:])
    want = O.xcorr_naive(f[: 1 << 16].astype(np.complex128), g.astype(np.complex128))[: (1 << 16) - M]
    err = O.rel_l2(res_out[: (1 << 16) - M], want)
    inplace()
    torch.cuda.synchronize()
    same = float(np.max(np.abs(work.to_complex() - res_out)))
    flops = blocks * (2 * 5 * N * math.log2(N) + 6 * N)
    algo_bytes = 16 * L + 8 * N
    res = {
        "metric": "fxcorr GFLOP/s (blocks*(2*5N*log2N+6N)), L=2^24, M=4096, N=16384",
        "value": round(flops / (ms_oop * 1e-3) / 1e9, 1), "unit": "GFLOP/s", "higher_is_better": True,
        "ms_out_of_place": round(ms_oop, 4), "ms_in_place": round(ms_in, 4), "blocks": blocks, "step": step,
        "msamples_per_s": round(L / (ms_oop * 1e-3) / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved": round(algo_bytes / (ms_oop * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(algo_bytes / (ms_oop * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": profile_figure("fxcorr_l16777216_m4096_n16384_trace_traffic.md",
                                               "traffic_bytes_per_launch"),
                     "kernel": "k_fxcorr (fused forward FFT, conj-multiply, inverse FFT per block)"},
        "peaks_index_exact": bool(np.array_equal(top, offs)), "rel_l2_vs_fp64_first_2^16": float(f"{err:.3e}"),
        "in_place_equals_out_of_place_maxabs": same, "precision": args.precision,
        "config": {"workload": "overlap-save cross-correlation, L=2^24, M=4096, N=16384, pack 8"},
    }
    if not args.no_cpu:
        # host-core baseline: the oracle's FP32 port on a 2^20-sample slice of the same signal
        threads = cpu_share()
        fs = f[: 1 << 20]
        O.port_fxcorr(fs, g, N, threads)
        t = time.perf_counter()
        _, nb = O.port_fxcorr(fs, g, N, threads)
        dt = time.perf_counter() - t
        cpu_flops = nb * (2 * 5 * N * math.log2(N) + 6 * N)
        res["cpu_baseline"] = {"value": round(cpu_flops / dt / 1e9, 2), "unit": "GFLOP/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                               "sample": f"2^20-sample slice ({nb} blocks), same kernel and block size, {threads} threads",
                               "note": CALIBRATION_NOTE}
    return res


def run_config4(args):
    """10,000 FP32 C2C transforms, sizes 2^6..2^20 drawn log-uniform (seeded), complex normal
    inputs, mixed directions: the work queue (every transform in one persistent launch, eager and
    as a hipGraph; then autotuned, where the tuner may move size classes to grouped launches after
    the persistent one) vs one launch per size class and one launch per transform."""
    import numpy as np
    import torch

    import packed_amd as pa
    from packed_amd._lib import check, lib
    from oracle import oracle as O

    rng = np.random.default_rng(0x5EED0004)
    lgs = rng.integers(6, 21, 10000)
    invs = rng.integers(0, 2, 10000)
    order = np.lexsort((invs, -lgs))  # layout grouped by (size, direction) so per-size batches work
    jobs, off = [], 0
    for k in order:
        n = 1 << int(lgs[k])
        jobs.append((int(lgs[k]), int(invs[k]), off, off))
        off += 2 * n
    src = torch.empty(off, dtype=torch.float32, device="cuda")
    dst = torch.empty(off, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    check(lib().pcx_generate(src.data_ptr(), 1, off // 2, 0x5EED0004, 1, math.sqrt(0.5), st.cuda_stream))
    torch.cuda.synchronize()
    q = pa.fft_queue(jobs)
    flops, algo_bytes = q.work()
    res = {"metric": "uneven batch GFLOP/s (sum 5N*log2N), 10,000 FFTs 2^6..2^20, mixed direction",
           "unit": "GFLOP/s", "higher_is_better": True, "jobs": len(jobs), "points": int(algo_bytes // 16),
           "config": {"workload": "10,000 C2C FFTs, log2 N ~ U{6..20} (seed 0x5EED0004), split re/im, "
                                  "natural order, out of place"}}

    def rate(ms):
        return {"ms": round(ms, 3), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                "gbps": round(algo_bytes / (ms * 1e-3) / 1e9, 1)}

    # the default queue: every transform in one persistent launch
    res["queue_eager"] = rate(time_steps(lambda: q.exec(dst.data_ptr(), src.data_ptr()), 5, 1))
    s1 = torch.cuda.Stream()
    q.capture(dst.data_ptr(), src.data_ptr(), s1)
    with torch.cuda.stream(s1):
        res["queue_graph_one_launch"] = rate(time_steps(lambda: q.launch(s1), 5, 1))
    # one launch per (size, direction) group
    units = {lg: pa.fft_unit(1 << lg) for lg in range(6, 21)}
    groups = []
    for lg, inv, so, do in jobs:
        if groups and groups[-1][0] == lg and groups[-1][1] == inv:
            groups[-1][3] += 1
        else:
            groups.append([lg, inv, so, 1])

    def per_size():
        for lg, inv, so, cnt in groups:
            n = 1 << lg
            units[lg].batch(dst.data_ptr() + 4 * so, src.data_ptr() + 4 * so, cnt, n, 2 * n, n, 2 * n,
                            inverse=bool(inv))

    res["one_launch_per_size"] = rate(time_steps(per_size, 3, 1))

    def per_transform():
        for lg, inv, so, do in jobs:
            n = 1 << lg
            units[lg].batch(dst.data_ptr() + 4 * so, src.data_ptr() + 4 * so, 1, n, 2 * n, n, 2 * n,
                            inverse=bool(inv))

    res["one_launch_per_transform"] = rate(time_steps(per_transform, 1, 1))
    cache = os.path.join(ROOT, "gpurun_out", "queue_tune.txt") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) \
        else None
    q.autotune(dst.data_ptr(), src.data_ptr(), cache)
    res["queue_autotuned"] = rate(time_steps(lambda: q.exec(dst.data_ptr(), src.data_ptr()), 5, 1))
    s = torch.cuda.Stream()
    q.capture(dst.data_ptr(), src.data_ptr(), s)
    with torch.cuda.stream(s):
        ms_graph = time_steps(lambda: q.launch(s), 5, 1)
    res["queue_graph"] = rate(ms_graph)
    # parity sample: the first and last transform of every size class
    out = dst.cpu().numpy()
    inp = src.cpu().numpy()
    worst = 0.0
    seen = {}
    for lg, inv, so, do in jobs:
        seen.setdefault((lg, inv), []).append(so)
    for (lg, inv), offs in seen.items():
        for so in (offs[0], offs[-1]):
            n = 1 << lg
            z = inp[so:so + n].astype(np.float64) + 1j * inp[so + n:so + 2 * n]
            want = O.fft64(z, bool(inv)) / (n if inv else 1)
            got = out[so:so + n] + 1j * out[so + n:so + 2 * n].astype(np.float64)
            worst = max(worst, O.rel_l2(got, want))
    res["max_rel_l2_vs_fp64_sampled"] = float(f"{worst:.3e}")
    res["value"] = res["queue_graph"]["gflops"]
    res["roofline"] = {"bound": "hbm", "achieved": res["queue_graph"]["gbps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(res["queue_graph"]["gbps"] / HBM_PEAK_GBS, 4)}
    return res


def run_config5(args):
    """Split-precision matrix-core forms (three-term BF16 operands, or hi + lo FP16 operands
    scaled per butterfly; FP32 accumulation) against the plain FP32 path: batched C2C N in {2^12, 2^16, 2^20} filling 2 GiB, plus 512-channel
    correlation of 2^20 samples with a 1024-tap kernel (block size chosen by timing)."""
    import numpy as np
    import torch

    import packed_amd as pa
    from packed_amd._lib import check, lib
    from oracle import oracle as O

    pts = 1 << 28
    x = torch.empty(2 * pts, dtype=torch.float32, device="cuda")
    y = torch.empty(2 * pts, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    check(lib().pcx_generate(x.data_ptr(), 1, pts, 0x5EED0005, 1, 1.0, st))
    res = {"metric": "split-BF16 / split-FP16 MFMA vs FP32 FFT, GFLOP/s (5N*log2N), 2 GiB batches",
           "unit": "GFLOP/s",
           "higher_is_better": True, "sizes": {},
           "config": {"workload": "batched C2C FFT forward, N in {2^12, 2^16, 2^20}, 2^28 points, split re/im"}}
    for lg in (12, 16, 20):
        n = 1 << lg
        batch = pts // n
        row = {}
        for prec in ("fp32", "bf16x3", "fp16x2"):
            u = pa.fft_unit(n, precision=prec)
            ms = time_steps(lambda: u.batch(y.data_ptr(), x.data_ptr(), batch, n, 2 * n, n, 2 * n), 10, 2)
            xin = x[:2 * n].cpu().numpy()
            z = xin[:n].astype(np.float64) + 1j * xin[n:]
            out = y[:2 * n].cpu().numpy()
            err = O.rel_l2(out[:n] + 1j * out[n:].astype(np.float64), np.fft.fft(z))
            row[prec] = {"ms": round(ms, 4), "gflops": round(5 * n * lg * batch / (ms * 1e-3) / 1e9, 1),
                         "rel_l2_vs_fp64": float(f"{err:.3e}")}
        row["speedup_bf16x3_over_fp32"] = round(row["fp32"]["ms"] / row["bf16x3"]["ms"], 3)
        row["speedup_fp16x2_over_fp32"] = round(row["fp32"]["ms"] / row["fp16x2"]["ms"], 3)
        res["sizes"][f"2^{lg}"] = row
    del x, y
    torch.cuda.empty_cache()
    # 512 channels x 2^20 samples, 1024-tap kernel
    C, L, M = 512, 1 << 20, 1024
    pack = 8
    dist = 2 * L
    sig = torch.empty(C * dist, dtype=torch.float32, device="cuda")
    out = torch.empty_like(sig)
    check(lib().pcx_generate(sig.data_ptr(), pack, C * L, 0x5EED0015, 1, math.sqrt(0.5), st))
    g = np.exp(1j * np.pi * (np.real(O.gen_uniform(0x5EED0016, M)).astype(np.float64) + 1.0)).astype(np.complex64)
    gv = pa.vector.from_complex(g, pack)
    xres = {}
    for prec in ("fp32", "bf16x3", "fp16x2"):
        best = None
        for N in (4096, 8192, 16384):
            xc = pa.fxcorr_unit(gv, N, precision=prec)
            ms = time_steps(lambda: xc.batch(out.data_ptr(), sig.data_ptr(), L, C, pack, dist), 5, 1)
            step = xc.step(pack)
            blocks = C * (-(-L // step))
            flops = blocks * (2 * 5 * N * math.log2(N) + 6 * N)
            cand = {"fft_size": N, "ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                    "msamples_per_s": round(C * L / (ms * 1e-3) / 1e6, 1)}
            if best is None or ms < best["ms"]:
                best = cand
                bestxc = xc
        bestxc.batch(out.data_ptr(), sig.data_ptr(), L, C, pack, dist)
        torch.cuda.synchronize()
        s0 = sig[:dist].cpu().numpy()
        f0 = O.from_packed(s0, 1 << 16, pack)
        want = O.xcorr_naive(f0, g.astype(np.complex128))[: (1 << 16) - M]
        got = O.from_packed(out[:dist].cpu().numpy(), (1 << 16) - M, pack)
        best["rel_l2_vs_fp64_ch0_first_2^16"] = float(f"{O.rel_l2(got, want):.3e}")
        xres[prec] = best
    res["fxcorr_512ch_2^20_M1024"] = xres
    res["value"] = res["sizes"]["2^16"]["bf16x3"]["gflops"]
    return res


CONFIGS = {1: run_config1, 3: run_config3, 4: run_config4, 5: run_config5}


if __name__ == "__main__":
    main()
