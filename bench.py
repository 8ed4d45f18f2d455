#!/usr/bin/env python3
"""PAR3 create throughput on one MI355X: GF(2^64) Cauchy recovery slices.

Headline workload: 1 GiB of seeded uniform random bytes as 1,024 input slices
x 1 MiB, R=8 recovery slices, Cauchy points (firstInput=0, firstRecovery=1024),
inputs already resident in HBM.  One step = the plain create call
(parhip_d_gf64_recovery) with the coefficient cache off, so every step builds
its coefficients on the device and runs the full contraction writing 8 x 1 MiB
recovery slices.

    python bench.py [--gpus 1] [--steps 200] [--warmup 300]

Prints one JSON line.  `value` is the headline rate; the same line carries the
rest of the metric, each measured in this run: the 262,144 x 4 KiB geometry,
the PAR2 GF(2^16) create (1 GiB, 512 recovery slices) with its CPU baseline,
the batched-vs-plain uneven create (200 jobs), the repair serving loop's
p50/p99 against a stated bound and the PAR2 backend from host memory
(PCIe-inclusive).  `roofline` is for the dominant kernel (the
contraction), timed with HIP events on the stream it runs on; `cpu_baseline`
times the reference's own CPU engine (oracle/_ref, built from the reference
sources) on a bounded sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GiB/s encoded (bit-exact), 1GiB R=8; batched vs plain GiB/s; repair p99 ms"
HBM_PEAK_GBS = 8000.0
# Stated bound of the repair loop: p99 latency at 4 requests/s (the default --rate)
REPAIR_P99_BOUND_MS = 120.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300,
                    help="untimed steps first; a few hundred let the GPU clocks settle "
                         "(3 warm-up steps measure ~7%% lower)")
    ap.add_argument("--slices", type=int, default=1024)
    ap.add_argument("--slice-bytes", type=int, default=1 << 20)
    ap.add_argument("--recovery", type=int, default=8)
    ap.add_argument("--method", default="auto",
                    choices=["auto", "dyadic", "dyadic8", "dyadic16", "dyadic32", "table", "table8", "mfma",
                             "mfma-i8"],
                    help="contraction strategy: auto (the XOR-block kernel for plain Cauchy "
                         "calls), nibble or byte tables in LDS, or the GF(2) bit-matrix on "
                         "fp4 / int8 MFMA")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--headline-only", action="store_true",
                    help="default workload: measure only the 1 GiB R=8 create")
    ap.add_argument("--workload", default="par3-create",
                    choices=["par3-create", "par2-create", "par2-host", "batch", "repair", "host"],
                    help="par3-create (default line, which also carries the other parts of the "
                         "metric); the others print a line of their own")
    ap.add_argument("--par2-slice-bytes", type=int, default=32768,
                    help="par2-host: slice size (even; 1 GiB of slices in total)")
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--rate", type=float, default=4.0, help="repair requests per second")
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--archives", type=int, default=8)
    ap.add_argument("--budget-gib", type=float, default=48.0)
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--p99-bound-ms", type=float, default=REPAIR_P99_BOUND_MS,
                    help="stated p99 latency bound of the repair loop at --rate")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed steps, run one elementwise copy of this many MiB "
                         "(a kernel of known traffic, for scaling rocprofv3 FETCH/WRITE_SIZE)")
    return ap.parse_args()


def gather_rate():
    """ns per wave-level LDS read per CU (16-byte, 8-byte) for random table-entry
    reads, measured on the device (parhip_measure_lds_gather)."""
    import ctypes
    import parparpar_amd as P
    a, b = ctypes.c_double(), ctypes.c_double()
    P.check(P.lib().parhip_measure_lds_gather(ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def device_cus():
    import torch
    return torch.cuda.get_device_properties(0).multi_processor_count


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": aff}


def cpu_threads():
    cpu = host_cpu()
    return min(cpu["affinity"], int(os.environ.get("OMP_NUM_THREADS", cpu["affinity"])), 128)


def timed_runs(run, budget_s, max_runs=50):
    run()  # warm-up
    times = []
    t_start = time.time()
    while (time.time() - t_start < budget_s and len(times) < max_runs) or len(times) < 3:
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2], times


def cpu_baseline(args, d_inputs, budget_s):
    """Reference CPU engine on a bounded sample (first S slices) + bit-exact check
    of the GPU engine on the same sample.  The only places bench.py uses the
    oracle are this leg and cpu_baseline_par2."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from parparpar_amd import device
    n_s = min(args.slices, 128)
    w = args.slice_bytes // 8
    r = args.recovery
    sample = d_inputs[:n_s].cpu().numpy().view(np.uint64)
    sample = np.ascontiguousarray(sample)
    threads = cpu_threads()
    out = np.zeros((r, w), dtype=np.uint64)
    if O.have_ref():
        R = O.ref()
        kind = "reference"

        def run():
            R.ref_compute_recovery(O.p64(sample), n_s, O.p64(out), r, w, 0, n_s, threads, 0)
    else:
        kind = "port"
        threads = 1
        co = O.cauchy(n_s, r, 0, n_s)

        def run():
            O.recovery_with_coeff(sample, co, out)
    med, times = timed_runs(run, budget_s)
    gib = n_s * args.slice_bytes / 2 ** 30
    # bit-exact: GPU engine on the same sample (points 0, n_s)
    g = device.recovery(d_inputs[:n_s], r, 0, n_s)
    torch.cuda.synchronize()
    exact = bool(np.array_equal(g.cpu().numpy().view(np.uint64), out))
    return {
        "value": round(gib / med, 4),
        "unit": "GiB/s",
        "cores": threads,
        "kind": kind,
        "sample": "%d x %d KiB slices, R=%d, points (0, %d); median of %d runs (%.1f s)"
                  % (n_s, args.slice_bytes // 1024, r, n_s, len(times), sum(times)),
        **host_cpu(),
        "bit_exact_vs_gpu": exact,
    }


def cpu_baseline_par2(x_dev, words, r, budget_s):
    """The project's split-nibble shuffle restatement of the reference's CPU
    PAR2 create (oracle/baseline_gf16.c) on a bounded sample of C3's slices,
    on this host's cores, + bit-exact check of the GPU contraction on it."""
    import numpy as np
    import torch
    import parparpar_amd as P
    from oracle import oracle as O
    n_s = 1024
    sample = np.ascontiguousarray(x_dev[:n_s].cpu().numpy().view(np.uint16))
    inums = np.arange(n_s, dtype=np.uint16)
    exps = np.arange(r, dtype=np.uint16)
    threads = cpu_threads()
    res = {}

    def run():
        res["out"], res["used"] = O.baseline_par2_create(sample, inums, exps, threads)
    med, times = timed_runs(run, budget_s, max_runs=20)
    L = P.lib()
    sp = torch.cuda.current_stream().cuda_stream
    co = torch.empty((r, n_s), dtype=torch.int16, device="cuda")
    out = torch.empty((r, words), dtype=torch.int16, device="cuda")
    P.check(L.parhip_d_gf16_coeffs(co.data_ptr(), inums.ctypes.data, n_s, exps.ctypes.data, r, sp))
    P.check(L.parhip_d_gf16_muladd(x_dev.data_ptr(), n_s, words, out.data_ptr(), r, words, words,
                                   co.data_ptr(), 0, sp))
    torch.cuda.synchronize()
    exact = bool(np.array_equal(out.cpu().numpy().view(np.uint16), res["out"]))
    return {
        "value": round(n_s * words * 2 / 2 ** 30 / med, 4),
        "unit": "GiB/s",
        "cores": res["used"],
        "kind": "port",
        "sample": "%d x %d KiB slices, %d recovery slices (exponents 0..%d); split-nibble "
                  "shuffle tables (AVX2), threads over outputs; median of %d runs (%.1f s)"
                  % (n_s, words * 2 // 1024, r, r - 1, len(times), sum(times)),
        **host_cpu(),
        "bit_exact_vs_gpu": exact,
    }


def measured_copy_GBps(mib=1024, reps=10):
    """Device-to-device copy of `mib` MiB (read + write), best of `reps`, GB/s."""
    import torch
    words = mib * (1 << 20) // 8
    src = torch.ones(words, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b) / 1e3)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * words * 8 / best / 1e9


def read_profile_value(name, key):
    """A figure recorded under profiles/ (rocprofv3 counter runs of this kernel)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1].strip())
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures a single GPU; --gpus %d is not supported "
              "(the path would shard slices by word range, one range per GPU)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "par2-create":
        return main_par2(args)
    if args.workload == "par2-host":
        return main_par2_host(args)
    if args.workload == "batch":
        return main_batch(args)
    if args.workload == "repair":
        return main_repair(args)
    if args.workload == "host":
        return main_host(args)
    return main_par3(args)


METHODS = {"auto": 0, "table": 1, "mfma": 2, "mfma-i8": 3, "table8": 4, "dyadic": 5, "dyadic8": 6,
           "dyadic16": 7, "dyadic32": 8}
MAC_NAME = {"mfma": "mac_mfma", "mfma-i8": "mac_mfma", "table8": "mac_tab8", "table": "mac_tab",
            "auto": "mac_dyadic", "dyadic": "mac_dyadic", "dyadic8": "mac_dyadic",
            "dyadic16": "mac_dyadic16", "dyadic32": "mac_dyadic32"}
METHOD_TEXT = {
    "auto": "XOR-block (dyadic) Cauchy contraction: 8x8 blocks as subset convolutions (27 products "
            "per block; 16x16 blocks, 81 products, or 32x32 blocks, 243 products, where they need "
            "fewer table reads), nibble tables in LDS read one table ahead (k_mac_dyadic / "
            "k_mac_dyadic16 / k_mac_dyadic32) + split-K XOR reduce",
    "dyadic8": "XOR-block Cauchy contraction, 8x8 blocks only (k_mac_dyadic) + split-K XOR reduce",
    "dyadic16": "XOR-block Cauchy contraction, 16x16 blocks (k_mac_dyadic16) + split-K XOR reduce",
    "dyadic32": "XOR-block Cauchy contraction, 32x32 blocks (k_mac_dyadic32) + split-K XOR reduce",
    "table8": "8-bit split tables in LDS (k_mac_tab8) + split-K XOR reduce",
    "mfma": "GF(2) bit-matrix on fp4 MFMA (k_mac_mfma)",
    "mfma-i8": "GF(2) bit-matrix on int8 MFMA, int32 parity (k_mac_mfma)",
    "table": "nibble split tables in LDS (k_mac_tab) + split-K XOR reduce",
}
METHOD_TEXT["dyadic"] = METHOD_TEXT["auto"]


def measure_par3(args, d_in, n, sb, r, steps, warmup):
    """Plain create of d_in (n slices of sb bytes) at R=r; coefficients rebuilt
    every step.  Returns (GiB/s, ms per step, per-kernel average ms, d_out)."""
    import torch
    import parparpar_amd as P
    w = sb // 8
    d_out = torch.empty((r, w), dtype=torch.int64, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    L = P.lib()
    mfma = args.method in ("mfma", "mfma-i8", "table8")
    d_coeff = torch.empty((r, n), dtype=torch.int64, device="cuda") if mfma else None

    def step():
        if mfma:  # these strategies take an explicit coefficient matrix
            P.check(L.parhip_d_gf64_build_cauchy(d_coeff.data_ptr(), n, r, 0, n, sp))
            P.check(L.parhip_d_gf64_recovery_with_coeff(d_in.data_ptr(), n, w, d_out.data_ptr(), r, w,
                                                        w, d_coeff.data_ptr(), 0, sp))
        else:
            P.check(L.parhip_d_gf64_recovery(d_in.data_ptr(), n, w, d_out.data_ptr(), r, w, w, 0, n,
                                              0, sp))
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    names = ("mac_dyadic", "mac_dyadic16", "mac_dyadic32", "mac_tab", "mac_tab8", "mac_mfma", "mfma_rows",
             "xor_reduce", "zero_rows", "dyadic_coeffs", "cauchy")
    # the contraction kernel this shape runs (the automatic choice may take
    # 16 x 16 or 32 x 32 XOR blocks): one probed step
    mac = MAC_NAME[args.method]
    P.prof_only(None)
    P.prof_enable(True)
    P.prof_reset()
    step()
    torch.cuda.synchronize()
    P.prof_enable(False)
    for k in ("mac_dyadic32", "mac_dyadic16", "mac_dyadic", "mac_tab", "mac_tab8", "mac_mfma"):
        if P.prof_get(k)[1]:
            mac = k
            break
    # the timed steps: HIP events around the contraction kernel only (two per
    # step; events around every kernel add ~20 us to a 0.54 ms step)
    P.prof_only(mac)
    P.prof_enable(True)
    P.prof_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    P.prof_enable(False)
    kern = {}
    ms, cnt = P.prof_get(mac)
    if cnt:
        kern[mac] = ms / cnt
    # the other kernels of a step, from a short bracket of their own
    P.prof_only(None)
    P.prof_enable(True)
    P.prof_reset()
    for _ in range(min(steps, 20)):
        step()
    torch.cuda.synchronize()
    P.prof_enable(False)
    for k in names:
        ms, cnt = P.prof_get(k)
        if cnt and k not in kern:
            kern[k] = ms / cnt
    return n * sb * steps / 2 ** 30 / elapsed, elapsed / steps * 1e3, kern, d_out, mac


def main_par3(args):
    import torch
    import parparpar_amd as P
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    L = P.lib()
    n, sb, r = args.slices, args.slice_bytes, args.recovery
    w = sb // 8
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x5EED)
    d_in = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda",
                         generator=gen)
    P.set_method(METHODS[args.method])
    P.check(L.parhip_set_coeff_cache(0))  # every step builds its own coefficients
    if args.copy_calibration_mib:
        words = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(words, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.add(src, 0, out=dst)  # reads and writes exactly words*8 bytes
        torch.cuda.synchronize()
        del src, dst
    value, ms_step, kern, d_out, mac_name = measure_par3(args, d_in, n, sb, r, args.steps, args.warmup)
    sp = torch.cuda.current_stream().cuda_stream
    # the step after create: R sealed PAR REC packets from the recovery blocks in HBM
    d_pk = torch.empty((r, 64 + sb), dtype=torch.uint8, device="cuda")
    isid = bytes(range(8))
    P.check(L.parhip_d_assemble_recovery_packets(d_out.data_ptr(), isid, n, r, sb, d_pk.data_ptr(), sp))
    torch.cuda.synchronize()  # untimed first call (workspace pool, code load)
    P.prof_enable(True)
    P.prof_reset()
    for _ in range(3):
        P.check(L.parhip_d_assemble_recovery_packets(d_out.data_ptr(), isid, n, r, sb, d_pk.data_ptr(), sp))
    torch.cuda.synchronize()
    pk_ms, pk_n = P.prof_get("rec_packets")
    P.prof_enable(False)
    del d_pk

    alg_bytes = n * sb + r * sb  # read every input slice once, write R recovery slices
    mac_avg_s = kern.get(mac_name, float("nan")) / 1e3
    achieved = alg_bytes / mac_avg_s / 1e9
    headline = (n, sb, r) == (1024, 1 << 20, 8)
    prof = "mac_dyadic_1024x1MiB_r8.txt" if mac_name == "mac_dyadic" else "traffic_mac_tab_1024x1MiB_r8.txt"
    traffic = read_profile_value(prof, "traffic_bytes_per_launch") if headline and args.method in (
        "auto", "dyadic", "table") else None
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GiB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform random bytes, resident in HBM)",
        "config": {
            "workload": "PAR3 create 1 GiB = %d slices x %d KiB, R=%d, GF(2^64) Cauchy points (0, %d)"
                        % (n, sb // 1024, r, n),
            "slices": n, "slice_bytes": sb, "recovery": r, "parallelism": "single GPU",
            "method": METHOD_TEXT[args.method],
            "step": "coefficients built on the device + contraction (coefficient cache off)",
        },
        "kernels_ms_avg": {k: round(v, 4) for k, v in kern.items()},
        "rec_packets_ms": round(pk_ms / max(pk_n, 1), 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": int(traffic) if traffic else None,
        },
        "cpu_baseline": None,
    }
    if headline:
        # a STREAM-like copy on the same device (1 GiB read + 1 GiB written,
        # HIP events on torch's stream): the HBM rate a plain kernel reaches here
        peak_meas = measured_copy_GBps()
        line["roofline"]["measured_copy_GBps"] = round(peak_meas, 1)
        line["roofline"]["frac_of_measured_copy"] = round(achieved / peak_meas, 4)
    if mac_name == "mac_dyadic" and headline:
        # the kernel's own limit: LDS delivery of the nibble-table reads
        busy = read_profile_value("mac_dyadic_1024x1MiB_r8.txt", "lds_busy")
        ceil_ms = read_profile_value("mac_dyadic_1024x1MiB_r8.txt", "lds_ceiling_ms")
        if busy is not None:
            line["roofline"]["lds_busy"] = busy
        if ceil_ms:
            line["roofline"]["lds_ceiling_ms"] = ceil_ms
            line["roofline"]["frac_of_lds_ceiling"] = round(ceil_ms / (mac_avg_s * 1e3), 4)
        if r <= 8 and args.method in ("auto", "dyadic", "dyadic8"):
            # the LDS gather rate measured live on this device: a word of an 8 x 8
            # block reads 13 x 16 16-byte and 16 8-byte table entries (one output
            # group at R <= 8); the time those reads take at that rate
            g128, g64 = gather_rate()
            nblocks = (n + 7) // 8
            reads128 = w * nblocks * 13 * 16 / 64.0
            reads64 = w * nblocks * 16 / 64.0
            gather_ms = (reads128 * g128 + reads64 * g64) / device_cus() / 1e6
            line["roofline"]["lds_gather_ns_per_read"] = {"b128": round(g128, 4), "b64": round(g64, 4)}
            line["roofline"]["lds_gather_bound_ms"] = round(gather_ms, 4)
            line["roofline"]["frac_of_lds_gather_rate"] = round(gather_ms / (mac_avg_s * 1e3), 4)
    if args.method in ("mfma", "mfma-i8"):
        peak = 10100.0 if args.method == "mfma" else 5000.0
        tops = 2.0 * 64 * 64 * n * r * w / mac_avg_s / 1e12
        line["roofline"] = {"bound": "mfma", "achieved": round(tops, 1), "peak": peak,
                            "unit": "TFLOP/s", "frac": round(tops / peak, 4), "traffic": None,
                            "hbm_gbs": round(achieved, 1)}
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(args, d_in, args.cpu_seconds)
        except Exception as e:  # reported, never silently replaced
            line["cpu_baseline"] = {"error": repr(e)}
    if headline and not args.headline_only:
        # the rest of the metric, each in its own bracket of this run
        v2, ms2, k2, o2, _ = measure_par3(args, d_in.view(262144, 512), 262144, 4096, 8, 50, 50)
        del o2
        line["par3_262144x4KiB_r8"] = {
            "GiBps": round(v2, 3), "ms_per_step": round(ms2, 4),
            "kernels_ms_avg": {k: round(v, 4) for k, v in k2.items()},
            "hbm_frac": round((1 << 30) / (k2.get(mac_name, float("nan")) / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        P.check(L.parhip_set_coeff_cache(1))
        P.set_method(0)
        line["par2_32768x32KiB_r512"] = measure_par2(args, d_in.view(torch.int16).view(32768, 16384))
        del d_in
        torch.cuda.empty_cache()
        line["batched_uneven_200_jobs"] = measure_batch(args)
        torch.cuda.empty_cache()
        line["repair_serving"] = measure_repair(args)
        torch.cuda.empty_cache()
        # the PAR2 backend from host memory, at its default configuration
        line["par2_host_32768x32KiB_r512"] = measure_par2_host(args, runs=(("default", 0, 2),))
    print(json.dumps(line), flush=True)


def measure_par2(args, x, steps=10, warmup=20, cpu=True):
    """PAR2 create, 1 GiB = 32,768 slices x 32 KiB (x: int16 device tensor of
    that shape), 512 recovery slices (exponents 0..511), inputs resident in HBM;
    one step = coefficients on the device + the GF(2^16) contraction."""
    import numpy as np
    import torch
    import parparpar_amd as P
    L = P.lib()
    n, words = x.shape
    sb = words * 2
    r = 512
    out = torch.empty((r, words), dtype=torch.int16, device="cuda")
    co = torch.empty((r, n), dtype=torch.int16, device="cuda")
    inums = np.arange(n, dtype=np.uint16)
    exps = np.arange(r, dtype=np.uint16)
    sp = torch.cuda.current_stream().cuda_stream
    mfma = args.method in ("mfma", "mfma-i8")
    kname = "gf16_mfma" if mfma else "gf16_tab"

    def step():
        P.check(L.parhip_d_gf16_coeffs(co.data_ptr(), inums.ctypes.data, n, exps.ctypes.data, r, sp))
        P.check(L.parhip_d_gf16_muladd(x.data_ptr(), n, words, out.data_ptr(), r, words, words,
                                       co.data_ptr(), 0, sp))
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    P.prof_enable(True)
    P.prof_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    P.prof_enable(False)
    tab_ms, tab_n = P.prof_get(kname)
    red_ms, red_n = P.prof_get("gf16_reduce")
    rows_ms, rows_n = P.prof_get("gf16_rows")
    k_s = tab_ms / max(tab_n, 1) / 1e3
    macs = n * r * words  # GF(2^16) multiply-accumulates per launch
    bit_ops = 2.0 * macs * 256  # as a GF(2) bit-matrix contraction (16x16 per MAC)
    peak = 5000.0 if args.method == "mfma-i8" else 10100.0
    res = {
        "GiBps": round(n * sb * steps / 2 ** 30 / el, 3),
        "ms_per_step": round(el / steps * 1e3, 3),
        "method": {"mfma": "GF(2) bit-matrix on fp4 MFMA (k_gf16_rows + k_gf16_fp4f for full 256-word "
                           "tiles, else k_gf16_mfma)",
                   "mfma-i8": "GF(2) bit-matrix on int8 MFMA (k_gf16_rows + k_gf16_mfma)"}.get(
            args.method, "GF(2^16) nibble split tables in LDS (k_gf16_tab) + split-K XOR reduce"),
        "kernels_ms_avg": {kname: round(tab_ms / max(tab_n, 1), 3),
                           "gf16_reduce": round(red_ms / max(red_n, 1), 3),
                           **({"gf16_rows": round(rows_ms / max(rows_n, 1), 3)} if mfma else {})},
        "roofline": {"bound": "mfma", "achieved": round(bit_ops / k_s / 1e12, 1), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(bit_ops / k_s / 1e12 / peak, 4),
                     "note": "the kernel's rate as GF(2) bit-matrix operations (2 x 256 per GF(2^16) "
                             "multiply-accumulate), against the dense MFMA peak of the type it uses "
                             "(int8 5.0 POP/s) or the densest one (fp4 10.1 POP/s)"},
        "hbm_GBps": round((n * sb + r * sb) / k_s / 1e9, 1),
        "cpu_baseline": None,
    }
    if cpu and not args.no_cpu:
        try:
            res["cpu_baseline"] = cpu_baseline_par2(x, words, r, min(args.cpu_seconds, 5.0))
        except Exception as e:
            res["cpu_baseline"] = {"error": repr(e)}
    return res


def main_par2(args):
    import torch
    import parparpar_amd as P
    torch.cuda.set_device(0)
    P.lib()
    if args.method in ("table8", "dyadic"):
        raise SystemExit("--method %s applies to the GF(2^64) contraction only" % args.method)
    P.set_method(METHODS[args.method])
    g = torch.Generator(device="cuda")
    g.manual_seed(0x5EED2)
    x = torch.randint(-32768, 32767, (32768, 16384), dtype=torch.int16, device="cuda", generator=g)
    res = measure_par2(args, x, steps=max(1, min(args.steps, 10)), warmup=min(args.warmup, 20))
    P.set_method(0)
    line = {
        "metric": METRIC, "value": res["GiBps"], "unit": "GiB/s",
        "n_gpus": 1, "steps": max(1, min(args.steps, 10)), "warmup": min(args.warmup, 20),
        "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16", "data": "synthetic (seeded uniform random bytes, resident in HBM)",
        "config": {"workload": "PAR2 create 1 GiB = 32768 slices x 32 KiB, 512 recovery slices (exponents 0..511)",
                   "method": res["method"]},
        "kernels_ms_avg": res["kernels_ms_avg"], "roofline": res["roofline"], "hbm_GBps": res["hbm_GBps"],
        "cpu_baseline": res["cpu_baseline"],
    }
    print(json.dumps(line), flush=True)


def measure_par2_host(args, runs=(("default", 0, 2), ("batch64", 64, 2), ("areas4", 0, 4))):
    """PAR2 create through the backend engine a PAR2 front end drives
    (parhip_par2_*): 1 GiB of host slices (32,768 x 32 KiB by default), 512
    recovery slices; each add claims a staging slot and the copy workers stage
    the slice (copy + checksum word) into pinned memory, batches go to HBM and
    through the GF(2^16) kernels, getOutput copies each recovery slice back and
    verifies its checksum.  PCIe-inclusive; the outputs are checked against the
    device-resident contraction of the same bytes.  runs: (name, inputs per
    batch (0: the engine's choice), staging areas)."""
    import numpy as np
    import torch
    import parparpar_amd as P
    from parparpar_amd.par2 import GfProc
    L = P.lib()
    r = 512
    sb = args.par2_slice_bytes
    n = (1 << 30) // sb
    words = sb // 2
    host = np.frombuffer(np.random.default_rng(0x5EED3).bytes(n * sb), dtype=np.uint8).reshape(n, sb)
    exps = np.arange(r, dtype=np.uint16)
    # the same create on HBM-resident data, for the byte check
    x = torch.from_numpy(host.view(np.int16).copy()).cuda()
    ref = torch.empty((r, words), dtype=torch.int16, device="cuda")
    co = torch.empty((r, n), dtype=torch.int16, device="cuda")
    inums = np.arange(n, dtype=np.uint16)
    sp = torch.cuda.current_stream().cuda_stream
    P.check(L.parhip_d_gf16_coeffs(co.data_ptr(), inums.ctypes.data, n, exps.ctypes.data, r, sp))
    P.check(L.parhip_d_gf16_muladd(x.data_ptr(), n, words, ref.data_ptr(), r, words, words,
                                   co.data_ptr(), 0, sp))
    ref_host = ref.cpu().numpy().view(np.uint8)
    del x, co, ref
    res = {}
    out = np.empty((r, sb), dtype=np.uint8)
    for name, batch, areas in runs:
        gp = GfProc(sb, input_batch=batch, staging_areas=areas)
        gp.set_recovery_slices(exps)
        nb = gp.input_batch_size
        for i in range(2 * nb):              # warm-up pass over a few batches
            gp.add(i, host[i])
        gp.end()
        gp.discard_output()
        # the raw C entry per slice, as a C++ caller makes it (no per-call array
        # wrapping): asynchronous adds, copied and checksummed by the engine's
        # copy workers; the slices stay in place until end_input returns
        add = L.parhip_par2_add_input_async
        h, base = gp._h, host.ctypes.data
        t0 = time.perf_counter()
        for i in range(n):
            add(h, base + i * sb, sb, i, 0, None, None)
        gp.end()
        valid = gp.get_all(out)
        el = time.perf_counter() - t0
        res[name] = {"GiBps": round(n * sb / 2 ** 30 / el, 3), "inputs_per_batch": nb,
                     "staging_areas": areas,
                     "bit_exact_and_checksums_valid": bool(valid and np.array_equal(out, ref_host))}
        gp.close()
    res["workload"] = ("PAR2 create 1 GiB = %d host slices x %d KiB, 512 recovery slices, asynchronous "
                       "adds (copy workers), end, get; PCIe-inclusive" % (n, sb >> 10))
    return res


def main_par2_host(args):
    import torch
    torch.cuda.set_device(0)
    res = measure_par2_host(args)
    d = res["default"]
    line = {
        "metric": "GiB/s encoded from host memory through the PAR2 backend (PCIe-inclusive), 1GiB R=512",
        "value": d["GiBps"], "unit": "GiB/s", "n_gpus": 1, "steps": 1,
        "higher_is_better": True, "dtype": "u16",
        "data": "synthetic (seeded uniform random bytes in host memory)",
        "config": {"workload": res["workload"] + "; %d inputs per batch (the engine's choice), %d staging "
                                                 "areas" % (d["inputs_per_batch"], d["staging_areas"])},
        "batch64_GiBps": res["batch64"]["GiBps"],
        "areas4_GiBps": res["areas4"]["GiBps"],
        "bit_exact_and_checksums_valid": all(v["bit_exact_and_checksums_valid"] for k, v in res.items()
                                             if k != "workload"),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def make_uneven_jobs(njobs, seed=0xC4):
    """200 jobs: size ~ LogUniform[4 MiB, 512 MiB], slice in {4, 8, 32, 1024} KiB,
    R ~ U{4..64}; last slice zero-padded. Returns (sizes, slice_bytes, R) lists."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lo, hi = np.log(4 << 20), np.log(512 << 20)
    sizes = (np.exp(rng.uniform(lo, hi, size=njobs)).astype(np.int64) // 8 * 8).tolist()
    slices = rng.choice([4 << 10, 8 << 10, 32 << 10, 1 << 20], size=njobs).tolist()
    recs = rng.integers(4, 65, size=njobs).tolist()
    return sizes, slices, recs


def measure_batch(args):
    """Batched uneven PAR3 create: persistent kernel + device queue + hipGraph
    (autotuned launch variants) vs plain per-job create calls, aggregate GiB/s;
    the batched outputs must equal the plain ones byte for byte."""
    import torch
    import parparpar_amd as P
    from parparpar_amd.batch import Batch
    L = P.lib()
    sizes, slices, recs = make_uneven_jobs(args.jobs)
    g = torch.Generator(device="cuda")
    g.manual_seed(0xC4)
    jobs = []
    for size, sb, r in zip(sizes, slices, recs):
        n = (size + sb - 1) // sb
        w = sb // 8
        x = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
        tail = n * w - size // 8
        if tail:
            x.view(-1)[size // 8:] = 0  # zero-padded last slice
        y = torch.empty((r, w), dtype=torch.int64, device="cuda")
        jobs.append((x, y, 0, n))
    total = sum(sizes)
    sp = torch.cuda.current_stream().cuda_stream

    def plain():
        # each job as its own plain create call (parhip_d_gf64_recovery), the
        # reference's one ComputeRecoveryBlocks call per job; coefficients built
        # per call like the batch's (200 jobs would only churn a 16-entry cache)
        for (x, y, fi, fr) in jobs:
            n, w = x.shape
            r = y.shape[0]
            P.check(L.parhip_d_gf64_recovery(x.data_ptr(), n, w, y.data_ptr(), r, w, w, fi, fr, 0, sp))
    P.check(L.parhip_set_coeff_cache(0))
    t0 = time.perf_counter()
    b = Batch(jobs, graph=True, autotune=True)
    plan_s = time.perf_counter() - t0
    steps = max(1, min(args.steps, 5))
    warm = max(1, min(args.warmup, 3))  # a step is ~0.1 s
    res = {}
    for name, fn in (("plain", plain), ("batched", b.run)):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps
    # bit-exact: batched outputs equal the plain outputs
    b.run()
    torch.cuda.synchronize()
    outs = [y.clone() for (_, y, _, _) in jobs]
    plain()
    torch.cuda.synchronize()
    exact = all(torch.equal(o, y) for o, (_, y, _, _) in zip(outs, jobs))
    P.check(L.parhip_set_coeff_cache(1))
    gib = total / 2 ** 30
    out = {
        "workload": "batched uneven PAR3 create: %d jobs, %.2f GiB total, size LogUniform[4 MiB, 512 MiB], "
                    "slice {4,8,32,1024} KiB, R U{4..64}" % (len(jobs), gib),
        "method": "persistent kernels, one per tile family (tables, 8x8, 16x16 and 32x32 XOR blocks), over "
                  "device tile queues, in one hipGraph; per shape: XOR-block tiles (block size by "
                  "table reads, tile width autotuned), or the GF(2) bit-matrix on MFMA where it wins",
        "plain": "one parhip_d_gf64_recovery call per job (the same XOR-block kernels where R >= 4)",
        "jobs_16x16_blocks": sum(b.job_config(k)[0] == 16 and b.job_method(k) == P.METHOD_DYADIC
                                 for k in range(len(jobs))),
        "jobs_32x32_blocks": sum(b.job_config(k)[0] == 32 and b.job_method(k) == P.METHOD_DYADIC
                                 for k in range(len(jobs))),
        "steps": steps,
        "plain_GiBps": round(gib / res["plain"], 3),
        "batched_GiBps": round(gib / res["batched"], 3),
        "batched_vs_plain": round(res["plain"] / res["batched"], 3),
        "plan_s": round(plan_s, 3),
        "batch_info": b.info(),
        "jobs_by_method": {"xor_block": sum(b.job_method(k) == P.METHOD_DYADIC for k in range(len(jobs))),
                           "table": sum(b.job_method(k) == P.METHOD_TABLE for k in range(len(jobs))),
                           "mfma": sum(b.job_method(k) == P.METHOD_MFMA for k in range(len(jobs))),
                           "mfma_candidates": sum(64 * y.shape[0] <= x.shape[1] for (x, y, _, _) in jobs)},
        "bit_exact_batched_vs_plain": bool(exact),
    }
    b.close()
    del jobs
    return out


def main_batch(args):
    import torch
    import parparpar_amd as P
    torch.cuda.set_device(0)
    P.lib()
    res = measure_batch(args)
    line = {
        "metric": METRIC, "value": res["batched_GiBps"], "unit": "GiB/s",
        "n_gpus": 1, "steps": res["steps"], "warmup": max(1, min(args.warmup, 3)),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded uniform random bytes, resident in HBM)",
        "config": {"workload": res["workload"], "method": res["method"]},
        **{k: v for k, v in res.items() if k not in ("workload", "method", "steps")},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def measure_repair(args):
    """PAR3 repair serving: Poisson arrivals of repair requests over
    device-resident archives (64 MiB - 4 GiB, slices 4 KiB - 1 MiB, R 8..1024);
    each request loses k ~ U{1..R} slices at random positions.  Reports p50/p99
    latency (submit -> repaired slices in HBM) against the stated bound,
    aggregate GiB/s of archive data consumed, and checks every repaired slice
    against the original."""
    import threading
    import numpy as np
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    from parparpar_amd.repair import RepairServer
    P.lib()
    rng = np.random.default_rng(0xC5)
    rate = args.rate
    nreq = args.requests
    sizes = np.exp(rng.uniform(np.log(64 << 20), np.log(4 << 30), size=args.archives)).astype(np.int64)
    sizes[0], sizes[-1] = 64 << 20, 4 << 30  # span the whole range
    archives = []
    g = torch.Generator(device="cuda")
    g.manual_seed(0xC5)
    for a, size in enumerate(sizes):
        sb = int(rng.choice([4 << 10, 16 << 10, 64 << 10, 256 << 10, 1 << 20]))
        n = int(max(16, size // sb))
        r = int(min(n - 1, np.exp(rng.uniform(np.log(8), np.log(1024)))))
        w = sb // 8
        data = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
        rec = device.recovery(data, r, 0, n)
        archives.append((data, rec, n, w, r, sb))
    torch.cuda.synchronize()
    budget = int(args.budget_gib * (1 << 30))
    srv = RepairServer(budget, max_inflight=args.inflight)
    for (data, rec, n, w, r, sb) in archives:
        srv.add_archive(data, rec)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    reqs = []
    for t in range(nreq):
        a = int(rng.integers(0, len(archives)))
        n, r = archives[a][2], archives[a][4]
        k = int(rng.integers(1, r + 1))
        reqs.append((a, np.sort(rng.choice(n, k, replace=False)), arrivals[t]))
    ids = [None] * nreq
    lat = [None] * nreq
    ok = [True]
    submitted = threading.Semaphore(0)

    def verifier():
        hip = None
        for t in range(nreq):
            submitted.acquire()
            lat[t] = srv.wait(ids[t])
            a, miss, _ = reqs[t]
            data, _, n, w, _, _ = archives[a]
            ptr, k = srv.result(ids[t])
            got = torch.empty((k, w), dtype=torch.int64, device="cuda")
            import ctypes as C
            if hip is None:
                hip = C.CDLL("libamdhip64.so")
                hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            hip.hipMemcpy(got.data_ptr(), ptr, k * w * 8, 3)
            if not torch.equal(got, data[torch.as_tensor(miss, device="cuda")]):
                ok[0] = False
            del got
            srv.release(ids[t])
    th = threading.Thread(target=verifier)
    th.start()
    t0 = time.perf_counter()
    for t, (a, miss, at) in enumerate(reqs):
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        ids[t] = srv.submit(a, miss)
        submitted.release()
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    cache = srv.cache_stats()
    srv.close()
    consumed = 0
    for (a, miss, _) in reqs:
        data, rec, n, w, r, sb = archives[a]
        consumed += n * sb  # (N-k) data slices + k recovery slices read
    lat_s = sorted(lat)
    p50 = lat_s[len(lat_s) // 2]
    p99 = lat_s[min(len(lat_s) - 1, int(round(0.99 * (len(lat_s) - 1))))]
    out = {
        "workload": "PAR3 repair serving: %d archives (%.2f GiB total), Poisson %.2f req/s, %d requests, "
                    "k ~ U{1..R}, HBM budget %.1f GiB, <= %d in flight"
                    % (len(archives), sum(int(x) for x in sizes) / 2 ** 30, rate, nreq,
                       args.budget_gib, args.inflight),
        "archives": [{"slices": x[2], "slice_bytes": x[5], "recovery": x[4]} for x in archives],
        "rate_per_s": rate,
        "p50_ms": round(p50, 2), "p99_ms": round(p99, 2), "max_ms": round(lat_s[-1], 2),
        "p99_bound_ms": args.p99_bound_ms,
        "p99_within_bound": bool(p99 <= args.p99_bound_ms),
        "server_p50_ms": round(st["p50_ms"], 2), "server_p99_ms": round(st["p99_ms"], 2),
        "peak_inflight": st["peak_inflight"], "peak_hbm_bytes": st["peak_bytes"],
        "inverse_cache": cache,
        "aggregate_GiBps": round(consumed / 2 ** 30 / wall, 3),
        "achieved_req_per_s": round(nreq / wall, 3),
        "wall_s": round(wall, 2),
        "bit_exact_all_requests": bool(ok[0]),
    }
    del archives
    return out


def main_repair(args):
    import torch
    torch.cuda.set_device(0)
    res = measure_repair(args)
    line = {
        "metric": METRIC, "value": res["p99_ms"], "unit": "ms (p99 repair latency)",
        "n_gpus": 1, "steps": args.requests, "warmup": 0,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded archives resident in HBM, seeded Poisson request stream)",
        "config": {"workload": res["workload"], "archives": res["archives"]},
        **{k: v for k, v in res.items() if k not in ("workload", "archives")},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main_host(args):
    """Host-buffer create (PCIe-inclusive): the reference's callers hand over
    host memory.  1 GiB pageable buffer -> parhip_gf64_compute_recovery_blocks
    streamed through the pinned ring, versus one pageable copy then compute;
    and the same bytes from a file via par3_create_streaming."""
    import tempfile
    import numpy as np
    import torch
    import parparpar_amd as P
    from parparpar_amd import binding
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n, sb, r = args.slices, args.slice_bytes, args.recovery
    w = sb // 8
    rng = np.random.default_rng(0x5EED)
    inp = np.frombuffer(rng.bytes(n * sb), dtype=np.uint64).reshape(n, w)
    out = np.zeros((r, w), dtype=np.uint64)
    res = {}
    steps = max(1, min(args.steps, 10))  # a step moves 1 GiB over PCIe
    for name, chunk in (("streamed", 0), ("single_copy", n * sb + 1)):
        binding.set_stream_chunk(chunk)
        binding.compute_recovery(inp, out, n, r, sb, 0, n)  # warm: staging ring, pools
        t0 = time.perf_counter()
        for _ in range(steps):
            binding.compute_recovery(inp, out, n, r, sb, 0, n)
        res[name] = n * sb * steps / 2 ** 30 / (time.perf_counter() - t0)
        if name == "streamed":
            ref = out.copy()
        else:
            assert np.array_equal(ref, out)
    binding.set_stream_chunk(0)
    with tempfile.NamedTemporaryFile(dir=os.environ.get("TMPDIR", "/tmp"), delete=True) as f:
        inp.tofile(f.name)
        opts = {"recoverySlices": r, "blockSize": sb, "firstInput": 0, "firstRecovery": n}
        binding.par3_create_streaming(f.name, opts)  # warm, file in page cache
        t0 = time.perf_counter()
        for _ in range(steps):
            fr = binding.par3_create_streaming(f.name, opts)
        file_rate = n * sb * steps / 2 ** 30 / (time.perf_counter() - t0)
        assert fr["recoveryBuffer"] == ref.tobytes()
    line = {
        "metric": "GiB/s encoded from host memory (PCIe-inclusive), 1GiB R=%d" % r,
        "value": round(res["streamed"], 3), "unit": "GiB/s", "n_gpus": 1, "steps": steps,
        "higher_is_better": True, "dtype": "u64",
        "data": "synthetic (seeded uniform random bytes in pageable host memory / page-cached file)",
        "config": {"workload": "host-pointer create %d x %d KiB, R=%d" % (n, sb // 1024, r)},
        "streamed_GiBps": round(res["streamed"], 3),
        "single_copy_GiBps": round(res["single_copy"], 3),
        "file_streaming_GiBps": round(file_rate, 3),
        "bit_exact": True,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
