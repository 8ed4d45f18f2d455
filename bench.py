#!/usr/bin/env python3
"""PAR3 create throughput on one MI355X: GF(2^64) Cauchy recovery slices.

Workload: 1 GiB of seeded uniform random bytes as 1,024 input slices x 1 MiB,
R=8 recovery slices, Cauchy points (firstInput=0, firstRecovery=1024), inputs
already resident in HBM.  One step = build the 8 x 1024 Cauchy coefficient
matrix on the device + the full contraction writing 8 x 1 MiB recovery slices.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Prints one JSON line.  `roofline` is for the dominant kernel (the contraction),
timed with HIP events on the stream it runs on; `cpu_baseline` times the
reference's own CPU engine (oracle/_ref, built from the reference sources) on a
bounded sample of the same workload on this host's cores.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slices", type=int, default=1024)
    ap.add_argument("--slice-bytes", type=int, default=1 << 20)
    ap.add_argument("--recovery", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    return ap.parse_args()


def cpu_baseline(args, d_inputs, budget_s):
    """Reference CPU engine on a bounded sample (first S slices) + bit-exact check
    of the GPU engine on the same sample.  Uses the checker only here."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from parparpar_amd import device
    n_s = min(args.slices, 128)
    w = args.slice_bytes // 8
    r = args.recovery
    sample = d_inputs[:n_s].cpu().numpy().view(np.uint64)
    sample = np.ascontiguousarray(sample)
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count() or 1
    threads = min(aff, int(os.environ.get("OMP_NUM_THREADS", aff)), 128)
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
    run()  # warm-up
    times = []
    t_start = time.time()
    while (time.time() - t_start < budget_s and len(times) < 50) or len(times) < 3:
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
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
        "bit_exact_vs_gpu": exact,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU; "
              "--gpus %d is not supported (the path would shard slices by word range)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    P.lib()

    n, sb, r = args.slices, args.slice_bytes, args.recovery
    w = sb // 8
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x5EED)
    d_in = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda",
                         generator=gen)
    d_out = torch.empty((r, w), dtype=torch.int64, device="cuda")
    d_coeff = torch.empty((r, n), dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    L = P.lib()

    def step():
        P.check(L.parhip_d_gf64_build_cauchy(d_coeff.data_ptr(), n, r, 0, n, sp))
        P.check(L.parhip_d_gf64_recovery_with_coeff(d_in.data_ptr(), n, w, d_out.data_ptr(), r, w, w,
                                                    d_coeff.data_ptr(), 0, sp))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    P.prof_enable(True)
    P.prof_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    P.prof_enable(False)
    mac_ms, mac_n = P.prof_get("mac_tab")
    red_ms, red_n = P.prof_get("xor_reduce")
    cau_ms, cau_n = P.prof_get("cauchy")

    total_bytes = n * sb * args.steps
    value = total_bytes / 2 ** 30 / elapsed
    alg_bytes = n * sb + r * sb  # read every input slice once, write R recovery slices
    mac_avg_s = (mac_ms / max(mac_n, 1)) / 1e3
    achieved = alg_bytes / mac_avg_s / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_mac_tab_1024x1MiB_r8.txt")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = int(line.split(":")[1].strip())

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GiB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform random bytes, resident in HBM)",
        "config": {
            "workload": "PAR3 create 1 GiB = %d slices x %d KiB, R=%d, GF(2^64) Cauchy points (0, %d)"
                        % (n, sb // 1024, r, n),
            "slices": n, "slice_bytes": sb, "recovery": r, "parallelism": "single GPU",
            "method": "nibble split tables in LDS (k_mac_tab) + split-K XOR reduce",
        },
        "kernels_ms_avg": {
            "mac_tab": round(mac_ms / max(mac_n, 1), 4),
            "xor_reduce": round(red_ms / max(red_n, 1), 4),
            "cauchy": round(cau_ms / max(cau_n, 1), 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(args, d_in, args.cpu_seconds)
        except Exception as e:  # reported, never silently replaced
            line["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
