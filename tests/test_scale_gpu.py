"""Batched create and repair serving at their benchmark scale.

C4: the 200-job uneven batch the bench measures (sizes LogUniform[4 MiB,
512 MiB], slices {4, 8, 32, 1024} KiB, R U{4..64}, about 20 GiB): the
persistent batched create must equal one plain create call per job byte for
byte, and three word columns of every job must equal the oracle's Cauchy
contraction (coefficients from the oracle's own point rule).

C5: the largest archive shape of the serving workload, 4 GiB with R = 1,024,
repaired with k = 1, 37 and 1,024 missing slices (k up to R) alone, then the
same requests mixed with requests on smaller archives in flight together;
every repaired slice must equal the original.
"""
import concurrent.futures

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_job_columns(oracle, xcols, r, n):
    """(r, 3) uint64: the oracle's contraction of the (n, 3) input columns with
    the Cauchy matrix of inputs 0..n-1 and recovery points n..n+r-1."""
    coeff = oracle.cauchy(n, r, 0, n)
    return oracle.recovery_columns(xcols, coeff, np.arange(xcols.shape[1]))


def test_c4_full_batch_equals_plain_and_oracle(gpu, oracle):
    import torch
    import bench
    from parparpar_amd.batch import Batch
    P = gpu
    L = P.lib()
    sizes, slices, recs = bench.make_uneven_jobs(200)
    g = torch.Generator(device="cuda")
    g.manual_seed(0xC4)
    jobs = []
    for size, sb, r in zip(sizes, slices, recs):
        n, w = (size + sb - 1) // sb, sb // 8
        x = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
        if n * w > size // 8:
            x.view(-1)[size // 8:] = 0  # zero-padded last slice
        jobs.append((x, torch.empty((r, w), dtype=torch.int64, device="cuda"), 0, n))
    print("C4: %d jobs, %.2f GiB" % (len(jobs), sum(sizes) / 2 ** 30), flush=True)
    b = Batch(jobs, graph=True, autotune=True)
    b.run()
    torch.cuda.synchronize()
    batched = [y.clone() for (_, y, _, _) in jobs]
    b.close()
    sp = torch.cuda.current_stream().cuda_stream
    P.check(L.parhip_set_coeff_cache(0))
    try:
        for (x, y, fi, fr) in jobs:
            n, w = x.shape
            P.check(L.parhip_d_gf64_recovery(x.data_ptr(), n, w, y.data_ptr(), y.shape[0], w, w,
                                             fi, fr, 0, sp))
        torch.cuda.synchronize()
    finally:
        P.check(L.parhip_set_coeff_cache(1))
    diff = [j for j, (o, (_, y, _, _)) in enumerate(zip(batched, jobs)) if not torch.equal(o, y)]
    assert not diff, "batched != plain for jobs %s" % diff[:10]
    print("C4: batched == plain on all jobs", flush=True)

    rng = np.random.default_rng(0xC4C4)
    work = []
    for j, (x, y, _, n) in enumerate(jobs):
        w = x.shape[1]
        cols = [0, int(rng.integers(1, w - 1)) if w > 2 else 0, w - 1]
        xcols = np.ascontiguousarray(x[:, cols].cpu().numpy().view(np.uint64))
        got = batched[j][:, cols].cpu().numpy().view(np.uint64)
        work.append((j, xcols, y.shape[0], n, got))
    del jobs, batched
    # the oracle releases the GIL inside its C calls: one job per worker
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        futs = {ex.submit(_oracle_job_columns, oracle, xc, r, n): (j, got) for (j, xc, r, n, got) in work}
        for f in concurrent.futures.as_completed(futs):
            j, got = futs[f]
            assert np.array_equal(f.result(), got), "job %d differs from the oracle" % j
    print("C4: oracle columns equal on all %d jobs" % len(work), flush=True)


def test_c5_largest_archive_alone_and_mixed(gpu):
    import torch
    from parparpar_amd import device
    from parparpar_amd.repair import RepairServer, copy_result
    g = torch.Generator(device="cuda")
    # (slices, words per slice, recovery slices): the 4 GiB archive at R = 1,024
    # (64 KiB slices), then two smaller ones of the serving mix
    shapes = [(65536, 8192, 1024), (8192, 1024, 8), (16384, 2048, 100)]
    archives = []
    for a, (n, w, r) in enumerate(shapes):
        g.manual_seed(0xC5 + a)
        data = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
        archives.append((data, device.recovery(data, r, 0, n)))
    torch.cuda.synchronize()
    print("C5: archives ready", flush=True)
    srv = RepairServer(64 << 30, max_inflight=4)
    for d, rc in archives:
        srv.add_archive(d, rc)
    rng = np.random.default_rng(0xC55)
    big = [np.sort(rng.choice(65536, k, replace=False)) for k in (1, 37, 1024)]
    # k = 1,024 = R: every recovery slice used
    for miss in big:
        rid = srv.submit(0, miss)
        srv.wait(rid)
        got = copy_result(srv, rid, 8192)
        srv.release(rid)
        assert torch.equal(got, archives[0][0][torch.as_tensor(miss, device="cuda")]), len(miss)
        print("C5: alone k=%d ok" % len(miss), flush=True)
    reqs = [(0, big[2]), (1, np.sort(rng.choice(8192, 8, replace=False))), (0, big[1]),
            (2, np.sort(rng.choice(16384, 100, replace=False))), (0, big[0]),
            (2, np.sort(rng.choice(16384, 3, replace=False)))]
    ids = [srv.submit(a, miss) for a, miss in reqs]
    for rid, (a, miss) in zip(ids, reqs):
        srv.wait(rid)
        got = copy_result(srv, rid, shapes[a][1])
        srv.release(rid)
        assert torch.equal(got, archives[a][0][torch.as_tensor(miss, device="cuda")]), (a, len(miss))
    print("C5: mixed requests ok", flush=True)
    srv.close()
