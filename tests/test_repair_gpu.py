"""PAR3 repair parity: repaired slices equal the dropped originals (the
repair round trip, the strongest oracle), equal the oracle's Gauss-Jordan
solution of the reference's system, and are the same alone and in a mix of
concurrent requests on the server."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _archive(torch, device, n, w, r, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    data = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
    rec = device.recovery(data, r, 0, n)
    torch.cuda.synchronize()
    return data, rec


@pytest.mark.parametrize("n,w,r,k", [(20, 64, 8, 5), (7, 1, 7, 7), (100, 513, 16, 1),
                                     (300, 4096, 64, 64), (1000, 128, 40, 33)])
def test_repair_round_trip(gpu, n, w, r, k):
    import torch
    from parparpar_amd import device, repair
    data, rec = _archive(torch, device, n, w, r, n * 31 + k)
    rng = np.random.default_rng(n + k)
    missing = np.sort(rng.choice(n, k, replace=False))
    damaged = data.clone()
    damaged[torch.as_tensor(missing, device="cuda")] = 0x0BAD  # lost slots never contribute
    out = repair.repair(damaged, rec, missing)
    torch.cuda.synchronize()
    assert torch.equal(out, data[torch.as_tensor(missing, device="cuda")])


def test_repair_matches_oracle_solve(gpu, oracle):
    """The reference's own system (rows = first k recovery blocks) solved by the
    oracle's Gauss-Jordan (src/par3_engine.cc:980-1048) gives the same bytes."""
    import torch
    from parparpar_amd import device, repair
    n, w, r, k = 12, 40, 6, 4
    data, rec = _archive(torch, device, n, w, r, 5)
    missing = [1, 4, 9, 11]
    known = [i for i in range(n) if i not in missing]
    D = data.cpu().numpy().view(np.uint64)
    REC = rec.cpu().numpy().view(np.uint64)
    co = oracle.cauchy(n, r, 0, n)
    A = np.ascontiguousarray(co[:k][:, missing])
    rhs = np.ascontiguousarray(REC[:k]).copy()
    part = oracle.recovery_with_coeff(np.ascontiguousarray(D[known]), np.ascontiguousarray(co[:k][:, known]))
    rhs ^= part
    rc, _, sol = oracle.solve_and_reconstruct(A, rhs)
    assert rc == 0
    out = repair.repair(data, rec, missing).cpu().numpy().view(np.uint64)
    assert np.array_equal(out, sol)
    assert np.array_equal(out, D[missing])


def test_repair_other_rows(gpu):
    import torch
    from parparpar_amd import device, repair
    n, w, r = 50, 300, 10
    data, rec = _archive(torch, device, n, w, r, 8)
    missing = [0, 13, 49]
    out = repair.repair(data, rec, missing, rows=[9, 2, 5])
    assert torch.equal(out, data[torch.as_tensor(missing, device="cuda")])


def test_server_alone_equals_mix(gpu):
    import torch
    from parparpar_amd import device
    from parparpar_amd.repair import RepairServer, copy_result
    archs = [(64, 2048, 16), (500, 512, 40), (40, 8192, 8)]
    srv = RepairServer(8 << 30, max_inflight=4)
    tensors = []
    for j, (n, w, r) in enumerate(archs):
        d, rc = _archive(torch, device, n, w, r, 100 + j)
        tensors.append((d, rc))
        assert srv.add_archive(d, rc) == j
    rng = np.random.default_rng(3)
    reqs = []
    for t in range(12):
        a = int(rng.integers(0, len(archs)))
        n, w, r = archs[a]
        k = int(rng.integers(1, r + 1))
        reqs.append((a, np.sort(rng.choice(n, k, replace=False))))
    alone = []
    for a, miss in reqs:
        rid = srv.submit(a, miss)
        srv.wait(rid)
        alone.append(copy_result(srv, rid, archs[a][1]))
        srv.release(rid)
    ids = [srv.submit(a, miss) for a, miss in reqs]
    mix = []
    for rid, (a, _) in zip(ids, reqs):
        srv.wait(rid)
        mix.append(copy_result(srv, rid, archs[a][1]))
        srv.release(rid)
    torch.cuda.synchronize()
    for (a, miss), x, y in zip(reqs, alone, mix):
        assert torch.equal(x, y)
        assert torch.equal(x, tensors[a][0][torch.as_tensor(miss, device="cuda")])
    st = srv.stats()
    assert st["completed"] == 24 and st["peak_inflight"] >= 2
    srv.close()


def test_server_budget_admission(gpu):
    """A budget that fits one request at a time keeps one in flight."""
    import torch
    from parparpar_amd import device, lib
    from parparpar_amd.repair import RepairServer
    n, w, r = 200, 4096, 16
    d, rc = _archive(torch, device, n, w, r, 9)
    one = lib().parhip_gf64_repair_workspace_bytes(n, w, r) + r * w * 8
    srv = RepairServer(int(one * 1.5), max_inflight=8)
    srv.add_archive(d, rc)
    ids = [srv.submit(0, list(range(i, i + r))) for i in range(0, 6 * r, r)]
    for rid in ids:
        srv.wait(rid)
        srv.release(rid)
    st = srv.stats()
    assert st["peak_inflight"] == 1 and st["peak_bytes"] <= int(one * 1.5)
    srv.close()


# Overlapping input and recovery ranges, the reference's "collision fallback"
# shape (test/par3-fenger-padded-engine.js:110-112): inputs at 0x10000 + i,
# recovery points 66300 + r.  Missing input 764 + r shares its point with
# recovery row r, so A holds the reference's 1 (src/par3_engine.cc:392-394)
# where a Cauchy entry would divide by zero.
COL_N, COL_R, COL_FI, COL_FR = 1000, 512, 0x10000, 66300


def _oracle_repair(oracle, D, REC, fi, fr, missing, rows):
    n, r = D.shape[0], REC.shape[0]
    co = oracle.cauchy(n, r, fi, fr)
    known = [i for i in range(n) if i not in set(missing)]
    A = np.ascontiguousarray(co[rows][:, missing])
    rhs = np.ascontiguousarray(REC[rows]).copy()
    rhs ^= oracle.recovery_with_coeff(np.ascontiguousarray(D[known]),
                                      np.ascontiguousarray(co[rows][:, known]))
    return oracle.solve_and_reconstruct(A, rhs)


@pytest.mark.parametrize("missing,rows", [
    ([764, 765, 800, 3, 999], None),          # two points shared with rows 0 and 1
    ([0, 2, 770], [6, 1, 2]),                 # a shared point on a chosen row
    ([764 + e for e in range(24)] + [5, 17], None),  # shared points on most rows
])
def test_repair_colliding_points(gpu, oracle, missing, rows):
    import torch
    from parparpar_amd import device, repair
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    data = torch.randint(-2 ** 63, 2 ** 63 - 1, (COL_N, 16), dtype=torch.int64, device="cuda",
                         generator=g)
    rec = device.recovery(data, COL_R, COL_FI, COL_FR)
    torch.cuda.synchronize()
    k = len(missing)
    use_rows = list(range(k)) if rows is None else rows
    D = data.cpu().numpy().view(np.uint64)
    REC = rec.cpu().numpy().view(np.uint64)
    rc, _, sol = _oracle_repair(oracle, D, REC, COL_FI, COL_FR, missing, use_rows)
    assert rc == 0
    assert np.array_equal(sol, D[missing])
    out = repair.repair(data, rec, missing, first_input=COL_FI, first_recovery=COL_FR, rows=rows)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint64), sol)


def test_repair_singular_returns_minus_one(gpu, oracle):
    """Missing inputs 764 and 765 on rows 0 and 1: A = [[1, 1], [1, 1]] (their
    points differ by 1), singular.  The reference's solve returns -1."""
    import torch
    from parparpar_amd import ParhipError, device, repair
    data = torch.arange(COL_N * 4, dtype=torch.int64, device="cuda").reshape(COL_N, 4)
    rec = device.recovery(data, COL_R, COL_FI, COL_FR)
    D = data.cpu().numpy().view(np.uint64)
    REC = rec.cpu().numpy().view(np.uint64)
    rc, _, _ = _oracle_repair(oracle, D, REC, COL_FI, COL_FR, [764, 765], [0, 1])
    assert rc == -1
    with pytest.raises(ParhipError) as ei:
        repair.repair(data, rec, [764, 765], first_input=COL_FI, first_recovery=COL_FR)
    assert ei.value.code == -1


def test_server_colliding_points(gpu):
    """The server gives the same bytes alone and in a mix for colliding points,
    and reports -1 for the singular request without disturbing the others."""
    import torch
    from parparpar_amd import ParhipError, device
    from parparpar_amd.repair import RepairServer, copy_result
    g = torch.Generator(device="cuda")
    g.manual_seed(78)
    data = torch.randint(-2 ** 63, 2 ** 63 - 1, (COL_N, 64), dtype=torch.int64, device="cuda",
                         generator=g)
    rec = device.recovery(data, COL_R, COL_FI, COL_FR)
    torch.cuda.synchronize()
    srv = RepairServer(4 << 30, max_inflight=4)
    srv.add_archive(data, rec, first_input=COL_FI, first_recovery=COL_FR)
    reqs = [[764, 765, 800, 3, 999], [766, 770, 771], [764, 765], [10, 767, 999, 500]]
    singular = 2
    alone = []
    for i, miss in enumerate(reqs):
        rid = srv.submit(0, miss)
        if i == singular:
            with pytest.raises(ParhipError) as ei:
                srv.wait(rid)
            assert ei.value.code == -1
            alone.append(None)
        else:
            srv.wait(rid)
            alone.append(copy_result(srv, rid, 64))
        srv.release(rid)
    ids = [srv.submit(0, miss) for miss in reqs]
    for i, (rid, miss) in enumerate(zip(ids, reqs)):
        if i == singular:
            with pytest.raises(ParhipError):
                srv.wait(rid)
        else:
            srv.wait(rid)
            got = copy_result(srv, rid, 64)
            assert torch.equal(got, alone[i])
            assert torch.equal(got, data[torch.as_tensor(miss, device="cuda")])
        srv.release(rid)
    srv.close()


# The two RHS routes: the create's XOR-block contraction over the whole archive
# with the lost inputs masked to zero (consecutive rows, k >= 4; blocks of 8,
# 16 or 32 points) and the explicit-coefficient tables (forced with the table
# method).  Points start
# off a multiple of 8, so virtual slots precede the first input and the first
# recovery point; lost inputs fall in one block of 8, across blocks and at the
# ends; the lost slots hold junk.
ROUTE_CASES = [
    # (n, w, r, fi, fr, missing, rows)
    (300, 513, 40, 0x10003, 0x10003 + 305, list(range(0, 9)) + [299], None),
    (300, 512, 40, 5, 311, [3, 150, 151, 152, 153, 154, 155, 156, 157, 158, 299], list(range(3, 14))),
    (64, 2048, 16, 0, 64, [7, 8, 15, 16, 63], None),
    (1000, 96, 64, 1, 1001, list(range(0, 1000, 17))[:58], None),
    (40, 1024, 8, 7, 47, [0, 1, 2, 3, 4, 5, 6, 39], None),
    # 16 x 16 blocks: 13 virtual slots in front, lost inputs inside one block
    # of 16 and across blocks, k = 16 and 33
    (500, 1024, 48, 13, 13 + 500 + 3, list(range(2, 18)), list(range(5, 21))),
    (333, 520, 40, 29, 29 + 333, [0, 1, 2] + list(range(100, 130)), None),
    # 32 x 32 blocks: 21 virtual slots in front, lost inputs inside one block
    # of 32 and across blocks, k = 24 (rows from 3) and 40
    (700, 1024, 64, 21, 21 + 700 + 5, list(range(10, 34)), list(range(3, 27))),
    (400, 512, 50, 45, 45 + 400, [0, 1] + list(range(60, 98)), None),
]


@pytest.mark.parametrize("case", ROUTE_CASES, ids=["%dx%d-r%d-fi%d-k%d" % (c[0], c[1], c[2], c[3], len(c[5]))
                                                    for c in ROUTE_CASES])
def test_repair_routes_agree(gpu, case):
    import torch
    import parparpar_amd as P
    from parparpar_amd import device, repair
    n, w, r, fi, fr, missing, rows = case
    g = torch.Generator(device="cuda")
    g.manual_seed(n * 7 + len(missing))
    data = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
    rec = device.recovery(data, r, fi, fr)
    damaged = data.clone()
    tm = torch.as_tensor(missing, device="cuda")
    damaged[tm] = torch.randint(-2 ** 63, 2 ** 63 - 1, (len(missing), w), dtype=torch.int64,
                                device="cuda", generator=g)
    outs = []
    try:
        for m in (P.METHOD_AUTO, P.METHOD_DYADIC8, P.METHOD_DYADIC16, P.METHOD_DYADIC32, P.METHOD_TABLE):
            P.set_method(m)
            outs.append(repair.repair(damaged, rec, missing, first_input=fi, first_recovery=fr, rows=rows))
    finally:
        P.set_method(P.METHOD_AUTO)
    torch.cuda.synchronize()
    for o in outs:
        assert torch.equal(o, data[tm])


def test_server_inverse_cache_hits_and_eviction(gpu):
    """The same lost set again takes the cached inverse (hits); a budget that
    holds one request and about one inverse evicts the least recently used
    entry for the next; every result equals the original slices and the
    server's memory never exceeds the budget."""
    import torch
    from parparpar_amd import device, lib
    from parparpar_amd.repair import RepairServer, copy_result
    n, w, r = 500, 512, 64
    d, rc = _archive(torch, device, n, w, r, 77)
    rng = np.random.default_rng(77)
    sets = [np.sort(rng.choice(n, 64, replace=False)) for _ in range(3)]
    srv = RepairServer(8 << 30, max_inflight=2)
    srv.add_archive(d, rc)
    for rep in range(3):
        for miss in sets:
            rid = srv.submit(0, miss)
            srv.wait(rid)
            got = copy_result(srv, rid, w)
            srv.release(rid)
            assert torch.equal(got, d[torch.as_tensor(miss, device="cuda")]), rep
    st = srv.cache_stats()
    assert st["misses"] == 3 and st["hits"] == 6 and st["entries"] == 3, st
    assert st["bytes"] == 3 * 64 * 64 * 8
    srv.close()
    # the server's peak for one request (its charge and its new inverse) plus
    # half an inverse: room for one cached inverse besides a request, not two
    srv = RepairServer(8 << 30, max_inflight=1)
    srv.add_archive(d, rc)
    rid = srv.submit(0, sets[0])
    srv.wait(rid)
    srv.release(rid)
    budget = srv.stats()["peak_bytes"] + 64 * 64 * 8 // 2
    srv.close()
    srv = RepairServer(budget, max_inflight=1)
    srv.add_archive(d, rc)
    for miss in sets + sets:
        rid = srv.submit(0, miss)
        srv.wait(rid)
        got = copy_result(srv, rid, w)
        srv.release(rid)
        assert torch.equal(got, d[torch.as_tensor(miss, device="cuda")])
    st = srv.cache_stats()
    assert st["entries"] <= 1 and st["evictions"] >= 4, st
    assert srv.stats()["peak_bytes"] <= budget
    srv.close()


def test_server_waits_for_release_with_nothing_in_flight(gpu):
    """Admission honours the budget even when nothing is in flight: a request
    that fits only once an earlier output is released waits for the release,
    then completes correctly."""
    import threading
    import torch
    from parparpar_amd import device, lib
    from parparpar_amd.repair import RepairServer, copy_result
    n, w, r, k = 200, 4096, 16, 16
    d, rc = _archive(torch, device, n, w, r, 12)
    # one request's peak (its charge and its cached inverse) plus half an
    # output: a second request fits only once the first output is released
    srv = RepairServer(8 << 30, max_inflight=4)
    srv.add_archive(d, rc)
    rid = srv.submit(0, list(range(2 * k, 3 * k)))
    srv.wait(rid)
    srv.release(rid)
    budget = srv.stats()["peak_bytes"] + k * w * 8 // 2
    srv.close()
    srv = RepairServer(budget, max_inflight=4)
    srv.add_archive(d, rc)
    a = srv.submit(0, list(range(k)))
    srv.wait(a)                       # done; its output stays charged
    b = srv.submit(0, list(range(k, 2 * k)))
    done = threading.Event()
    threading.Thread(target=lambda: (srv.wait(b), done.set())).start()
    assert not done.wait(0.5), "admitted beyond the budget"
    got_a = copy_result(srv, a, w)
    srv.release(a)
    assert done.wait(30), "not admitted after the release"
    got_b = copy_result(srv, b, w)
    srv.release(b)
    assert torch.equal(got_a, d[:k]) and torch.equal(got_b, d[k:2 * k])
    srv.close()
