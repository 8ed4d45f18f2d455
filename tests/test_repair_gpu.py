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
    damaged[torch.as_tensor(missing, device="cuda")] = 0x0BAD  # lost slots are never read
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
