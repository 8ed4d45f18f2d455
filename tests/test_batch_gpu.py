"""Batched create: every job of a persistent-kernel batch equals the plain
per-job computation (and the oracle on sampled columns), graph replays equal
eager runs, and the autotuner's choice does not change any byte."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(rng, shapes, torch):
    jobs = []
    for (n, w, r, fi, fr) in shapes:
        g = torch.Generator(device="cuda")
        g.manual_seed(int(rng.integers(0, 2 ** 31)))
        x = torch.randint(-2 ** 63, 2 ** 63 - 1, (n, w), dtype=torch.int64, device="cuda", generator=g)
        y = torch.full((r, w), 0x55, dtype=torch.int64, device="cuda")
        jobs.append((x, y, fi, fr))
    return jobs


SHAPES = [
    (1, 1, 1, 0, 1), (3, 5, 2, 0, 3), (64, 8192, 8, 0, 64), (1000, 512, 4, 0, 1000),
    (5000, 512, 33, 7, 9000), (16, 131072, 8, 0, 16), (77, 1024, 64, 0, 77),
    (300, 4096, 17, 2, 302), (2, 33, 5, 100, 0), (2000, 512, 32, 5, 4096),
]


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_equals_plain(gpu, oracle, autotune):
    import torch
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(31 + autotune)
    jobs = _jobs(rng, SHAPES, torch)
    b = Batch(jobs, graph=True, autotune=autotune)
    info = b.info()
    assert info["tiles"] > 0
    b.run()
    torch.cuda.synchronize()
    for k, (x, y, fi, fr) in enumerate(jobs):
        ref = device.recovery(x, y.shape[0], fi, fr)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), "job %d (%s) cfg %s" % (k, SHAPES[k], b.job_config(k))
        n, w = x.shape
        cols = np.unique(np.array([0, w - 1, w // 2], dtype=np.int64))
        tc = torch.as_tensor(cols, device="cuda")
        samp = np.ascontiguousarray(x[:, tc].cpu().numpy().view(np.uint64))
        exp = oracle.recovery_with_coeff(samp, oracle.cauchy(n, y.shape[0], fi, fr))
        assert np.array_equal(y[:, tc].cpu().numpy().view(np.uint64), exp)
    b.close()


def test_graph_replays_match_eager(gpu):
    """Replay the captured graph several times with eager work and host
    synchronisation in between; every replay must reproduce the eager bytes."""
    import torch
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(77)
    shapes = [(5000, 512, 33, 7, 9000), (64, 8192, 8, 0, 64), (300, 4096, 17, 2, 302)]
    jobs = _jobs(rng, shapes, torch)
    refs = [device.recovery(x, y.shape[0], fi, fr) for (x, y, fi, fr) in jobs]
    b = Batch(jobs, graph=True, autotune=False)
    for rep in range(4):
        for (_, y, _, _) in jobs:
            y.fill_(rep + 1)          # eager work between replays
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for k, (_, y, _, _) in enumerate(jobs):
            assert torch.equal(y, refs[k]), "replay %d job %d" % (rep, k)
    b.close()


def test_batch_skips_empty_jobs(gpu):
    """Jobs with no inputs or no outputs are skipped; job indices still refer to
    the caller's array."""
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(5)
    shapes = [(3, 5, 2, 0, 3), (0, 64, 4, 0, 0), (10, 64, 0, 0, 10), (8, 256, 3, 0, 8)]
    jobs = _jobs(rng, shapes, torch)
    b = Batch(jobs, graph=False, autotune=False)
    b.run()
    torch.cuda.synchronize()
    for k in (0, 3):
        x, y, fi, fr = jobs[k]
        assert torch.equal(y, device.recovery(x, y.shape[0], fi, fr)), k
        assert b.job_config(k)[0] > 0
    for k in (1, 2):
        with pytest.raises(P.ParhipError):
            b.job_config(k)
    with pytest.raises(P.ParhipError):
        b.job_config(4)
    b.close()


def _eligible(shape):
    n, w, r = shape[:3]
    return 64 * r <= w


def _dyadic(shape):
    """XOR-block tiles: R >= 4 and slices of a multiple of 512 words (the
    tensors here are aligned, at an even stride)."""
    n, w, r = shape[:3]
    return r >= 4 and w % 512 == 0


@pytest.mark.parametrize("graph", [False, True])
def test_batch_mfma_strategy(gpu, oracle, graph):
    """Jobs whose bit-matrix rows fit (64 R <= words per slice) run on the MFMA
    strategy when asked; the others stay on the tables; every byte equals the
    plain computation, replay after replay."""
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(404)
    jobs = _jobs(rng, SHAPES, torch)
    refs = [device.recovery(x, y.shape[0], fi, fr) for (x, y, fi, fr) in jobs]
    b = Batch(jobs, graph=graph, autotune=False, mfma=True)
    for k, s in enumerate(SHAPES):
        want = P.METHOD_MFMA if _eligible(s) else (P.METHOD_DYADIC if _dyadic(s) else P.METHOD_TABLE)
        assert b.job_method(k) == want, (k, s)
    assert 0 < sum(_eligible(s) for s in SHAPES) < len(SHAPES)
    for rep in range(3):
        for (_, y, _, _) in jobs:
            y.fill_(rep + 7)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for k, (_, y, _, _) in enumerate(jobs):
            assert torch.equal(y, refs[k]), "replay %d job %d %s" % (rep, k, SHAPES[k])
    x, y, fi, fr = jobs[2]
    cols = np.array([0, 4095, 8191], dtype=np.int64)
    tc = torch.as_tensor(cols, device="cuda")
    samp = np.ascontiguousarray(x[:, tc].cpu().numpy().view(np.uint64))
    exp = oracle.recovery_with_coeff(samp, oracle.cauchy(x.shape[0], y.shape[0], fi, fr))
    assert np.array_equal(y[:, tc].cpu().numpy().view(np.uint64), exp)
    b.close()


def test_batch_all_mfma_has_no_tiles(gpu):
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(405)
    shapes = [(64, 8192, 8, 0, 64), (33, 1024, 3, 5, 40), (1000, 512, 4, 0, 1000)]
    jobs = _jobs(rng, shapes, torch)
    b = Batch(jobs, graph=True, autotune=False, mfma=True)
    assert b.info()["tiles"] == 0
    assert all(b.job_method(k) == P.METHOD_MFMA for k in range(len(shapes)))
    b.run()
    torch.cuda.synchronize()
    for (x, y, fi, fr) in jobs:
        assert torch.equal(y, device.recovery(x, y.shape[0], fi, fr))
    b.close()


def test_autotune_picks_a_strategy_per_shape(gpu):
    """With autotuning, shapes that qualify for MFMA are timed on both
    strategies; whichever wins, the cache reports it and the bytes match."""
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    L = P.lib()
    P.check(L.parhip_autotune_clear())
    rng = np.random.default_rng(406)
    shapes = [(64, 8192, 8, 0, 64), (77, 1024, 64, 0, 77)]
    jobs = _jobs(rng, shapes, torch)
    b = Batch(jobs, graph=True, autotune=True)
    b.run()
    torch.cuda.synchronize()
    for k, (x, y, fi, fr) in enumerate(jobs):
        assert torch.equal(y, device.recovery(x, y.shape[0], fi, fr))
        m = b.job_method(k)
        assert m in (P.METHOD_DYADIC, P.METHOD_MFMA)  # both shapes qualify for XOR-block tiles
        n, w, r = shapes[k][:3]
        assert L.parhip_autotune_lookup_method(n, r, w) == m
    assert b.job_method(1) == P.METHOD_DYADIC    # 64 R > W: MFMA not offered
    b.close()


def test_batch_dyadic_and_table_tiles_mixed(gpu, oracle):
    """XOR-block tiles (split over atomic tiles, unaligned points, partial
    output groups) and table tiles (R < 4, slices not a multiple of 512 words)
    in one persistent launch, graph replayed; every job equals the oracle."""
    import torch
    import parparpar_amd as P
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(407)
    shapes = [(3000, 512, 13, 5, 3), (9, 1024, 4, 0, 9), (2000, 1536, 8, 1, 2001),
              (100, 700, 9, 0, 100), (40, 512, 2, 0, 40), (513, 2048, 21, 3, 7)]
    jobs = _jobs(rng, shapes, torch)
    b = Batch(jobs, graph=True, autotune=False)
    for k, s in enumerate(shapes):
        assert b.job_method(k) == (P.METHOD_DYADIC if _dyadic(s) else P.METHOD_TABLE), (k, s)
    assert b.info()["atomic_jobs"] >= 1
    for rep in range(3):
        for (_, y, _, _) in jobs:
            y.fill_(rep + 3)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for k, (x, y, fi, fr) in enumerate(jobs):
            n, w = x.shape
            if rep == 0:
                got = y.cpu().numpy().view(np.uint64)
                ref = oracle.recovery(x.cpu().numpy().view(np.uint64), y.shape[0], fi, fr)
                assert np.array_equal(got, ref), (k, shapes[k])
                if k == 0:
                    first = y.clone()
            elif k == 0:
                assert torch.equal(y, first), rep
    b.close()


def test_batch_block_sizes_follow_the_plain_rule(gpu):
    """XOR-block jobs take 8 x 8, 16 x 16 or 32 x 32 blocks by the plain
    create's rule (fewest table reads: R = 8 -> 8, 16 and 40 -> 16, 32 and 64
    -> 32 with aligned points), and each job equals the plain call."""
    import torch
    import parparpar_amd as P
    from parparpar_amd import device
    from parparpar_amd.batch import Batch
    rng = np.random.default_rng(409)
    shapes = [(64, 512, 8, 0, 64), (64, 512, 16, 0, 64), (64, 512, 32, 0, 64),
              (64, 512, 40, 0, 64), (64, 1024, 64, 0, 64), (100, 512, 32, 3, 4096)]
    points = [8, 16, 32, 16, 32, 32]
    jobs = _jobs(rng, shapes, torch)
    b = Batch(jobs, graph=True, autotune=False)
    for k, s in enumerate(shapes):
        assert b.job_method(k) == P.METHOD_DYADIC, (k, s)
        assert b.job_config(k)[0] == points[k], (k, s, b.job_config(k))
    b.run()
    torch.cuda.synchronize()
    for k, (x, y, fi, fr) in enumerate(jobs):
        ref = device.recovery(x, y.shape[0], fi, fr)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), (k, shapes[k])
    b.close()
