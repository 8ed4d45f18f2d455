"""Several host threads calling the engine at once (the addon's libuv pool and
main thread do): per-thread streams and staging rings, the shared coefficient
cache and the shared copy-thread pool must give every caller its own exact
bytes.  Expected results are computed by the oracle beforehand."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_words(rng, *shape):
    return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)


def test_threads_share_the_engine(gpu, oracle):
    from parparpar_amd import binding, par2
    rng = np.random.default_rng(1234)
    jobs = []
    # (kind, args, expected)
    for t in range(6):
        n, w, r = 20 + 7 * t, 512 + 64 * t, 2 + t
        inp = rand_words(rng, n, w)
        fi, fr = t, n + t  # distinct Cauchy points per thread -> distinct cache entries
        jobs.append(("recovery", (inp, n, w, r, fi, fr), oracle.recovery(inp, r, fi, fr)))
    n, w, r = 33, 700, 4
    inp = rand_words(rng, n, w)
    co = rand_words(rng, r, n)
    jobs.append(("with_coeff", (inp, n, w, r, co), oracle.recovery_with_coeff(inp, co)))
    n, w, r = 40, 1024, 3
    inp = rand_words(rng, n, w)
    jobs.append(("streamed", (inp, n, w, r), oracle.recovery(inp, r, 0, n)))
    valid = np.ones(300, dtype=bool)
    valid[[3, 77, 150, 299]] = False
    ok, used, F = oracle.gf16_rec_matrix(valid, np.arange(10, dtype=np.uint16))
    jobs.append(("par2_matrix", (valid,), (used, F)))

    errors = []

    def worker(job, reps=4):
        kind, args, expect = job
        try:
            for _ in range(reps):
                if kind == "recovery":
                    inp, n, w, r, fi, fr = args
                    out = np.zeros((r, w), dtype=np.uint64)
                    binding.compute_recovery(inp, out, n, r, w * 8, fi, fr)
                    assert np.array_equal(out, expect), kind
                elif kind == "with_coeff":
                    inp, n, w, r, co = args
                    out = np.zeros((r, w), dtype=np.uint64)
                    binding.compute_recovery_with_coeff(co, inp, out, n, r, w * 8)
                    assert np.array_equal(out, expect), kind
                elif kind == "streamed":
                    inp, n, w, r = args
                    binding.set_stream_chunk(3 * w * 8)  # this thread only: 14 chunks
                    out = np.zeros((r, w), dtype=np.uint64)
                    binding.compute_recovery(inp, out, n, r, w * 8, 0, n)
                    assert np.array_equal(out, expect), kind
                else:
                    ok, used, F = par2.rec_matrix_compute(args[0], np.arange(10, dtype=np.uint16))
                    assert ok and np.array_equal(used, expect[0]) and np.array_equal(F, expect[1]), kind
        except Exception as e:  # reported after join
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(j,)) for j in jobs]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a caller thread did not finish"
    assert not errors, errors
