"""GPU correctness of the batched and served configurations at their full
size, with the benchmark's own seeded inputs (fast_mul_ntt_amd.workloads):

* the batch of 4,096 products (operands 64..65,536 limbs drawn
  independently, 25 % squares), planned once and replayed as a HIP graph
  twice with different operand values;
* the 20,000-request serving stream (operands 256..4,194,304 limbs, 20 %
  squares) through the request loop with a 64 GiB pool.

Every output is checked by its residue modulo random 61-bit primes against
the residues of its operands, and by its lowest limb; a seeded subset is
compared word for word with the oracle (the per-product semantics of
src/fast_mul.c:308-460; the checking method of fft_small/fsd_test.c:304-369).
"""

import time

import numpy as np
import pytest

from checks import pool, primes61
from fast_mul_ntt_amd.workloads import (batch_operand, batch_products, serve_requests,
                                        serve_source)

pytestmark = pytest.mark.gpu

M64 = 0xFFFFFFFFFFFFFFFF


@pytest.mark.timeout(900)
def test_batch_of_4096_full_size_graph_replays(fm, orc):
    l1, l2, sq = batch_products()
    count = l1.size
    orc.sqr(np.ones(1, np.uint64))  # the oracle's constant tables, before its calls run on threads
    qs = primes61(2, 4001)
    rng = np.random.default_rng(4002)
    subset = set(int(i) for i in rng.choice(count, size=256, replace=False))
    bufs, items = [], []
    for i in range(count):
        da = fm.DeviceBuffer(int(l1[i]))
        db = None if sq[i] else fm.DeviceBuffer(int(l2[i]))
        do = fm.DeviceBuffer(int(2 * l1[i] if sq[i] else l1[i] + l2[i]))
        bufs += [d for d in (da, db, do) if d is not None]
        items.append((da, int(l1[i]), db, 0 if sq[i] else int(l2[i]), do))
    try:
        with fm.Batch(items) as batch, fm.Stream() as s, pool() as ex:
            for rep in range(2):
                # replay 1 runs on new operand values in the same buffers
                ops = []
                for i, (da, na, db, nb, do) in enumerate(items):
                    a = batch_operand(i + rep * count, 0, na)
                    b = None if db is None else batch_operand(i + rep * count, 1, nb)
                    da.upload(a)
                    if db is not None:
                        db.upload(b)
                    ops.append((a, b))
                print(f"batch replay {rep}: {count} products", flush=True)
                batch.run(s)
                s.synchronize()

                def check(i, out):
                    a, b = ops[i]
                    b = a if b is None else b
                    assert out.size == a.size + b.size, i
                    assert int(out[0]) == (int(a[0]) * int(b[0])) & M64, (rep, i)
                    for q in qs:
                        assert orc.mod_q(out, q) == orc.mulmod_q(orc.mod_q(a, q), orc.mod_q(b, q), q), (rep, i, q)
                    if i in subset:
                        want = orc.sqr(a) if ops[i][1] is None else orc.mul(a, b)
                        assert np.array_equal(out, want), ("oracle", rep, i)

                futs = [ex.submit(check, i, do.download(na + (nb if db is not None else na)))
                        for i, (da, na, db, nb, do) in enumerate(items)]
                for f in futs:
                    f.result()
    finally:
        for d in bufs:
            d.free()


@pytest.mark.timeout(900)
def test_serve_20000_requests_full_size(fm, orc):
    l1, l2, sq, o1, o2, _gaps, src_len = serve_requests()
    count = l1.size
    src_host = serve_source(src_len)
    orc.sqr(np.ones(1, np.uint64))  # the oracle's constant tables, before its calls run on threads
    (q,) = primes61(1, 5001)
    pre = orc.prefix_mod_q(src_host, q)
    rng = np.random.default_rng(5002)
    big = np.flatnonzero(np.where(sq, 2 * l1, l1 + l2) >= (1 << 22))
    subset = set(int(i) for i in rng.choice(count, size=22, replace=False)) | set(int(i) for i in big[:2])

    def operands(i):
        a = (int(o1[i]), int(l1[i]))
        b = a if sq[i] else (int(o2[i]), int(l2[i]))
        return a, b

    src = fm.DeviceBuffer.from_host(src_host)
    srv = fm.Server(pool_bytes=64 << 30, streams=4)
    try:
        ids = {}
        for i in range(count):
            (oa, na), (ob, nb) = operands(i)
            rid = srv.submit(src.view(oa, na), na, None if sq[i] else src.view(ob, nb), 0 if sq[i] else nb)
            ids[rid] = i
        print(f"submitted {count} requests", flush=True)

        def check(i, out):
            (oa, na), (ob, nb) = operands(i)
            assert out.size == na + nb, i
            assert int(out[0]) == (int(src_host[oa]) * int(src_host[ob])) & M64, i
            ra, rb = orc.slice_mod_q(pre, oa, na, q), orc.slice_mod_q(pre, ob, nb, q)
            assert orc.mod_q(out, q) == ra * rb % q, i
            if i in subset:
                a, b = src_host[oa:oa + na], src_host[ob:ob + nb]
                want = orc.sqr(a) if sq[i] else orc.mul(a, b)
                assert np.array_equal(out, want), ("oracle", i)

        done = 0
        futs = []
        with pool() as ex:
            while done < count:
                got = srv.poll()
                if not got:
                    time.sleep(2e-4)
                for rid in got:
                    i = ids.pop(rid)
                    out = srv.result(rid)  # copied out before the pool memory is released
                    srv.release(rid)
                    futs.append(ex.submit(check, i, out))
                    done += 1
                    if done % 2000 == 0:
                        print(f"served {done} of {count}", flush=True)
                # bound the outputs held by pending checks
                while len(futs) > 64:
                    futs.pop(0).result()
            for f in futs:
                f.result()
        st = srv.stats()
        assert st["completed"] == count and st["failed"] == 0, st
    finally:
        srv.close()
        src.free()
