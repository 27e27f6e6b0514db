"""GPU tests of the request loop (fmh_server_*): continuous batching over a
device pool.  Every served product must equal the same product computed
alone (fmh_mul_device / fmh_sqr_device) and the oracle, bit for bit."""

import random
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log_uniform(rng, lo, hi):
    import math
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def test_served_equals_alone_under_load(fm, orc):
    """Requests arrive from a client thread at random gaps while earlier ones
    run; the pool (24 MiB) holds only part of them, so admission waits for
    memory.  Each result equals the product computed alone."""
    rng = random.Random(5001)
    n_req = 160
    specs = []
    with fm.Stream() as s:
        ops, bufs = [], []
        for i in range(n_req):
            n1 = _log_uniform(rng, 16, 120000)
            sq = rng.random() < 0.2
            n2 = n1 if sq else _log_uniform(rng, 16, 120000)
            a = orc.splitmix(50_000 + 2 * i, n1)
            b = None if sq else orc.splitmix(50_001 + 2 * i, n2)
            da = fm.DeviceBuffer.from_host(a, s)
            db = None if sq else fm.DeviceBuffer.from_host(b, s)
            bufs += [da] + ([db] if db else [])
            ops.append((a, b, da, db))
        # products computed alone
        alone = []
        for a, b, da, db in ops:
            with fm.DeviceBuffer(a.size + (a.size if b is None else b.size)) as do:
                if b is None:
                    fm.sqr_device(da, a.size, do, s)
                else:
                    fm.mul_device(da, a.size, db, b.size, do, s)
                s.synchronize()
                alone.append(do.download())
        with fm.Server(pool_bytes=24 << 20, streams=4, max_group=32) as srv:
            ids = [None] * n_req
            err = []

            def client():
                try:
                    for i, (a, b, da, db) in enumerate(ops):
                        ids[i] = srv.submit(da, a.size, db, 0 if b is None else b.size)
                        time.sleep(rng.expovariate(1 / 0.0004))
                except Exception as e:  # surfaced below
                    err.append(e)

            t = threading.Thread(target=client)
            t.start()
            served = {}

            def collect():
                # outputs are copied out and released as they complete, which
                # returns their pool memory to waiting requests
                for rid in srv.poll():
                    served[rid] = srv.result(rid)
                    srv.release(rid)

            while t.is_alive() or len(served) < n_req:
                collect()
                time.sleep(0.001)
                assert not err, err
            t.join()
            st = srv.stats()
            assert st["completed"] == n_req and st["failed"] == 0, st
            assert st["pool_peak_bytes"] <= 24 << 20
            for i, rid in enumerate(ids):
                assert np.array_equal(served[rid], alone[i]), i
        # spot-check the alone results themselves against the oracle
        for i in range(0, n_req, 16):
            a, b, _, _ = ops[i]
            want = orc.sqr(a) if b is None else orc.mul(a, b)
            assert np.array_equal(alone[i], want), i
        for d in bufs:
            d.free()


def test_served_spectrum_requests(fm, orc):
    """Requests against one resident spectrum, mixed with plain requests."""
    n1 = 30000
    a = orc.splitmix(777, n1)
    with fm.Stream() as s, fm.DeviceBuffer.from_host(a, s) as da, fm.Spectrum(da, n1, 0, s) as sp, \
            fm.Server(pool_bytes=256 << 20, streams=3) as srv:
        s.synchronize()
        want, ids, bufs = [], [], []
        for i in range(24):
            n2 = 1000 + 1100 * i
            b = orc.splitmix(900 + i, n2)
            db = fm.DeviceBuffer.from_host(b, s)
            bufs.append(db)
            if i % 3 == 0:
                ids.append(srv.submit(db, n2, da, n1))
            else:
                ids.append(srv.submit_spectrum(sp, db, n2))
            want.append(orc.mul(b, a))
        ids.append(srv.submit_spectrum(sp))
        want.append(orc.sqr(a))
        srv.drain()
        for rid, w in zip(ids, want):
            assert np.array_equal(srv.result(rid), w), rid
            srv.release(rid)
        for d in bufs:
            d.free()


def test_server_reports_a_pool_held_by_outputs(fm, orc):
    """Unreleased outputs that leave no room for the next request: drain
    reports it instead of blocking, and releasing lets the request run."""
    x = orc.splitmix(5, 100000)
    with fm.Server(pool_bytes=16 << 20, streams=2) as srv, fm.DeviceBuffer.from_host(x) as d:
        # 8 outputs of 1 MiB each stay in the pool (6 MiB of residues while running)
        held = [srv.submit(d, 65536, d, 65535) for _ in range(8)]
        for rid in held:
            srv.wait(rid)
        # 12 MiB of residues + 1.5 MiB of output: fits the pool, not next to the held outputs
        blocked = srv.submit(d, 100000, d, 99999)
        with pytest.raises(fm.FmhError, match="unreleased"):
            srv.drain()
        for rid in held:
            srv.release(rid)
        srv.wait(blocked)
        assert np.array_equal(srv.result(blocked), orc.mul(x, x[:99999]))
        srv.release(blocked)


def test_server_refuses_what_cannot_fit(fm, orc):
    with fm.Server(pool_bytes=8 << 20) as srv, fm.DeviceBuffer(1 << 19) as d:
        with pytest.raises(fm.FmhError):
            srv.submit(d, 1 << 19, d, 1 << 19)  # needs ~50 MiB of residues
        rid = srv.submit(d, 100, d, 100)
        srv.wait(rid)
        srv.release(rid)
