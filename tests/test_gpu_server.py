"""The request server's lanes: batches of different padded sizes run on the
device at the same time, each result equals the same request solved alone (bit
for bit), and a batch that fails completes its requests with the error."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grids(wm, G, ctx, d, tag):
    pb, ps = os.path.join(d, f"lane_big_{tag}.geo"), os.path.join(d, f"lane_small_{tag}.geo")
    G.planar_grid(pb, 4096, seed=31)
    G.planar_grid(ps, 512, seed=32)
    # fresh geometry objects: no refinement history from other servers
    return wm.Geometry.load(ctx, pb), wm.Geometry.load(ctx, ps)


def _drain(srv, ids):
    got = {}
    while len(got) < len(ids):
        srv.step(wait=True)
        for k in srv.done():
            got[k] = srv.result(k)
    return [got[i] for i in ids]


def test_server_lanes_overlap_bit_identical(wm, geometries, ctx, geo_dir):
    fb = np.array([40.0, 30.0, 20.0, 10.0, 5.0, 3.0], np.float32)
    fs = np.array([25.0, 12.0, 2.0], np.float32)
    # alone: one lane, one request at a time
    gb, gs = _grids(wm, geometries, ctx, geo_dir, "a")
    one = wm.Server(ctx, "f32", size_policy=False, streams=1)
    alone_b = _drain(one, [one.submit(gb, fb, [3, 900])])[0]
    alone_s = _drain(one, [one.submit(gs, fs, [7])])[0]
    one.close()
    # mix: two lanes, both requests at once (the small one on the second lane
    # while the large one is factored on the first)
    gb2, gs2 = _grids(wm, geometries, ctx, geo_dir, "b")
    two = wm.Server(ctx, "f32", size_policy=False, streams=2)
    ids = [two.submit(gb2, fb, [3, 900]), two.submit(gs2, fs, [7])]
    mix_b, mix_s = _drain(two, ids)
    batches, busy, over = two.overlap()
    two.close()
    print(f"lanes: {batches} batches, busy {busy:.3f} s, overlapped {over:.3f} s")
    assert batches >= 2 and over > 0.0
    for a, m in ((alone_b, mix_b), (alone_s, mix_s)):
        assert np.array_equal(a[0], m[0])
        assert np.array_equal(a[1], m[1])


def test_server_failed_batch_completes_with_error(wm, geometries, ctx, geo_dir):
    """A batch whose solve fails (forced on the host before any launch) completes
    its request with the error instead of leaving it running; the server serves
    the next request normally."""
    L = wm.lib()
    L.wmx_server_inject_failure.argtypes = [C.c_void_p, C.c_int]
    p = os.path.join(geo_dir, "srv_fail.geo")
    geometries.planar_grid(p, 512, seed=41)
    g = wm.Geometry.load(ctx, p)
    srv = wm.Server(ctx, "f32", size_policy=False, streams=2)
    assert L.wmx_server_inject_failure(srv.h, 1) == 0
    rid = srv.submit(g, [10.0], [5])
    while True:
        srv.step(wait=True)
        if rid in srv.done():
            break
    with pytest.raises(wm.WireMoMError) as e:
        srv.result(rid)
    assert e.value.code == wm.ERR_ARG
    q, r, c, used = srv.status()
    assert (q, r, c, used) == (0, 0, 0, 0)
    ok = _drain(srv, [srv.submit(g, [10.0], [5])])[0]
    z64, _, _ = wm.solve_sweep(g, [10.0], [5], "f64")
    assert abs(ok[0][0, 0] - z64[0, 0]) <= 1e-10 * abs(z64[0, 0])
    srv.close()


def test_serving_replay_c5_scale(wm, geometries, ctx, geo_dir):
    """A Poisson stream at the serving configuration's scale: 12 requests, N up to
    8192, 1-32 frequencies in 10 kHz - 50 MHz, 1-8 feeds, FP32 factors + FP64
    refinement on two lanes.  Every Zin is checked against its own structure's
    FP64 sweep: within 1e-10, or within kappa_1 u64 where the matrix's condition
    number makes two FP64 solutions differ by more (kappa_1 from the FP64 sweep's
    estimate)."""
    import wiremom.serving as S
    reqs = S.poisson_requests(12, 6.0, seed=9, sizes=(512, 1024, 2048, 4096, 8192), fmin=0.01, fmax=50.0,
                              max_freqs=32, max_rhs=8)
    geoms, feeds = [], []
    for i, r in enumerate(reqs):
        p = os.path.join(geo_dir, f"c5_{i}.geo")
        geometries.planar_grid(p, r["n"], seed=r["seed"])
        geoms.append(wm.Geometry.load(ctx, p))
        feeds.append(np.random.RandomState(r["seed"]).choice(geoms[-1].npls, r["nrhs"], replace=False) + 1)
    srv = wm.Server(ctx, "f32")
    out = S.replay(srv, reqs, geoms, feeds)
    batches, busy, over = srv.overlap()
    srv.close()
    print(f"c5 replay: {len(reqs)} requests, {sum(len(r['freqs']) for r in reqs)} freq-solves, "
          f"makespan {out['makespan']:.2f} s, batches {batches}, overlap {over:.2f} of {busy:.2f} s")
    worst = 0.0
    for r, g, fd, (zin, I, st) in zip(reqs, geoms, feeds, out["results"]):
        z64, _, s64 = wm.solve_sweep(g, r["freqs"], fd, "f64", cond_est=True)
        rel = (np.abs(zin - z64) / np.abs(z64)).max(axis=1)
        for f, e, s_, k in zip(r["freqs"], rel, st, s64):
            bound = max(1e-10, k["cond1"] * 2.0 ** -53)
            worst = max(worst, e / bound)
            assert e <= bound, (r["n"], float(f), float(e), k["cond1"], s_)
    print(f"worst error / bound = {worst:.3f}")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_server_uneven_sizes_share_a_batch(wm, geometries, ctx, geo_dir, prec):
    """Requests of padded sizes 2048 and 1536 go into one batch (the smaller
    matrices padded with the identity; the factorization skips the padding).
    FP64: each result equals the same request solved alone, bit for bit (the
    skipped work is exactly the identity's).  FP32 + refinement: within 1e-10 of
    the FP64 sweep of its structure."""
    gs = []
    for i, n in enumerate((2000, 1500)):
        p = os.path.join(geo_dir, f"uneven_{prec}_{i}.geo")
        geometries.planar_grid(p, n, seed=50 + i)
        gs.append(wm.Geometry.load(ctx, p))
    assert [g.ld for g in gs] == [2048, 1536]
    freqs = [np.array([30.0, 12.0, 4.0], np.float32), np.array([25.0, 8.0], np.float32)]
    feeds = [[3, 400], [9]]
    srv = wm.Server(ctx, prec, size_policy=False, streams=2)
    ids = [srv.submit(g, f, fd) for g, f, fd in zip(gs, freqs, feeds)]
    res = _drain(srv, ids)
    batches, _, _ = srv.overlap()
    srv.close()
    assert batches == 1
    for g, f, fd, (zin, I, st, lat) in zip(gs, freqs, feeds, res):
        z64, I64, _ = wm.solve_sweep(g, f, fd, "f64")
        if prec == "f64":
            assert np.array_equal(zin, z64) and np.array_equal(I, I64)
        else:
            rel = np.abs(zin - z64) / np.abs(z64)
            assert (rel <= 1e-10).all(), rel
