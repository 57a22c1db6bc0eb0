"""Benchmark: freq-solves/s (fill + LU + solve) at N = 8192 on one MI355X.

Workload (the lightning-protection structure): 4 down-conductors 20 m tall, a
10 x 10 m roof mesh at 1 m pitch and a ground ring, radius 5 mm, segmented to
N = 8192 unknowns, delta-gap feed just below the top corner, 64 frequencies
log-spaced over 10 kHz - 10 MHz.  Synthetic geometry generated in-process.

One step = one sweep over the 64 frequencies: fill, factor, solve and Zin for
every frequency, with inputs (geometry) resident in HBM.  `value` is the
mixed-precision path (FP32 matrix-core LU + FP64 iterative refinement, or the
precision given by --prec); the plain FP64 GPU path is timed beside it and
every Zin is compared with it.

Also reported: the roofline of the dominant kernel class (HIP events on the
launch stream, algorithmic flops per launch, MI355X datasheet peaks), a
serving leg (a Poisson stream of requests on the continuous-batching server,
p50/p99 latency and solves/s) and the CPU oracle timed on a bounded sample of
the same workload on the host cores.

    python bench.py [--steps K] [--warmup W] [--prec f32|f64] [--nfreq F] [--no-cpu]
                    [--no-serving] [--requests R] [--rate PER_S]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X datasheet peaks (dense)
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16x3": 2516.6 / 3.0}
PEAK_FP64_VALU = 78.6
PEAK_HBM_GBS = 8000.0
# kernel classes timed by the library profiler: (bound, peak)
CLASSES = {
    "fill": ("fp64", PEAK_FP64_VALU),
    "panel": ("fp64", 78.6), "swap_trsm": ("fp64", 78.6), "gemm": ("mfma", 78.6),
    "panel_f32": ("fp32", 157.3), "swap_trsm_f32": ("fp32", 157.3), "gemm_f32": ("mfma", 157.3),
    "trsv": ("hbm", PEAK_HBM_GBS), "resid": ("hbm", PEAK_HBM_GBS),
}
# HBM bytes / algorithmic bytes of k_gemm_f64, from FETCH_SIZE + WRITE_SIZE passes
# calibrated on a plain copy (profiles/gemm_f64_traffic_c3_n8192.md)
GEMM_F64_HBM_RATIO = 1.88
REF_BEST_PUBLISHED = 0.148  # freq-solves/s at N=8192, reference's best variant (GPU fill + MKL LU, 6 procs)


def _roctx():
    """roctx range markers (visible to rocprofv3 --marker-trace), or None."""
    import ctypes as C
    for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
        try:
            L = C.CDLL(name)
            L.roctxRangePushA.argtypes = [C.c_char_p]
            return L
        except OSError:
            continue
    return None


def _load():
    import __graft_entry__ as ge
    return ge._load_package()


def cpu_baseline(geo: str, freq: float, threads: int):
    """Oracle (C restatement of the reference CPU path) on a bounded sample:
    the fill of 1024 of the 8192 columns (scaled x8) and an FP64 LU + solve of
    a 4096 x 4096 complex system (scaled x8 for N^3)."""
    from oracle import oracle as O
    O.set_threads(threads)
    t = O.load(geo)
    ka = O.wavenumber(freq)
    n = t.npls
    cols = 1024
    t0 = time.perf_counter()
    O.zmatrix(t, ka, cols=(0, cols))
    t_fill = (time.perf_counter() - t0) * n / cols
    m = 4096
    rng = np.random.RandomState(0)
    A = rng.randn(m, m) + 1j * rng.randn(m, m)
    t0 = time.perf_counter()
    LU, piv, _ = O.zgetrf(A)
    O.zgetrs(LU, piv, np.ones(m, complex))
    t_lu = (time.perf_counter() - t0) * (n / m) ** 3
    per = t_fill + t_lu
    return {"value": 1.0 / per, "unit": "freq-solves/s", "cores": threads, "kind": "port",
            "sample": f"oracle fill of {cols}/{n} columns (x{n // cols}) + zgetrf/zgetrs at N={m} (x{(n / m) ** 3:.0f}); "
                      f"fill {t_fill:.1f} s + LU {t_lu:.1f} s per freq-solve (extrapolated)"}


def serving_leg(wm, ctx, tmp, count, rate):
    """Poisson request stream against the continuous-batching server: planar
    grids with N uniform in {512, ..., 8192}, 1-32 frequencies uniform in
    10 kHz - 50 MHz, 1-8 feeds, FP32 factors + FP64 refinement."""
    import wiremom.geometries as G
    import wiremom.serving as S
    reqs = S.poisson_requests(count, rate, seed=5)
    geoms, feeds = [], []
    for i, r in enumerate(reqs):
        p = os.path.join(tmp, f"req{i}.geo")
        G.planar_grid(p, r["n"], seed=r["seed"])
        g = wm.Geometry.load(ctx, p)
        geoms.append(g)
        feeds.append(np.random.RandomState(r["seed"]).choice(g.npls, r["nrhs"], replace=False) + 1)
    srv = wm.Server(ctx, "f32")
    out = S.replay(srv, reqs, geoms, feeds)
    srv.close()
    lat = np.array(out["latency"])
    nsolve = sum(len(r["freqs"]) for r in reqs)
    st = [s for res in out["results"] for s in res[2]]
    return {"requests": count, "arrival_rate_per_s": rate, "freq_solves": nsolve,
            "rhs_solves": int(sum(len(r["freqs"]) * r["nrhs"] for r in reqs)),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "max_latency_s": float(lat.max()), "solves_per_s": nsolve / out["makespan"],
            "makespan_s": out["makespan"], "fell_back": int(sum(s["fell_back"] == 1 for s in st)),
            "routed_f64": int(sum(s["fell_back"] == 2 for s in st)), "prec": "f32",
            "workload": "Poisson stream; planar grids N in {512,1024,2048,4096,8192}, 1-32 freqs "
                        "U(10 kHz, 50 MHz), 1-8 feeds"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--prec", default="f32")
    ap.add_argument("--nfreq", type=int, default=64)
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-bf16x3", action="store_true")
    ap.add_argument("--no-band", action="store_true", help="skip the 2.5 - 10 MHz band where every frequency refines")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch factorizations directly instead of replaying captured HIP graphs "
                         "(for rocprofv3, whose tracing of graph capture faults on this ROCm)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (three times), the reference for HBM counters")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rate", type=float, default=4.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(frequencies would shard across GPUs, one process per GPU, with no exchange "
              "beyond the geometry)", file=sys.stderr)
        sys.exit(2)

    wm = _load()
    import wiremom.geometries as G
    tmp = tempfile.mkdtemp(prefix="wm_bench_")
    geo = os.path.join(tmp, "c3.geo")
    n_target = 8192
    G.lightning(geo, n_target)
    freqs = G.freqs_c3()[: a.nfreq]
    ctx = wm.Context(0)
    if a.no_graphs:
        ctx.set_graphs(False)
    name, ncu, hbm = ctx.info()
    g = wm.Geometry.load(ctx, geo, os.path.join(tmp, "configwire.txt"))
    t = wm.configwire(geo, cwd=tmp)
    gap = G.nearest_basis(t, (0.0, 0.0, 19.9))

    def sweep(prec):
        return wm.solve_sweep(g, freqs, [gap], prec, batch=a.batch)

    for _ in range(a.warmup):
        sweep(a.prec)
    ctx.sync()
    ctx.prof_enable(True)
    rx = _roctx()
    if rx:
        rx.roctxRangePushA(b"timed")
    t0 = time.perf_counter()
    for _ in range(a.steps):
        zin, I, st = sweep(a.prec)
    ctx.sync()
    dt = time.perf_counter() - t0
    if rx:
        rx.roctxRangePop()
    prof = {k: ctx.prof_get(k) for k in CLASSES}
    ctx.prof_enable(False)
    value = a.steps * len(freqs) / dt

    # plain FP64 GPU path on the same sweep (the drop-in precision and the accuracy reference),
    # after one untimed sweep that captures its factorization graphs
    if a.prec != "f64":
        sweep("f64")
        ctx.sync()
        t0 = time.perf_counter()
        z64, _, _ = sweep("f64")
        dt64 = time.perf_counter() - t0
        fp64 = len(freqs) / dt64
        rel = np.abs(zin[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
    else:
        z64, fp64, rel = zin, value, np.zeros(len(freqs))

    # the same sweep with split-BF16x3 updates (precision ladder: BF16x3, FP32, FP64 per frequency)
    other = None
    if not a.no_bf16x3 and a.prec != "bf16x3":
        sweep("bf16x3")
        ctx.sync()
        t0 = time.perf_counter()
        zb, _, stb = sweep("bf16x3")
        dtb = time.perf_counter() - t0
        relb = np.abs(zb[:, 0] - z64[:, 0]) / np.abs(z64[:, 0])
        names = {0: "f64", 1: "f32", 2: "bf16x3"}
        other = {"prec": "bf16x3", "solves_per_s": len(freqs) / dtb, "speedup_vs_fp64": len(freqs) / dtb / fp64,
                 "zin_rel_err_vs_fp64_max": float(relb.max()),
                 "factor_prec_counts": {v: int(sum(s["prec"] == k for s in stb)) for k, v in names.items()},
                 "fell_back": int(sum(s["fell_back"] == 1 for s in stb)),
                 "ir_iters_mean": float(np.mean([s["iters"] for s in stb]))}

    # A band where every frequency refines (2.5 - 10 MHz on the same structure, 32 log-spaced
    # frequencies): the speedup of complex64 factors with FP64 refinement over FP64 factors that
    # the conditioning allows, beside the full band's (below about 1.4 MHz refinement cannot finish)
    band = None
    if not a.no_band:
        fb = np.logspace(np.log10(2.5), 1.0, 32).astype(np.float32)
        runs = {}
        for prec in ("f64", "f32", "bf16x3"):
            wm.solve_sweep(g, fb, [gap], prec, batch=a.batch)
            ctx.sync()
            t0 = time.perf_counter()
            zb_, _, stb_ = wm.solve_sweep(g, fb, [gap], prec, batch=a.batch)
            runs[prec] = (zb_, stb_, time.perf_counter() - t0)
        z64b, _, dt64b = runs["f64"]
        band = {"freqs_mhz": "32 log-spaced, 2.5 - 10 MHz", "fp64_solves_per_s": len(fb) / dt64b}
        for prec in ("f32", "bf16x3"):
            zb_, stb_, dtb_ = runs[prec]
            band[prec] = {"solves_per_s": len(fb) / dtb_, "speedup_vs_fp64": dt64b / dtb_,
                          "zin_rel_err_vs_fp64_max": float((np.abs(zb_[:, 0] - z64b[:, 0]) / np.abs(z64b[:, 0])).max()),
                          "routed_f64": int(sum(x["prec"] == 0 for x in stb_)),
                          "ir_iters_mean": float(np.mean([x["iters"] for x in stb_]))}

    # roofline of the dominant kernel class (largest event time)
    peak_mfma = PEAK_TFLOPS.get(a.prec, 157.3)
    dom = max(prof, key=lambda k: prof[k][0])
    ms, nl, fl, by = prof[dom]
    bound, peak = CLASSES[dom]
    if bound == "hbm":
        ach = by / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": peak, "unit": "GB/s", "frac": ach / peak}
    else:
        # flops of the kernel's algorithm: the complex128 and complex64 products
        # take three real products per complex multiply-add (6 flops); the same
        # update counted at the standard 8 flops per complex multiply-add is
        # reported beside as complex_equivalent
        ach = fl / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak}
        if dom in ("gemm", "gemm_f32"):
            roof["flops_per_complex_madd"] = 6
            roof["complex_equivalent_achieved"] = ach * 8.0 / 6.0
    traffic = GEMM_F64_HBM_RATIO * by / max(nl, 1) if dom == "gemm" else None
    roof.update({"kernel": dom, "launches": nl, "avg_launch_us": 1e3 * ms / max(nl, 1), "traffic": traffic,
                 "algorithmic_bytes_per_launch": by / max(nl, 1)})
    if traffic is not None:
        roof["traffic_source"] = "rocprofv3 FETCH_SIZE+WRITE_SIZE ratio, profiles/gemm_f64_traffic_c3_n8192.md"

    # whole-solve roofline fraction: t_ideal / t_measured per freq-solve, each
    # frequency's LU (8/3 N^3 real flops, the complex getrf count) priced at the
    # peak of the precision its factors used
    N = g.ld
    iters = float(np.mean([s["iters"] for s in st]))
    peak_of = {0: PEAK_TFLOPS["f64"], 1: PEAK_TFLOPS["f32"], 2: PEAK_TFLOPS["bf16x3"]}
    t_lu = float(np.mean([8 / 3 * N ** 3 / (peak_of[s["prec"]] * 1e12) for s in st]))
    t_ideal = (t_lu + 220 * N ** 2 / (PEAK_FP64_VALU * 1e12)
               + (16 * N ** 2 + iters * 24 * N ** 2) / (PEAK_HBM_GBS * 1e9))
    solve_frac = t_ideal / (dt / (a.steps * len(freqs)))

    serving = None if a.no_serving else serving_leg(wm, ctx, tmp, a.requests, a.rate)

    cpu = None
    if not a.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_baseline(geo, float(freqs[-1]), threads)

    out = {
        "metric": "freq-solves/s (fill+LU) at N=8192, FP32-MFMA+IR vs FP64; p99 request latency",
        "value": value,
        "unit": "freq-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "c64 LU + f64 IR", "f64": "c128", "bf16x3": "bf16x3 LU + f64 IR"}.get(a.prec, a.prec),
        "data": "synthetic lightning-protection geometry (generated), 64 log-spaced frequencies",
        "config": {"workload": "lightning structure N=8192, 64 freqs 10 kHz-10 MHz, 1 delta-gap feed",
                   "N": int(g.npls), "ld": int(g.ld), "nfreq": len(freqs), "prec": a.prec,
                   "parallelism": "1 GPU, frequencies batched"},
        "fp64_solves_per_s": fp64,
        "speedup_vs_fp64": value / fp64,
        "zin_rel_err_vs_fp64_max": float(rel.max()),
        "ir_iters_mean": iters,
        "fell_back": int(sum(s["fell_back"] == 1 for s in st)),
        "routed_f64": int(sum(s["fell_back"] == 2 for s in st)),
        "rel_err_estimate_max": float(max(s["dx_rel"] * s["rate"] / (1 - s["rate"]) if 0 <= s["rate"] < 1 else 0.0
                                          for s in st)),
        "solve_roofline_frac": solve_frac,
        "kernel_ms": {k: round(v[0] / a.steps, 3) for k, v in prof.items()},
        "kernel_algo": {k: {"launches": v[1], "flops": v[2], "bytes": v[3]} for k, v in prof.items() if v[1]},
        "roofline": roof,
        "bf16x3_sweep": other,
        "refinable_band": band,
        "serving": serving,
        "cpu_baseline": cpu,
        "reference_best_published_solves_per_s": REF_BEST_PUBLISHED,
        "device": f"{name} x{ncu} CU",
        "hip_graphs": not a.no_graphs,
    }
    if a.calibrate_copy:
        # plain device-to-device copies of known size (the runtime's copy kernel)
        import ctypes as C
        L = wm.lib()
        nbytes = a.calibrate_copy * 2 ** 20
        src, dst = C.c_void_p(), C.c_void_p()
        wm._check(L.wm_dev_alloc(ctx.h, nbytes, C.byref(src)))
        wm._check(L.wm_dev_alloc(ctx.h, nbytes, C.byref(dst)))
        for _ in range(3):
            wm._check(L.wm_memcpy_d2d(ctx.h, dst, src, nbytes))
        wm._check(L.wm_dev_free(ctx.h, src))
        wm._check(L.wm_dev_free(ctx.h, dst))
        out["calibrate_copy_bytes"] = nbytes
    print(json.dumps(out))


if __name__ == "__main__":
    main()
