"""Benchmark: freq-solves/s (fill + LU + solve) at N = 8192 on one MI355X.

Workload (the lightning-protection structure): 4 down-conductors 20 m tall, a
10 x 10 m roof mesh at 1 m pitch and a ground ring, radius 5 mm, segmented to
N = 8192 unknowns, delta-gap feed just below the top corner, 64 frequencies
log-spaced over 10 kHz - 10 MHz.  Synthetic geometry generated in-process.

One step = one sweep over the 64 frequencies: fill, factor, solve and Zin for
every frequency, with the geometry resident in HBM.  `value` is the sweep in
the configuration given by --prec/--recycle (default: recycled, i.e. FP64
factors only at anchor frequencies and FP64 GMRES refinement with the nearest
anchor's factors at the others); it is timed with the profiler off and the
factorizations replayed from HIP graphs.  The plain FP64 GPU path (the drop-in:
one LU per frequency) is timed the same way beside it, and every Zin is compared
with it together with a per-frequency condition estimate.

Also reported: the mixed-precision sweeps without recycling (FP32 and
split-BF16x3 factors + FP64 refinement), a band where every frequency refines
in complex64, the reference's own 127-frequency grid (1 kHz - 10 MHz) recycled
on the same structure, per-kernel times and the roofline of the dominant kernel (one
extra profiled step: HIP events around every launch on its stream), the other
configurations (C1 dipole, C2 planar grid, C4 3-D cage), a serving leg and the
CPU oracle on bounded samples of the same workloads.

    python bench.py [--steps K] [--warmup W] [--prec f32|f64|bf16x3] [--recycle R]
                    [--no-cpu] [--no-serving] [--no-configs] [--serving-spec]
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
PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2516.6}
PEAK_HBM_GBS = 8000.0
# kernel classes timed by the library profiler: (bound, peak, unit)
CLASSES = {
    "fill": ("fp64", 78.6), "panel": ("fp64", 78.6), "swap_trsm": ("hbm", PEAK_HBM_GBS),
    "gemm": ("mfma", 78.6), "panel_f32": ("fp32", 157.3), "swap_trsm_f32": ("hbm", PEAK_HBM_GBS),
    "gemm_f32": ("mfma", 157.3),
    # split-BF16x3 update: flops counted FP32-equivalent (8 per complex multiply-add), each real
    # product issued as three bf16 products, so its peak is the bf16 peak / 3
    "gemm_bf16x3": ("mfma", 2516.6 / 3),
    "trsv": ("hbm", PEAK_HBM_GBS), "resid": ("hbm", PEAK_HBM_GBS),
}
# HBM bytes / algorithmic bytes of k_gemm_f64, from FETCH_SIZE + WRITE_SIZE passes
# calibrated on a plain copy (profiles/gemm_f64_traffic_c3_n8192.md)
GEMM_F64_HBM_RATIO = 1.91
# the same for the FP64 matrix products of GMRES (k_resid_part), over every launch
# of the value sweep (profiles/resid_traffic_c3_n8192.md)
RESID_HBM_RATIO = 1.00
HBM_RATIO = {"gemm": (GEMM_F64_HBM_RATIO, "profiles/gemm_f64_traffic_c3_n8192.md"),
             "resid": (RESID_HBM_RATIO, "profiles/resid_traffic_c3_n8192.md")}
# tuning cache written by `bench.py --tune` on the benchmark's device (JSON lines, wm_tune_save)
TUNE_CACHE = os.path.join(ROOT, "praca-magisterska_amd", "tune_cache.json")
REF_BEST_PUBLISHED = 0.148  # freq-solves/s at N=8192, reference's best variant (GPU fill + MKL LU, 6 procs)
U64 = 2.0 ** -53


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


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "threads_used": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "cpu_model": model}


_T0 = time.perf_counter()


def progress(stage: str):
    """One line on stderr per finished stage (the run is minutes long)."""
    print(f"bench: {stage} done at {time.perf_counter() - _T0:.0f} s", file=sys.stderr, flush=True)


def timed(ctx, fn, steps, warmup):
    """Mean wall time of fn over `steps` calls after `warmup` untimed ones, the
    device synchronised on both sides; returns (seconds per call, last result)."""
    for _ in range(warmup):
        fn()
    ctx.sync()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    ctx.sync()
    return (time.perf_counter() - t0) / steps, out


def profiled(ctx, fn):
    """One call of fn with every launch of the timed kernel classes bracketed by
    HIP events on its stream (factorizations launched directly, not from graphs).
    Returns {class: (ms, launches, flops, bytes)}."""
    ctx.sync()
    ctx.prof_enable(True)
    fn()
    ctx.sync()
    prof = {k: ctx.prof_get(k) for k in CLASSES}
    ctx.prof_enable(False)
    return prof


def roofline(prof):
    """Roofline of the kernel class with the largest event time: achieved =
    algorithmic flops (or bytes) of its launches / their summed event time."""
    dom = max(prof, key=lambda k: prof[k][0])
    ms, nl, fl, by = prof[dom]
    bound, peak = CLASSES[dom]
    if bound == "hbm":
        ach = by / (ms * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": ach, "peak": peak, "unit": "GB/s", "frac": ach / peak}
    else:
        ach = fl / (ms * 1e-3) / 1e12
        r = {"bound": "mfma" if bound == "mfma" else bound, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
             "frac": ach / peak}
        if dom in ("gemm", "gemm_f32"):
            # the complex products take three real products per complex multiply-add
            # (6 flops); the same update at the standard 8 flops is reported beside
            r["flops_per_complex_madd"] = 6
            r["complex_equivalent_achieved"] = ach * 8.0 / 6.0
    ratio, src = HBM_RATIO.get(dom, (None, None))
    traffic = ratio * by / max(nl, 1) if ratio else None
    r.update({"kernel": dom, "launches": nl, "avg_launch_us": 1e3 * ms / max(nl, 1), "traffic": traffic,
              "algorithmic_bytes_per_launch": by / max(nl, 1), "algorithmic_flops_per_launch": fl / max(nl, 1)})
    if traffic is not None:
        r["traffic_source"] = "rocprofv3 FETCH_SIZE+WRITE_SIZE ratio, " + src
    return r


def stats_summary(st):
    names = {0: "f64", 1: "f32", 2: "bf16x3"}
    out = {"factor_prec_counts": {v: int(sum(s["prec"] == k for s in st)) for k, v in names.items()},
           "fell_back": int(sum(s["fell_back"] == 1 for s in st)),
           "routed_f64": int(sum(s["fell_back"] == 2 for s in st)),
           "iters_mean": float(np.mean([s["iters"] for s in st]))}
    anchors = sum(1 for s in st if s["anchor_mhz"] > 0 and s["iters"] == 0)
    if any(s["anchor_mhz"] > 0 for s in st):
        out["anchors"] = anchors
        out["gmres_iters_max"] = int(max(s["iters"] for s in st))
    return out


def compare(z, z64, kappa):
    """Zin of a sweep against the FP64 sweep, split by the condition number:
    where kappa_1 u64 <= 1e-11 the bar is 1e-10; elsewhere two backward-stable
    FP64 solutions may differ by about kappa_1 u64 (reported as a ratio)."""
    rel = np.abs(z - z64) / np.abs(z64)
    ku = np.asarray(kappa) * U64
    good = ku <= 1e-11
    out = {"zin_rel_err_vs_fp64_max": float(rel.max()),
           "zin_rel_err_vs_fp64_max_where_kappa_u_le_1e-11": float(rel[good].max()) if good.any() else None,
           "freqs_with_kappa_u_le_1e-11": int(good.sum())}
    if (~good).any():
        out["max_rel_err_over_kappa_u_elsewhere"] = float((rel[~good] / ku[~good]).max())
    return out


def cpu_baseline(geo, freq, threads, fill_cols=(64, 1024), lu_m=4096):
    """Oracle (C restatement of the reference CPU path) on a bounded sample of the
    workload: the fill of a few columns single-threaded (as the reference's
    zmatrix) and on all host threads, each scaled to all N columns, and an FP64
    LU + solve of an lu_m x lu_m system scaled by (N / lu_m)^3; scipy's LAPACK
    zgetrf (if present) is timed on the same system as a library-grade stand-in
    for the reference's MKL."""
    from oracle import oracle as O
    t = O.load(geo)
    ka = O.wavenumber(freq)
    n = t.npls
    out = {}
    for label, th, cols in (("fill_1thread_s", 1, fill_cols[0]), ("fill_allthreads_s", threads, fill_cols[1])):
        cols = min(cols, n)
        O.set_threads(th)
        t0 = time.perf_counter()
        O.zmatrix(t, ka, cols=(0, cols))
        out[label] = (time.perf_counter() - t0) * n / cols
    O.set_threads(threads)
    m = min(lu_m, n)
    rng = np.random.RandomState(0)
    A = rng.randn(m, m) + 1j * rng.randn(m, m)
    t0 = time.perf_counter()
    LU, piv, _ = O.zgetrf(A)
    O.zgetrs(LU, piv, np.ones(m, complex))
    out["lu_oracle_s"] = (time.perf_counter() - t0) * (n / m) ** 3
    try:
        import scipy.linalg as sl
        t0 = time.perf_counter()
        sl.lu_solve(sl.lu_factor(A), np.ones(m, complex))
        out["lu_lapack_s"] = (time.perf_counter() - t0) * (n / m) ** 3
    except ImportError:
        pass
    per = out["fill_allthreads_s"] + out["lu_oracle_s"]
    return {"value": 1.0 / per, "unit": "freq-solves/s", "cores": threads, "kind": "port",
            "sample": f"N={n}: oracle fill of {fill_cols[0]} columns on 1 thread and {min(fill_cols[1], n)} on "
                      f"{threads}, zgetrf/zgetrs at N={m}, scaled to N (fill x N/cols, LU x (N/{m})^3); "
                      f"value = 1 / (fill on all threads + oracle LU)",
            **{k: round(v, 3) for k, v in out.items()}, **host_info()}


def serving_leg(wm, ctx, tmp, count, rate, warm=0, sizes=(512, 1024, 2048, 4096, 8192), progress_every=0):
    """Poisson request stream against the continuous-batching server: planar
    grids with N uniform in `sizes`, 1-32 frequencies uniform in 10 kHz - 50 MHz,
    1-8 feeds, FP32 factors + FP64 refinement.  The first `warm` requests are
    not counted in the latency statistics."""
    import wiremom.geometries as G
    import wiremom.serving as S
    reqs = S.poisson_requests(count + warm, rate, seed=5, sizes=sizes)
    geoms, feeds = [], []
    for i, r in enumerate(reqs):
        p = os.path.join(tmp, f"req{i}.geo")
        G.planar_grid(p, r["n"], seed=r["seed"])
        g = wm.Geometry.load(ctx, p)
        geoms.append(g)
        feeds.append(np.random.RandomState(r["seed"]).choice(g.npls, r["nrhs"], replace=False) + 1)
    srv = wm.Server(ctx, "f32")
    out = S.replay(srv, reqs, geoms, feeds, progress_every=progress_every)
    batches, busy, over = srv.overlap()
    srv.close()
    for g in geoms:
        g.close()
    lat = np.array(out["latency"][warm:])
    meas = reqs[warm:]
    nsolve = sum(len(r["freqs"]) for r in meas)
    st = [s for res in out["results"][warm:] for s in res[2]]
    span = out["makespan"] - (reqs[warm]["t"] if warm else 0.0)
    return {"requests": count, "warmup_requests": warm, "arrival_rate_per_s": rate, "freq_solves": nsolve,
            "rhs_solves": int(sum(len(r["freqs"]) * r["nrhs"] for r in meas)),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "max_latency_s": float(lat.max()), "solves_per_s": nsolve / span,
            "makespan_s": span, "fell_back": int(sum(s["fell_back"] == 1 for s in st)),
            "routed_f64": int(sum(s["fell_back"] == 2 for s in st)), "prec": "f32",
            "lanes": 2, "batches": batches, "lanes_busy_s": busy, "lanes_overlapped_s": over,
            "workload": f"Poisson stream; planar grids N in {list(sizes)}, 1-32 freqs U(10 kHz, 50 MHz), 1-8 feeds"}


def config_line(wm, ctx, tmp, name, make, freqs, feeds_of, recycle, cpu, threads, steps=2):
    """Solves/s of one configuration: the plain FP64 path, FP32 and split-BF16x3
    factors + refinement and the recycled sweep, each timed after one untimed
    call, with its Zin compared with the FP64 sweep's and the roofline of its
    dominant kernel."""
    geo = os.path.join(tmp, f"{name}.geo")
    make(geo)
    g = wm.Geometry.load(ctx, geo)
    t = wm.configwire(geo, cwd=tmp)
    gaps = feeds_of(t)
    res = {"N": int(g.npls), "nfreq": len(freqs), "nrhs": len(gaps)}
    runs = {}
    kinds = {"f64": dict(prec="f64"), "f32_ir": dict(prec="f32"), "bf16x3_ir": dict(prec="bf16x3"),
             "recycled": dict(prec="f64", recycle=recycle)}
    for label, kw in kinds.items():
        dt, (z, _, st) = timed(ctx, lambda: wm.solve_sweep(g, freqs, gaps, **kw), steps, 1)
        runs[label] = (z, st, dt)
        res[label] = {"solves_per_s": len(freqs) / dt, "ms_per_sweep": 1e3 * dt}
        progress(f"{name} {label}")
    _, _, s64 = wm.solve_sweep(g, freqs, gaps, "f64", cond_est=True)
    kappa = [s["cond1"] for s in s64]
    res["kappa1_range"] = [float(min(kappa)), float(max(kappa))]
    z64 = runs["f64"][0]
    for label in ("f32_ir", "bf16x3_ir", "recycled"):
        z, st, dt = runs[label]
        res[label].update(compare(z[:, 0] if z.ndim > 1 else z, z64[:, 0], kappa))
        res[label].update(stats_summary(st))
        res[label]["speedup_vs_fp64"] = runs["f64"][2] / dt
    best = max(kinds, key=lambda k: res[k]["solves_per_s"])
    prof = profiled(ctx, lambda: wm.solve_sweep(g, freqs, gaps, **kinds[best]))
    res["roofline"] = roofline(prof)
    res["roofline"]["of"] = best
    if cpu:
        res["cpu_baseline"] = cpu_baseline(geo, float(freqs[-1]), threads,
                                           fill_cols=(min(64, g.npls), min(1024, g.npls)),
                                           lu_m=min(4096, g.npls))
        progress(f"{name} CPU baseline")
    g.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--prec", default="f64", help="factor precision of the value sweep (f64 with --recycle: "
                                                   "FP64 anchor factors + FP64 GMRES refinement)")
    ap.add_argument("--recycle", type=float, default=2.5,
                    help="value sweep: anchors within this frequency factor (0: one LU per frequency)")
    ap.add_argument("--nfreq", type=int, default=64)
    ap.add_argument("--refinable", action="store_true",
                    help="with --profile-only: sweep the refinable band (32 frequencies 2.5 - 10 MHz) instead")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the C1, C2 and C4 lines")
    ap.add_argument("--no-band", action="store_true", help="skip the 2.5 - 10 MHz band where every frequency refines")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch factorizations directly instead of replaying captured HIP graphs")
    ap.add_argument("--no-overlap", action="store_true",
                    help="recycled sweeps: fill on the launch stream only (every kernel's time its own)")
    ap.add_argument("--serving-spec", action="store_true",
                    help="only the serving stream at its specification: 1000 requests after 50 warm-up at "
                         "0.8 x the measured capacity")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the warmup and timed steps of the value sweep (for rocprofv3), then a short JSON line")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (three times), the reference for HBM counters")
    ap.add_argument("--tune", action="store_true",
                    help="run the tuners (fill, and the factorization depth per precision) on the benchmark "
                         "structure and write the cache to --tune-cache")
    ap.add_argument("--tune-cache", default=TUNE_CACHE,
                    help="tuning cache (JSON lines): loaded if present, written by --tune")
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
    ctx = wm.Context(0)
    if a.no_graphs:
        ctx.set_graphs(False)
    if a.no_overlap:
        ctx.set_fill_overlap(False)
    name, ncu, hbm = ctx.info()
    if not name.split("(")[0].strip() and "gfx950" in name:
        name = "AMD Instinct MI355X " + name.strip()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    if a.serving_spec:
        print(json.dumps(serving_spec(wm, ctx, tmp, name)))
        return

    geo = os.path.join(tmp, "c3.geo")
    G.lightning(geo, 8192)
    freqs = G.freqs_c3()[: a.nfreq]
    if a.refinable:
        freqs = np.logspace(np.log10(2.5), 1.0, 32).astype(np.float32)
    g = wm.Geometry.load(ctx, geo, os.path.join(tmp, "configwire.txt"))
    tuning = {"cache": None, "entries_applied": 0}
    if a.tune:
        tuned = {"fill": ctx.tune_fill(g)}
        for prec in ("f64", "f32", "bf16x3"):
            tuned[prec] = ctx.tune_getrf(prec, g.ld)
        tuning = {"cache": a.tune_cache, "entries_written": ctx.tune_save(a.tune_cache),
                  "tuned_ms": {k: round(v[1], 3) for k, v in tuned.items()}}
        print(f"tuned: {tuned}", file=sys.stderr, flush=True)
    elif a.tune_cache and os.path.exists(a.tune_cache):
        tuning = {"cache": os.path.relpath(a.tune_cache, ROOT), "entries_applied": ctx.tune_load(a.tune_cache)}
    tuning["getrf_block"] = {p_: ctx.getrf_block(p_, g.ld) for p_ in ("f64", "f32", "bf16x3")}
    tuning["fill_freqs_per_launch"] = ctx.fill_freqs(1, g.ld) or 4
    t = wm.configwire(geo, cwd=tmp)
    gap = G.nearest_basis(t, (0.0, 0.0, 19.9))

    def sweep(prec, recycle=0.0, fb=None):
        return wm.solve_sweep(g, freqs if fb is None else fb, [gap], prec, batch=a.batch, recycle=recycle)

    # value: the timed steps, profiler off, factorizations replayed from graphs
    rx = _roctx()
    for _ in range(a.warmup):
        sweep(a.prec, a.recycle)
    ctx.sync()
    if rx:
        rx.roctxRangePushA(b"timed")
    t0 = time.perf_counter()
    for _ in range(a.steps):
        zin, I, st = sweep(a.prec, a.recycle)
    ctx.sync()
    dt = time.perf_counter() - t0
    if rx:
        rx.roctxRangePop()
    value = a.steps * len(freqs) / dt
    progress(f"value sweep ({value:.1f} freq-solves/s)")
    if a.profile_only:
        prof = profiled(ctx, lambda: sweep(a.prec, a.recycle))
        out = {"value": value, "unit": "freq-solves/s", "steps": a.steps, "warmup": a.warmup,
               "recycle": a.recycle, "prec": a.prec, "hip_graphs": not a.no_graphs, "fill_overlap": not a.no_overlap, "tuning": tuning,
               "roofline": roofline(prof), "kernel_ms_profiled_step": {k: round(v[0], 3) for k, v in prof.items()}}
        if a.calibrate_copy:
            out["calibrate_copy_bytes"] = calibrate_copy(wm, ctx, a.calibrate_copy)
        print(json.dumps(out))
        return

    # the plain FP64 GPU path (one LU per frequency, the drop-in) timed the same way
    ref_steps = max(2, a.steps // 2)
    dt64, (z64, _, _) = timed(ctx, lambda: sweep("f64"), ref_steps, 1)
    fp64 = len(freqs) / dt64
    _, _, s64 = wm.solve_sweep(g, freqs, [gap], "f64", cond_est=True)
    kappa = [s["cond1"] for s in s64]
    progress("FP64 sweep and condition estimates")

    # mixed precision without recycling: FP32 and split-BF16x3 factors + FP64 refinement (ladder)
    mixed = {}
    for prec in ("f32", "bf16x3"):
        dtm, (zm, _, stm) = timed(ctx, lambda: sweep(prec), 2, 1)
        mixed[prec] = {"solves_per_s": len(freqs) / dtm, "speedup_vs_fp64": dt64 / dtm,
                       **compare(zm[:, 0], z64[:, 0], kappa), **stats_summary(stm)}
        progress(f"mixed sweep {prec}")

    # a band where every frequency refines in complex64 (2.5 - 10 MHz, 32 log-spaced)
    band = None
    if not a.no_band:
        fb = np.logspace(np.log10(2.5), 1.0, 32).astype(np.float32)
        runs = {}
        for prec in ("f64", "f32", "bf16x3"):
            d_, (zb_, _, stb_) = timed(ctx, lambda: sweep(prec, fb=fb), 2, 1)
            runs[prec] = (zb_, stb_, d_)
            progress(f"refinable band {prec}")
        d_, (zb_, _, stb_) = timed(ctx, lambda: sweep("f64", a.recycle, fb=fb), 2, 1)
        runs["recycled"] = (zb_, stb_, d_)
        z64b, _, dt64b = runs["f64"]
        band = {"freqs_mhz": "32 log-spaced, 2.5 - 10 MHz", "fp64_solves_per_s": len(fb) / dt64b}
        for prec in ("f32", "bf16x3", "recycled"):
            zb_, stb_, dtb_ = runs[prec]
            band[prec] = {"solves_per_s": len(fb) / dtb_, "speedup_vs_fp64": dt64b / dtb_,
                          "zin_rel_err_vs_fp64_max": float((np.abs(zb_[:, 0] - z64b[:, 0]) / np.abs(z64b[:, 0])).max()),
                          **stats_summary(stb_)}

    # the reference's own lightning grid ("wersja referencyjna/freq1.txt": 127
    # frequencies, 1 kHz - 10 MHz, piecewise uniform), recycled like the value
    fr = np.concatenate([np.arange(1, 11) * 1e-3, np.arange(2, 11) * 1e-2, 0.1 + np.arange(1, 19) * 0.05,
                         1.0 + np.arange(1, 11) * 0.1, 2.0 + np.arange(1, 81) * 0.1]).astype(np.float32)
    dtr, (_, _, str_) = timed(ctx, lambda: sweep(a.prec, a.recycle, fb=fr), 2, 1)
    ref_grid = {"freqs_mhz": "127, 1 kHz - 10 MHz (the reference's freq1.txt grid)", "solves_per_s": len(fr) / dtr,
                **stats_summary(str_)}
    progress("reference grid")

    # one profiled step of the value sweep (events around every launch; no graphs)
    prof = profiled(ctx, lambda: sweep(a.prec, a.recycle))
    roof = roofline(prof)
    prof64 = profiled(ctx, lambda: sweep("f64"))
    progress("profiled steps")

    # whole-sweep roofline fraction: t_ideal / t_measured per freq-solve, each
    # factored frequency's LU (8/3 N^3 real flops) at the peak of its precision,
    # the fill at the FP64 peak, and the refinement's matrix reads at HBM speed
    N = g.ld
    peak_of = {0: PEAK["f64"], 1: PEAK["f32"], 2: PEAK["bf16"] / 3}
    own = [s for s in st if s["anchor_mhz"] == 0 or s["iters"] == 0 or s["fell_back"]]
    t_ideal = (sum(8 / 3 * N ** 3 / (peak_of[s["prec"]] * 1e12) for s in own)
               + len(st) * 220 * N ** 2 / (PEAK["f64"] * 1e12)
               + sum(16 * N ** 2 * (1 + s["iters"]) for s in st) / (PEAK_HBM_GBS * 1e9)) / len(st)
    solve_frac = t_ideal / (dt / (a.steps * len(freqs)))

    configs = None
    if not a.no_configs:
        configs = {
            "C1_dipole_N511_150MHz": config_line(
                wm, ctx, tmp, "c1", lambda p: G.dipole(p, 512, 1.0, 0.001), np.array([150.0], np.float32),
                lambda t_: [256], a.recycle, not a.no_cpu, threads, steps=5),
            "C2_planar_grid_N4096_12freqs": config_line(
                wm, ctx, tmp, "c2", lambda p: G.planar_grid(p, 4096, seed=2), G.freqs_c2(),
                lambda t_: [G.nearest_basis(t_, (5.0, 5.0, 0.0))], a.recycle, not a.no_cpu, threads),
            "C4_cage_N14336_16freqs_8rhs": config_line(
                wm, ctx, tmp, "c4", lambda p: G.cage(p, 14336, seed=4), G.freqs_c4(),
                lambda t_: list(np.random.RandomState(4).choice(t_.npls, 8, replace=False) + 1), a.recycle,
                not a.no_cpu, threads, steps=1),
        }

    serving = None if a.no_serving else serving_leg(wm, ctx, tmp, a.requests, a.rate)
    progress("serving leg")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(geo, float(freqs[-1]), threads)
        progress("CPU baseline")

    recyc = a.recycle > 1.0
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
        "dtype": ("c128 LU at anchor freqs + f64 GMRES-IR" if recyc else
                  {"f32": "c64 LU + f64 IR", "f64": "c128", "bf16x3": "bf16x3 LU + f64 IR"}.get(a.prec, a.prec)),
        "data": "synthetic lightning-protection geometry (generated), 64 log-spaced frequencies",
        "config": {"workload": "lightning structure N=8192, 64 freqs 10 kHz-10 MHz, 1 delta-gap feed",
                   "N": int(g.npls), "ld": int(g.ld), "nfreq": len(freqs), "prec": a.prec,
                   "recycle": a.recycle, "parallelism": "1 GPU, frequencies batched"},
        "fp64_solves_per_s": fp64,
        "speedup_vs_fp64": value / fp64,
        **compare(zin[:, 0], z64[:, 0], kappa),
        "kappa1_per_freq": [float("%.3g" % k) for k in kappa],
        **stats_summary(st),
        "solve_roofline_frac": solve_frac,
        "kernel_ms_profiled_step": {k: round(v[0], 3) for k, v in prof.items()},
        "kernel_ms_fp64_profiled_step": {k: round(v[0], 3) for k, v in prof64.items()},
        "roofline": roof,
        "roofline_fp64_path": roofline(prof64),
        "mixed_no_recycling": mixed,
        "refinable_band": band,
        "reference_grid_127": ref_grid,
        "configs": configs,
        "serving": serving,
        "cpu_baseline": cpu,
        "reference_best_published_solves_per_s": REF_BEST_PUBLISHED,
        "device": f"{name} x{ncu} CU",
        "hip_graphs": not a.no_graphs, "fill_overlap": not a.no_overlap,
        "tuning": tuning,
        "timing": "value and every solves/s figure: wall time of whole sweeps, profiler off; per-kernel times "
                  "and rooflines from one separate profiled step",
    }
    if a.calibrate_copy:
        out["calibrate_copy_bytes"] = calibrate_copy(wm, ctx, a.calibrate_copy)
    print(json.dumps(out))


def calibrate_copy(wm, ctx, mib):
    """Three plain device-to-device copies of mib MiB (the runtime's copy kernel),
    the reference for HBM byte counters."""
    import ctypes as C
    L = wm.lib()
    nbytes = mib * 2 ** 20
    src, dst = C.c_void_p(), C.c_void_p()
    wm._check(L.wm_dev_alloc(ctx.h, nbytes, C.byref(src)))
    wm._check(L.wm_dev_alloc(ctx.h, nbytes, C.byref(dst)))
    for _ in range(3):
        wm._check(L.wm_memcpy_d2d(ctx.h, dst, src, nbytes))
    wm._check(L.wm_dev_free(ctx.h, src))
    wm._check(L.wm_dev_free(ctx.h, dst))
    return nbytes


def serving_spec(wm, ctx, tmp, name):
    """C5 at its specification: capacity measured by a saturated stream (all
    requests at once), then 1000 requests after 50 warm-up at 0.8 x capacity."""
    cap = serving_leg(wm, ctx, tmp, 60, 1e6, progress_every=20)
    req_per_s = cap["requests"] / cap["makespan_s"]
    rate = 0.8 * req_per_s
    print(f"serving spec: capacity {req_per_s:.2f} requests/s, replaying at {rate:.2f}/s", file=sys.stderr, flush=True)
    res = serving_leg(wm, ctx, tmp, 1000, rate, warm=50, progress_every=50)
    return {"metric": "C5 serving: p50/p99 request latency at 0.8 x measured capacity", "device": name,
            "capacity_requests_per_s": req_per_s, "capacity_solves_per_s": cap["solves_per_s"],
            "arrival_rate_per_s": rate, **res}


if __name__ == "__main__":
    main()
