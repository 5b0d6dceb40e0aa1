"""Benchmark: CI ground-state solves/s, p50/p99 latency and SpMV roofline on one MI355X.

Default workload (--workload stream): the served stream of rotated-integral
ground-state solves -- 70% DOCI with L uniform in [12, 20] and 30% FCI with L
uniform in [8, 11], half filling, integrals from seed 1000+k rotated by
U = exp(K), K antisymmetric with N(0, 0.05^2) entries.  Each request's
precision is CF_PREC_AUTO (--served-precision): FP32 values + FP64 polish
where its operator stores matrix values, FP64 where it does not (every route
of this mix: the on-the-fly sigma, the block-staged and one-launch DOCI
products), and it stops at a Kato-Temple energy estimate of 1e-9 Eh.  The timed stream holds
--requests (4,096) requests; one step = one slice of requests/steps
consecutive requests, and the K slices are served back to back.
  * value: solves/s of the whole stream submitted at once (closed batch) to
    the continuously batched server (--slots worker streams), first
    submission to last completion,
  * p50/p99: latency (submit -> result) of Poisson arrivals at 70% of that
    rate, over the first 1,024 requests,
  * one_at_a_time_fp64: the plain path (one request at a time, FP64, ARPACK
    tolerance 1e-12) on every 16th request of each (kind, L) class (256+
    requests), its class-weighted rate over the whole stream, and max |dE|
    against the served energies,
  * roofline: the stored-CRS SpMV of FCI L=12 6a6b (1,553,018,544 nonzeros),
    timed with HIP events on the library stream during one FP64 solve.
--workload fci12: FCI L=12 6a6b seed 1 build + solve per step.
--workload spmv_dets: the stored SpMV of a CAS-like determinant list (--L 12:
101,640 determinants; --L 16: 1,029,600).

Prints ONE JSON line.  cpu_baseline times the oracle's restatement of the
reference CPU path on a bounded sample and extrapolates (see cpu_baseline_*).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "CI solves/s + p50/p99 latency at |dE|<=1e-8 Eh; SpMV HBM GB/s (% of peak)"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=["stream", "fci12", "doci24", "direct16", "spmv", "spmv_doci", "spmv_dets",
                                          "rotate", "sigma", "batch"],
                   default="stream")
    p.add_argument("--steps", type=int, default=None,
                   help="stream: slices of the --requests stream (16); fci12: solves (4)")
    p.add_argument("--warmup", type=int, default=None, help="stream: warm-up slices (2); fci12: solves (1)")
    p.add_argument("--requests", type=int, default=4096, help="stream: requests in the timed stream")
    p.add_argument("--latency-requests", type=int, default=1024, help="stream: requests in the Poisson run")
    p.add_argument("--slots", type=int, default=8)
    p.add_argument("--load", type=float, default=0.7, help="Poisson rate as a fraction of the measured rate")
    p.add_argument("--rate", type=float, default=0.0, help="Poisson rate (solves/s); overrides --load")
    p.add_argument("--baseline-every", type=int, default=16,
                   help="one-at-a-time FP64 baseline on every n-th request of each class (0: skip)")
    p.add_argument("--no-latency", action="store_true", help="skip the Poisson latency run")
    p.add_argument("--L", type=int, default=12)
    p.add_argument("--sigma-L", default="8,9,10,11,12,16", help="sigma: orbital counts (FCI, half filling)")
    p.add_argument("--batch-class", default="FCI11", help="batch: request class (FCI8..FCI11, DOCI17..DOCI20)")
    p.add_argument("--lanes", type=int, default=32, help="batch: requests per lock-step batch")
    p.add_argument("--served-precision", choices=["auto", "f32", "f64"], default="auto",
                   help="stream: per-request precision (auto: FP32 values + FP64 polish only where the "
                        "operator stores matrix values, FP64 elsewhere)")
    p.add_argument("--batch-precision", choices=["f32", "f64"], default="f32",
                   help="batch: FP32 operator + FP64 polish (f32) or FP64 throughout (f64)")
    p.add_argument("--nel", type=int, default=6)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--precision", choices=["f64", "f32"], default="f64")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--no-fp64", action="store_true", help="direct16: skip the FP64 solve")
    p.add_argument("--cpu-lines", type=int, default=6, help="alpha lines in the FCI CPU-baseline sample")
    p.add_argument("--solver", choices=["auto", "lanczos", "davidson"], default="auto",
                   help="stream: served solver (auto: DOCI up to 13,312 determinants in batched one-launch "
                        "solves, Davidson for FCI and larger DOCI)")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baselines (oracle = the restatement of the reference CPU path)
# ---------------------------------------------------------------------------
def _oracle_threads():
    from oracle import oracle as O
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or O.num_threads()


def _sample_build_mv(kind, h, V, nel, ranges, n, threads):
    """Build the CRS rows of the picked line ranges (build_parallel + PushToRow)
    and run the serial upper-triangle mvprod over them; returns (t_build,
    t_mvprod, sampled rows, sampled upper nnz)."""
    from oracle import oracle as O
    x = np.random.default_rng(0).standard_normal(n)
    y = np.zeros(n)
    tb = tm = 0.0
    rows = nnz = 0
    for s0, s1 in ranges:
        t0 = time.perf_counter()
        m = O.build_crs(kind, h, V, nel, nel, start=int(s0), end=int(s1), threads=threads)
        tb += time.perf_counter() - t0
        rows += m.n
        nnz += m.nnz
        t0 = time.perf_counter()
        O.lib().or_mvprod_sym(m.n, m.row.ctypes.data_as(O.u64p), m.col.ctypes.data_as(O.u32p),
                              m.data.ctypes.data_as(O.dp), x.ctypes.data_as(O.dp), y.ctypes.data_as(O.dp))
        tm += time.perf_counter() - t0
    return tb, tm, rows, nnz


def _line_ranges(total, count, width=1):
    """`count` evenly spaced ranges of `width` lines in [0, total)."""
    width = max(1, min(width, total // max(count, 1)))
    starts = np.linspace(0, total - width, min(count, total)).round().astype(int)
    return [(int(a), int(a) + width) for a in starts]


def cpu_baseline_fci(h, V, L, nel, nnz_upper, n, n_matvec, lines):
    """FCI::construct_CI_matrix with PushToRow (build_parallel over the host
    threads) for `lines` evenly spaced alpha strings, plus the serial
    upper-triangle SparseMatrix_CRS_Sym::mvprod (the reference's g++ path) on
    that slice; extrapolated per nonzero to one solve = build + n_matvec products."""
    from oracle import oracle as O
    threads = _oracle_threads()
    da = O.binom(L, nel)
    tb, tm, rows, nnz_s = _sample_build_mv("FCI", h, V, nel, _line_ranges(da, lines), n, threads)
    scale = nnz_upper / nnz_s
    t_solve = tb * scale + n_matvec * tm * scale
    return {
        "value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": (f"oracle FCI::construct_CI_matrix (PushToRow, build_parallel on {threads} threads) + serial "
                   f"upper-triangle mvprod over {lines} alpha lines ({rows} rows, {nnz_s} upper nnz); "
                   f"extrapolated by nnz to one solve = build + {n_matvec} mvprod"),
        "build_s_est": tb * scale, "mvprod_s_est": tm * scale, "sample_s": tb + tm,
    }


def _median3(f):
    """Median wall time of three calls of f (the timing rule of the CPU baseline)."""
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        f()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline_stream(reqs, nmv_by_class, lines=48):
    """Per (kind, L) class of the stream, the reference's one-at-a-time CPU path
    on a bounded sample: the oracle's rotation of the integrals
    (OrbitalTransform::fillHamCI restated), the CRS build of `lines` evenly
    spaced alpha lines (DOCI: 16 blocks of <= 4096 rows) with PushToRow and
    build_parallel on all host threads, and two products over those rows:
    the serial upper-triangle SparseMatrix_CRS_Sym::mvprod (the reference's
    g++ path, SparseMatrix_CRS.cpp:712-720) and an OpenMP full-row gather (the
    open stand-in for the Intel build's threaded mkl_cspblas_dcsrsymv, :705-707;
    timed on the sampled upper rows and scaled per nonzero to the full matrix).
    Every stage is the median of three runs.  Extrapolated by rows (gather: by
    nonzeros) to the class's full matrix and its mean FP64 matvec count on the
    GPU (the one-at-a-time FP64 path).  Stream CPU time = sum over the requests
    of their class time."""
    from math import comb

    from ciflow_amd import stream as S
    from oracle import oracle as O
    threads = _oracle_threads()
    lib = O.lib()
    per, per_omp = {}, {}
    t0_all = time.perf_counter()
    for key in sorted({(r.kind, r.L) for r in reqs}):
        r = next(q for q in reqs if (q.kind, q.L) == key)
        kind = "DOCI" if r.kind == 1 else "FCI"
        h, V, U = S.request_arrays(r)
        box = {}
        t_rot = _median3(lambda: box.__setitem__("hv", O.rotate_integrals(U, h, V)))
        hr, Vr = box["hv"]
        d1 = comb(r.L, r.n)
        n = d1 if kind == "DOCI" else d1 * d1
        ranges = _line_ranges(d1, lines) if kind == "FCI" else _line_ranges(d1, 16, 4096)
        mats = []

        def build():
            mats.clear()
            for s0, s1 in ranges:
                mats.append(O.build_crs(kind, hr, Vr, r.n, r.n, start=int(s0), end=int(s1), threads=threads))
        tb = _median3(build)
        rows = sum(m.n for m in mats)
        nnz_s = sum(m.nnz for m in mats)
        x = np.random.default_rng(0).standard_normal(n)
        y = np.zeros(n)

        def mv_sym():
            for m in mats:
                lib.or_mvprod_sym(m.n, m.row.ctypes.data_as(O.u64p), m.col.ctypes.data_as(O.u32p),
                                  m.data.ctypes.data_as(O.dp), x.ctypes.data_as(O.dp), y.ctypes.data_as(O.dp))

        def mv_omp():
            for m in mats:
                lib.or_mvprod_full(m.n, m.row.ctypes.data_as(O.u64p), m.col.ctypes.data_as(O.u32p),
                                   m.data.ctypes.data_as(O.dp), x.ctypes.data_as(O.dp), y.ctypes.data_as(O.dp),
                                   threads)
        tm = _median3(mv_sym)
        to = _median3(mv_omp)
        # full nonzeros of the class: 2 * upper - n (diagonal stored once)
        if kind == "FCI":
            da = d1
            nnz_full = n * (1 + 2 * r.n * (r.L - r.n) + 2 * comb(r.n, 2) * comb(r.L - r.n, 2)
                            + (r.n * (r.L - r.n)) ** 2)
        else:
            nnz_full = n * (1 + r.n * (r.L - r.n))
        scale = n / rows
        nmv = nmv_by_class.get(key, 0.0)
        per[key] = t_rot + tb * scale + nmv * tm * scale
        per_omp[key] = t_rot + tb * scale + nmv * to * nnz_full / nnz_s
    total = sum(per[(r.kind, r.L)] for r in reqs)
    total_omp = sum(per_omp[(r.kind, r.L)] for r in reqs)
    return {
        "value": len(reqs) / total, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": (f"per (kind, L) class ({len(per)} classes, {len(reqs)} requests): oracle integral rotation + "
                   f"CRS build (PushToRow, build_parallel on {threads} threads) of {lines} alpha lines (DOCI: 16 "
                   f"blocks of <= 4096 rows) + the serial upper-triangle mvprod (reference g++ path) on them, "
                   f"median of 3 runs each, extrapolated by rows to the full matrix and to the class's mean FP64 "
                   f"matvec count"),
        "variants": {"omp_full_row_gather_solves_per_s": len(reqs) / total_omp,
                     "note": ("same build; products as an OpenMP full-row gather on "
                              f"{threads} threads (stand-in for threaded MKL dcsrsymv), scaled per nonzero")},
        "sample_s": time.perf_counter() - t0_all,
    }


def stratified_sample(reqs, every):
    """Every `every`-th request of each (kind, L) class in stream order (each
    class contributes ceil(count / every) requests)."""
    by = {}
    for i, r in enumerate(reqs):
        by.setdefault((r.kind, r.L), []).append(i)
    return sorted(i for idx in by.values() for i in idx[::every])


# ---------------------------------------------------------------------------
# SpMV roofline on the stored CRS (FCI L=12 6a6b)
# ---------------------------------------------------------------------------
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "spmv_ell_fci_L12_traffic.txt")


def measured_traffic(prec_name):
    """HBM bytes per launch of the FCI L=12 position-major SpMV from the two
    rocprofv3 --pmc passes summarised in profiles/ (FETCH_SIZE and WRITE_SIZE,
    scaled by a 2 GiB copy); None when the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch"):
                    kv = dict(t.split("=") for t in ln.split()[1:])
                    return float(kv[prec_name])
    except (OSError, KeyError, ValueError):
        return None
    return None


def spmv_roofline(CA, ctx, L=12, nel=6, seed=1):
    h, V = CA.gen_integrals(L, seed)
    ham = CA.Hamiltonian(h, V, nel, nel, 0.0, ctx)
    cim = CA.FCI(ham, value_prec=CA.CF_F64)
    m = cim.get_mat()
    n, nnz_full, _ = m.info()
    x = np.random.default_rng(0).standard_normal(n)
    m.mvprod(x)  # warm
    tab = m.autotune(CA.CF_F64, reps=3)  # installs the fastest launch shape
    best = min(tab, key=lambda t: t[3])
    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.reset_timing()
    info = cim.solve(precision=CA.CF_F64)
    ctx.synchronize()
    nl, ms = ctx.kernel_time("spmv_f64")
    ctx.enable_timing(False)
    avg_ms = ms / max(nl, 1)
    kern, algo = m.spmv_bytes(CA.CF_F64)
    ach = algo / (avg_ms * 1e-3) / 1e9
    names = {1: "k_spmv (CRS row gather)", 3: "k_spmv_v4 (CRS row gather, vectorised streams)",
             4: "k_spmv_ell2 (FCI position-major complete rows, double-buffered LDS x blocks, FP64 values)",
             5: "k_spmv_doci (DOCI position-major pair-hop columns, pair elements in LDS)",
             6: "k_spmv_doci_blk (DOCI block-staged, x blocks and pair elements in LDS)",
             7: "k_spmv_sell (sliced rows: 64-row chunks, interleaved entries)"}
    crs_bytes = nnz_full * 12 + (n + 1) * 8 + 16 * n
    out = {"bound": "hbm", "kernel": names.get(kern, str(kern)), "crs_equivalent_bytes": crs_bytes,
           "crs_equivalent_GBps": crs_bytes / (avg_ms * 1e-3) / 1e9,
           "workload": f"FCI L={L} {nel}a{nel}b seed {seed}: n={n}, nnz={nnz_full}, one FP64 Lanczos solve",
           "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
           "traffic": measured_traffic("f64") if (L, nel, kern) == (12, 6, 4) else None,
           "traffic_source": "profiles/spmv_ell_fci_L12_traffic.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
           "avg_launch_ms": avg_ms, "launches": nl, "algorithmic_bytes": algo,
           "autotune": {"candidates": len(tab), "kernel": best[0], "p1": best[1], "p2": best[2],
                        "ms": best[3]},
           "solve_energy": cim.get_ci_energy(), "solve_n_matvec": info.n_matvec}
    del cim, m
    return out


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def class_service_ms(reqs, res):
    """Mean service time (start -> end, ms) and count per (kind, L) class; a
    batched request's service time is its batch's."""
    by = {}
    for r, d in zip(reqs, res):
        by.setdefault(f"{'DOCI' if r.kind == 1 else 'FCI'}{r.L}", []).append((d["t_end"] - d["t_start"]) * 1e3)
    return {k: {"n": len(v), "mean_ms": float(np.mean(v))} for k, v in sorted(by.items())}


def latency_by_class(reqs, res):
    """p50 / p99 latency (submit -> result, ms) and count per (kind, L) class."""
    by = {}
    for r, d in zip(reqs, res):
        by.setdefault(f"{'DOCI' if r.kind == 1 else 'FCI'}{r.L}", []).append((d["t_end"] - d["t_submit"]) * 1e3)
    return {k: {"n": len(v), "p50_ms": float(np.percentile(v, 50)), "p99_ms": float(np.percentile(v, 99))}
            for k, v in sorted(by.items())}


def _slices(total, k):
    """Split range(total) into k contiguous slices whose sizes differ by at most one."""
    edges = np.linspace(0, total, k + 1).round().astype(int)
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:])]


def bench_stream(a, CA):
    """The served stream at its stated size: --requests (4,096) requests, one
    step = one slice of requests/steps of them.  The K slices are submitted
    back to back to one continuously batched server (a closed batch of the whole
    stream) and the timed region runs from the first submission to the last
    completion; the W warm-up steps are slices of the same size from a
    warm-up stream of the same mix (other seeds), served before the timed run."""
    from ciflow_amd import stream as S
    K = a.steps if a.steps is not None else 16
    W = a.warmup if a.warmup is not None else 2
    NREQ = a.requests
    if K < 1 or K > NREQ:
        raise SystemExit(f"--steps must be in [1, {NREQ}] (one step = a slice of the {NREQ}-request stream)")
    slices = _slices(NREQ, K)
    slice_sz = NREQ / K
    nwarm = int(round(W * slice_sz))
    method = {"auto": CA.CF_SOLVER_AUTO, "lanczos": CA.CF_LANCZOS, "davidson": CA.CF_DAVIDSON}[a.solver]
    sprec = {"auto": CA.CF_PREC_AUTO, "f32": CA.CF_F32, "f64": CA.CF_F64}[a.served_precision]
    served = dict(precision=sprec, check_every=2, tol_energy=1e-9, direct_fci=True, method=method)
    reqs = S.make_stream(NREQ, rate=0.0)
    warm = S.make_stream(nwarm, rate=0.0, first_seed=500000, mix_seed=99) if nwarm > 0 else []
    t0 = time.perf_counter()
    arrs = S.prepare(reqs)
    warr = S.prepare(warm) if warm else []
    log(f"prepared {NREQ}+{nwarm} requests in {time.perf_counter() - t0:.1f} s")
    srv = CA.SolveServer(slots=a.slots)
    try:
        if warm:
            S.run_stream(warm, warr, server=srv, timeout=900, **served)
            log(f"warm-up: {W} steps ({nwarm} requests)")
        log(f"timed: {K} steps of {slices[0][1] - slices[0][0]}-{max(b - a_ for a_, b in slices)} requests "
            f"({NREQ} requests) ...")
        out = S.run_stream(reqs, arrs, server=srv, timeout=1500, **served)
        rate = out["solves_per_s"]
        log(f"closed batch: {rate:.2f} solves/s ({out['wall_s']:.2f} s)")
        res = out["results"]
        # the slice boundaries in completion time: when the last request of slice j finished
        t_first = min(d["t_submit"] for d in res)
        step_end = [max(res[i]["t_end"] for i in range(s0, s1)) - t_first for s0, s1 in slices]
        nlat = min(NREQ, a.latency_requests)
        lam = a.rate if a.rate > 0 else a.load * rate
        lat = {"p50_ms": None, "p99_ms": None, "solves_per_s": None, "mean_ms": None}
        if not a.no_latency:
            lreqs = S.with_arrivals(reqs[:nlat], lam)
            log(f"Poisson arrivals at {lam:.1f}/s over {nlat} requests ...")
            lat = S.run_stream(lreqs, arrs[:nlat], server=srv, timeout=1500, **served)
            log(f"Poisson {lam:.1f}/s: p50 {lat['p50_ms']:.1f} ms p99 {lat['p99_ms']:.1f} ms")
            lat["by_class"] = latency_by_class(lreqs, lat["results"])
    finally:
        srv.close()
    # the plain path: one request at a time, FP64 values, tolerance 1e-12 (ARPACK semantics),
    # on a class-stratified sample of the same requests
    base = None
    dE = np.zeros(1)
    nmv_by_class = {}
    sub = stratified_sample(reqs, a.baseline_every) if a.baseline_every > 0 else []
    if sub:
        breqs = [reqs[i] for i in sub]
        log(f"one-at-a-time FP64 over {len(sub)} requests ...")
        base = S.run_stream(breqs, [arrs[i] for i in sub], slots=1, precision=CA.CF_F64, timeout=1500)
        log(f"one-at-a-time FP64: {base['solves_per_s']:.2f} solves/s over {len(sub)} requests")
        Eserved = np.array([res[i]["energy"] for i in sub])
        Ebase = np.array([d["energy"] for d in base["results"]])
        dE = np.abs(Eserved - Ebase)
        for r, d in zip(breqs, base["results"]):
            nmv_by_class.setdefault((r.kind, r.L), []).append(d["n_matvec"])
        nmv_by_class = {k: float(np.mean(v)) for k, v in nmv_by_class.items()}
        # the one-at-a-time rate of the whole stream: class mix weighted by each class's mean time
        t_cls = {}
        for r, d in zip(breqs, base["results"]):
            t_cls.setdefault((r.kind, r.L), []).append(d["t_end"] - d["t_start"])
        t_stream = sum(float(np.mean(t_cls[(r.kind, r.L)])) for r in reqs)
        base_stream_rate = NREQ / t_stream
    roof = None
    if not a.no_roofline:
        log("SpMV roofline (FCI L=12) ...")
        roof = spmv_roofline(CA, CA.Context(0))
        log(f"SpMV {roof['achieved']:.0f} GB/s ({100 * roof['frac']:.1f}%)")
    cpu = None
    if not a.no_cpu_baseline and nmv_by_class:
        log("cpu baseline ...")
        cpu = cpu_baseline_stream(reqs, nmv_by_class)
        log(f"cpu baseline: {cpu['value']:.3f} solves/s (sample {cpu['sample_s']:.1f} s)")
    kinds = np.array([r.kind for r in reqs])
    n_f32 = sum(1 for d in res if d["n_matvec_f32"] > 0)
    dtype = ("f64" if n_f32 == 0 else "f32-values/f64-accumulate+f64-polish" if n_f32 == len(res)
             else f"f64, f32-values/f64-accumulate+f64-polish on {n_f32} stored-CRS requests")
    prec_note = {"auto": "per-request precision CF_PREC_AUTO (FP32 values + FP64 polish where the operator "
                         "stores matrix values; none of this mix does, so FP64 throughout)",
                 "f32": "FP32 operator + FP64 polish", "f64": "FP64"}[a.served_precision]
    return {
        "metric": METRIC,
        "value": rate,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": out["wall_s"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": ("served stream of rotated-integral ground-state solves: 70% DOCI L~U[12,20], "
                                "30% FCI L~U[8,11] (on-the-fly sigma), half filling, seeds 1000+k, U=exp(K) "
                                f"K~N(0,0.05^2), {prec_note}, Kato-Temple energy stop 1e-9 Eh"),
                   "solver": {"auto": ("DOCI <= 13,312 determinants: batched FP64 Lanczos, one workgroup per "
                                       "request (cf_solve_doci_batch); FCI and larger DOCI: Davidson"),
                              "lanczos": "thick-restart Lanczos", "davidson": "Davidson"}[a.solver],
                   "requests": NREQ, "step": f"one slice of {NREQ}/{K} = {slice_sz:.1f} consecutive requests",
                   "warmup_requests": nwarm,
                   "doci": int((kinds == 1).sum()), "fci": int((kinds == 0).sum()),
                   "slots": a.slots, "closed_batch": True,
                   "inputs": "h, V, U copied from host memory at submission (1.3 MB per L=20 request)"},
        "step_completion_s": step_end,
        "p50_ms": lat["p50_ms"],
        "p99_ms": lat["p99_ms"],
        "latency_run": {"arrivals": "Poisson", "rate_per_s": lam, "requests": nlat,
                        "achieved_solves_per_s": lat["solves_per_s"], "mean_ms": lat["mean_ms"],
                        "by_class": lat.get("by_class")},
        "closed_batch_latency_ms": {"p50": out["p50_ms"], "p99": out["p99_ms"]},
        "one_at_a_time_fp64": ({"solves_per_s": base["solves_per_s"], "requests": len(sub),
                                "stream_weighted_solves_per_s": base_stream_rate,
                                "sample": f"every {a.baseline_every}th request of each (kind, L) class",
                                "tol": 1e-12} if base else None),
        "speedup_vs_one_at_a_time_fp64": rate / base_stream_rate if base else None,
        "max_abs_dE_vs_fp64": float(dE.max()) if base else None,
        "mean_matvec": {"f32": float(np.mean([d["n_matvec_f32"] for d in res])),
                        "f64": float(np.mean([d["n_matvec"] for d in res]))},
        "class_service_ms": class_service_ms(reqs, res),
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def bench_fci12(a, CA):
    """One CI space built and solved per step: FCI L=12 6a6b seed 1 (config 2) or,
    with --workload doci24, DOCI L=24 12 pairs seed 2 (config 3)."""
    K = a.steps if a.steps is not None else 4
    W = a.warmup if a.warmup is not None else 1
    doci = a.workload == "doci24"
    if doci:
        a.L, a.nel, a.seed = 24, 12, 2
    cls = CA.DOCI if doci else CA.FCI
    ctx = CA.Context(0)
    h, V = CA.gen_integrals(a.L, a.seed)
    ham = CA.Hamiltonian(h, V, a.nel, a.nel, 0.0, ctx)
    prec = CA.CF_F32 if a.precision == "f32" else CA.CF_F64
    method = {"auto": CA.CF_SOLVER_AUTO, "lanczos": CA.CF_LANCZOS, "davidson": CA.CF_DAVIDSON}[a.solver]
    ctx.synchronize()

    def step():
        cim = cls(ham, value_prec=prec)
        info = cim.solve(precision=prec, method=method)
        return cim, info

    for i in range(W):
        cim, info = step()
        log(f"warmup {i}: E={cim.get_ci_energy():.12f} matvec={info.n_matvec}+{info.n_matvec_f32} ms={info.ms:.1f}")
        del cim
    ctx.enable_timing(True)
    ctx.reset_timing()
    lat, infos = [], []
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        s0 = time.perf_counter()
        cim, info = step()
        ctx.synchronize()
        lat.append(time.perf_counter() - s0)
        infos.append(info.as_dict())
        energy = cim.get_ci_energy()
        log(f"step {i}: {lat[-1] * 1e3:.1f} ms E={energy:.12f} matvec={info.n_matvec}+{info.n_matvec_f32}")
        if i + 1 < K:
            del cim
    ctx.synchronize()
    total = time.perf_counter() - t0
    # outside the timed region: a DOCI matrix on the block-staged kernel builds
    # its CRS arrays only here, when they are read
    n, nnz_full, _ = cim.get_mat().info()
    nl64, ms64 = ctx.kernel_time("spmv_f64")
    nl32, ms32 = ctx.kernel_time("spmv_f32")
    ncount, ms_count = ctx.kernel_time("crs_count")
    nfill, ms_fill = ctx.kernel_time("crs_fill")
    nhop, ms_hop = ctx.kernel_time("doci_ell_fill")
    if ms32 > ms64:
        fam, nl, ms, sval = "spmv_f32", nl32, ms32, 4
    else:
        fam, nl, ms, sval = "spmv_f64", nl64, ms64, 8
    avg_ms = ms / max(nl, 1)
    nnz_upper = cim.get_mat().datasize()
    kern, algo_bytes = cim.get_mat().spmv_bytes(CA.CF_F32 if sval == 4 else CA.CF_F64)
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    crs_eq = nnz_full * (sval + 4) + (n + 1) * 8 + 16 * n  # what the stored-CRS product would stream
    ctx.enable_timing(False)
    r0 = time.perf_counter()
    cim.construct_density(trdm=True)
    rdm_ms = (time.perf_counter() - r0) * 1e3
    tr = sum(cim.get_one_rdm(s, i, i) for s in range(2) for i in range(a.L))
    cpu = None
    if not a.no_cpu_baseline and not doci:
        log("cpu baseline ...")
        cpu = cpu_baseline_fci(h, V, a.L, a.nel, nnz_upper, n, infos[-1]["n_matvec"], a.cpu_lines)
        log(f"cpu baseline: {cpu['value']:.3e} solves/s (sample {cpu['sample_s']:.1f} s)")
    lat_ms = np.array(lat) * 1e3
    return {
        "metric": METRIC,
        "value": K / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": total / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == CA.CF_F64 else "f32-values/f64-accumulate+f64-polish",
        "data": "synthetic",
        "config": {"workload": (f"DOCI L={a.L} {a.nel} pairs" if doci else f"FCI L={a.L} {a.nel}a{a.nel}b")
                   + f" seed {a.seed}: GPU CRS build + ground state",
                   "determinants": n, "nnz_full": nnz_full, "nnz_upper": nnz_upper, "tol": 1e-12,
                   "precision": a.precision,
                   "solver": {"auto": "Davidson (spaces of >= 2,000 determinants)", "lanczos": "thick-restart Lanczos",
                              "davidson": "Davidson"}[a.solver]},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "energy": energy,
        "n_matvec": infos[-1]["n_matvec"],
        "n_matvec_f32": infos[-1]["n_matvec_f32"],
        "resid": infos[-1]["resid"],
        "build_ms": {"count": ms_count / max(ncount, 1), "fill": ms_fill / max(nfill, 1),
                     "doci_hop_columns": ms_hop / max(nhop, 1) if nhop else None},
        "rdm_ms": rdm_ms,
        "rdm_trace": tr,
        "roofline": {"bound": "hbm" if kern < 6 else "l2+lds (block-staged: HBM sees x, y, diag once)",
                     "kernel": f"{fam} (kernel {kern})", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "crs_equivalent_bytes": crs_eq,
                     "crs_equivalent_GBps": crs_eq / (avg_ms * 1e-3) / 1e9,
                     "traffic": (measured_traffic("f32" if sval == 4 else "f64")
                                 if (not doci and a.L == 12 and a.nel == 6 and kern == 4) else None),
                     "avg_launch_ms": avg_ms,
                     "launches": nl, "algorithmic_bytes": algo_bytes},
        "cpu_baseline": cpu,
    }


def bench_direct16(a, CA):
    """FCI L=16 8a8b seed 3 (165,636,900 determinants) on the on-the-fly sigma:
    one step = one solve in --precision (f64, the default: FP64 sigma
    throughout, measured faster here than the FP32 phase, whose sigma saves
    only 12% and whose FP64 refinement restarts the basis; f32: FP32 sigma +
    FP64 refinement), by Davidson (--solver auto/davidson: residual
    corrections preconditioned by the diagonal, at most 24 basis vectors) or
    by Lanczos (--solver lanczos), then the 1-RDM of the ground state and,
    unless --no-fp64, the plain FP64 Lanczos solve (ARPACK semantics)."""
    L, nel, seed = 16, 8, 3
    ctx = CA.Context(0)
    h, V = CA.gen_integrals(L, seed)
    ham = CA.Hamiltonian(h, V, nel, nel, 0.0, ctx)
    t0 = time.perf_counter()
    d = CA.FCI_Direct(ham, value_prec=CA.CF_F32)
    ctx.synchronize()
    t_setup = time.perf_counter() - t0
    n = d.n
    ctx.enable_timing(True)
    ctx.reset_timing()
    prec = CA.CF_F32 if a.precision == "f32" else CA.CF_F64
    log(f"direct FCI L={L}: n={n}, setup {t_setup:.2f} s; {'FP32+polish' if prec == CA.CF_F32 else 'FP64'} "
        f"solve ...")
    def prog(nmv, th, r):
        log(f"  {nmv} products, theta {th:.10f}, residual {r:.2e}, {time.perf_counter() - t0:.1f} s")

    method = CA.CF_LANCZOS if a.solver == "lanczos" else CA.CF_DAVIDSON
    t0 = time.perf_counter()
    i32 = d.solve(precision=prec, ncv=39, progress=prog, method=method)
    t32 = time.perf_counter() - t0
    e32 = d.get_ci_energy()
    log(f"  E={e32:.12f} matvec {i32.n_matvec}+{i32.n_matvec_f32} (FP64+FP32) {t32:.1f} s")
    n64, ms64 = ctx.kernel_time("sigma_f64")
    n32, ms32 = ctx.kernel_time("sigma_f32")
    t0 = time.perf_counter()
    one, _ = d.construct_density(trdm=False)
    t_rdm = time.perf_counter() - t0
    tr = float(sum(d.get_one_rdm(s_, i, i) for s_ in range(2) for i in range(L)))
    out64 = None
    if not a.no_fp64:
        log("  FP64 solve ...")
        ctx.reset_timing()
        t0 = time.perf_counter()
        i64 = d.solve(precision=CA.CF_F64, ncv=39, progress=prog)
        t64 = time.perf_counter() - t0
        e64 = d.get_ci_energy()
        n64b, ms64b = ctx.kernel_time("sigma_f64")
        out64 = {"energy": e64, "n_matvec": i64.n_matvec, "seconds": t64, "sigma_ms": ms64b / max(n64b, 1),
                 "dE_vs_fp64_lanczos": abs(e32 - e64)}
        log(f"  E={e64:.12f} matvec {i64.n_matvec} {t64:.1f} s")
    ctx.enable_timing(False)
    return {
        "metric": METRIC, "value": 1.0 / t32, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t32 * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-sigma/f64-accumulate+f64-polish" if prec == CA.CF_F32 else "f64", "data": "synthetic",
        "config": {"workload": f"direct FCI L={L} {nel}a{nel}b seed {seed}, on-the-fly sigma, ncv 39",
                   "determinants": n, "solver": "lanczos" if method == CA.CF_LANCZOS else "davidson"},
        "energy": e32, "n_matvec": i32.n_matvec, "n_matvec_f32": i32.n_matvec_f32,
        "sigma_ms": {"f32": ms32 / max(n32, 1), "f64": ms64 / max(n64, 1)},
        "setup_s": t_setup, "rdm1_s": t_rdm, "rdm1_trace": tr, "fp64": out64,
        "roofline": None, "cpu_baseline": None,
    }


def bench_spmv(a, CA, doci=False):
    """Stored-matrix SpMV alone on FCI L=12 6a6b (config 2), or with doci=True on
    DOCI L=24 12 pairs seed 2 (config 3): 3 FP64 and 3 FP32-value products with
    the automatic kernel, then three 2 GiB device-to-device copies (k_copy) --
    the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings
    to bytes."""
    ctx = CA.Context(0)
    if doci:
        h, V = CA.gen_integrals(24, 2)
        cim = CA.DOCI(CA.Hamiltonian(h, V, 12, 12, 0.0, ctx), value_prec=CA.CF_F32)
    else:
        h, V = CA.gen_integrals(12, 1)
        cim = CA.FCI(CA.Hamiltonian(h, V, 6, 6, 0.0, ctx), value_prec=CA.CF_F32)
    m = cim.get_mat()
    n, nnz, _ = m.info()
    x = np.random.default_rng(0).standard_normal(n)
    out = {}
    for prec, fam, sval in ((CA.CF_F64, "spmv_f64", 8), (CA.CF_F32, "spmv_f32", 4)):
        m.mvprod(x, prec)
        ctx.enable_timing(True)
        ctx.reset_timing()
        for _ in range(3):
            m.mvprod(x, prec)
        nl, ms = ctx.kernel_time(fam)
        ctx.enable_timing(False)
        kern, algo = m.spmv_bytes(prec)
        out[fam] = {"kernel": kern, "avg_launch_ms": ms / nl, "algorithmic_bytes": algo,
                    "GBps": algo / (ms / nl * 1e-3) / 1e9}
        log(f"{fam}: {ms / nl:.3f} ms {out[fam]['GBps']:.0f} GB/s")
    copy_bytes = 2 << 30
    CA._check(CA.lib().cf_ctx_stream_copy(ctx.ptr, copy_bytes, 3))
    out["copy_bytes_each_way"] = copy_bytes
    wl = "DOCI L=24 12 pairs SpMV" if doci else "FCI L=12 6a6b stored-CRS SpMV"
    return {"metric": METRIC, "workload": wl, "n": n, "nnz": nnz, **out}


def bench_spmv_dets(a, CA):
    """Stored-matrix SpMV over a determinant list (selected CI, FCI_File): a
    CAS-like list of --L orbitals at half filling, the first N alpha strings
    times every beta string in (up, down) order -- N = 110 at L=12 (101,640
    determinants), 80 at L=16 (1,029,600), else about 10^5 determinants --
    seeded integrals (seed 1).  Per value precision, --steps products (after
    one warm-up) on the automatic kernel (the sliced rows) and on the
    vectorised CRS row gather (16 lanes per row, unroll 2), HIP events on the
    library stream; then three 2 GiB device-to-device copies (the reference
    kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)."""
    from math import comb
    L = a.L
    ne = L // 2
    d = comb(L, ne)
    nalpha = {12: 110, 16: 80}.get(L, max(1, min(d, round(1e5 / d))))
    ctx = CA.Context(0)
    s = ctx.unrank(L, ne, 0, d)
    up = np.repeat(s[:nalpha], d)
    down = np.tile(s, nalpha)
    h, V = CA.gen_integrals(L, 1)
    t0 = time.perf_counter()
    m = CA.FCI_File(CA.Hamiltonian(h, V, ne, ne, 0.0, ctx), (up, down), value_prec=CA.CF_F32).get_mat()
    ctx.synchronize()
    build_s = time.perf_counter() - t0
    n, nnz, _ = m.info()
    x = np.random.default_rng(0).standard_normal(n)
    steps = a.steps or 5
    names = {3: "k_spmv_v4", 7: "k_spmv_sell"}
    out = {}
    for prec, fam in ((CA.CF_F64, "spmv_f64"), (CA.CF_F32, "spmv_f32")):
        for label, shape in (("auto", None), ("row_gather", (3, 16, 2))):
            if shape:
                m.set_spmv_shape(prec, *shape)
            m.mvprod(x, prec)
            ctx.enable_timing(True)
            ctx.reset_timing()
            for _ in range(steps):
                m.mvprod(x, prec)
            nl, ms = ctx.kernel_time(fam)
            ctx.enable_timing(False)
            kern, algo = m.spmv_bytes(prec)
            gbs = algo / (ms / nl * 1e-3) / 1e9
            out[f"{fam}_{label}"] = {"kernel": names.get(kern, str(kern)), "avg_launch_ms": ms / nl,
                                     "algorithmic_bytes": algo, "GBps": gbs, "frac": gbs / HBM_PEAK_GBS}
            log(f"{fam} {label} ({names.get(kern, kern)}): {ms / nl:.4f} ms {gbs:.0f} GB/s "
                f"({gbs / HBM_PEAK_GBS * 100:.1f}%)")
    copy_bytes = 2 << 30  # reference copies for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes
    CA._check(CA.lib().cf_ctx_stream_copy(ctx.ptr, copy_bytes, 3))
    out["copy_bytes_each_way"] = copy_bytes
    return {"metric": METRIC, "workload": (f"determinant list: first {nalpha} alpha strings x all {d} beta "
                                           f"strings of L={L}, half filling, seed 1"),
            "n": n, "nnz": nnz, "build_s": build_s, "steps": steps, **out}


def bench_sigma(a, CA):
    """On-the-fly FCI sigma alone (CI_Big::mvprod's role): for each L in
    --sigma-L, FCI floor(L/2) a + b electrons with seeded integrals (seed 3),
    --steps applications (after one warm-up) per value precision and alpha-beta
    kernel (2: kernel 2; 3: packed-offset kernel 3); sigma time from HIP events
    on the library stream (the whole product: same-spin parts and alpha-beta
    kernel)."""
    from math import comb
    ctx = CA.Context(0)
    reps = a.steps if a.steps is not None else 3
    out = []
    for L in [int(v) for v in a.sigma_L.split(",")]:
        nel = L // 2
        h, V = CA.gen_integrals(L, 3)
        d = CA.FCI_Direct(CA.Hamiltonian(h, V, nel, nel, 0.0, ctx), value_prec=CA.CF_F32)
        x = np.random.default_rng(0).standard_normal(d.get_dim())
        sa = nel * (L - nel) + 1
        terms = d.get_dim() * sa * nel * (L - nel + 1)
        for kern in (2, 3):
            d.set_kernel(kern)
            for prec, fam in ((CA.CF_F32, "sigma_f32"), (CA.CF_F64, "sigma_f64")):
                d.mvprod(x, precision=prec)
                ctx.enable_timing(True)
                ctx.reset_timing()
                for _ in range(reps):
                    d.mvprod(x, precision=prec)
                nl, ms = ctx.kernel_time(fam)
                ctx.enable_timing(False)
                row = {"L": L, "n": d.get_dim(), "kernel": kern, "prec": fam[-3:], "ms": ms / nl,
                       "alpha_beta_terms": terms, "terms_per_s": terms / (ms / nl * 1e-3)}
                log(f"sigma L={L} kernel {kern} {fam[-3:]}: {ms / nl:.3f} ms")
                out.append(row)
        del d
    return {"metric": METRIC, "workload": "on-the-fly FCI sigma, seeded integrals (seed 3), half filling",
            "sigma": out}


def bench_rotate(a, CA):
    """Batched FP64-MFMA integral rotation (OrbitalTransform::fillHamCI): 64
    Hamiltonians at L=20 per call, 8 L^5 + 4 L^3 flop each; kernel time from HIP
    events on the library stream, 1 warm-up + 5 timed calls.  At L/8 flop per
    byte the transform is bound by HBM: the roofline is the bytes of its six
    passes (each reads and writes its array once)."""
    ctx = CA.Context(0)
    L, B = 20, 64
    h1, V1 = CA.gen_integrals(L, 5)
    h = np.repeat(h1[None], B, axis=0)
    V = np.repeat(V1[None], B, axis=0)
    U = np.stack([CA.expm_skew(CA.gen_rotation(L, 100 + b, 0.05)) for b in range(B)])
    CA.rotate_integrals(h, V, U, ctx)
    ctx.enable_timing(True)
    ctx.reset_timing()
    for _ in range(5):
        CA.rotate_integrals(h, V, U, ctx)
    nl, ms = ctx.kernel_time("rotate")
    ctx.enable_timing(False)
    flop = B * (8.0 * L ** 5 + 4.0 * L ** 3)
    # four two-body and two one-body passes, each reading and writing its array once
    nbytes = B * (4 * 2 * L ** 4 + 2 * 2 * L ** 2) * 8.0
    sec = ms / nl * 1e-3
    tf, gbs = flop / sec / 1e12, nbytes / sec / 1e9
    log(f"rotate L={L} batch {B}: {ms / nl:.3f} ms per call, {gbs:.0f} GB/s, {tf:.1f} TFLOP/s")
    return {"metric": METRIC, "workload": f"batched integral rotation, L={L}, {B} Hamiltonians per call",
            "avg_call_ms": ms / nl, "flop_per_call": flop, "bytes_per_call": nbytes, "TFLOPs": tf,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "mfma_frac": tf / 78.6}}


def bench_batch(a, CA):
    """Lock-step batched Davidson solves of one stream class (cf_solve_batch):
    --lanes requests of class --batch-class (the stream's rotated seeded
    integrals, --batch-precision: FP32 operator + FP64 polish or FP64, energy
    stop 1e-9 Eh) per batch, one warm-up batch, then --steps timed batches of
    fresh requests."""
    from ciflow_amd import stream as S
    cls = a.batch_class.upper()
    kind = CA.CF_DOCI if cls.startswith("DOCI") else CA.CF_FCI
    L = int(cls[4:] if kind == CA.CF_DOCI else cls[3:])
    K = a.steps if a.steps is not None else 3
    B = a.lanes
    prec = CA.CF_F64 if a.batch_precision == "f64" else CA.CF_F32
    ctx = CA.Context(0)

    def batch(first):
        reqs = []
        for k in range(B):
            r = S.StreamRequest(k, kind, L, L // 2, first + k, 0.0)
            h, V, U = S.request_arrays(r)
            reqs.append(dict(kind=kind, h=h, V=V, nup=L // 2, ndown=L // 2, U=U, precision=prec,
                             tol_energy=1e-9, tag=first + k))
        return reqs

    CA.solve_batch(batch(900000), ctx)
    sets = [batch(910000 + 1000 * i) for i in range(K)]
    ctx.synchronize()
    t0 = time.perf_counter()
    res = [CA.solve_batch(q, ctx) for q in sets]
    dt = time.perf_counter() - t0
    flat = [r for rs in res for r in rs]
    assert all(r["status"] == 0 for r in flat)
    rate = len(flat) / dt
    n32 = float(np.mean([r["n_matvec_f32"] for r in flat]))
    n64 = float(np.mean([r["n_matvec"] for r in flat]))
    log(f"batch {cls} x {B}: {dt / K * 1e3:.1f} ms per batch, {rate:.1f} solves/s, products {n32:.1f} + {n64:.1f}")
    return {"metric": "batched solves/s", "value": rate, "unit": "solves/s", "higher_is_better": True,
            "workload": f"{B} {cls} requests per lock-step batch (cf_solve_batch), "
                        f"{'FP32 + FP64 polish' if prec == CA.CF_F32 else 'FP64'}, "
                        f"energy stop 1e-9 Eh", "steps": K, "ms_per_batch": dt / K * 1e3,
            "mean_products": {"f32": n32, "f64": n64}, "n": flat[0]["n"]}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import ciflow_amd as CA

    run = {"stream": bench_stream, "fci12": bench_fci12, "doci24": bench_fci12, "direct16": bench_direct16,
           "spmv": bench_spmv, "spmv_doci": lambda a, CA: bench_spmv(a, CA, doci=True),
           "spmv_dets": bench_spmv_dets,
           "rotate": bench_rotate, "sigma": bench_sigma, "batch": bench_batch}[a.workload]
    out = run(a, CA)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
