"""Mixed DOCI/FCI stream of rotated-integral ground-state solves.

The stream stands in for what the reference's orbital optimisers issue one
call at a time: every step of SimulatedAnnealing::optimize
(SimulatedAnnealing.cpp:306-390) and LocalMinimizer::optimize
(LocalMinimizer.cpp:73-235) rotates the integrals by U = exp(K)
(UnitaryMatrix::updateUnitary, OrbitalTransform::fillHamCI) and solves the CI
ground state again (OrbitalOptimization::update_cim, SimulatedAnnealing.cpp:83-87).

Request k: kind DOCI with probability 0.7 (L uniform in [12, 20]) or FCI
(L uniform in [8, 11]), nalpha = nbeta = floor(L/2) (the reference requires
nalpha == nbeta for DOCI, Hamiltonian.cpp:106,170), integrals from seed
1000+k, K antisymmetric with N(0, 0.05^2) entries, Poisson arrivals at a fixed
rate.  Everything is a deterministic function of the stream seed.
"""
from __future__ import annotations

import concurrent.futures as cf
import threading
import time
from dataclasses import dataclass, replace

import numpy as np

from . import CF_DOCI, CF_F32, CF_F64, CF_FCI, SolveServer, expm_skew, gen_integrals, gen_rotation


@dataclass
class StreamRequest:
    k: int
    kind: int
    L: int
    n: int          # electrons per spin
    seed: int       # integral and rotation seed
    t_arrive: float  # seconds after stream start


def make_stream(count: int, rate: float = 0.0, frac_doci: float = 0.7, doci_L=(12, 20), fci_L=(8, 11),
                first_seed: int = 1000, mix_seed: int = 4) -> list[StreamRequest]:
    """rate <= 0: every request arrives at t = 0 (closed batch)."""
    rng = np.random.default_rng(mix_seed)
    is_doci = rng.random(count) < frac_doci
    Ld = rng.integers(doci_L[0], doci_L[1] + 1, count)
    Lf = rng.integers(fci_L[0], fci_L[1] + 1, count)
    gaps = rng.exponential(1.0 / rate, count) if rate > 0 else np.zeros(count)
    t = np.cumsum(gaps) - (gaps[0] if count else 0.0)
    out = []
    for k in range(count):
        L = int(Ld[k] if is_doci[k] else Lf[k])
        out.append(StreamRequest(k, CF_DOCI if is_doci[k] else CF_FCI, L, L // 2, first_seed + k, float(t[k])))
    return out


def with_arrivals(reqs: list[StreamRequest], rate: float, seed: int = 7) -> list[StreamRequest]:
    """The same requests with Poisson arrival times at `rate` per second (first at t = 0)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate > 0 else np.zeros(len(reqs))
    t = np.cumsum(gaps) - (gaps[0] if len(reqs) else 0.0)
    return [replace(r, t_arrive=float(ti)) for r, ti in zip(reqs, t)]


def request_arrays(r: StreamRequest, sigma: float = 0.05):
    """(h, V, U) of request r: seeded integrals and U = exp(K)."""
    h, V = gen_integrals(r.L, r.seed)
    U = expm_skew(gen_rotation(r.L, r.seed, sigma))
    return h, V, U


def prepare(reqs: list[StreamRequest], threads: int = 8) -> list[tuple]:
    """Generate every request's arrays up front (host work outside the timed stream)."""
    with cf.ThreadPoolExecutor(threads) as ex:
        return list(ex.map(request_arrays, reqs))


def run_stream(reqs: list[StreamRequest], arrays: list[tuple], slots: int = 4, precision: int = CF_F32,
               hbm_budget: float = 0.0, tol: float = 0.0, server: SolveServer | None = None,
               timeout: float = 0.0, check_every: int = 0, tol_energy: float = 0.0, direct_fci: bool = False,
               ncv: int = 0, method: int = 0) -> dict:
    """Submit the requests at their arrival times and collect the results.

    Returns {"results": [dict per request, ordered by k], "wall_s": first
    arrival to last completion, "solves_per_s", "p50_ms", "p99_ms"}; latency =
    completion - submission (queueing included)."""
    own = server is None
    srv = server or SolveServer(slots=slots, hbm_budget=hbm_budget)
    results: dict[int, dict] = {}
    err: list[BaseException] = []

    def submitter(t0: float):
        try:
            for r, (h, V, U) in zip(reqs, arrays):
                dt = t0 + r.t_arrive - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                srv.submit(r.k, r.kind, h, V, r.n, r.n, U=U, precision=precision, ncv=ncv, tol=tol,
                           check_every=check_every, tol_energy=tol_energy, direct=direct_fci and r.kind == CF_FCI,
                           method=method)
        except BaseException as e:  # surfaced by the caller
            err.append(e)

    try:
        t0 = time.perf_counter()
        th = threading.Thread(target=submitter, args=(t0,), daemon=True)
        th.start()
        deadline = time.perf_counter() + timeout if timeout > 0 else None
        while len(results) < len(reqs):
            for d in srv.poll():
                results[d["tag"]] = d
            if err:
                raise err[0]
            if deadline is not None and time.perf_counter() > deadline:
                raise TimeoutError(f"stream: {len(results)} of {len(reqs)} finished before the time limit")
            if len(results) < len(reqs):
                time.sleep(0.0005)
        th.join()
    finally:
        if own:
            srv.close()
    res = [results[r.k] for r in reqs]
    bad = [d for d in res if d["status"] != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} request(s) failed, first: {bad[0]['error']}")
    t_first = min(d["t_submit"] for d in res)
    t_last = max(d["t_end"] for d in res)
    lat = np.array([d["t_end"] - d["t_submit"] for d in res]) * 1e3
    wall = t_last - t_first
    return {"results": res, "wall_s": wall, "solves_per_s": len(res) / wall if wall > 0 else float("inf"),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


__all__ = ["StreamRequest", "make_stream", "with_arrivals", "request_arrays", "prepare", "run_stream", "CF_F32",
           "CF_F64"]
