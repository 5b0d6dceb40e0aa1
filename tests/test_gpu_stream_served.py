"""GPU test of the served stream exactly as bench.py serves it.

The requests are the first two of every (kind, L) class of the 4,096-request
stream (DOCI L = 12..20, FCI L = 8..11; 26 requests), served with the bench's
settings: FP32 values (FP32 sigma for FCI) + FP64 polish, convergence tested
every 2 steps, Kato-Temple energy stop 1e-9 Eh, FCI on the on-the-fly sigma
(cf_solve_direct), solver CF_SOLVER_AUTO.  This is the loop the reference's
orbital optimisers run one request at a time (OrbitalOptimization::update_cim
-> set_ham -> build_ci -> solve, SimulatedAnnealing.cpp:83-87,
CIMethod.cpp:117-126,241-268).

* alone == in the mix: each request served by itself on a one-slot server and
  all of them at once on eight worker streams give bitwise the same energy
  and the same FP32/FP64 product counts;
* every served energy is within 1e-8 Eh of the oracle's FP64 CRS + Lanczos
  (tol 1e-12) energy of the same rotated integrals (tests/golden/golden.json
  "stream", written by tests/golden/make_golden.py --stream).
"""
import json
import os
import sys

import pytest

import ciflow_amd as CA
from ciflow_amd import stream as S

pytestmark = pytest.mark.gpu

SERVED = dict(precision=CA.CF_F32, check_every=2, tol_energy=1e-9, direct_fci=True, method=CA.CF_SOLVER_AUTO)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def golden_requests(golden_dir):
    g = json.load(open(os.path.join(golden_dir, "golden.json")))["stream"]
    stream = S.make_stream(4096)
    reqs = []
    for e in g:
        r = stream[e["k"]]
        assert (r.kind, r.L, r.n, r.seed) == (e["kind"], e["L"], e["n"], e["seed"])
        reqs.append(r)
    return reqs, [e["energy"] for e in g]


def test_golden_covers_every_stream_class(golden_dir):
    reqs, _ = golden_requests(golden_dir)
    classes = {(r.kind, r.L) for r in reqs}
    assert classes == {(CA.CF_DOCI, L) for L in range(12, 21)} | {(CA.CF_FCI, L) for L in range(8, 12)}


def test_bench_stream_alone_equals_mix_and_matches_oracle(golden_dir):
    reqs, E_ref = golden_requests(golden_dir)
    arrs = S.prepare(reqs)
    alone = []
    with CA.SolveServer(slots=1) as srv:
        for r, a in zip(reqs, arrs):
            say(f"alone: request {r.k} kind {r.kind} L {r.L}")
            alone.append(S.run_stream([r], [a], server=srv, timeout=120, **SERVED)["results"][0])
    say("mix: all requests on 8 worker streams")
    mix = S.run_stream(reqs, arrs, slots=8, timeout=300, **SERVED)["results"]
    for r, d0, d1, E in zip(reqs, alone, mix, E_ref):
        for d in (d0, d1):
            assert d["tag"] == r.k and d["status"] == 0 and d["solve_status"] == 0, (r, d["error"])
        assert d1["energy"] == d0["energy"], (r, d1["energy"], d0["energy"])
        assert (d1["n_matvec"], d1["n_matvec_f32"]) == (d0["n_matvec"], d0["n_matvec_f32"]), r
        assert abs(d1["energy"] - E) <= 1e-8, (r, d1["energy"], E)
    overlapped = any(a["t_start"] < b["t_end"] and b["t_start"] < a["t_end"]
                     for i, a in enumerate(mix) for b in mix[i + 1:])
    assert overlapped
