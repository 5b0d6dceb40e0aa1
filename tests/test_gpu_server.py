"""GPU tests of the served stream of rotated-integral solves.

* every result is the same alone and in the mix: each request solved by
  itself (FCI/DOCI classes, one at a time) and through the server with
  several requests in flight gives the identical energy and matvec count;
* served FP32-value + FP64-polish energies agree with the oracle's FP64
  CRS + Lanczos on the same rotated integrals within 1e-8 Eh;
* admission by HBM footprint: with a budget that fits one request at a time
  the stream still completes, with the same results, and the budget in use
  returns to zero;
* a request that fails (DOCI with nalpha != nbeta) comes back as an error result
  without disturbing the others.
"""
import sys

import numpy as np
import pytest

import ciflow_amd as CA
from ciflow_amd import stream as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_stream(count=12):
    return S.make_stream(count, rate=0.0, doci_L=(12, 14), fci_L=(8, 9))


def solve_alone(r, arrays, precision):
    h, V, U = arrays
    ham = CA.Hamiltonian(h, V, r.n, r.n).rotate(U)
    cls = CA.DOCI if r.kind == CA.CF_DOCI else CA.FCI
    cim = cls(ham, value_prec=precision)
    info = cim.solve(precision=precision)
    return cim.get_ci_energy(), info.n_matvec, info.n_matvec_f32


def oracle_energy(r, arrays):
    h, V, U = arrays
    hr, Vr = O.rotate_integrals(U, h, V)
    m = O.build_crs("DOCI" if r.kind == CA.CF_DOCI else "FCI", hr, Vr, r.n, r.n)
    E, _, info = O.lanczos(O.full_from_upper(m), tol=1e-12)
    assert info["status"] == 0
    return E


@pytest.mark.parametrize("precision", [CA.CF_F32, CA.CF_F64])
def test_alone_equals_in_mix(precision):
    reqs = small_stream()
    arrs = S.prepare(reqs)
    say("alone", precision, [(r.kind, r.L) for r in reqs])
    alone = [solve_alone(r, a, precision) for r, a in zip(reqs, arrs)]
    say("mix")
    out = S.run_stream(reqs, arrs, slots=4, precision=precision, timeout=240)
    for r, (E, nmv, nmv32), d in zip(reqs, alone, out["results"]):
        assert d["tag"] == r.k and d["status"] == 0 and d["solve_status"] == 0
        assert d["energy"] == E, (r, d["energy"], E)
        assert d["n_matvec"] == nmv and d["n_matvec_f32"] == nmv32
    # overlap actually happened: some request started before another ended
    res = out["results"]
    overlapped = any(a["t_start"] < b["t_end"] and b["t_start"] < a["t_end"]
                     for i, a in enumerate(res) for b in res[i + 1:])
    assert overlapped


def test_served_fp32_energies_match_fp64_oracle():
    reqs = small_stream(8)
    arrs = S.prepare(reqs)
    out = S.run_stream(reqs, arrs, slots=4, precision=CA.CF_F32, timeout=240)
    for r, a, d in zip(reqs, arrs, out["results"]):
        say("oracle", r)
        E = oracle_energy(r, a)
        assert abs(d["energy"] - E) <= 1e-8, (r, d["energy"], E)


def test_admission_by_footprint_serialises_and_releases():
    reqs = small_stream(6)
    arrs = S.prepare(reqs)
    # no two requests fit together; one larger than the budget runs only when nothing else does
    budget = int(1.5 * min(CA.SolveServer.footprint(r.kind, r.L, r.n, r.n, CA.CF_F32) for r in reqs))
    ref = S.run_stream(reqs, arrs, slots=4, precision=CA.CF_F32, timeout=240)
    with CA.SolveServer(slots=4, hbm_budget=float(budget)) as srv:
        out = S.run_stream(reqs, arrs, precision=CA.CF_F32, server=srv, timeout=240)
        st = srv.stats()
        assert st["hbm_in_use"] == 0 and st["finished"] == len(reqs) and st["hbm_budget"] == budget
    res = out["results"]
    # one at a time: no two runs overlap
    spans = sorted((d["t_start"], d["t_end"]) for d in res)
    assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))
    assert [d["energy"] for d in res] == [d["energy"] for d in ref["results"]]


def test_failed_request_reports_error_and_others_finish():
    reqs = small_stream(4)
    arrs = S.prepare(reqs)
    with CA.SolveServer(slots=2) as srv:
        for r, (h, V, U) in zip(reqs, arrs):
            srv.submit(r.k, r.kind, h, V, r.n, r.n, U=U)
        h, V, U = arrs[0]
        srv.submit(99, CA.CF_DOCI, h, V, 3, 2, U=U)   # DOCI needs nalpha == nbeta
        srv.drain(240)
        res = {d["tag"]: d for d in srv.poll()}
    assert set(res) == {r.k for r in reqs} | {99}
    assert res[99]["status"] != 0 and res[99]["error"]
    assert all(res[r.k]["status"] == 0 for r in reqs)


def test_precision_auto_resolves_per_route():
    """CF_PREC_AUTO: FP32 values + FP64 polish for a stored-CRS request, FP64
    for the on-the-fly sigma (lock-step batch and single solve), the one-launch
    DOCI batch and the block-staged lock-step batch; every energy within 1e-8
    Eh of the oracle's FP64 solve, and the admission footprint that of the
    resolved precision."""
    import time

    cases = [  # kind, L, direct, method, FP32 phase expected
        (CA.CF_FCI, 8, False, CA.CF_LANCZOS, True),
        (CA.CF_FCI, 8, True, CA.CF_SOLVER_AUTO, False),
        (CA.CF_FCI, 9, True, CA.CF_LANCZOS, False),
        (CA.CF_DOCI, 12, False, CA.CF_SOLVER_AUTO, False),
        (CA.CF_DOCI, 17, False, CA.CF_SOLVER_AUTO, False),
    ]
    reqs, arrs = [], []
    for k, (kind, L, _, _, _) in enumerate(cases):
        reqs.append(S.StreamRequest(k, kind, L, L // 2, 7000 + k, 0.0))
        arrs.append(S.request_arrays(reqs[-1]))
    results = {}
    with CA.SolveServer(slots=2) as srv:
        for r, (h, V, U), (_, _, direct, method, _) in zip(reqs, arrs, cases):
            srv.submit(r.k, r.kind, h, V, r.n, r.n, U=U, precision=CA.CF_PREC_AUTO, tol_energy=1e-9, check_every=2,
                       direct=direct, method=method)
        deadline = time.perf_counter() + 240
        while len(results) < len(cases):
            for d in srv.poll():
                results[d["tag"]] = d
            assert time.perf_counter() < deadline, "served requests did not finish"
            time.sleep(0.01)
    for r, a, (kind, L, direct, method, f32) in zip(reqs, arrs, cases):
        d = results[r.k]
        say("auto", kind, L, direct, method, d["n_matvec"], d["n_matvec_f32"])
        assert d["status"] == 0, d["error"]
        assert (d["n_matvec_f32"] > 0) == f32, (kind, L, direct, method)
        assert abs(d["energy"] - oracle_energy(r, a)) <= 1e-8
    fp = CA.SolveServer.footprint
    assert fp(CA.CF_FCI, 10, 5, 5, precision=CA.CF_PREC_AUTO) == fp(CA.CF_FCI, 10, 5, 5, precision=CA.CF_F32)
    assert fp(CA.CF_FCI, 10, 5, 5, precision=CA.CF_PREC_AUTO, direct=True) == \
        fp(CA.CF_FCI, 10, 5, 5, precision=CA.CF_F64, direct=True)
