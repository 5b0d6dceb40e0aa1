"""GPU tests of cf_solve_batch: same-shape Davidson solves advanced in lock
step (the loop the reference runs one solve at a time:
OrbitalOptimization::update_cim -> set_ham -> build_ci -> solve,
SimulatedAnnealing.cpp:83-87; CIMethod::solve, CIMethod.cpp:241-268).

* FP64 batched energies match the oracle's FP64 CRS + Lanczos on the same
  rotated integrals within 1e-10 Eh (FCI with nalpha == nbeta and !=, DOCI),
  and the single-request Davidson solver (cf_solve_direct / cf_solve) within
  1e-10 Eh;
* FP32 operator + FP64 polish with the served energy stop lands within
  1e-8 Eh of the oracle;
* a request gives bitwise the same energy and product counts alone and in a
  batch of mixed classes and precisions, and in a batch split over two
  launches (more than 32 requests of one shape);
* requests the batch does not take come back with an error and leave the
  rest; a basis restart happens (more than 24 products).
"""
import sys

import numpy as np
import pytest

import ciflow_amd as CA
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def req(kind, L, nup, ndown, seed, precision=CA.CF_F64, rot=0.05, **kw):
    h, V = CA.gen_integrals(L, seed)
    U = CA.expm_skew(CA.gen_rotation(L, seed, rot)) if rot > 0 else None
    d = dict(kind=kind, h=h, V=V, nup=nup, ndown=ndown, U=U, precision=precision, tag=seed)
    d.update(kw)
    return d


def oracle_energy(q):
    h, V = q["h"], q["V"]
    if q.get("U") is not None:
        h, V = O.rotate_integrals(q["U"], h, V)
    kind = "DOCI" if q["kind"] == CA.CF_DOCI else "FCI"
    m = O.build_crs(kind, h, V, q["nup"], q["ndown"])
    E, _, info = O.lanczos(O.full_from_upper(m), tol=1e-12)
    assert info["status"] == 0
    return E


def single_davidson(q):
    ham = CA.Hamiltonian(q["h"], q["V"], q["nup"], q["ndown"])
    if q.get("U") is not None:
        ham = ham.rotate(q["U"])
    if q["kind"] == CA.CF_FCI:
        cim = CA.FCI_Direct(ham, value_prec=q["precision"])
    else:
        cim = CA.DOCI(ham, value_prec=q["precision"])
    info = cim.solve(precision=q["precision"], method=CA.CF_DAVIDSON, tol_energy=q.get("tol_energy", 0.0))
    return info


def test_fp64_batch_matches_oracle_and_single_davidson():
    reqs = [req(CA.CF_FCI, 8, 4, 4, 11), req(CA.CF_FCI, 8, 4, 4, 12), req(CA.CF_FCI, 8, 5, 4, 13),
            req(CA.CF_DOCI, 12, 6, 6, 14), req(CA.CF_DOCI, 12, 6, 6, 15), req(CA.CF_FCI, 6, 3, 3, 16)]
    res = CA.solve_batch(reqs)
    for q, r in zip(reqs, res):
        assert r["status"] == 0 and r["solve_status"] == 0, (q["tag"], r["error"])
        E = oracle_energy(q)
        s = single_davidson(q)
        say(q["tag"], q["kind"], q["nup"], q["ndown"], r["n"], r["n_matvec"], r["energy"] - E,
            r["energy"] - s.eigval, s.n_matvec)
        assert abs(r["energy"] - E) <= 1e-10, (q["tag"], r["energy"], E)
        assert abs(r["energy"] - s.eigval) <= 1e-10
    assert max(r["n_matvec"] for r in res) > 24  # at least one basis restart


def test_fp32_polish_with_energy_stop_matches_oracle():
    reqs = [req(CA.CF_FCI, 8, 4, 4, 21, CA.CF_F32, tol_energy=1e-9),
            req(CA.CF_FCI, 9, 4, 4, 22, CA.CF_F32, tol_energy=1e-9),
            req(CA.CF_DOCI, 14, 7, 7, 23, CA.CF_F32, tol_energy=1e-9),
            req(CA.CF_FCI, 8, 4, 4, 24, CA.CF_F64, tol_energy=1e-9)]
    res = CA.solve_batch(reqs)
    for q, r in zip(reqs, res):
        assert r["status"] == 0 and r["solve_status"] == 0, (q["tag"], r["error"])
        E = oracle_energy(q)
        say(q["tag"], r["n_matvec_f32"], r["n_matvec"], r["energy"] - E)
        assert abs(r["energy"] - E) <= 1e-8, (q["tag"], r["energy"], E)
        if q["precision"] == CA.CF_F32:
            assert r["n_matvec_f32"] > 0 and r["n_matvec"] > 0
        else:
            assert r["n_matvec_f32"] == 0


def test_alone_equals_in_batch_bitwise():
    reqs = [req(CA.CF_FCI, 8, 4, 4, 31, CA.CF_F32, tol_energy=1e-9),
            req(CA.CF_FCI, 8, 4, 4, 32, CA.CF_F64),
            req(CA.CF_FCI, 8, 4, 4, 33, CA.CF_F32),
            req(CA.CF_DOCI, 16, 8, 8, 34, CA.CF_F32, tol_energy=1e-9),
            req(CA.CF_DOCI, 16, 8, 8, 35, CA.CF_F64),
            req(CA.CF_FCI, 9, 4, 4, 36, CA.CF_F32, tol_energy=1e-9)]
    mix = CA.solve_batch(reqs)
    for q, r in zip(reqs, mix):
        alone = CA.solve_batch([q])[0]
        say(q["tag"], r["energy"], alone["energy"], r["n_matvec_f32"], r["n_matvec"])
        assert r["status"] == 0 and alone["status"] == 0
        assert alone["energy"] == r["energy"], q["tag"]
        assert (alone["n_matvec"], alone["n_matvec_f32"]) == (r["n_matvec"], r["n_matvec_f32"]), q["tag"]


def test_more_than_one_launch_of_a_shape():
    reqs = [req(CA.CF_DOCI, 10, 5, 5, 100 + k, CA.CF_F32, tol_energy=1e-9) for k in range(34)]
    res = CA.solve_batch(reqs)
    for k in (0, 31, 32, 33):
        alone = CA.solve_batch([reqs[k]])[0]
        assert res[k]["status"] == 0 and alone["energy"] == res[k]["energy"], k
    for q, r in zip(reqs[::11], res[::11]):
        assert abs(r["energy"] - oracle_energy(q)) <= 1e-8


def test_rejected_requests_leave_the_rest():
    good = req(CA.CF_FCI, 8, 4, 4, 41)
    reqs = [dict(good, tag=1, method=CA.CF_LANCZOS),
            dict(req(CA.CF_DOCI, 10, 5, 4, 42), tag=2),
            dict(good, tag=3)]
    res = CA.solve_batch(reqs)
    assert res[0]["status"] != 0 and "Davidson" in res[0]["error"]
    assert res[1]["status"] != 0 and "nalpha == nbeta" in res[1]["error"]
    assert res[2]["status"] == 0 and abs(res[2]["energy"] - oracle_energy(good)) <= 1e-10


def test_server_refills_a_running_batch():
    """Requests of one shape that arrive while their shape is being served
    join the running batch (one worker slot, so overlapping service means a
    shared batch): every result equals the request solved alone, bitwise."""
    from ciflow_amd import stream as S
    reqs = [S.StreamRequest(k, CA.CF_FCI, 8, 4, 5000 + k, 0.004 * k) for k in range(10)]
    arrs = S.prepare(reqs)
    served = dict(precision=CA.CF_F32, check_every=2, tol_energy=1e-9, direct_fci=True, method=CA.CF_SOLVER_AUTO)
    with CA.SolveServer(slots=1) as srv:
        out = S.run_stream(reqs, arrs, server=srv, timeout=120, **served)["results"]
    for r, (h, V, U), d in zip(reqs, arrs, out):
        alone = CA.solve_batch([dict(kind=CA.CF_FCI, h=h, V=V, nup=4, ndown=4, U=U, precision=CA.CF_F32,
                                     tol_energy=1e-9, tag=r.k)])[0]
        say(r.k, d["t_start"] - out[0]["t_start"], d["t_end"] - out[0]["t_start"], d["energy"], alone["energy"])
        assert d["status"] == 0 and alone["energy"] == d["energy"], r.k
        assert (alone["n_matvec"], alone["n_matvec_f32"]) == (d["n_matvec"], d["n_matvec_f32"]), r.k
    joined = [b for b in out if any(a["t_start"] < b["t_start"] < a["t_end"] for a in out)]
    assert joined, "no request joined a running batch"
