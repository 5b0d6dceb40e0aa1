"""GPU tests of cf_solve_doci_batch: many small DOCI solves in one launch,
one workgroup per request (the loop the reference runs one solve at a time:
OrbitalOptimization::update_cim -> set_ham -> build_ci -> solve,
SimulatedAnnealing.cpp:83-87; DOCI::construct_CI_matrix, CIMethod.cpp:611-660).

* every request of a mixed batch (L = 7 ... 16, rotated integrals, a RedBCS
  model, a non-half-filled space and the reference's BeH2 KAT) matches the
  oracle's FP64 CRS + Lanczos (or the KAT) within 1e-10 Eh at tol 1e-12, and
  within 1e-8 Eh with the served energy stop;
* a request solved alone (batch of one) and inside the batch gives bitwise the
  same energy and product count;
* requests that do not qualify come back with an error and leave the rest.
"""
import json
import os
import sys

import numpy as np
import pytest

import ciflow_amd as CA
from oracle import oracle as O
from oracle.richardson import redbcs_integrals

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def batch_requests():
    reqs = []
    for k, (L, npair) in enumerate([(8, 4), (10, 5), (12, 6), (13, 6), (14, 7), (15, 7), (16, 8), (14, 3), (12, 6)]):
        h, V = CA.gen_integrals(L, 300 + k)
        U = CA.expm_skew(CA.gen_rotation(L, 300 + k, 0.05))
        reqs.append({"tag": k, "h": h, "V": V, "npair": npair, "U": U})
    h, V = redbcs_integrals(np.arange(12.0), -0.5)
    reqs.append({"tag": 100, "h": h, "V": V, "npair": 6})
    d = np.load(os.path.join(GOLD, "beh2_sto3g.npz"))
    reqs.append({"tag": 101, "h": d["h"], "V": d["V"], "npair": int(d["nalpha"]), "econst": float(d["econst"])})
    return reqs


def oracle_energy(q):
    h, V = q["h"], q["V"]
    if q.get("U") is not None:
        h, V = O.rotate_integrals(q["U"], h, V)
    m = O.build_crs("DOCI", h, V, q["npair"], q["npair"])
    E, _, info = O.lanczos(O.full_from_upper(m), tol=1e-12)
    assert info["status"] == 0
    return E + q.get("econst", 0.0)


def test_batched_doci_matches_oracle_and_kat():
    reqs = batch_requests()
    say("batch of", len(reqs))
    res = CA.solve_doci_batch(reqs)
    kat = json.load(open(os.path.join(GOLD, "golden.json")))["beh2"]["doci_energy"]
    for q, r in zip(reqs, res):
        assert r["status"] == 0 and r["solve_status"] == 0, (q["tag"], r["error"])
        E = kat if q["tag"] == 101 else oracle_energy(q)
        say(q["tag"], r["n"], r["n_matvec"], r["energy"], E, r["energy"] - E)
        assert abs(r["energy"] - E) <= 1e-10, (q["tag"], r["energy"], E)
    assert abs(res[-1]["energy"] - (-15.225362047972723)) <= 1e-10   # TestDOCI.cpp:31


def test_batched_doci_alone_equals_in_batch_and_energy_stop():
    reqs = batch_requests()
    served = [dict(q, tol_energy=1e-9, check_every=2) for q in reqs]
    mix = CA.solve_doci_batch(served)
    for q, r in zip(served, mix):
        alone = CA.solve_doci_batch([q])[0]
        assert alone["energy"] == r["energy"] and alone["n_matvec"] == r["n_matvec"], q["tag"]
        if q["tag"] < 100:
            assert abs(r["energy"] - oracle_energy(q)) <= 1e-8, q["tag"]


def test_batched_doci_rejects_what_does_not_fit():
    h, V = CA.gen_integrals(18, 5)   # C(18, 9) = 48,620 determinants: too large
    hs, Vs = CA.gen_integrals(10, 6)
    reqs = [{"tag": 1, "h": h, "V": V, "npair": 9}, {"tag": 2, "h": hs, "V": Vs, "npair": 5},
            {"tag": 3, "h": hs, "V": Vs, "npair": 0}]
    res = CA.solve_doci_batch(reqs)
    assert res[0]["status"] != 0 and "too large" in res[0]["error"]
    assert res[2]["status"] != 0
    assert res[1]["status"] == 0 and abs(res[1]["energy"] - oracle_energy(reqs[1])) <= 1e-10


def test_context_holding_batch_tables_is_destroyed_cleanly():
    """A context that ran batched solves caches the strings and hop columns of
    each (L, pairs) shape; destroying it releases them before its stream
    (regression: they were released after, on a destroyed stream)."""
    import gc
    for _ in range(2):
        ctx = CA.Context(0)
        h, V = CA.gen_integrals(10, 8)
        r = CA.solve_doci_batch([{"h": h, "V": V, "npair": 5}], ctx)[0]
        assert r["status"] == 0
        del ctx
        gc.collect()
