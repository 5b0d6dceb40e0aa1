"""Generate tests/golden/golden.json with the CPU oracle (run in the build
container; the JSON is committed and the GPU tests read only the JSON).

Contents:
  integrals   SHA-256 of the seeded (h, V) arrays for the (L, seed) pairs used
              by the workloads, so a change of the generator is caught.
  config1     FCI L=8 4a4b seed 0 (4,900 determinants): upper-triangle nnz,
              SHA-256 of the reference-layout row/col arrays and of the value
              sign bits, sum |H_ij|, and the ground-state energy from a dense
              numpy eigh of the oracle's CRS.
  beh2        FCI and DOCI energies of the reference fixture (eigh).

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

INTEGRALS = [(8, 0), (12, 1), (24, 2), (16, 3), (4, 5), (6, 7), (10, 11)]


def sha(*arrays) -> str:
    m = hashlib.sha256()
    for a in arrays:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


def crs_digest(m):
    signs = np.signbit(m.data).astype(np.uint8)
    return {
        "n": int(m.n),
        "nnz_upper": int(m.nnz),
        "row_col_sha256": sha(m.row.astype(np.uint64), m.col.astype(np.uint32)),
        "sign_sha256": sha(signs),
        "abs_sum": float(np.abs(m.data).sum()),
    }


def main():
    out = {"integrals": {}}
    for L, seed in INTEGRALS:
        h, V = O.gen_integrals(L, seed)
        out["integrals"][f"{L}_{seed}"] = sha(h, V)
    h, V = O.gen_integrals(8, 0)
    m = O.build_crs("FCI", h, V, 4, 4)
    H = O.dense_from_upper(m)
    w = np.linalg.eigvalsh(H)
    out["config1"] = crs_digest(m)
    out["config1"]["energy"] = float(w[0])
    out["config1"]["gap"] = float(w[1] - w[0])
    d = np.load(os.path.join(HERE, "beh2_sto3g.npz"))
    h, V, na, nb, ec = d["h"], d["V"], int(d["nalpha"]), int(d["nbeta"]), float(d["econst"])
    mf = O.build_crs("FCI", h, V, na, nb)
    md = O.build_crs("DOCI", h, V, na, na)
    out["beh2"] = {
        "fci_energy": float(np.linalg.eigvalsh(O.dense_from_upper(mf))[0] + ec),
        "doci_energy": float(np.linalg.eigvalsh(O.dense_from_upper(md))[0] + ec),
        "fci_crs": crs_digest(mf),
        "doci_crs": crs_digest(md),
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
