"""Extract the reference's own golden test data into small fixtures.

Reads, as text, two data files the reference's unit tests hold and writes
the numbers they contain as fixtures next to this script:

  smith_nh_golden.json  <- mochi_core/test/data/material_test_data.h,
                           namespace smith_neo_hookean_test_data (5 FEniCS
                           cases at E=1, nu=0.4: F, energy, PK1, tangent),
                           used by BatchedMaterials.SmithNeoHookeanGoldenData
                           (mochi_core/test/batched_materials_test.cpp:413, rel tol 1e-4 at :154).
  krylov_psd100.npz     <- mochi_core/test/data/krylov_solver_test_data.h,
                           kPsdMatrix* (100x100 sparse PSD system, rhs,
                           solution, "0.4 + 0.6/diag" vector), used by
                           KrylovSolver.Pcg (mochi_core/test/pcg_test.cpp:276).

Run in a container that has the reference at /root/reference; the fixtures
are committed, so nothing reads the reference at test time.
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

REF = "/root/reference/superdex_physics/libraries/mochi/mochi_core/test/data"
HERE = os.path.dirname(os.path.abspath(__file__))
_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _numbers(text: str) -> list[float]:
    # literals carry a "_r" suffix; identifiers like Real3/NdArray<real, 3, 3> hold integers
    # that are not data, so only floating literals (with '.' or exponent) are taken
    return [float(m) for m in re.findall(r"[-+]?(?:\d+\.\d*|\.\d+)(?:[eE][-+]?\d+)?(?=_r)", text)]


def smith_cases() -> list[dict]:
    text = open(os.path.join(REF, "material_test_data.h")).read()
    start = text.index("namespace smith_neo_hookean_test_data")
    end = text.index("} // namespace smith_neo_hookean_test_data") if "} // namespace smith_neo_hookean_test_data" in text else len(text)
    block = _strip_comments(text[start:end])
    nums = _numbers(block)
    per = 9 + 1 + 9 + 81
    n = len(nums) // per
    if n * per != len(nums) or n != 5:
        raise RuntimeError(f"unexpected layout: {len(nums)} numbers")
    cases = []
    for c in range(n):
        v = nums[c * per:(c + 1) * per]
        cases.append({"F": v[0:9], "energy": v[9], "P": v[10:19], "tangent": v[19:100]})
    return cases


def krylov_system():
    text = _strip_comments(open(os.path.join(REF, "krylov_solver_test_data.h")).read())

    def arr(name, kind=float):
        m = re.search(r"constexpr\s+\w+\s+" + name + r"\[\]\s*=\s*\{(.*?)\};", text, flags=re.S)
        body = m.group(1)
        if kind is int:
            return np.array([int(x) for x in re.findall(r"-?\d+", body)], dtype=np.int32)
        return np.array([float(x) for x in re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?(?=_r)", body)])

    rows = arr("kPsdMatrixRows", int)
    cols = arr("kPsdMatrixCols", int)
    vals = arr("kPsdMatrixVals")
    return {
        "rows": rows, "cols": cols, "vals": vals,
        "pc": arr("kPsdMatrixInvDiag"), "rhs": arr("kPsdMatrixRhs"), "x": arr("kPsdMatrixOut"),
    }


def main() -> int:
    cases = smith_cases()
    with open(os.path.join(HERE, "smith_nh_golden.json"), "w") as f:
        json.dump({"source": "mochi_core/test/data/material_test_data.h smith_neo_hookean_test_data",
                   "youngs_modulus": 1.0, "poisson_ratio": 0.4, "rel_tol": 1e-4, "cases": cases}, f, indent=1)
    k = krylov_system()
    assert len(k["rows"]) == len(k["cols"]) == len(k["vals"]) == 3718, [len(v) for v in k.values()]
    assert len(k["rhs"]) == len(k["x"]) == len(k["pc"]) == 100
    np.savez_compressed(os.path.join(HERE, "krylov_psd100.npz"), **k)
    print("wrote", len(cases), "material cases and the 100x100 Krylov system")
    return 0


if __name__ == "__main__":
    sys.exit(main())
