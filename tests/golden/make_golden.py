"""Write tests/golden/ref_scaling.npz: input/output vectors of the reference's own
build_scaling, estimate_spectral_norm and SparseMatrix::multiply / transpose_multiply
(src/la/scaling.cpp, src/la/sparse.cpp), compiled by oracle/Makefile into oracle/_ref.

Run here, where /root/reference exists:  python tests/golden/make_golden.py
The fixtures are data only (inputs and the reference's outputs); tests compare the oracle
and the GPU engine against them without needing the reference at run time.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402
from tests.mps_ref import read_mps  # noqa: E402
from tests.lp_cases import dense_to_lp  # noqa: E402


def random_lp(seed: int, m: int, n: int, density: float, spread: float) -> dict:
    rng = np.random.default_rng(seed)
    a = np.where(rng.random((m, n)) < density,
                 np.exp(rng.uniform(-spread, spread, (m, n))) * rng.choice([-1.0, 1.0], (m, n)), 0.0)
    a[m // 2, :] = 0.0  # an empty row
    a[:, n // 3] = 0.0  # an empty column
    lo = np.where(rng.random(m) < 0.5, -np.inf, rng.standard_normal(m))
    hi = np.where(rng.random(m) < 0.3, np.inf, lo + rng.random(m) * 5)
    hi = np.where(np.isinf(lo) & np.isinf(hi), 1.0, hi)
    cu = np.where(rng.random(n) < 0.3, np.inf, rng.uniform(1, 10, n))
    cu[0] = 1e30  # MPS-scale infinity
    return dense_to_lp(a, lo, hi, rng.standard_normal(n), cu)


def cases() -> dict[str, dict]:
    import sih26_amd as sk
    out = {"c1_seed1": sk.Lp.generate(1, 1).export(),
           "c3_member30": sk.Lp.generate(3, 30).export(),
           "afiro": read_mps(ROOT / "tests" / "golden" / "netlib" / "afiro.mps")}
    out["random_a"] = random_lp(1, 40, 70, 0.15, 3.0)
    out["random_b"] = random_lp(2, 90, 60, 0.08, 7.0)
    for v in out.values():
        v.pop("row_names", None)
        v.pop("col_names", None)
    return out


def main() -> None:
    oracle.build(ref=True)
    if oracle.ref() is None:
        raise SystemExit("oracle/_ref was not built: /root/reference is required")
    blob = {}
    for name, lp in cases().items():
        a = oracle.arrays(lp)
        s = oracle.build_scaling(a, use_ref=True)
        rng = np.random.default_rng(123)
        x = rng.standard_normal(a["num_cols"])
        y = rng.standard_normal(a["num_rows"])
        for k in ("col_starts", "row_indices", "values", "cost", "col_lower", "col_upper",
                  "row_lower", "row_upper"):
            blob[f"{name}/in/{k}"] = a[k]
        blob[f"{name}/in/dims"] = np.array([a["num_rows"], a["num_cols"], a["sense"]], np.int64)
        blob[f"{name}/in/x"] = x
        blob[f"{name}/in/y"] = y
        for k in ("row", "column", "values", "cost", "col_lower", "col_upper", "row_lower",
                  "row_upper"):
            blob[f"{name}/scaling/{k}"] = s[k]
        blob[f"{name}/scaling/minmax"] = np.array([s["min_abs"], s["max_abs"]])
        blob[f"{name}/norm_seed1"] = np.array([oracle.spectral_norm(a, s["values"], seed=1, use_ref=True)])
        blob[f"{name}/ax"] = oracle.multiply(a, x, use_ref=True)
        blob[f"{name}/aty"] = oracle.multiply(a, y, transpose=True, use_ref=True)
    path = Path(__file__).resolve().parent / "ref_scaling.npz"
    np.savez_compressed(path, **blob)
    print(f"wrote {path} ({path.stat().st_size} bytes, {len(blob)} arrays)")


if __name__ == "__main__":
    main()
