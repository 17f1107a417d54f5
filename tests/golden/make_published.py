"""Write tests/golden/pdhg_published.json: the reference's own published PDHG results on the
nine Netlib fixtures (bench/results/pdhg-79ec7f7.csv, measured with solve()'s defaults:
presolve on, pdhg_tolerance 1e-4, restarts on, pdhg_stop_at_request off), one row per
instance: iterations, the 17-digit objective, the status.

Run here, where /root/reference exists:  python tests/golden/make_published.py
The file is data only (the numbers the CSV holds); no reference code is run.
"""
from __future__ import annotations

import csv
import json
from pathlib import Path

SOURCE = Path("/root/reference/bench/results/pdhg-79ec7f7.csv")
OUT = Path(__file__).resolve().parent / "pdhg_published.json"


def main() -> None:
    rows = {}
    with SOURCE.open() as f:
        for r in csv.DictReader(f):
            if r["algorithm"] != "pdhg" or r["tolerance"] != "0.0001" or r["restarts_enabled"] != "1":
                continue
            rows.setdefault(r["instance"], {"iterations": int(r["iterations"]),
                                            "objective": r["objective"], "status": r["status"],
                                            "published_optimal": r["published_objective"]})
    OUT.write_text(json.dumps({"source": "bench/results/pdhg-79ec7f7.csv", "machine": "Windows-AMD64",
                               "instances": rows}, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT} ({len(rows)} instances)")


if __name__ == "__main__":
    main()
