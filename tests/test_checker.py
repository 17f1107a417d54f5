"""The answer checker (oracle/checker.py) against the reference checker's own verdicts on the
same model/answer pairs (tests/golden/checker/verdicts.json, written by
tests/golden/make_checker_fixtures.py). The answers were written by sk_write_solution, so the
reference checker accepting the good ones also pins the .sol writer."""
import json
from pathlib import Path

import pytest

from oracle import checker
from tests.mps_ref import read_mps

GOLD = Path(__file__).resolve().parent / "golden"
VERDICTS = json.loads((GOLD / "checker" / "verdicts.json").read_text())


@pytest.mark.parametrize("case", sorted(VERDICTS))
def test_checker_matches_the_reference_verdict(case):
    v = VERDICTS[case]
    model = read_mps(GOLD / v["model"])
    report = checker.verify(model, checker.read_sol(GOLD / v["sol"]))
    assert [[n, ok] for n, ok in report.outcome()] == v["checks"], report.lines
    assert report.verified == v["verified"]


def test_good_answers_verify_and_bad_ones_do_not():
    good = [c for c in VERDICTS if c.endswith(".good")]
    assert len(good) >= 7 and all(VERDICTS[c]["verified"] for c in good)
    assert not any(VERDICTS[c]["verified"] for c in VERDICTS if not c.endswith(".good"))


def test_sol_reader_handles_quoted_names(tmp_path):
    p = tmp_path / "q.sol"
    p.write_text('status optimal\nbegin columns 1\n"DEDO3 11" 1.5 -0.25 basic\nend columns\n'
                 'begin rows 1\n"A\\"B" 2 3 basic\nend rows\n')
    sol = checker.read_sol(p)
    assert sol.columns == {"DEDO3 11": (1.5, -0.25)}
    assert sol.rows == {'A"B': (2.0, 3.0)}
