"""Writing LPs out (sk_lp_write_mps, sk_lp_save / sk_lp_load): what is written is read back,
by the library's MPS reader and by the independent Python reader, as the same arrays bit for
bit; the instance cache refuses damaged files."""
import numpy as np
import pytest

import sih26_amd as sk
from tests.mps_ref import read_mps
from tests.lp_cases import dense_to_lp
from tests.test_mps import ARRAYS, FIXTURES

INF = np.inf


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def same(a, b):
    for k in ARRAYS:
        np.testing.assert_array_equal(bits(a[k]), bits(b[k]), err_msg=k)
    assert (a["sense"], a["offset"]) == (b["sense"], b["offset"])


def from_dict(d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1),
                             d.get("offset", 0.0))


def every_kind():
    """Every row and bound form the writer distinguishes, a maximisation with an offset, and
    two empty columns (one with a cost)."""
    rows = [[1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
            [0.0, 1.0, -3.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
            [1.0, 0.0, 0.0, 1e-3, 0.0, 0.0, 0.0, 0.0, 0.0],
            [0.0, 0.0, 1.0, 0.0, 7.25, 0.0, 0.0, 0.0, 0.0],
            [0.0, 0.0, 0.0, 1.0, 0.0, 2.0, 0.0, 0.0, 0.0],
            [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]]
    #        E     G      L     ranged      free   ranged (negative bounds)
    lo = [4.0, 1.0 / 3.0, -INF, -2.5, -INF, -7.0]
    hi = [4.0, INF, 10.0, 0.1, INF, -1.0 / 7.0]
    cost = [1.0, -2.0, 0.0, 0.5, 3.0, 1e-9, 0.25, 0.0, 2.0]
    col_lower = [0.0, -INF, -INF, 2.0, -1.5, -3.0, 5.0, 0.0, 0.0]
    col_upper = [INF, INF, 4.0, 2.0, INF, -0.5, 9.0, 1.0, INF]
    d = dense_to_lp(rows, lo, hi, cost, col_upper, sense=-1, offset=12.5, col_lower=col_lower)
    assert d["col_starts"][-1] == d["col_starts"][-3]  # columns 7 and 8 are empty
    return d


CASES = {
    "every_kind": lambda: from_dict(every_kind()),
    "config1": lambda: sk.Lp.generate(1, 1),
    "batch_member": lambda: sk.Lp.generate(3, 17),
    "power_law": lambda: sk.Lp.generate_params(**dict(sk.Lp.preset(2, 11), num_rows=2000,
                                                      num_cols=5000, num_equality_rows=1200,
                                                      target_nnz=25000)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_mps_round_trip_is_exact(tmp_path, name):
    lp = CASES[name]()
    path = tmp_path / f"{name}.mps"
    lp.write_mps(path)
    want = lp.export()
    same(sk.Lp.read_mps(path).export(), want)  # the library's reader
    # the independent Python reader keeps a written 1e30 as 1e30; |v| >= 1e30 is infinite in
    # the model's own rule (types.hpp:32-39), so compare under that rule
    py = read_mps(path)
    for k in ("row_lower", "row_upper", "col_lower", "col_upper"):
        v = np.asarray(py[k], float)
        py[k] = np.where(np.abs(v) >= 1e30, np.copysign(INF, v), v)
    same(py, want)


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
def test_fixtures_survive_read_write_read(tmp_path, path):
    lp = sk.Lp.read_mps(path)
    out = tmp_path / "again.mps"
    lp.write_mps(out)
    again = sk.Lp.read_mps(out)
    same(again.export(), lp.export())
    assert again.names() == lp.names()  # names are kept


@pytest.mark.parametrize("name", sorted(CASES))
def test_instance_cache_round_trip_is_exact(tmp_path, name):
    lp = CASES[name]()
    path = tmp_path / f"{name}.sklp"
    lp.save(path)
    same(sk.Lp.load(path).export(), lp.export())


def refused(path, *words):
    with pytest.raises(sk.SkError) as e:
        sk.Lp.load(path)
    assert e.value.status == 10
    msg = sk.lib().sk_gpu_last_error().decode()
    for w in words:
        assert w in msg, msg


def test_instance_cache_refuses_damaged_files(tmp_path):
    path = tmp_path / "good.sklp"
    sk.Lp.generate(1, 2).save(path)
    data = bytearray(path.read_bytes())
    flipped = tmp_path / "flipped.sklp"
    flipped.write_bytes(data[:200] + bytes([data[200] ^ 0x10]) + data[201:])
    refused(flipped, "checksum")
    short = tmp_path / "short.sklp"
    short.write_bytes(data[:-9])
    refused(short, "size")
    other = tmp_path / "other.sklp"
    other.write_bytes(b"NOTANLP!" + data[8:])
    refused(other, "not an instance cache")
    refused(tmp_path / "missing.sklp", "cannot open")


def test_writers_report_unwritable_paths(tmp_path):
    lp = sk.Lp.generate(1, 1)
    for call in (lp.write_mps, lp.save):
        with pytest.raises(sk.SkError) as e:
            call(tmp_path / "no_such_dir" / "x")
        assert e.value.status == 10
        assert "cannot open" in sk.lib().sk_gpu_last_error().decode()
