"""Tuning-table file handling (host only: the table needs no device)."""
import ctypes

import pytest

import ligesis_amd
from ligesis_amd import LigesisError


@pytest.fixture(autouse=True)
def _clean_table():
    ligesis_amd.tune_clear()
    yield
    ligesis_amd.tune_clear()


def _entries(path):
    return [ln.split() for ln in open(path) if ln.strip() and not ln.lstrip().startswith("#")]


def test_load_save_roundtrip(tmp_path):
    src = tmp_path / "in.txt"
    src.write_text("# comment line\n"
                   "12 12 1 0 8 2 7 2 0.925\n"
                   "\n"
                   "  10 14 2 0 0 0 8 1 1.5\n")
    assert ligesis_amd.tune_load(str(src)) == 2
    out = tmp_path / "out.txt"
    assert ligesis_amd.tune_save(str(out)) == 2
    got = _entries(out)
    assert [g[:8] for g in got] == [["10", "14", "2", "0", "0", "0", "8", "1"],
                                    ["12", "12", "1", "0", "8", "2", "7", "2"]]
    assert float(got[1][8]) == pytest.approx(0.925)
    # a second load merges by shape: the newer entry replaces the older one
    upd = tmp_path / "upd.txt"
    upd.write_text("12 12 1 3 16 0 6 4 0.9\n")
    assert ligesis_amd.tune_load(str(upd)) == 1
    ligesis_amd.tune_save(str(out))
    assert [g[:8] for g in _entries(out)][1] == ["12", "12", "1", "3", "16", "0", "6", "4"]


@pytest.mark.parametrize("line", [
    "12 12 1 0 8 2 7 1",              # missing field
    "12 12 1 0 8 2 7 1 0.9 extra",    # trailing field
    "12 12 1 13 8 2 7 1 0.9",         # hash_impl out of range
    "12 12 1 0 6 2 7 1 0.9",          # ksplit not a power of two
    "12 12 1 0 8 6 7 1 0.9",          # ntt_impl out of range
    "12 12 1 0 8 2 9 1 0.9",          # merkle_lg out of range
    "12 12 1 0 8 2 7 3 0.9",          # pipeline not 1, 2, 4 or 8
    "21 12 1 0 8 2 7 1 0.9",          # log_m out of range
    "12 12 x 0 8 2 7 1 0.9",          # not a number
])
def test_malformed_file_rejected_whole(tmp_path, line):
    p = tmp_path / "bad.txt"
    p.write_text("10 10 1 0 4 0 8 1 0.1\n" + line + "\n")
    with pytest.raises(LigesisError) as ei:
        ligesis_amd.tune_load(str(p))
    assert ei.value.code == ligesis_amd.LGS_ERR_ARG
    out = tmp_path / "out.txt"
    assert ligesis_amd.tune_save(str(out)) == 0  # nothing of the bad file was merged


def test_missing_file_and_null_args(tmp_path):
    with pytest.raises(LigesisError):
        ligesis_amd.tune_load(str(tmp_path / "absent.txt"))
    L = ligesis_amd.lib()
    assert L.lgs_tune_load(None, None, None) == ligesis_amd.LGS_ERR_ARG
    assert L.lgs_autotune(None, None, 0, None, None) == ligesis_amd.LGS_ERR_ARG
    assert L.lgs_tune_get(None, 12, 12, None) == ligesis_amd.LGS_ERR_ARG
