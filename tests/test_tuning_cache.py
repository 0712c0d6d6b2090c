"""Process-wide tuning cache (rdu_tune_cache_*): file round trip, merge and rejection, on the CPU.

These entry points touch no device: they run without a GPU.
"""
import pytest

from rendu_amd import tuning
from rendu_amd._lib import RduError

HEADER = "rdu-tune 2\n"


@pytest.fixture(autouse=True)
def _empty_cache():
    tuning.clear()
    yield
    tuning.clear()


def test_round_trip(tmp_path):
    src = tmp_path / "a.txt"
    src.write_text(HEADER + "20 11 0 23 1 24 128 6\n" + "13 5 3 16 8 0 0 0\n" + "\n" + "13 5 3 16 2 64 512 1\n")
    tuning.load(src)
    assert tuning.size() == 3
    out = tmp_path / "b.txt"
    tuning.save(out)
    lines = out.read_text().splitlines()
    assert lines[0] == HEADER.strip()
    # Entries come back sorted by key, with their values.
    assert sorted(lines[1:]) == sorted(["20 11 0 23 1 24 128 6", "13 5 3 16 8 0 0 0", "13 5 3 16 2 64 512 1"])
    tuning.clear()
    assert tuning.size() == 0
    tuning.load(out)
    assert tuning.size() == 3


def test_merge_file_entries_win(tmp_path):
    a = tmp_path / "a.txt"
    a.write_text(HEADER + "20 11 0 23 1 24 128 6\n")
    b = tmp_path / "b.txt"
    b.write_text(HEADER + "20 11 0 23 1 8 64 0\n10 4 0 10 0 16 0 0\n")
    tuning.load(a)
    tuning.load(b)
    assert tuning.size() == 2
    out = tmp_path / "c.txt"
    tuning.save(out)
    assert "20 11 0 23 1 8 64 0" in out.read_text().splitlines()


@pytest.mark.parametrize("body", [
    "",                                 # no header
    "rdu-tune 1\n",                     # the format of 128-thread workgroups
    HEADER + "20 11 0 23 1 24 128\n",   # a field missing
    HEADER + "20 11 0 23 1 24 128 6 7\n",  # a field too many
    HEADER + "20 11 0 23 1 65 128 6\n",  # refill above 64
    HEADER + "20 11 0 23 1 24 8 6\n",    # chunk below 16
    HEADER + "20 11 0 23 16 24 64 6\n",  # unknown class bits
    HEADER + "x 11 0 23 1 24 64 6\n",    # not a number
])
def test_malformed_file_changes_nothing(tmp_path, body):
    good = tmp_path / "good.txt"
    good.write_text(HEADER + "12 4 0 12 0 16 0 0\n")
    tuning.load(good)
    bad = tmp_path / "bad.txt"
    bad.write_text(body + "13 5 3 16 8 0 0 0\n" if body == "" else body)
    with pytest.raises(RduError) as e:
        tuning.load(bad)
    assert e.value.code == -1
    assert tuning.size() == 1


def test_missing_file(tmp_path):
    with pytest.raises(RduError):
        tuning.load(tmp_path / "absent.txt")
    with pytest.raises(RduError):
        tuning.save(tmp_path / "no_dir" / "x.txt")
