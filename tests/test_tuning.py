"""Tuning table plumbing on the host (no GPU): install, read back, reject
invalid entries, JSON round trip."""
import json

import pytest

from tetragpu_amd import tuning
from tetragpu_amd._native import NativeError


def test_install_and_read_back(tmp_path):
    table = [{"max_cells": 1000, "variant": 1, "lane_cap": 32, "lane_block": 64, "bv_log": 7,
              "tets_per_chunk": 8192},
             {"max_cells": 1 << 62, "variant": 2, "lane_cap": 32, "lane_block": 128, "bv_log": 8,
              "tets_per_chunk": 4096}]
    tuning.install(list(reversed(table)))
    assert tuning.current() == table  # sorted by max_cells
    p = tmp_path / "t.json"
    tuning.save(table, p)
    assert tuning.load(p) == table and json.loads(p.read_text()) == table
    tuning.install([])
    assert tuning.current() == []


@pytest.mark.parametrize("bad", [{"variant": 5}, {"variant": 0}, {"lane_cap": 48}, {"bv_log": 9},
                                 {"tets_per_chunk": 100}, {"lane_cap": 64, "lane_block": 256}])
def test_invalid_entries_rejected(bad):
    e = {"max_cells": 10, "variant": 2, "lane_cap": 32, "lane_block": 128, "bv_log": 7,
         "tets_per_chunk": 8192}
    e.update(bad)
    with pytest.raises(NativeError):
        tuning.install([e])
