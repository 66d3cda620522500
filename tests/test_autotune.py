"""The tile-configuration choice (isoext_amd/autotune.py) on the CPU: the
committed table's size classes, the library's static rule for classes the
table lacks, and the per-shape memo (which a rewritten table clears)."""

from __future__ import annotations

import json

from isoext_amd import autotune


def test_table_classes_and_static_rule():
    table = json.loads(autotune.TABLE.read_text())
    shape = (1024, 1024, 1024)
    assert autotune.size_class(shape) == "huge-zbig"
    assert autotune.mc_config(shape) == table["huge-zbig"]
    # a class the table does not hold: the library's static rule (host code)
    odd = (5, 7, 3000)
    assert autotune.size_class(odd) not in table
    assert 0 <= autotune.mc_config(odd) < len(autotune.CONFIGS)


def test_memo_holds_until_the_table_changes(monkeypatch):
    shape = (300, 300, 70)
    cls = autotune.size_class(shape)
    first = autotune.mc_config(shape)
    assert autotune._by_shape[shape] == first
    other = (first + 1) % len(autotune.CONFIGS)
    table = dict(autotune._load())
    table[cls] = other
    monkeypatch.setattr(autotune, "_table", table)
    # the memo answers until it is cleared, as tune() does when it writes the table
    assert autotune.mc_config(shape) == first
    autotune._by_shape.clear()
    assert autotune.mc_config(shape) == other
    monkeypatch.undo()
    autotune._by_shape.clear()
    assert autotune.mc_config(shape) == first
