"""Autotuner host logic (no GPU): subsamples, the table's lookups and JSON form."""
from ndyn_omp_amd import autotune as AT
from ndyn_omp_amd import workloads as W


def test_subsample_deterministic_and_bounded():
    mix = W.replica_mix(64, seed=7)
    total = sum(s.np for s in mix)
    a = AT.subsample(mix, total // 4)
    b = AT.subsample(mix, total // 4)
    assert [s.np for s in a] == [s.np for s in b]
    assert 0 < len(a) < len(mix)
    assert abs(sum(s.np for s in a) - total // 4) < 0.5 * total // 4
    assert AT.subsample(mix, 10 * total) == list(mix)


def test_table_lookup_json_and_edit():
    t = AT.TuneTable([(2048, 2, 0.75), (8192, 1, 1.0), (1 << 26, 1, 0.5)])
    assert t.skin_for(100) == 0.75
    assert t.skin_for(2049) == 1.0
    assert t.skin_for(10**6) == 0.5
    assert t.skin_for(1 << 27, default=0.3) == 0.3
    assert t.tpa_classes() == [(2048, 2), (8192, 1), (1 << 26, 1)]
    u = AT.TuneTable.from_json(t.to_json())
    assert u.classes == t.classes
    # tables saved without a cell-block shape read as the linear order (1 x 1)
    assert t.classes[1] == (8192, 1, 1.0, 1, 1)
    assert t.block_classes() == [(2048, 1, 1), (8192, 1, 1), (1 << 26, 1, 1)]
    v = t.with_class(1, skin=1.25)
    assert v.classes[1] == (8192, 1, 1.25, 1, 1) and t.classes[1] == (8192, 1, 1.0, 1, 1)
    w = v.with_class(2, block=(2, 1), tpa=2)
    assert w.classes[2] == (1 << 26, 2, 0.5, 2, 1)
    assert w.summary()[f"np<={1 << 26}"] == {"tpa": 2, "skin": 0.5, "cell_block": "1x2x1"}
    assert AT.TuneTable.from_json(w.to_json()).classes == w.classes
