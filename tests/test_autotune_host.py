"""Host side of the launch-shape autotuner (no GPU needed): the committed
store holds one valid candidate per shape for each benchmark workload, a
missing workload or store gives None (the library's defaults), and the
candidate table lists every candidate with one choice per shape."""
import json

import shell_amd as S
from shell_amd import autotune


def test_store_holds_valid_choices():
    for name in ("warm1024", "sweep512"):
        shp = autotune.load_shapes(name)
        assert shp is not None and set(shp) == set(S.SHAPE_FIELDS)
        for k, v in shp.items():
            assert v in autotune.CANDIDATES[k]
    assert autotune.load_shapes("no such workload") is None
    assert autotune.load_shapes("warm1024", path="/nonexistent/shapes.json") is None
    assert set(S.DEFAULT_SHAPES) == set(autotune.CANDIDATES)
    for k, v in S.DEFAULT_SHAPES.items():
        assert v in autotune.CANDIDATES[k]


def test_candidate_table():
    with open(autotune.STORE) as f:
        d = json.load(f)
    rows = [(c["shape"], c["candidate"], c["kernel"], c["launches"], c["avg_ms"], c["total_ms"], c["chosen"])
            for c in d["warm1024"]["candidates"]]
    md = autotune.markdown(rows).splitlines()
    assert len(md) == 2 + sum(len(v) for v in autotune.CANDIDATES.values())
    for dim, cands in autotune.CANDIDATES.items():
        mine = [r for r in rows if r[0] == dim]
        assert [r[1] for r in mine] == list(cands)
        assert sum(r[6] for r in mine) == 1
        chosen = [r for r in mine if r[6]][0]
        assert chosen[1] == d["warm1024"]["shapes"][dim]
        assert chosen[5] == min(r[5] for r in mine)
