"""CPU checks of the committed tuning table (matflat_amd/tuning/gfx950.json):
every entry is a choice the library accepts, and the exact-Cholesky shapes
the autotuner considers are the ones the library builds."""
import json
import os

from matflat_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_table_entries_are_valid_choices():
    t = json.load(open(os.path.join(ROOT, "matflat_amd", "tuning", "gfx950.json")))
    assert t["arch"] == "gfx950"
    for rb, w in t["fast_waves"].items():
        rb = int(rb)
        assert 1 <= rb <= 17 and w in (1, 2, 4) and not (w == 2 and rb < 2) and not (w == 4 and rb < 4)
    for b, (g, s) in t["chol_shape"].items():
        assert (g, s) in autotune.chol_shapes_for(int(b))
    assert 1 <= t["bucket_streams"] <= 4
    # every choice recorded was measured
    for rb, w in t["fast_waves"].items():
        times = t["measured_ms"]["fast_waves"][rb]
        assert str(w) in times


def test_shape_candidates_cover_their_bucket():
    for b in range(autotune.NBUCKETS):
        rows = 16 * (b + 1)
        shapes = autotune.chol_shapes_for(b)
        assert shapes, b
        assert all(g * s >= rows for g, s in shapes)


def test_builtin_defaults_match_the_table():
    """The library's compiled-in choices (used by callers that never load
    the table) are the table's: exact-Cholesky shapes, FAST wavefronts and
    the bucket-stream count."""
    import re
    t = json.load(open(os.path.join(ROOT, "matflat_amd", "tuning", "gfx950.json")))
    src = open(os.path.join(ROOT, "matflat_amd", "csrc", "mfh_exact.hip")).read()
    body = src[src.index("CholShape chol_shape(int b)"):]
    body = body[:body.index("}\n}")]
    cases = {int(a): (int(g), int(s)) for a, g, s in re.findall(r"case (\d+): return \{(\d+), (\d+)\}", body)}
    default = tuple(int(v) for v in re.search(r"default: return \{(\d+), (\d+)\}", body).groups())
    for b, (g, s) in t["chol_shape"].items():
        assert cases.get(int(b), default) == (g, s), b
    fast = open(os.path.join(ROOT, "matflat_amd", "csrc", "mfh_fast.hip")).read()
    lo, hi = (int(v) for v in re.search(r"rb >= (\d+) && rb <= (\d+)\) \? 2 : 1", fast).groups())
    for rb, w in t["fast_waves"].items():
        assert w == (2 if lo <= int(rb) <= hi else 1), rb
    internal = open(os.path.join(ROOT, "matflat_amd", "csrc", "mfh_internal.h")).read()
    assert int(re.search(r"int bucket_streams = (\d+);", internal).group(1)) == t["bucket_streams"]
