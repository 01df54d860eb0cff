"""The GEMM autotuner's host side: the split-K rounding it assumes matches
launch_gemm's (chunks of ceil(K/s) rounded up to the K step of 16), and the
table it writes is the C initializer list the library compiles."""
import re

from mpsamd.autotune import _effective_splits, write_table


def test_effective_splits_follow_the_chunk_rounding():
    for K in (1, 9, 16, 17, 81, 128, 729, 1134, 1152, 1792):
        for s in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32):
            chunk = max(16, -(-(-(-K // s)) // 16) * 16)
            e = _effective_splits(K, s)
            assert e == max(1, -(-K // chunk))
            assert (e - 1) * chunk < K <= e * chunk or K <= 16


def test_table_lines_are_initializers(tmp_path):
    entries = [dict(M=9, N=126, K=126, batch=81, nb_class=0, cmp=0, tm=32, splits=2, t_rule=27.6, t_best=16.7,
                    rule=(64, 4), label="w"),
               dict(M=9, N=126, K=126, batch=81, nb_class=0, cmp=0, tm=32, splits=3, t_rule=27.6, t_best=20.0,
                    rule=(64, 4), label="w2")]
    p = tmp_path / "t.inc"
    assert write_table(entries, p) == 1          # one entry per key, the larger gain kept
    rows = [ln for ln in p.read_text().splitlines() if ln.strip().startswith("{")]
    assert len(rows) == 1
    m = re.match(r"\s*\{(\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\},\s*//", rows[0])
    assert m and [int(x) for x in m.groups()] == [9, 126, 126, 81, 0, 0, 32, 2]


def test_committed_table_parses():
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    inc = os.path.join(here, "..", "matrixproductstates.jl_amd", "csrc", "gemm_tuned.inc")
    rows = [ln for ln in open(inc).read().splitlines() if ln.strip().startswith("{")]
    assert rows
    for ln in rows:
        m = re.match(r"\s*\{(\d+), (\d+), (\d+), (\d+), (\d+), ([01]), (32|64), (\d+)\},", ln)
        assert m, ln
        assert 1 <= int(m.group(8)) <= 32
