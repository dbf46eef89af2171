"""The trace timeline summary on a hand-made kernel table: union of overlapping intervals, the
window start of --after, per-kernel totals."""
import sqlite3

from maxwell_solver_amd.prof_timeline import load, summarize


def _db(tmp_path, rows):
    p = tmp_path / "t_results.db"
    con = sqlite3.connect(p)
    con.execute("create table kernels (start integer, end integer, name text)")
    con.executemany("insert into kernels values (?, ?, ?)", rows)
    con.commit()
    con.close()
    return str(p)


def test_union_and_concurrency(tmp_path):
    # warm-up, then [1000,3000) and [2000,4000) overlap, a gap, [5000,6000)
    rows = [(0, 100, "warm"), (1000, 3000, "k_rows_batch<a>"), (2000, 4000, "k_rows_batch<b>"),
            (5000, 6000, "k_sheets_batch")]
    ks = load(_db(tmp_path, rows))
    r = summarize(ks, after="batch")
    assert r["span_ms"] * 1e6 == 5000
    assert round(r["busy_ms"] * 1e6) == 4000
    assert round(r["kernel_sum_ms"] * 1e6) == 5000
    assert abs(r["mean_in_flight"] - 1.25) < 1e-12
    names = [k[0] for k in r["kernels"]]
    assert "warm" not in names and len(names) == 3
    r = summarize(ks)
    assert round(r["busy_ms"] * 1e6) == 4100 and r["span_ms"] * 1e6 == 6000
