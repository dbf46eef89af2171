"""Timeline summary of a rocprofv3 kernel trace (``rocprofv3 --kernel-trace -o NAME``, SQLite
results database ``NAME_results.db``): the span from the first to the last kernel of the window,
the GPU's busy time (union of kernel intervals), the sum of kernel durations, the mean number of
kernels in flight while busy, and calls / average / total per kernel.

    python3 -m maxwell_solver_amd.prof_timeline PATH_results.db [--after SUBSTRING] [--top N]

--after starts the window at the first kernel whose name holds SUBSTRING (so warm-up launches
before a run's first batched kernel are left out)."""
import argparse
import sqlite3
from collections import defaultdict


def load(db):
    con = sqlite3.connect(db)
    try:
        return con.execute("select start, end, name from kernels order by start").fetchall()
    finally:
        con.close()


def summarize(ks, after=None):
    """ks: (start_ns, end_ns, name) sorted by start.  Returns the summary dict."""
    if after:
        first = next((q for q, k in enumerate(ks) if after in k[2]), None)
        if first is None:
            raise SystemExit(f"no kernel name holds {after!r}")
        ks = ks[first:]
    if not ks:
        raise SystemExit("no kernels in the trace")
    t0 = ks[0][0]
    t1 = max(e for _, e, _ in ks)
    busy, cs, ce = 0, None, None
    for s, e, _ in ks:
        if ce is None or s > ce:
            if ce is not None:
                busy += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    total = sum(e - s for s, e, _ in ks)
    per = defaultdict(lambda: [0, 0])
    for s, e, n in ks:
        per[n][0] += 1
        per[n][1] += e - s
    return {
        "span_ms": (t1 - t0) / 1e6,
        "busy_ms": busy / 1e6,
        "busy_frac": busy / (t1 - t0) if t1 > t0 else 1.0,
        "kernel_sum_ms": total / 1e6,
        "mean_in_flight": total / busy if busy else 0.0,
        "kernels": sorted(((n, c, t / c / 1e3, t / 1e6) for n, (c, t) in per.items()), key=lambda x: -x[3]),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("db")
    ap.add_argument("--after", default=None)
    ap.add_argument("--top", type=int, default=10)
    a = ap.parse_args(argv)
    r = summarize(load(a.db), a.after)
    print(f"span {r['span_ms']:.1f} ms, busy {r['busy_ms']:.1f} ms ({r['busy_frac']:.2%}), "
          f"kernel sum {r['kernel_sum_ms']:.1f} ms, mean in flight {r['mean_in_flight']:.2f}")
    print(f"{'calls':>8} {'avg us':>10} {'total ms':>10}  kernel")
    for n, c, avg, tot in r["kernels"][:a.top]:
        print(f"{c:8d} {avg:10.2f} {tot:10.1f}  {n[:100]}")


if __name__ == "__main__":
    main()
