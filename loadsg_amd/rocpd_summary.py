"""Summaries of rocprofv3 result databases (rocpd SQLite: tables `kernels` and
`counters_collection`).

    python3 -m loadsg_amd.rocpd_summary trace DB [--kernel RE --tail N --group K] [--after RE]
        per-kernel table (calls, total ms, average / max us, share); with --kernel the
        last N dispatches whose name matches RE are also summed and divided into
        groups of K dispatches (K = apply launches of one matvec -> time per matvec);
        with --after only the dispatches after the last one matching RE are counted (and
        the span from the first to the last of them is printed).
    python3 -m loadsg_amd.rocpd_summary sequence DB --kernel RE --tail N
        average duration per position of the dispatch sequences that start at each
        dispatch matching RE (e.g. one MLP training step), over the last N sequences.
    python3 -m loadsg_amd.rocpd_summary pmc DB --kernel RE --tail N --group K --calib RE
        per-counter sum over the last N dispatches matching --kernel, per group of K,
        and the counter of the last dispatch matching --calib (the copy kernel).
"""
import argparse
import re
import sqlite3
from collections import defaultdict


def _short(name):
    name = name.replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", name)


def trace(db, kernel=None, tail=0, group=1, top=20, after=None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    if after:
        marks = [k for k, (name, _, _) in enumerate(rows) if re.search(after, name)]
        if marks:
            rows = rows[marks[-1] + 1:]
        if rows:
            print("window after the last %r dispatch: %d dispatches, %.3f ms from the first start to the last end"
                  % (after, len(rows), (rows[-1][2] - rows[0][1]) / 1e6))
    agg = defaultdict(lambda: [0, 0, 0])
    for name, s, e in rows:
        a = agg[_short(name)]
        a[0] += 1
        a[1] += e - s
        a[2] = max(a[2], e - s)
    tot = sum(a[1] for a in agg.values()) or 1
    print("%-58s %7s %11s %10s %10s %6s" % ("kernel", "calls", "total_ms", "avg_us", "max_us", "pct"))
    for k, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        print("%-58s %7d %11.3f %10.2f %10.2f %6.2f" % (k[:58], a[0], a[1] / 1e6, a[1] / a[0] / 1e3,
                                                        a[2] / 1e3, 100.0 * a[1] / tot))
    if kernel:
        sel = [(s, e) for name, s, e in rows if re.search(kernel, name)]
        sel = sel[-tail:] if tail else sel
        d = [e - s for s, e in sel]
        ng = max(1, len(sel) // group)
        print("window: last %d dispatches matching %r: sum %.3f ms, average %.2f us per dispatch, "
              "%.4f ms per group of %d (%d groups)" % (len(sel), kernel, sum(d) / 1e6, sum(d) / max(1, len(d)) / 1e3,
                                                       sum(d) / 1e6 / ng, group, ng))


def sequence(db, start, tail):
    """Split the dispatch list at every dispatch whose name matches `start` and average the
    duration of each position over the sequences among the last `tail` that have the most
    common length there (one training step = one sequence)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    seqs, cur = [], []
    for k, (name, s, e) in enumerate(rows):
        if re.search(start, name) and cur:
            seqs.append(cur)
            cur = []
        gap = rows[k + 1][1] - e if k + 1 < len(rows) else 0
        cur.append((_short(name), e - s, gap))
    seqs.append(cur)
    lengths = defaultdict(int)
    for q in (seqs[-tail:] if tail else seqs):     # the most common length among the last `tail`
        lengths[len(q)] += 1
    L = max(lengths, key=lambda k: lengths[k])
    sel = [q for q in seqs if len(q) == L][-tail:] if tail else [q for q in seqs if len(q) == L]
    tot = gaps = 0.0
    print("%d sequences of %d dispatches (starting at %r); average per position "
          "(duration, then the idle gap to the next dispatch):" % (len(sel), L, start))
    for i in range(L):
        v = sum(q[i][1] for q in sel) / len(sel) / 1e3
        g = sum(q[i][2] for q in sel) / len(sel) / 1e3
        tot += v
        if i + 1 < L:
            gaps += g
        print("%3d %-58s %10.2f us %10.2f us" % (i, sel[0][i][0][:58], v, g))
    print("sum %.2f us of kernels per sequence, %.2f us of gaps between them (first to last dispatch: "
          "%.2f us)" % (tot, gaps, tot + gaps))


def pmc(db, kernel, tail, group, calib):
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection "
                       "order by dispatch_id").fetchall()
    per = defaultdict(lambda: defaultdict(float))
    names = {}
    for did, kn, cn, v in rows:
        per[did][cn] += v
        names[did] = kn
    order = sorted(per)
    sel = [d for d in order if re.search(kernel, names[d])]
    sel = sel[-tail:] if tail else sel
    ng = max(1, len(sel) // group)
    tot = defaultdict(float)
    for d in sel:
        for c, v in per[d].items():
            tot[c] += v
    for c, v in sorted(tot.items()):
        print("%s: %d dispatches matching %r, sum %.9g, per group of %d: %.9g" % (c, len(sel), kernel, v, group, v / ng))
    cal = [d for d in order if re.search(calib, names[d])]
    if cal:
        for c, v in sorted(per[cal[-1]].items()):
            print("calibration %s (%s): %.9g" % (c, _short(names[cal[-1]])[:60], v))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("trace", "pmc", "sequence"))
    ap.add_argument("db")
    ap.add_argument("--kernel")
    ap.add_argument("--tail", type=int, default=0)
    ap.add_argument("--group", type=int, default=1)
    ap.add_argument("--calib", default="elementwise")
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--after")
    a = ap.parse_args()
    if a.mode == "trace":
        trace(a.db, a.kernel, a.tail, a.group, a.top, a.after)
    elif a.mode == "sequence":
        sequence(a.db, a.kernel, a.tail)
    else:
        pmc(a.db, a.kernel, a.tail, a.group, a.calib)


if __name__ == "__main__":
    main()
