"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) into text
tables: per kernel name, dispatches, total and average duration; with
--rounds, the engine's rounds grouped by face count.  Used to write the
summaries under profiles/; reads only the database it is given.

    python3 rocprof_summary.py DB [TOP] [--rounds]
"""
import sqlite3
import sys


def summary(db, top=30):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(end-start), avg(end-start) from kernels group by name "
                       "order by sum(end-start) desc").fetchall()
    tot = sum(r[2] for r in rows)
    out = ["total kernel time %.2f ms over %d dispatches" % (tot / 1e6, sum(r[1] for r in rows)), "",
           "| kernel | dispatches | total ms | avg us | share |", "|---|---|---|---|---|"]
    for name, n, s, a in rows[:top]:
        short = name.split("(")[0]
        out.append("| %s | %d | %.2f | %.2f | %.1f%% |" % (short, n, s / 1e6, a / 1e3, 100.0 * s / tot))
    return "\n".join(out)


BUCKETS = [(1 << 30, 1 << 20), (1 << 20, 1 << 18), (1 << 18, 65536), (65536, 16384), (16384, 0)]


def rounds(db, top=3):
    """Engine rounds (a round starts at k_round_init) grouped by the round's
    face count, taken from the fan-count launch's grid (one lane per face,
    rounded up to the 256-lane workgroup).  Span = from the round's first
    launch to the next round's, so it includes the gaps between launches."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end, grid_x from kernels order by start").fetchall()
    rs, cur = [], None
    for name, a, b, grid in rows:
        if "k_round_init" in name:
            cur = dict(start=a, end=b, n=0, ksum=0, F=None, names={})
            rs.append(cur)
        if cur is None:
            continue
        cur["n"] += 1
        cur["ksum"] += b - a
        cur["end"] = max(cur["end"], b)
        short = name.split("(")[0].split("::")[-1]
        cur["names"][short] = cur["names"].get(short, 0) + (b - a)
        if "k_fan_count_snap" in name and cur["F"] is None:
            cur["F"] = grid
    for i, r in enumerate(rs):
        r["span"] = (rs[i + 1]["start"] if i + 1 < len(rs) else r["end"]) - r["start"]
    out = ["| faces in the round | rounds | span ms | kernel ms | dispatches per round | largest (ms) |",
           "|---|---|---|---|---|---|"]
    for hi, lo in BUCKETS:
        sel = [r for r in rs if r["F"] is not None and lo <= r["F"] < hi]
        if not sel:
            continue
        agg = {}
        for r in sel:
            for k, v in r["names"].items():
                agg[k] = agg.get(k, 0) + v
        big = sorted(agg.items(), key=lambda x: -x[1])[:top]
        out.append("| [%d, %d) | %d | %.2f | %.2f | %.1f | %s |" % (
            lo, hi, len(sel), sum(r["span"] for r in sel) / 1e6, sum(r["ksum"] for r in sel) / 1e6,
            sum(r["n"] for r in sel) / len(sel), ", ".join("%s %.2f" % (k, v / 1e6) for k, v in big)))
    return "\n".join(out)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    print(summary(args[0], int(args[1]) if len(args) > 1 else 30))
    if "--rounds" in sys.argv:
        print()
        print(rounds(args[0]))
