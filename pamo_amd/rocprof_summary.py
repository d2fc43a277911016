"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) into a text
table: per kernel name, dispatches, total and average duration.  Used to
write the summaries under profiles/; reads only the database it is given."""
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


if __name__ == "__main__":
    print(summary(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 30))
