"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) into text
tables: per kernel name, dispatches, total and average duration; with
--rounds, the engine's rounds grouped by face count.  Used to write the
summaries under profiles/; reads only the database it is given.

    python3 rocprof_summary.py DB [TOP] [--rounds] [--match=SUBSTRING]
    python3 rocprof_summary.py --pmc AGENT_CSV COUNTER_CSV [COUNTER_CSV ...]
"""
import sqlite3
import sys


def summary(db, top=30, match=None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(end-start), avg(end-start) from kernels group by name "
                       "order by sum(end-start) desc").fetchall()
    if match:
        rows = [r for r in rows if match in r[0]]
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


def _agent(agent_csv):
    import csv
    for row in csv.DictReader(open(agent_csv)):
        if row["Agent_Type"] == "GPU":
            return dict(cu=int(row["Cu_Count"]), xcc=int(row["Num_Xcc"]), waves_per_cu=int(row["Max_Waves_Per_Cu"]))
    raise ValueError("no GPU agent in %s" % agent_csv)


def pmc(counter_csvs, agent_csv):
    """Per kernel, from rocprofv3 --pmc counter CSVs (one or more passes of
    the same run): dispatches, summed counters and the derived metrics
    with AMD's expressions.  GRBM_GUI_ACTIVE is summed over the XCCs in the
    CSV, so the kernel's own cycles are that sum / XCC count.
      occupancy %       = 4 SQ_WAVE_CYCLES / (cycles x CUs x max waves per CU)
                          (SQ_WAVE_CYCLES is in quad-cycles)
      VALU busy %       = SQ_ACTIVE_INST_VALU / CUs / cycles
      VALU utilization % = SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64)
      LDS bank conflict = SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT)
                          (conflict cycles per conflict-free LDS cycle)"""
    import csv
    ag = _agent(agent_csv)
    k = {}
    for path in counter_csvs:
        for row in csv.DictReader(open(path)):
            name = row["Kernel_Name"].split("(")[0]
            d = k.setdefault(name, dict(dispatch=set(), c={}, vgpr=row["VGPR_Count"], agpr=row["Accum_VGPR_Count"],
                                        lds=row["LDS_Block_Size"], wg=row["Workgroup_Size"]))
            d["dispatch"].add(row["Dispatch_Id"])
            cn = row["Counter_Name"]
            key = (cn, path)   # the same counter in two passes is kept apart
            d["c"][key] = d["c"].get(key, 0.0) + float(row["Counter_Value"])
    out = ["| kernel | dispatches | VGPR | LDS B | occupancy % | VALU busy % | VALU util % | VALU inst / wave | "
           "LDS bank conflict | wait / wave cycles |", "|---|---|---|---|---|---|---|---|---|---|"]
    for name, d in sorted(k.items()):
        def c(cn):
            vals = [v for (n, _), v in d["c"].items() if n == cn]
            return vals[0] if vals else None
        def ratio(a, b, scale=1.0):
            return None if a is None or not b else scale * a / b
        cycles = {p: v / ag["xcc"] for (n, p), v in d["c"].items() if n == "GRBM_GUI_ACTIVE"}
        cyc_for = lambda cn: next((cycles[p] for (n, p) in d["c"] if n == cn and p in cycles), None)
        occ = ratio(c("SQ_WAVE_CYCLES"), (cyc_for("SQ_WAVE_CYCLES") or 0) * ag["cu"] * ag["waves_per_cu"], 400.0)
        busy = ratio(c("SQ_ACTIVE_INST_VALU"), (cyc_for("SQ_ACTIVE_INST_VALU") or 0) * ag["cu"], 100.0)
        util = ratio(c("SQ_THREAD_CYCLES_VALU"), (c("SQ_ACTIVE_INST_VALU") or 0) * 64, 100.0)
        ipw = ratio(c("SQ_INSTS_VALU"), c("SQ_WAVES"))
        bc, idx = c("SQ_LDS_BANK_CONFLICT"), c("SQ_LDS_IDX_ACTIVE")
        lbc = None if bc is None or idx is None or idx - bc <= 0 else bc / (idx - bc)
        wait = ratio(c("SQ_WAIT_INST_ANY"), c("SQ_WAVE_CYCLES"))
        f = lambda x, fmt="%.1f": "-" if x is None else fmt % x
        out.append("| %s | %d | %s | %s | %s | %s | %s | %s | %s | %s |" % (
            name.split("::")[-1], len(d["dispatch"]), d["vgpr"], d["lds"], f(occ), f(busy), f(util), f(ipw, "%.0f"),
            f(lbc, "%.3f"), f(wait, "%.2f")))
    return "\n".join(out)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    if "--pmc" in sys.argv:   # --pmc AGENT_CSV COUNTER_CSV [COUNTER_CSV ...]
        print(pmc(args[1:], args[0]))
        sys.exit(0)
    match = next((a.split("=", 1)[1] for a in sys.argv if a.startswith("--match=")), None)
    print(summary(args[0], int(args[1]) if len(args) > 1 else 30, match))
    if "--rounds" in sys.argv:
        print()
        print(rounds(args[0]))
