"""Summarise rocprofv3 SQLite outputs (kernel-trace stats and per-dispatch PMC counters).

    python -m t2x_py.rocprof_summary trace <run_results.db>
    python -m t2x_py.rocprof_summary pmc <run_results.db> <COUNTER> [kernel-substring]
"""
import sqlite3
import sys


def trace(db, top=25):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels").fetchall()
    # the top_kernels view reports microseconds
    out = ["| kernel | calls | total ms | avg ms | % |", "|---|---|---|---|---|"]
    for name, calls, tot, avg, pct in rows[:top]:
        short = name.split("(")[0].replace("void ", "")
        out.append(f"| `{short}` | {calls} | {tot / 1e3:.3f} | {avg / 1e3:.3f} | {pct:.2f} |")
    return "\n".join(out)


def pmc(db, counter, sub=None):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, value, duration from counters_collection "
                       "where counter_name = ? order by id", (counter,)).fetchall()
    out = [f"| kernel | {counter} (KiB) | duration ms |", "|---|---|---|"]
    for name, v, d in rows:
        if sub and not any(s in name for s in sub.split("|")):
            continue
        if v < 1024 and "gemm" not in name:
            continue
        # counters_collection durations are nanoseconds
        out.append(f"| `{name.split('(')[0].replace('void ', '')}` | {v:.1f} | {d / 1e6:.3f} |")
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2]))
    else:
        print(pmc(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else None))
