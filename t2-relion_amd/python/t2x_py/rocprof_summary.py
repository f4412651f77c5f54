"""Summarise rocprofv3 SQLite outputs (kernel-trace stats and per-dispatch PMC counters).

    python -m t2x_py.rocprof_summary trace <run_results.db>
    python -m t2x_py.rocprof_summary pmc <run_results.db> <COUNTER> [kernel-substring|...]
    python -m t2x_py.rocprof_summary traffic <fetch.db> <write.db> <kernel-substring|...> <n_last>
    python -m t2x_py.rocprof_summary last <run_results.db> <kernel-substring|...> <n_last>

`traffic` calibrates FETCH_SIZE and WRITE_SIZE (KiB per dispatch) against the 1 GiB
device-to-device copy kernel k_copy16 in the same run (1 GiB read + 1 GiB written), sums the
matching kernels dispatched per batch and averages the last n_last batches (the timed step).
"""
import json
import sqlite3
import sys

GIB = 1 << 30


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


def last(db, sub, n_last):
    """Mean duration (ms) of the last n_last dispatches of each kernel matching one of the
    '|'-separated substrings, in dispatch order (kernel-trace database)."""
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    name = "name" if "name" in cols else "kernel_name"
    dur = "duration" if "duration" in cols else "(end - start)"
    order = "start" if "start" in cols else "rowid"
    rows = con.execute(f"select {name}, {dur} from kernels order by {order}").fetchall()
    out = {}
    for s in sub.split("|"):
        d = [r[1] for r in rows if s in r[0]]
        tail = d[-n_last:]
        out[s] = {"dispatches": len(d), "mean_ms_last": sum(tail) / len(tail) / 1e6 if tail else None}
    return out


def dispatches(db, counter):
    con = sqlite3.connect(db)
    return con.execute("select kernel_name, value, duration from counters_collection "
                       "where counter_name = ? order by id", (counter,)).fetchall()


def pmc(db, counter, sub=None):
    out = [f"| kernel | {counter} (KiB) | duration ms |", "|---|---|---|"]
    for name, v, d in dispatches(db, counter):
        if sub and not any(s in name for s in sub.split("|")):
            continue
        # counters_collection durations are nanoseconds
        out.append(f"| `{name.split('(')[0].replace('void ', '')}` | {v:.1f} | {d / 1e6:.3f} |")
    return "\n".join(out)


def traffic(fetch_db, write_db, sub, n_last):
    subs = sub.split("|")
    res = {}
    for counter, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        rows = dispatches(db, counter)
        copy = [v for name, v, d in rows if "k_copy16" in name]
        factor = GIB / (copy[0] * 1024.0)
        # group consecutive matching dispatches into batches: a batch starts at the first kernel
        # of the list (e.g. the dense GEMM), later kernels of the list join it
        batches = []
        for name, v, d in rows:
            hit = [i for i, s in enumerate(subs) if s in name]
            if not hit or "k_copy16" in name:
                continue
            if hit[0] == 0 or not batches:
                batches.append({"bytes": 0.0, "ms": 0.0, "kernels": []})
            batches[-1]["bytes"] += v * 1024.0 * factor
            batches[-1]["ms"] += d / 1e6
            batches[-1]["kernels"].append(name.split("(")[0])
        last = batches[-n_last:]
        res[counter] = {"calibration_factor": factor,
                        "bytes_per_batch": sum(b["bytes"] for b in last) / len(last),
                        "ms_per_batch": sum(b["ms"] for b in last) / len(last),
                        "kernels": sorted(set(k for b in last for k in b["kernels"]))}
    res["traffic_bytes_per_batch"] = res["FETCH_SIZE"]["bytes_per_batch"] + res["WRITE_SIZE"]["bytes_per_batch"]
    return res


if __name__ == "__main__":
    if sys.argv[1] == "last":
        print(json.dumps(last(sys.argv[2], sys.argv[3], int(sys.argv[4])), indent=1))
    elif sys.argv[1] == "trace":
        print(trace(sys.argv[2]))
    elif sys.argv[1] == "pmc":
        print(pmc(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else None))
    else:
        print(json.dumps(traffic(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5])), indent=1))
