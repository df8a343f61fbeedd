"""Text summaries of rocprofv3 output, and the copy kernel the HBM counters
are calibrated against.

    python -m webgpuraytracer_amd.rocprof_summary kernels RUN_results.db
    python -m webgpuraytracer_amd.rocprof_summary pmc RUN_results.db COUNTER
    python -m webgpuraytracer_amd.rocprof_summary copy [--mib 1024]
    python -m webgpuraytracer_amd.rocprof_summary overlap RUN_results.db

`kernels` prints the top-kernels table of a `--kernel-trace --stats` run
(durations in microseconds).  `pmc` prints a `--pmc` counter per dispatch,
grouped by kernel (FETCH_SIZE / WRITE_SIZE are in KiB per dispatch).  `copy`
launches three device-to-device copies of a known size: the FETCH_SIZE and
WRITE_SIZE those dispatches read give the factors that turn the counters
into HBM bytes.  `overlap` measures how much kernels ran side by side: the
sum of kernel durations against the time at least one kernel was running,
and the hardware queues the dispatches used.
"""
import argparse
import sqlite3
import sys


def kernels(db):
    c = sqlite3.connect(db)
    print("| kernel | calls | total us | average us | % |")
    print("|---|---|---|---|---|")
    for name, calls, total, avg, pct in c.execute(
            "select name,total_calls,total_duration,average,percentage from top_kernels"):
        print("| `%s` | %d | %.1f | %.1f | %.2f |" % (name[:100], calls, total, avg, pct))


def pmc(db, counter):
    c = sqlite3.connect(db)
    rows = c.execute(
        "select kernel_name, count(distinct dispatch_id), sum(value), avg(end - start) from counters_collection "
        "where counter_name = ? group by kernel_name order by sum(value) desc", (counter,))
    print("| kernel | dispatches | %s per dispatch | avg duration ms |" % counter)
    print("|---|---|---|---|")
    for name, n, total, dur in rows:
        print("| `%s` | %d | %.4f | %.3f |" % (name[:100], n, total / n, (dur or 0) / 1e6))


def overlap(db):
    c = sqlite3.connect(db)
    iv = sorted(c.execute("select start, end, queue_id from kernels"))
    if not iv:
        print("no kernels")
        return
    total = sum(e - s for s, e, _ in iv)
    union, cur_s, cur_e = 0, iv[0][0], iv[0][1]
    for s, e, _ in iv[1:]:
        if s > cur_e:
            union += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    union += cur_e - cur_s
    # largest number of kernels running at once
    ev = sorted([(s, 1) for s, _, _ in iv] + [(e, -1) for _, e, _ in iv])
    run = peak = 0
    for _, d in ev:
        run += d
        peak = max(peak, run)
    queues = len({q for _, _, q in iv})
    span = iv[-1][1] - iv[0][0]
    print("| dispatches | kernel time ms | busy (union) ms | span ms | kernel time / busy | peak concurrent | queues |")
    print("|---|---|---|---|---|---|---|")
    print("| %d | %.3f | %.3f | %.3f | %.2f | %d | %d |" % (len(iv), total / 1e6, union / 1e6, span / 1e6,
                                                         total / max(union, 1), peak, queues))


def copy(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 3 x %d bytes (%d KiB per copy)" % (n * 4, n * 4 // 1024), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    p = sub.add_parser("pmc")
    p.add_argument("db")
    p.add_argument("counter")
    o = sub.add_parser("overlap")
    o.add_argument("db")
    cp = sub.add_parser("copy")
    cp.add_argument("--mib", type=int, default=1024)
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        kernels(a.db)
    elif a.cmd == "pmc":
        pmc(a.db, a.counter)
    elif a.cmd == "overlap":
        overlap(a.db)
    else:
        copy(a.mib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
