"""Prints a markdown summary of a rocprofv3 rocpd database (top kernels by
total duration; durations in microseconds as stored by rocprofv3)."""
import sqlite3
import sys


def main(db):
    c = sqlite3.connect(db)
    print("| kernel | calls | total us | average us | % |")
    print("|---|---|---|---|---|")
    for name, calls, total, avg, pct in c.execute(
            "select name,total_calls,total_duration,average,percentage from top_kernels"):
        print("| `%s` | %d | %.1f | %.1f | %.2f |" % (name[:100], calls, total, avg, pct))


if __name__ == "__main__":
    main(sys.argv[1])
