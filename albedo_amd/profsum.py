"""Writes a short Markdown summary of a rocprofv3 --stats kernel table.

    python -m albedo_amd.profsum STATS_CSV OUT_MD "TITLE" "COMMAND" [TOP]

STATS_CSV is the *_kernel_stats.csv that `rocprofv3 --kernel-trace --stats
--output-format csv` writes. The summary keeps the TOP kernels by total time
(default 20) with calls, total, average, min and max durations.
"""
import csv
import sys


def short(name):
    name = name.replace("(anonymous namespace)::", "")
    cut = name.find("(")
    return name[:cut] if cut > 0 else name


def main(argv):
    stats, out, title, command = argv[1:5]
    top = int(argv[5]) if len(argv) > 5 else 20
    with open(stats) as fh:
        rows = list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    with open(out, "w") as fh:
        fh.write("# %s\n\nCommand (repository root, ROCm 7.2 rocprofv3, one MI355X):\n\n    %s\n\n" % (title, command))
        fh.write("GPU kernel time in the run: %.1f ms.\n\n" % (total / 1e6))
        fh.write("| kernel | calls | total ms | % | avg us | min us | max us |\n|---|---|---|---|---|---|---|\n")
        for r in rows[:top]:
            fh.write("| `%s` | %s | %.2f | %.1f | %.1f | %.1f | %.1f |\n" % (
                short(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e6,
                100 * float(r["TotalDurationNs"]) / total, float(r["AverageNs"]) / 1e3,
                float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3))


if __name__ == "__main__":
    main(sys.argv)
