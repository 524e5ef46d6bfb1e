"""Turns rocprofv3 counter CSVs into HBM traffic per kernel launch.

Usage (after two counter passes of bench.py --trace-only --copy-probe, one with
--pmc FETCH_SIZE and one with --pmc WRITE_SIZE, each with --output-format csv):

    python -m albedo_amd.counters FETCH_DIR WRITE_DIR OUT_JSON OUT_MD

FETCH_SIZE and WRITE_SIZE are read per dispatch (KiB). Each is scaled by the
factor that makes the same counter read true on a plain device copy of known
size in the same run: the 256 MiB device-to-device copy that --copy-probe
issues (the copy dispatches whose WRITE_SIZE is 256 MiB; dispatch ids match
between the two passes because the program is the same). The closest-hit
dispatches alternate between the coherent and the incoherent ray set.
"""
import csv
import glob
import json
import os
import sys

COPY_BYTES = 256 << 20
KERNEL = "kQuery"
RAYS = 1 << 24
ALGORITHMIC_PER_RAY = 48


def _rows(folder, counter):
    files = glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % folder)
    out = {}
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    out[int(r["Dispatch_Id"])] = (r["Kernel_Name"], float(r["Counter_Value"]))
    return out


def summarize(fetch_dir, write_dir):
    f = _rows(fetch_dir, "FETCH_SIZE")
    w = _rows(write_dir, "WRITE_SIZE")
    copy_ids = [d for d, (name, v) in w.items()
                if "copyBuffer" in name and abs(v * 1024 - COPY_BYTES) < 0.01 * COPY_BYTES and d in f]
    if not copy_ids:
        raise SystemExit("no 256 MiB copy found among the dispatches")
    copy_fetch = sum(f[d][1] for d in copy_ids) / len(copy_ids)
    copy_write = sum(w[d][1] for d in copy_ids) / len(copy_ids)
    fetch_scale = COPY_BYTES / (copy_fetch * 1024)
    write_scale = COPY_BYTES / (copy_write * 1024)
    trace = sorted(d for d, (name, _) in f.items() if KERNEL in name and d in w)
    sets = {"coherent": trace[0::2], "incoherent": trace[1::2]}
    per_set = {}
    for k, ids in sets.items():
        fb = sum(f[d][1] for d in ids) / len(ids) * 1024 * fetch_scale
        wb = sum(w[d][1] for d in ids) / len(ids) * 1024 * write_scale
        per_set[k] = {"fetch_bytes": round(fb), "write_bytes": round(wb), "bytes": round(fb + wb)}
    mean = sum(v["bytes"] for v in per_set.values()) / len(per_set)
    return {
        "kernel": KERNEL,
        "command": "rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE, over python3 bench.py --trace-only "
                   "--copy-probe --steps 2 --warmup 1 --no-cpu",
        "copy_dispatches": copy_ids,
        "copy_fetch_kib": copy_fetch, "copy_write_kib": copy_write,
        "fetch_scale": fetch_scale, "write_scale": write_scale,
        "per_set": per_set,
        "bytes_per_launch": round(mean),
        "algorithmic_bytes_per_launch": ALGORITHMIC_PER_RAY * RAYS,
    }


def main(argv):
    fetch_dir, write_dir, out_json, out_md = argv[1:5]
    s = summarize(fetch_dir, write_dir)
    with open(out_json, "w") as fh:
        json.dump(s, fh, indent=1)
    with open(out_md, "w") as fh:
        fh.write("# HBM traffic of the closest-hit kernel kQuery (trefoil tube, 2^24 rays per launch)\n\n")
        fh.write("Commands (repository root, ROCm 7.2 rocprofv3, one MI355X):\n\n")
        fh.write("    rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -o run -- "
                 "python3 bench.py --trace-only --copy-probe --steps 2 --warmup 1 --no-cpu\n")
        fh.write("    rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write -o run -- "
                 "python3 bench.py --trace-only --copy-probe --steps 2 --warmup 1 --no-cpu\n")
        fh.write("    python -m albedo_amd.counters gpurun_out/pmc_fetch gpurun_out/pmc_write "
                 "profiles/trace_closest_trefoil_traffic.json profiles/trace_closest_trefoil_traffic.md\n\n")
        fh.write("Calibration: the 256 MiB device copy reads FETCH_SIZE %.0f KiB and WRITE_SIZE %.0f KiB per "
                 "dispatch, so FETCH_SIZE is scaled by %.4f and WRITE_SIZE by %.4f.\n\n"
                 % (s["copy_fetch_kib"], s["copy_write_kib"], s["fetch_scale"], s["write_scale"]))
        fh.write("| ray set | fetched bytes | written bytes | total | algorithmic (48 B/ray) |\n|---|---|---|---|---|\n")
        for k, v in s["per_set"].items():
            fh.write("| %s | %d | %d | %d | %d |\n" % (k, v["fetch_bytes"], v["write_bytes"], v["bytes"],
                                                      s["algorithmic_bytes_per_launch"]))
        fh.write("\nMean over the two sets: %d bytes per launch (%.2f x algorithmic).\n"
                 % (s["bytes_per_launch"], s["bytes_per_launch"] / s["algorithmic_bytes_per_launch"]))
    print(json.dumps(s))


if __name__ == "__main__":
    main(sys.argv)
