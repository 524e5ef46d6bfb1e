"""Memory-pipeline and issue counters of the closest-hit kernel per ray set, from
rocprofv3 counter passes (CSV output) of `bench.py --trace-only`.

Usage:

    python -m albedo_amd.setpmc OUT_MD TITLE COMMANDS_FILE PASS_DIR...

The kQuery dispatches alternate between the coherent and the incoherent ray
set; each counter is averaged over each set's dispatches, whatever pass it came
from. GRBM_GUI_ACTIVE sums the 8 XCDs, so a dispatch's cycles are
GRBM_GUI_ACTIVE / 8. Formulas (MI355X: 256 CUs, one TA and one TD per CU):

  VALU busy        = SQ_ACTIVE_INST_VALU / 256 / cycles;
  VALU utilisation = SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64);
  TA busy          = TA_BUSY_avr / cycles;
  TD busy          = TD_TD_BUSY_sum / 256 / cycles;
  lines per load   = TCP_TOTAL_CACHE_ACCESSES_sum / SQ_INSTS_VMEM_RD;
  tag-conflict stall = TCP_READ_TAGCONFLICT_STALL_CYCLES_sum / 256 / cycles.

COMMANDS_FILE holds the commands that made the passes (one a line), quoted in
the summary.
"""
import csv
import glob
import os
import sys
from collections import defaultdict

CUS = 256
XCDS = 8
KERNEL = "kQuery"


def _per_dispatch(folder):
    files = glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % folder)
    rows = defaultdict(dict)
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if KERNEL in r["Kernel_Name"]:
                    rows[int(r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
    return rows


def summarize(folders):
    sets = {"coherent": defaultdict(list), "incoherent": defaultdict(list)}
    for folder in folders:
        rows = _per_dispatch(folder)
        ids = sorted(rows)
        for name, sel in (("coherent", ids[0::2]), ("incoherent", ids[1::2])):
            for i in sel:
                for k, v in rows[i].items():
                    sets[name][k].append(v)
    out = {}
    for name, vals in sets.items():
        m = {k: sum(v) / len(v) for k, v in vals.items()}
        cyc = m["GRBM_GUI_ACTIVE"] / XCDS
        out[name] = {
            "ms": cyc / 2.4e6,
            "valu_busy": m["SQ_ACTIVE_INST_VALU"] / CUS / cyc,
            "valu_util": m["SQ_THREAD_CYCLES_VALU"] / (m["SQ_ACTIVE_INST_VALU"] * 64),
            "ta_busy": m["TA_BUSY_avr"] / cyc,
            "td_busy": m["TD_TD_BUSY_sum"] / CUS / cyc,
            "lines_per_load": m["TCP_TOTAL_CACHE_ACCESSES_sum"] / m["SQ_INSTS_VMEM_RD"],
            "tag_stall": m["TCP_READ_TAGCONFLICT_STALL_CYCLES_sum"] / CUS / cyc,
            "valu_insts": m["SQ_INSTS_VALU"],
            "vmem_insts": m["SQ_INSTS_VMEM_RD"],
        }
    return out


def main(argv):
    out_md, title, commands = argv[1:4]
    folders = argv[4:]
    s = summarize(folders)
    with open(commands) as fh:
        cmds = [l.strip() for l in fh if l.strip()]
    with open(out_md, "w") as fh:
        fh.write("# %s\n\nCommands (repository root, ROCm 7.2 rocprofv3, one MI355X):\n\n" % title)
        for c in cmds:
            fh.write("    %s\n" % c)
        fh.write("\nEach counter is averaged over the ray set's kQuery dispatches. Cycles = "
                 "GRBM_GUI_ACTIVE / 8; VALU busy = SQ_ACTIVE_INST_VALU / 256 / cycles; VALU utilisation = "
                 "SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64); TA busy = TA_BUSY_avr / cycles; TD busy = "
                 "TD_TD_BUSY_sum / 256 / cycles; cache lines per load = TCP_TOTAL_CACHE_ACCESSES_sum / "
                 "SQ_INSTS_VMEM_RD; tag-conflict stall = TCP_READ_TAGCONFLICT_STALL_CYCLES_sum / 256 / cycles.\n\n")
        fh.write("| ray set | cycles (ms at 2.4 GHz) | VALU busy % | VALU utilisation % | TA busy % | TD busy % "
                 "| cache lines per load | tag-conflict stall % | VALU insts | VMEM read insts |\n"
                 "|---|---|---|---|---|---|---|---|---|---|\n")
        for name in ("coherent", "incoherent"):
            r = s[name]
            fh.write("| %s | %.2f | %.1f | %.1f | %.1f | %.1f | %.1f | %.1f | %.3g | %.3g |\n"
                     % (name, r["ms"], 100 * r["valu_busy"], 100 * r["valu_util"], 100 * r["ta_busy"],
                        100 * r["td_busy"], r["lines_per_load"], 100 * r["tag_stall"], r["valu_insts"],
                        r["vmem_insts"]))
    print("wrote", out_md)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
