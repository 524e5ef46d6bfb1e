"""Per-kernel VALU utilisation, VALU busy, occupancy and L2 hit rate from two
rocprofv3 counter passes (CSV output) of the same program.

Usage:

    python -m albedo_amd.pmcsum VALU_DIR L2_DIR OUT_MD TITLE COMMAND1 COMMAND2

VALU_DIR holds a pass with --pmc SQ_THREAD_CYCLES_VALU SQ_ACTIVE_INST_VALU
SQ_WAVE_CYCLES GRBM_GUI_ACTIVE, L2_DIR one with --pmc TCC_HIT_sum TCC_MISS_sum.
Counters are summed over each kernel's dispatches. GRBM_GUI_ACTIVE reads the
sum over the 8 XCDs (a 1.62 ms dispatch reads 29.6 M, 7.6 x its 3.9 M cycles at
2.4 GHz), so the kernel's GPU cycles are GRBM_GUI_ACTIVE / 8. Formulas (MI355X:
256 CUs, 4 SIMDs of 8 wave slots each per CU):

  VALU utilisation = THREAD_CYCLES_VALU / (ACTIVE_INST_VALU x 64), the share of
                     lanes active per VALU instruction;
  VALU busy        = ACTIVE_INST_VALU / 256 / cycles;
  occupancy        = 4 x WAVE_CYCLES / cycles / 256 / 32;
  L2 hit           = TCC_HIT / (TCC_HIT + TCC_MISS).
"""
import csv
import glob
import os
import re
import sys
from collections import defaultdict

CUS = 256
XCDS = 8


def _sums(folder):
    files = glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % folder)
    sums = defaultdict(lambda: defaultdict(float))
    dispatches = defaultdict(set)
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                k = r["Kernel_Name"]
                sums[k][r["Counter_Name"]] += float(r["Counter_Value"])
                dispatches[k].add(int(r["Dispatch_Id"]))
    return sums, dispatches


def short(name):
    n = re.sub(r"\(anonymous namespace\)::", "", name)
    n = re.sub(r"^void ", "", n)
    return n.split("(")[0]


def main(argv):
    valu_dir, l2_dir, out_md, title, cmd1, cmd2 = argv[1:7]
    v, disp = _sums(valu_dir)
    l2, _ = _sums(l2_dir)
    rows = []
    for k, c in v.items():
        gui = c.get("GRBM_GUI_ACTIVE", 0.0) / XCDS
        inst = c.get("SQ_ACTIVE_INST_VALU", 0.0)
        if gui <= 0 or k.startswith("__amd"):
            continue
        util = 100.0 * c.get("SQ_THREAD_CYCLES_VALU", 0.0) / (inst * 64) if inst > 0 else 0.0
        busy = 100.0 * inst / CUS / gui
        occ = 100.0 * 4 * c.get("SQ_WAVE_CYCLES", 0.0) / gui / CUS / 32
        h = l2.get(k, {})
        hits, miss = h.get("TCC_HIT_sum", 0.0), h.get("TCC_MISS_sum", 0.0)
        hit = 100.0 * hits / (hits + miss) if hits + miss > 0 else 0.0
        rows.append((gui, short(k), len(disp[k]), util, busy, occ, hit))
    rows.sort(reverse=True)
    with open(out_md, "w") as fh:
        fh.write("# %s\n\nCommands (repository root, ROCm 7.2 rocprofv3, one MI355X):\n\n" % title)
        fh.write("    %s\n    %s\n\n" % (cmd1, cmd2))
        fh.write("Formulas: cycles = GRBM_GUI_ACTIVE / 8 (the counter sums the 8 XCDs); VALU utilisation = "
                 "THREAD_CYCLES_VALU / (ACTIVE_INST_VALU x 64) (share of lanes active per VALU instruction); "
                 "VALU busy = ACTIVE_INST_VALU / 256 CUs / cycles; occupancy = 4 x WAVE_CYCLES / cycles / 256 "
                 "/ 32 (of 32 waves per CU); L2 hit = TCC_HIT / (TCC_HIT + TCC_MISS). Sums over each kernel's "
                 "dispatches.\n\n")
        fh.write("| kernel | dispatches | GPU cycles (sum) | VALU utilisation % | VALU busy % | occupancy % "
                 "| L2 hit % |\n|---|---|---|---|---|---|---|\n")
        for gui, name, n, util, busy, occ, hit in rows:
            fh.write("| `%s` | %d | %.2e | %.1f | %.1f | %.1f | %.1f |\n" % (name, n, gui, util, busy, occ, hit))
    print("wrote", out_md)


if __name__ == "__main__":
    main(sys.argv)
