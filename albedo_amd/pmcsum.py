"""Per-kernel utilisation from rocprofv3 counter passes.

    python -m albedo_amd.pmcsum VALU_DIR L2_DIR OUT_MD "COMMAND_VALU" "COMMAND_L2"

VALU_DIR holds a pass with --pmc SQ_THREAD_CYCLES_VALU SQ_ACTIVE_INST_VALU
SQ_WAVE_CYCLES GRBM_GUI_ACTIVE, L2_DIR one with --pmc TCC_HIT_sum TCC_MISS_sum
(each --output-format csv). Per kernel, summed over its dispatches, with the
formulas of rocprofv3's derived metrics on MI355X (256 CUs, 32 waves per CU):
  VALU utilisation = THREAD_CYCLES_VALU / (ACTIVE_INST_VALU * 64)  (active lanes)
  VALU busy        = ACTIVE_INST_VALU / CUs / GUI_ACTIVE
  occupancy        = 4 * WAVE_CYCLES / GUI_ACTIVE / CUs / 32       (of the maximum)
  L2 hit rate      = TCC_HIT / (TCC_HIT + TCC_MISS)
"""
import collections
import csv
import glob
import os
import re
import sys

CUS = 256


def load(folder):
    files = glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % folder)
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    calls = collections.Counter()
    for f in files:
        with open(f) as fh:
            seen = set()
            for r in csv.DictReader(fh):
                k = short(r["Kernel_Name"])
                per[k][r["Counter_Name"]] += float(r["Counter_Value"])
                if r["Dispatch_Id"] not in seen:
                    seen.add(r["Dispatch_Id"])
                    calls[k] += 1
    return per, calls


def short(name):
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    m = re.match(r"([\w:]+(<[^>]*>)?)", name)
    return m.group(1) if m else name[:40]


def main(argv):
    valu_dir, l2_dir, out = argv[1:4]
    cmd_valu, cmd_l2 = argv[4:6]
    v, calls = load(valu_dir)
    l2, _ = load(l2_dir)
    rows = []
    for k, c in v.items():
        if not k.startswith("ab::"):
            continue
        act = c.get("SQ_ACTIVE_INST_VALU", 0)
        thr = c.get("SQ_THREAD_CYCLES_VALU", 0)
        # The CSV sums GRBM_GUI_ACTIVE over the 8 XCDs; the derived metrics use one XCD's count.
        gui = c.get("GRBM_GUI_ACTIVE", 0) / 8
        wav = c.get("SQ_WAVE_CYCLES", 0)
        h = l2.get(k, {}).get("TCC_HIT_sum", 0)
        m = l2.get(k, {}).get("TCC_MISS_sum", 0)
        rows.append((gui, k, calls[k], 100 * thr / (act * 64) if act else 0, 100 * act / CUS / gui if gui else 0,
                     400 * wav / gui / CUS / 32 if gui else 0, 100 * h / (h + m) if h + m else 0))
    rows.sort(reverse=True)
    with open(out, "w") as fh:
        fh.write("# VALU utilisation, occupancy and L2 hit rate per kernel\n\n")
        fh.write("Commands (repository root, ROCm 7.2 rocprofv3, one MI355X):\n\n    %s\n    %s\n\n" % (cmd_valu, cmd_l2))
        fh.write("Formulas: VALU utilisation = THREAD_CYCLES_VALU / (ACTIVE_INST_VALU x 64) (share of lanes "
                 "active per VALU instruction); VALU busy = ACTIVE_INST_VALU / 256 CUs / GRBM_GUI_ACTIVE; "
                 "occupancy = 4 x WAVE_CYCLES / GRBM_GUI_ACTIVE / 256 / 32 (of 32 waves per CU); "
                 "L2 hit = TCC_HIT / (TCC_HIT + TCC_MISS). Sums over each kernel's dispatches.\n\n")
        fh.write("| kernel | dispatches | GPU cycles (sum) | VALU utilisation % | VALU busy % | occupancy % | L2 hit % |\n")
        fh.write("|---|---|---|---|---|---|---|\n")
        for gui, k, n, util, busy, occ, hit in rows:
            fh.write("| `%s` | %d | %.3g | %.1f | %.1f | %.1f | %.1f |\n" % (k, n, gui, util, busy, occ, hit))
    print(open(out).read())


if __name__ == "__main__":
    main(sys.argv)
