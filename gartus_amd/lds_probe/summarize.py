"""Markdown summary of an LDS probe run.

usage: python3 summarize.py <run_counter_collection.csv> <trdump_out.txt>

Per dispatch, SQ_LDS_BANK_CONFLICT counts the cycles lost to bank conflicts
and SQ_LDS_IDX_ACTIVE the cycles the LDS was busy with the wave's accesses;
the ratio conflict / (active - conflict) is the extra time per useful cycle
(0: conflict free, 1: every access takes twice as long).
"""
import collections
import csv
import json
import sys

KIND = {0: "ds_read_b128", 1: "ds_write_b64", 2: "ds_write_b16", 3: "ds_read_b64", 4: "ds_read_b64_tr_b16",
        5: "ds_write_b128"}


def ranges(lanes):
    out, start, prev = [], None, None
    for l in sorted(lanes):
        if start is None:
            start = prev = l
        elif l == prev + 1:
            prev = l
        else:
            out.append((start, prev))
            start = prev = l
    if start is not None:
        out.append((start, prev))
    return ", ".join(f"{a}" if a == b else f"{a}-{b}" for a, b in out)


def main(csv_path, tr_path):
    names = json.load(open("patterns.json"))
    per = collections.defaultdict(dict)
    for r in csv.DictReader(open(csv_path)):
        if "k_pat" in r["Kernel_Name"]:
            per[int(r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
    ids = sorted(per)
    assert len(ids) == len(names), (len(ids), len(names))
    res = {n[0]: (per[i]["SQ_LDS_BANK_CONFLICT"], per[i]["SQ_LDS_IDX_ACTIVE"], per[i]["SQ_INSTS_LDS"])
           for i, n in zip(ids, names)}
    print("## Lanes served together, by access width\n")
    print("Only lanes a and b active, at addresses in the same banks: b conflicts with a when the two are served "
          "in the same pass.\n")
    print("| access | a | lanes that conflict with a |")
    print("|---|---:|---|")
    for kind in range(6):
        for a in (0, 4, 16, 32):
            single = res[f"{KIND[kind]} single a={a}"][0]
            conf = [b for b in range(64) if b != a and res[f"{KIND[kind]} pair a={a} b={b}"][0] > single]
            print(f"| {KIND[kind]} | {a} | {ranges(conf)} |")
    print("\n## ds_read_b64_tr_b16\n")
    print("Lane l reads 8 bytes at row (l >> 2) & 3 (64-byte rows), columns 4 (l & 3) .. + 3, of its 16-lane "
          "group's block; the LDS holds element i at element i. What lanes 0-3 and 16 received:\n")
    rows = [line.split() for line in open(tr_path)]
    for r in rows[:4] + rows[16:17]:
        print(f"- lane {r[0]}: elements {', '.join(r[2:6])}")
    ok = all([int(v) for v in r[2:6]] == [256 * (int(r[0]) >> 4) + 32 * k + (int(r[0]) & 15) for k in range(4)]
             for r in rows)
    print(f"\nEvery lane L received column L & 15 of its group's 4 rows: {'yes' if ok else 'NO'}.\n")
    print("## The kernels' LDS accesses\n")
    print("| accesses | instruction | dispatch instructions | conflict / useful cycles | busy cycles per "
          "instruction |")
    print("|---|---|---:|---:|---:|")
    for n, kind, npat in names:
        if n.startswith("k_"):
            c, a, k = res[n]
            print(f"| {n} | {KIND[kind]} | {k:.0f} | {c / max(a - c, 1):.2f} | {a / k:.2f} |")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
