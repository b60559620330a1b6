"""Summaries of rocprofv3 CSV output for the profiles kept with the sources.

  python -m shyhip.profsum stats  --trace T.csv --stats S.csv --out F.md --command "..."
      per-kernel calls, average / total duration, share, VGPRs, scratch
  python -m shyhip.profsum traffic --fetch F.csv --write W.csv --copy-bytes N --out F.md
                                   --command-fetch "..." --command-write "..."
      FETCH_SIZE / WRITE_SIZE per dispatch (KiB) of each kernel, each scaled
      by what the counter reads for a plain copy of N bytes (the largest
      __amd_rocclr_copyBuffer dispatch of the run); the result is written as
      a table and as a JSON block that bench.py reads.

CSV column names are those of rocprofv3 --output-format csv.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re

COPY_KERNEL = "__amd_rocclr_copyBuffer"
JSON_BEGIN = "<!-- traffic-json"
JSON_END = "-->"


def short_name(name: str) -> str:
    """`void shy::k_bicg_a<float, true>(shy::SolvArgs<float>, int)` -> `k_bicg_a<float, true>`."""
    n = name.replace("shy::", "")
    if n.startswith("void "):
        n = n[5:]
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def stats(args) -> None:
    regs = {}
    for r in csv.DictReader(open(args.trace)):
        regs.setdefault(r["Kernel_Name"], (r.get("VGPR_Count", ""), r.get("Scratch_Size", "")))
    rows = sorted(csv.DictReader(open(args.stats)), key=lambda r: -float(r["TotalDurationNs"]))
    with open(args.out, "w") as f:
        f.write(f"# {args.title}\n\nCommand (repository root on the GPU host):\n\n```\n{args.command}\n```\n\n")
        f.write(f"Summary: `python -m shyhip.profsum stats --trace {args.trace} --stats {args.stats} ...`\n\n")
        f.write("rocprofv3 kernel statistics over the whole profiled process:\n\n")
        f.write("| kernel | calls | avg us | total ms | % | VGPR | scratch B |\n|---|---:|---:|---:|---:|---:|---:|\n")
        for r in rows[: args.top]:
            v = regs.get(r["Name"], ("", ""))
            f.write(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                    f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} | {v[0]} | {v[1]} |\n")


def _counters(path):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[(r["Kernel_Name"], int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in d.items()}, {k: len(v) for k, v in d.items()}


def traffic(args) -> None:
    fch, cnt = _counters(args.fetch)
    wrt, _ = _counters(args.write)
    copies = [k for k in fch if k[0].startswith(COPY_KERNEL)]
    if not copies:
        raise SystemExit("no copy kernel in the fetch run")
    ck = max(copies, key=lambda k: k[1])
    kib = args.copy_bytes / 1024.0
    ff = kib / fch[ck]
    fw = kib / wrt[ck] if wrt.get(ck) else None
    rows = []
    for key, fv in fch.items():
        if key[0].startswith("__amd"):
            continue
        wv = wrt.get(key)
        rows.append((short_name(key[0]), key[1], cnt[key], fv, wv))
    rows.sort(key=lambda r: -(r[3] * r[2]))
    table = {}
    for name, grid, n, fv, wv in rows:
        if name in table:
            continue
        tb = None if wv is None or fw is None else (fv * ff + wv * fw) * 1024.0
        table[name] = {"grid": grid, "dispatches": n, "fetch_kib": round(fv, 1),
                       "write_kib": None if wv is None else round(wv, 1),
                       "traffic_bytes": None if tb is None else int(tb)}
    with open(args.out, "w") as f:
        f.write(f"# {args.title}\n\nCommands (repository root on the GPU host), one counter per run:\n\n")
        f.write(f"```\n{args.command_fetch}\n{args.command_write}\n```\n\n")
        f.write(f"Calibration: the largest `{COPY_KERNEL}` dispatch copies {args.copy_bytes} bytes "
                f"({kib:.0f} KiB); FETCH_SIZE reads {fch[ck]:.1f} KiB (factor {ff:.4f}), WRITE_SIZE reads "
                f"{wrt.get(ck, float('nan')):.1f} KiB (factor {fw if fw is None else round(fw, 4)}).\n"
                "Traffic per dispatch = FETCH_SIZE x fetch factor + WRITE_SIZE x write factor.  FETCH_SIZE "
                "counts what L2 fetches from beyond it (MALL or HBM); data a kernel reads again from L2 "
                "is not in it.\n\n")
        f.write("| kernel | grid | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic MB |\n"
                "|---|---:|---:|---:|---:|---:|\n")
        for name, t in list(table.items())[: args.top]:
            tb = "" if t["traffic_bytes"] is None else f"{t['traffic_bytes'] / 1e6:.3f}"
            f.write(f"| `{name}` | {t['grid']} | {t['dispatches']} | {t['fetch_kib']} | {t['write_kib']} | {tb} |\n")
        f.write(f"\n{JSON_BEGIN}\n")
        f.write(json.dumps({"copy_bytes": args.copy_bytes, "fetch_factor": ff, "write_factor": fw,
                            "kernels": table}, indent=0))
        f.write(f"\n{JSON_END}\n")


def read_traffic(path: str):
    """The JSON block of a traffic summary, or None."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(re.escape(JSON_BEGIN) + r"(.*?)" + re.escape(JSON_END), txt, re.S)
    return json.loads(m.group(1)) if m else None


def main() -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("--trace", required=True)
    s.add_argument("--stats", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--title", default="Kernel statistics")
    s.add_argument("--command", required=True)
    s.add_argument("--top", type=int, default=40)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--title", default="HBM traffic per dispatch")
    t.add_argument("--command-fetch", required=True)
    t.add_argument("--command-write", required=True)
    t.add_argument("--top", type=int, default=30)
    a = ap.parse_args()
    stats(a) if a.cmd == "stats" else traffic(a)


if __name__ == "__main__":
    main()
