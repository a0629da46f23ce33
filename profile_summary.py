"""Summaries of rocprofv3 output for the update path, written under profiles/.

    python3 profile_summary.py kernels RESULTS_DB OUT_MD --command "..." [--last N]
        per-kernel launch durations (a --kernel-trace sqlite database) over
        the update batches, without the bulk loads' dispatches; with --last,
        over the last N batches whatever their sort (a workload whose timed
        batches follow its bulk loads, such as the batched form's rounds).

    python3 profile_summary.py traffic OUT_JSON OUT_MD --config JSON --command "..." \
        FETCH_DIR WRITE_DIR RDREQ_DIR WRREQ_DIR
        HBM bytes per launch from four --pmc passes (FETCH_SIZE; WRITE_SIZE;
        TCC_EA0_RDREQ_{32B,64B,128B} with TCC_EA0_RDREQ; TCC_EA0_WRREQ_64B with
        TCC_EA0_WRREQ), per kernel over the same update batches.  Bytes are
        counted per request size (32 B, 64 B, 128 B reads; 64 B and 32 B
        writes); a device copy of known size in the same program checks the
        method.

An update batch is the run of dispatches from its first kernel to the next
batch's: k_batch_begin, or the sample-split sort's k_ss_split where that
carries the batch's first work (a lone PMA's batch of up to 2^17 keys).  A
batch whose sort took the large-batch radix kernels is a bulk load (the
update batches sort with the sample-split kernels) and is left out.
"""
import argparse
import csv
import glob
import json
import os
import re
import sqlite3
import statistics
from collections import defaultdict

BULK_KERNELS = {"k_pack_group", "k_rs_hist", "k_rs_scatter", "k_pack_count", "k_os_pass"}
# the structural check a harness runs between batches: part of no batch
NOT_BATCH = {"k_check"}
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short(name):
    n = re.sub(r"\(.*", "", name)
    n = re.sub(r"<.*", "", n)
    return n.replace("void ", "").replace("pma::", "").strip()


def batches(dispatches):
    """Group (short name, payload) dispatches in order into batches; returns
    (update batches, bulk batches, dispatches before the first batch)."""
    segs, cur, pre = [], [], []
    for d in dispatches:
        first = d[0] == "k_batch_begin" or (d[0] == "k_ss_split" and not (cur and cur[-1][0] == "k_batch_begin"))
        if first:
            if cur:
                segs.append(cur)
            cur = [d]
        elif cur:
            cur.append(d)
        else:
            pre.append(d)
    if cur:
        segs.append(cur)
    upd, bulk = [], []
    for sg in segs:
        names = {k for k, _ in sg}
        (bulk if (names & BULK_KERNELS) and "k_ss_split" not in names else upd).append(sg)
    return upd, bulk, pre


def cmd_kernels(a):
    c = sqlite3.connect(a.db)
    rows = [(short(n), (s, e)) for n, s, e in c.execute("select name, start, end from kernels order by start")]
    rows = [r for r in rows if r[0] not in NOT_BATCH]
    upd, bulk, _ = batches(rows)
    if a.last:
        # the last N batches in trace order, whatever their sort
        allb = sorted(upd + bulk, key=lambda sg: sg[0][1][0])
        upd, bulk = allb[-a.last:], allb[:-a.last]
    per = defaultdict(list)
    spans = []
    for sg in upd:
        spans.append((sg[-1][1][1] - sg[0][1][0]) / 1e3)
        for k, (s, e) in sg:
            per[k].append((e - s) / 1e3)
    tot = sum(sum(v) for v in per.values())
    which = (f"the last {len(upd)} batches (each from its first kernel, `k_batch_begin` or the sort's "
             f"`k_ss_split`, to the next batch's); the {len(bulk)} batches before them are left out"
             if a.last else
             f"{len(upd)} update batches (each from its first kernel, `k_batch_begin` or the sort's "
             f"`k_ss_split`, to the next batch's); the {len(bulk)} bulk loads' batches are left out")
    lines = [f"# {a.title}", "", "Command (repo root, one MI355X):", "", "```", a.command, "```", "",
             f"Launch durations from the kernel trace over {which}.  "
             "`k_check` (a structural check run between batches) is left out.  "
             "The trace slows the launch chain (gaps between kernels), not the kernels.", "",
             "| kernel | launches | avg us | median us | p99 us | max us | total ms | share |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        s = sorted(v)
        p99 = s[min(len(s) - 1, int(0.99 * len(s)))]
        lines.append(f"| `{k}` | {len(v)} | {sum(v) / len(v):.2f} | {statistics.median(v):.2f} | {p99:.2f} | "
                     f"{s[-1]:.2f} | {sum(v) / 1e3:.3f} | {sum(v) / tot:.3f} |")
    s = sorted(spans)
    tail = "; the last batch's span also holds what ran after it" if a.last else ""
    lines += ["", f"Batch span (first kernel start to last kernel end, under the trace): median "
                  f"{statistics.median(s):.1f} us, p99 {s[min(len(s) - 1, int(0.99 * len(s)))]:.1f} us{tail}.", ""]
    with open(a.out_md, "w") as f:
        f.write("\n".join(lines))


def read_pmc(d):
    """{dispatch id: (short kernel name, {counter: value})} of a --pmc pass."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    out = {}
    for path in files:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                did = int(r["Dispatch_Id"])
                name = short(r["Kernel_Name"])
                ent = out.setdefault(did, (name, {}))
                ent[1][r["Counter_Name"]] = ent[1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def per_kernel(passes):
    """Average of every counter of every pass per kernel over the update
    batches, plus the copy calibration dispatch (the largest copy)."""
    acc = defaultdict(lambda: defaultdict(list))
    copy = {}
    for p in passes:
        items = [(name, cnt) for _, (name, cnt) in sorted(p.items())]
        upd, _, _ = batches(items)
        for sg in upd:
            for k, cnt in sg:
                for cn, v in cnt.items():
                    acc[k][cn].append(v)
        # the calibration copy runs after the last batch: the largest copy
        copies = [cnt for name, cnt in items if name == COPY_KERNEL]
        if copies:
            big = max(copies, key=lambda c: max(c.values()))
            copy.update(big)
    return acc, copy


def req_bytes(c):
    """(read bytes, write bytes) from per-request-size counts."""
    rd = 32 * c.get("TCC_EA0_RDREQ_32B", 0) + 64 * c.get("TCC_EA0_RDREQ_64B", 0) + 128 * c.get("TCC_EA0_RDREQ_128B", 0)
    w64 = c.get("TCC_EA0_WRREQ_64B", 0)
    wr = 64 * w64 + 32 * (c.get("TCC_EA0_WRREQ", 0) - w64)
    return rd, wr


def cmd_traffic(a):
    passes = [read_pmc(d) for d in a.dirs]
    acc, copy = per_kernel(passes)
    calib = None
    if copy:
        rd, wr = req_bytes(copy)
        calib = {"copy_bytes": a.copy_bytes, "fetch_size_bytes": copy.get("FETCH_SIZE", 0) * 1024,
                 "write_size_bytes": copy.get("WRITE_SIZE", 0) * 1024, "read_request_bytes": rd,
                 "write_request_bytes": wr, "counters": copy}
    kern = {}
    for k, cs in sorted(acc.items()):
        avg = {cn: sum(v) / len(v) for cn, v in cs.items()}
        rd, wr = req_bytes(avg)
        kern[k] = {"dispatches": max(len(v) for v in cs.values()), "fetch_size_bytes": avg.get("FETCH_SIZE", 0) * 1024,
                   "write_size_bytes": avg.get("WRITE_SIZE", 0) * 1024, "read_request_bytes": rd,
                   "write_request_bytes": wr, "hbm_bytes_per_launch": rd + wr,
                   "read_requests": {s: avg.get("TCC_EA0_RDREQ_" + s, 0) for s in ("32B", "64B", "128B")},
                   "write_requests": {"64B": avg.get("TCC_EA0_WRREQ_64B", 0),
                                      "32B": avg.get("TCC_EA0_WRREQ", 0) - avg.get("TCC_EA0_WRREQ_64B", 0)}}
    out = {"workload": a.workload, "config": json.loads(a.config), "commands": a.command,
           "window": a.window,
           "method": "HBM bytes per launch = 32*RDREQ_32B + 64*RDREQ_64B + 128*RDREQ_128B (reads) + "
                     "64*WRREQ_64B + 32*(WRREQ - WRREQ_64B) (writes), TCC_EA0 counters summed over the "
                     "L2 channels, averaged over the kernel's dispatches in the update batches",
           "calibration": calib, "kernels": kern}
    with open(a.out_json, "w") as f:
        json.dump(out, f, indent=1)
    lines = [f"# HBM traffic per launch: {a.workload}", "", "Four counter passes, one MI355X, repo root:", "", "```"]
    lines += a.command + ["```", "", f"Window: {a.window}.", "", out["method"] + "."]
    if calib:
        lines += ["", f"Check: the calibration copy of {a.copy_bytes / 2**20:.0f} MiB reads "
                      f"{calib['read_request_bytes'] / 2**20:.1f} MiB and writes "
                      f"{calib['write_request_bytes'] / 2**20:.1f} MiB by request size "
                      f"(FETCH_SIZE {calib['fetch_size_bytes'] / 2**20:.1f} MiB, WRITE_SIZE "
                      f"{calib['write_size_bytes'] / 2**20:.1f} MiB)."]
    lines += ["", "| kernel | dispatches | read MB | write MB | HBM MB per launch | reads 32/64/128 B | "
                  "FETCH_SIZE MB | WRITE_SIZE MB |", "|---|---:|---:|---:|---:|---|---:|---:|"]
    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["hbm_bytes_per_launch"]):
        rq = v["read_requests"]
        lines.append(f"| `{k}` | {v['dispatches']} | {v['read_request_bytes'] / 1e6:.2f} | "
                     f"{v['write_request_bytes'] / 1e6:.2f} | {v['hbm_bytes_per_launch'] / 1e6:.2f} | "
                     f"{rq['32B']:.0f} / {rq['64B']:.0f} / {rq['128B']:.0f} | {v['fetch_size_bytes'] / 1e6:.2f} | "
                     f"{v['write_size_bytes'] / 1e6:.2f} |")
    with open(a.out_md, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("out_md")
    k.add_argument("--title", default="Kernel profile")
    k.add_argument("--command", required=True)
    k.add_argument("--last", type=int, default=0)
    t = sub.add_parser("traffic")
    t.add_argument("out_json")
    t.add_argument("out_md")
    t.add_argument("dirs", nargs=4)
    t.add_argument("--config", required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--window", required=True)
    t.add_argument("--command", action="append", required=True)
    t.add_argument("--copy-bytes", type=int, default=256 << 20)
    a = ap.parse_args()
    (cmd_kernels if a.cmd == "kernels" else cmd_traffic)(a)


if __name__ == "__main__":
    main()
