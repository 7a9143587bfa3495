"""Summaries of rocprofv3 output for this library's kernels (the figures
the profile notes quote).

    python3 -m wavelets_amd.tracestats kernels  <kernel_trace.csv> [top]
    python3 -m wavelets_amd.tracestats submit   <kernel_trace.csv>
    python3 -m wavelets_amd.tracestats busy     <kernel_trace.csv> <seconds>
    python3 -m wavelets_amd.tracestats counters <fetch_dir> <write_dir> [copy_bytes]
    python3 -m wavelets_amd.tracestats pmc      <counter_dir>

kernels   per-kernel calls, total and average duration, share of the total;
          the select's radix rounds split by round (launch order mod 3);
submit    the batched mixed stream's last timed submit (the one just before
          the plain path starts): per part, the transform launches' span and
          the select's span; the quantize span; the whole submit;
busy      the last <seconds> of the trace: time with at least one kernel
          running, kernels in flight while busy, the busiest kernels;
counters  FETCH_SIZE / WRITE_SIZE per dispatch (two --pmc runs), scaled by
          what they read for the library's own copy kernel (wv_copy_stream)
          of copy_bytes bytes, and the DWT passes' traffic per launch;
pmc       every counter of one --pmc run, mean per dispatch, for the DWT
          passes, the collect, the quantize and the copy kernel.
"""
import collections
import csv
import glob
import os
import sys


def _rows(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return rows


def _us(r):
    return (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3


def kernels(path, top=40):
    rows = _rows(path)
    agg = collections.OrderedDict()
    for r in rows:
        a = agg.setdefault(r["Kernel_Name"], [0, 0.0])
        a[0] += 1
        a[1] += _us(r)
    tot = sum(a[1] for a in agg.values())
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for k, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        kk = k if len(k) < 110 else k[:107] + "..."
        out.append("| `%s` | %d | %.2f | %.1f | %.1f%% |" % (kk, n, t / 1e3, t / n, 100 * t / tot))
    sr = [_us(r) for r in rows if r["Kernel_Name"].startswith("wv::sel_round")]
    for k in range(3):
        v = sr[k::3]
        if v:
            out.append("sel_round %d: %d launches, avg %.1f us, min %.1f us" %
                       (k, len(v), sum(v) / len(v), min(v)))
    return "\n".join(out)


def submit(path):
    rows = _rows(path)
    names = [r["Kernel_Name"] for r in rows]
    S = lambda i: int(rows[i]["Start_Timestamp"])
    E = lambda i: int(rows[i]["End_Timestamp"])
    fb = next(i for i, n in enumerate(names) if "bquant_hist" in n)
    plain = next(i for i in range(fb, len(names)) if names[i].startswith("wv::sel_collect"))
    bq = [i for i in range(plain) if "bquant_hist" in names[i]]
    bsel_f = [i for i in range(plain) if "bsel_finish" in names[i]][-2:]
    bsel_s = [i for i in range(plain) if "bsel_sample" in names[i]][-2:]
    prev_end = bq[-3]           # the previous submit's last quantize
    out = []
    starts = [prev_end, bsel_f[0]]
    first = None
    for p in range(2):
        idx = [i for i in range(starts[p] + 1, bsel_s[p])
               if "dwt_pass_kernel" in names[i] or "dwt_plane_kernel" in names[i]]
        first = idx[0] if first is None else first
        span = (max(E(i) for i in idx) - min(S(i) for i in idx)) / 1e6
        sel = (E(bsel_f[p]) - S(bsel_s[p])) / 1e6
        out.append("part %d: %d transform launches over %.3f ms; select %.3f ms" %
                   (p, len(idx), span, sel))
    out.append("quantize (both parts): %.3f ms" % ((E(bq[-1]) - S(bq[-2])) / 1e6))
    out.append("submit: %.3f ms from the first transform launch to the last quantize" %
               ((E(bq[-1]) - S(first)) / 1e6))
    return "\n".join(out)


def busy(path, seconds):
    rows = _rows(path)
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                for r in rows)
    t_end = max(e for _, e, _ in iv)
    sel = [x for x in iv if x[0] >= t_end - seconds * 1e9]
    b = 0
    cs = ce = None
    for s, e, _ in sel:
        if ce is None or s > ce:
            if ce is not None:
                b += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    if ce is not None:
        b += ce - cs
    span = sel[-1][1] - sel[0][0]
    dur = collections.Counter()
    cnt = collections.Counter()
    for s, e, k in sel:
        kk = k.split("(")[0]
        dur[kk] += e - s
        cnt[kk] += 1
    tot = sum(dur.values())
    out = ["last %.1f s: %d kernels, span %.3f s, busy %.3f s (%.1f%%), %.0f launches/s" %
           (seconds, len(sel), span / 1e9, b / 1e9, 100 * b / span, len(sel) / (span / 1e9)),
           "kernel time %.3f s = %.2f kernels in flight while busy" % (tot / 1e9, tot / max(b, 1))]
    for k, d in dur.most_common(12):
        out.append("%8d  %10.2f ms  %s" % (cnt[k], d / 1e6, k[:100]))
    return "\n".join(out)


def _counter(d, name):
    f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(f)):
        if r["Counter_Name"] == name:
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return agg


def counters(fetch_dir, write_dir, copy_bytes=1 << 30):
    fe = _counter(fetch_dir, "FETCH_SIZE")
    wr = _counter(write_dir, "WRITE_SIZE")
    mean = lambda v: sum(v) / len(v) if v else 0.0
    out = ["| kernel | dispatches (fetch / write run) | FETCH_SIZE KiB | WRITE_SIZE KiB |",
           "|---|---|---|---|"]
    keep = ("dwt_pass", "copy_kernel", "quant_hist", "sel_collect")
    for k in sorted(set(fe) | set(wr)):
        if any(t in k for t in keep):
            out.append("| `%s` | %d / %d | %.1f | %.1f |" % (
                k.split("(")[0][:80], len(fe.get(k, [])), len(wr.get(k, [])),
                mean(fe.get(k, [])), mean(wr.get(k, []))))
    cp = [k for k in fe if "copy_kernel" in k]
    fs = mean([x for k in cp for x in fe[k]]) * 1024 / copy_bytes
    ws = mean([x for k in cp for x in wr.get(k, [])]) * 1024 / copy_bytes
    out.append("scale (counter / true bytes of the copy): FETCH_SIZE %.4f, WRITE_SIZE %.4f" % (fs, ws))
    dw = [k for k in fe if "dwt_pass" in k]
    if dw and fs and ws:
        t = [mean(fe[k]) * 1024 / fs + mean(wr.get(k, [])) * 1024 / ws for k in dw]
        out.append("traffic_bytes_per_launch = %d" % round(sum(t) / len(t)))
    return "\n".join(out)


def pmc(counter_dir):
    f = glob.glob(os.path.join(counter_dir, "*counter_collection.csv"))[0]
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(f)):
        agg[r["Kernel_Name"].split("(")[0]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for d in agg.values() for c in d})
    out = ["| kernel | dispatches | " + " | ".join(names) + " |",
           "|---|---|" + "---|" * len(names)]
    keep = ("dwt_pass", "dwt_plane", "copy_kernel", "quant_hist", "sel_collect")
    for k in sorted(agg):
        if any(t in k for t in keep):
            d = agg[k]
            out.append("| `%s` | %d | %s |" % (
                k[:80], max(len(v) for v in d.values()),
                " | ".join("%.0f" % (sum(d[c]) / len(d[c])) if d.get(c) else "-" for c in names)))
    return "\n".join(out)


def main(argv):
    if len(argv) < 3:
        print(__doc__)
        return 2
    cmd = argv[1]
    if cmd == "kernels":
        print(kernels(argv[2], int(argv[3]) if len(argv) > 3 else 40))
    elif cmd == "submit":
        print(submit(argv[2]))
    elif cmd == "busy":
        print(busy(argv[2], float(argv[3])))
    elif cmd == "pmc":
        print(pmc(argv[2]))
    elif cmd == "counters":
        print(counters(argv[2], argv[3], int(argv[4]) if len(argv) > 4 else 1 << 30))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
