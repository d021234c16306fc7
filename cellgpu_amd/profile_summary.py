"""Summaries of rocprofv3 runs of bench.py for the profiles/ folder.

    python -m cellgpu_amd.profile_summary trace  <dir> <out.md> "<command>" ["<title>"] [phases=a-b,c-d] [steps=K]
    python -m cellgpu_amd.profile_summary counters <fetch_dir> <write_dir> <copy_fetch_dir> <copy_write_dir> \
        <out.md> <out.json> "<command>"
    python -m cellgpu_amd.profile_summary issue <sq_dir> <out.md> "<command>" ["<title>"]

`trace` reads <dir>/*_kernel_stats.csv and *_kernel_trace.csv (rocprofv3
--kernel-trace --stats --output-format csv) and tabulates per-kernel times, over
the whole run and by phase of the run (the first steps after the random initial
condition repair far more than the relaxed steady state).

`counters` reads FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch) of the bench
and of the plain copy probe (`bench.py --copy-probe 256`, a device-to-device copy
of 256 MiB). The probe fixes the factor that turns each counter into bytes:
expected bytes / counted bytes. The JSON holds, per kernel, the median scaled
HBM traffic per launch over the last dispatches; bench.py reports it as
roofline.traffic for the dominant kernel.

`issue` reads one rocprofv3 --pmc pass of SQ counters (SQ_WAVES, SQ_INSTS_VALU,
SQ_INSTS_SALU, SQ_INSTS_VMEM_RD, SQ_BUSY_CYCLES, ...) and tabulates per kernel
the instructions per wavefront and the VALU issue fraction: a wavefront's VALU
instruction occupies its SIMD for 4 cycles (64 lanes on 16-lane ALUs), so
4 SQ_INSTS_VALU / (SIMDs x cycles of the dispatch) is the share of the
machine's VALU issue the kernel used. SQ_BUSY_CYCLES is summed over the shader
arrays (agent Array_Count), so SQ_BUSY_CYCLES / arrays is the dispatch's
length in cycles.
"""
import csv
import glob
import json
import os
import sys

import numpy as np


def _name(raw):
    n = raw.split("(cgpu")[0].replace("(anonymous namespace)::", "").replace("void ", "")
    return n.strip()


def _one(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, pattern)))
    if not hits:
        raise FileNotFoundError(os.path.join(d, pattern))
    return hits[0]


def trace(d, out, cmd, title="Voronoi SPV ensemble (512 systems, 833 607 cells): kernel trace",
          phases="0-20,20-60,60-120,120-210", steps=None):
    """Kernel-trace summary. phases: dispatch-index windows "a-b,..." (a
    dispatch index = a step for the once-per-step kernels); steps: the number of
    steps of the run, to align kernels that also ran before the first step
    (their last `steps` dispatches are the steps)."""
    stats = list(csv.DictReader(open(_one(d, "*kernel_stats.csv"))))
    tr = list(csv.DictReader(open(_one(d, "*kernel_trace.csv"))))
    tr.sort(key=lambda r: int(r["Start_Timestamp"]))
    per = {}
    for r in tr:
        per.setdefault(_name(r["Kernel_Name"]), []).append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    lines = ["# " + title, "", "Command (repository root, MI355X):", "", "```", cmd, "```", "",
             "## rocprofv3 --stats (whole run)", "",
             "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in stats[:24]:
        lines.append("| %s | %s | %.2f | %.2f | %.2f | %.2f | %.2f |" % (
            _name(r["Name"]), r["Calls"], int(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
            int(r["MinNs"]) / 1e3, int(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    phases = [tuple(int(x) for x in p.split("-")) for p in phases.split(",")]
    need = steps if steps else phases[-1][1]
    lines += ["", "## mean us per dispatch by phase of the run (dispatch index = step)", "",
              "| kernel | " + " | ".join("%d-%d" % p for p in phases) + " | last 100 |",
              "|---|" + "---|" * (len(phases) + 1)]
    tot = np.zeros(len(phases) + 1)
    for n, v in sorted(per.items()):
        v = np.array(v)
        if len(v) < need:
            continue
        if steps:
            v = v[-steps:]
        row = [v[a:b].mean() for a, b in phases] + [v[-100:].mean()]
        tot += row
        lines.append("| %s | " % n + " | ".join("%.1f" % x for x in row) + " |")
    lines.append("| **sum** | " + " | ".join("%.1f" % x for x in tot) + " |")
    open(out, "w").write("\n".join(lines) + "\n")
    return out


def _counter(d, cname):
    vals = {}
    for r in csv.DictReader(open(_one(d, "*counter_collection.csv"))):
        if r["Counter_Name"] != cname:
            continue
        vals.setdefault(_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return vals


def _short(k):
    if "rocprim" in k:  # library sort kernels: their configuration name
        for tag in ("radix_sort_block_sort", "merge_sort_block_merge", "merge_mergepath_partition",
                    "merge_mergepath", "radix_sort"):
            if tag in k:
                return "rocprim " + tag
        return "rocprim"
    return k.split("(Words")[0]


def counters(fd, wd, cfd, cwd, out_md, out_json, cmd, probe_mib=256, last=10):
    f, w = _counter(fd, "FETCH_SIZE"), _counter(wd, "WRITE_SIZE")
    cf, cw = _counter(cfd, "FETCH_SIZE"), _counter(cwd, "WRITE_SIZE")
    expect_kib = probe_mib * 1024.0
    # the probe's copy kernel reads the most (the fill that creates the source
    # tensor writes as much but reads nothing)
    ck = max(cf, key=lambda k: np.median(cf[k]))
    f_scale = expect_kib / np.median(cf[ck])
    w_scale = expect_kib / np.median(cw[ck])
    res = {"copy_probe": {"kernel": ck, "mib": probe_mib, "fetch_kib": float(np.median(cf[ck])),
                          "write_kib": float(np.median(cw[ck])), "fetch_scale": float(f_scale),
                          "write_scale": float(w_scale)},
           "kernels": {}}
    lines = ["# HBM traffic per launch from FETCH_SIZE / WRITE_SIZE", "", "Commands (repository root, MI355X):",
             "", "```", cmd, "```", "",
             "Copy probe (%s, %d MiB): FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB per dispatch -> scale factors "
             "%.3f (fetch), %.3f (write)." % (ck, probe_mib, np.median(cf[ck]), np.median(cw[ck]), f_scale, w_scale),
             "", "Median over the last %d dispatches of each kernel." % last, "",
             "| kernel | FETCH KiB | WRITE KiB | traffic MB (scaled) |", "|---|---|---|---|"]
    for k in f:
        fv = np.median(f[k][-last:])
        wv = np.median(w.get(k, [0.0])[-last:])
        mb = (fv * f_scale + wv * w_scale) * 1024 / 1e6
        res["kernels"][k] = {"fetch_kib": float(fv), "write_kib": float(wv), "traffic_bytes": mb * 1e6}
        lines.append("| %s | %.0f | %.0f | %.1f |" % (_short(k), fv, wv, mb))
    open(out_md, "w").write("\n".join(lines) + "\n")
    json.dump(res, open(out_json, "w"), indent=1)
    return res


def issue(d, out, cmd, title="SQ instruction counters per kernel", last=10):
    agent = [r for r in csv.DictReader(open(_one(d, "*agent_info.csv"))) if r.get("Name", "").startswith("gfx")][0]
    simds, arrays = int(agent["Simd_Count"]), int(agent["Array_Count"])
    clock_mhz = float(agent["Max_Engine_Clk_Fcompute"])
    rows = list(csv.DictReader(open(_one(d, "*counter_collection.csv"))))
    per = {}
    for r in rows:
        k = _short(_name(r["Kernel_Name"]))
        per.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(
            (int(r["Dispatch_Id"]), float(r["Counter_Value"]),
             (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    lines = ["# " + title, "", "Command (repository root, MI355X):", "", "```", cmd, "```", "",
             "Median over the last %d dispatches of each kernel; %d SIMDs in %d shader arrays, %.0f MHz." % (
                 last, simds, arrays, clock_mhz), "",
             "| kernel | waves | VALU / wave | SALU / wave | VMEM reads / wave | busy us | VALU issue |",
             "|---|---|---|---|---|---|---|"]
    res = {}
    for k, cs in sorted(per.items()):
        if "SQ_INSTS_VALU" not in cs or "SQ_WAVES" not in cs:
            continue

        def med(c):
            v = [x[1] for x in sorted(cs[c], key=lambda t: t[0])[-last:]] if c in cs else []
            return float(np.median(v)) if v else float("nan")
        waves, valu = med("SQ_WAVES"), med("SQ_INSTS_VALU")
        busy = med("SQ_BUSY_CYCLES") / arrays if "SQ_BUSY_CYCLES" in cs else float("nan")
        frac = 4.0 * valu / (simds * busy) if busy == busy and busy > 0 else float("nan")
        res[k] = dict(waves=waves, valu_per_wave=valu / waves, issue=frac)
        lines.append("| %s | %.0f | %.0f | %.0f | %.0f | %.1f | %.2f |" % (
            k, waves, valu / waves, med("SQ_INSTS_SALU") / waves, med("SQ_INSTS_VMEM_RD") / waves,
            busy / clock_mhz, frac))
    open(out, "w").write("\n".join(lines) + "\n")
    json.dump(res, open(os.path.splitext(out)[0] + ".json", "w"), indent=1)
    return res


if __name__ == "__main__":
    a = sys.argv[1:]
    if a and a[0] == "trace":
        kw = dict(kv.split("=", 1) for kv in a[5:])
        if "steps" in kw:
            kw["steps"] = int(kw["steps"])
        if len(a) > 4:
            trace(a[1], a[2], a[3], a[4], **kw)
        else:
            trace(a[1], a[2], a[3])
    elif a and a[0] == "counters":
        counters(a[1], a[2], a[3], a[4], a[5], a[6], a[7])
    elif a and a[0] == "issue":
        issue(*a[1:5])
    else:
        print(__doc__)
        sys.exit(2)
