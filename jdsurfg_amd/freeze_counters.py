"""Freeze rocprofv3 counter runs of the grid256 bench into the per-kernel traffic
and hardware-FP64 table that bench.py's roofline reads
(jdsurfg_amd/traffic_model.json), and write the Markdown summary of the same
runs.

Usage (each rocprofv3 run writes a directory of CSV files):

  python3 -m jdsurfg_amd.freeze_counters copy
      a 768 MiB device copy, three times: run under
      rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate the two counters
  python3 -m jdsurfg_amd.freeze_counters freeze TRACE SQ FETCH WRITE CFETCH CWRITE OUT.md
      TRACE: --kernel-trace --stats run of the bench; SQ: the FP64 and VALU
      instruction counters; FETCH, WRITE: FETCH_SIZE and WRITE_SIZE runs of the
      bench; CFETCH, CWRITE: the same counters over the copy.  Writes
      jdsurfg_amd/traffic_model.json and OUT.md.
  python3 -m jdsurfg_amd.freeze_counters stats TRACE OUT.md
      appends the kernel table of a --kernel-trace --stats run to OUT.md.

HBM bytes per dispatch are FETCH_SIZE (KiB) and WRITE_SIZE (KiB) times 1024,
each scaled so that the copy reads and writes its 768 MiB; hardware FP64
flops are 64 lanes x (ADD + MUL + 2 FMA + TRANS) wave instructions.
"""
import collections
import csv
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
COPY_BYTES = 768 << 20


def _short(name):
    n = name.split("(")[0]
    return n[5:] if n.startswith("void ") else n


def _counters(d):
    """{kernel: {counter: summed value}}, {kernel: dispatches}"""
    f = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)[0]
    tot = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for r in csv.DictReader(open(f)):
        k = _short(r["Kernel_Name"])
        tot[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k].add(r["Dispatch_Id"])
    return tot, {k: len(v) for k, v in disp.items()}


def _stats(d):
    f = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)[0]
    out = {}
    for r in csv.DictReader(open(f)):
        out[_short(r["Name"])] = (int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6)
    return out


def copy_main():
    import torch
    a = torch.empty(COPY_BYTES, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied", 3 * COPY_BYTES, "bytes", flush=True)


def _copy_factor(d, counter):
    tot, n = _counters(d)
    k = [k for k in tot if "copy" in k.lower() or "elementwise" in k.lower()]
    per = max(tot[x][counter] / n[x] for x in k) * 1024.0
    return COPY_BYTES / per


def freeze_main(trace, sq, fetch, write, cfetch, cwrite, out_md):
    st = _stats(trace)
    tsq, nsq = _counters(sq)
    tf, nf = _counters(fetch)
    tw, nw = _counters(write)
    ff, fw = _copy_factor(cfetch, "FETCH_SIZE"), _copy_factor(cwrite, "WRITE_SIZE")
    src = (f"{os.path.relpath(out_md, os.path.dirname(HERE))} (rocprofv3 FETCH_SIZE x{ff:.2f}, "
           f"WRITE_SIZE x{fw:.2f} against a 768 MiB device copy; SQ_INSTS_VALU_*_F64 x64 lanes, "
           "FMA x2; grid256 bench)")
    model = {"source": src, "kernels": {}}
    rows = []
    for k, (calls, ms) in sorted(st.items(), key=lambda kv: -kv[1][1]):
        if k not in tsq or k not in tf or k not in tw:
            continue
        # the library's kernels, and others that take a millisecond or more
        if not k.startswith("jdsg_") and ms < 1.0:
            continue
        q = tsq[k]
        n = nsq[k]
        f64 = {c: q.get(f"SQ_INSTS_VALU_{c}_F64", 0.0) / n for c in ("ADD", "MUL", "FMA", "TRANS")}
        hw = 64.0 * (f64["ADD"] + f64["MUL"] + 2.0 * f64["FMA"] + f64["TRANS"])
        valu = q.get("SQ_INSTS_VALU", 0.0) / n
        rd = tf[k]["FETCH_SIZE"] / nf[k] * 1024.0 * ff
        wr = tw[k]["WRITE_SIZE"] / nw[k] * 1024.0 * fw
        avg = ms / calls
        model["kernels"][k] = {"dispatches": calls, "avg_ms": avg,
                               "hbm_read_bytes_per_dispatch": rd,
                               "hbm_write_bytes_per_dispatch": wr,
                               "hw_fp64_flops_per_dispatch": hw,
                               "valu_wave_instructions_per_dispatch": valu}
        # VALU issue: wave instructions x 4 cycles over the 1024 SIMDs' cycles at 2.4 GHz
        issue = valu * 4.0 / (avg * 1e-3 * 2.4e9 * 1024)
        rows.append(f"| `{k}` | {calls} | {ms:.1f} | {avg:.3f} | {valu:.3g} | "
                    f"{f64['ADD']:.2g}/{f64['MUL']:.2g}/{f64['FMA']:.2g}/{f64['TRANS']:.2g} | "
                    f"{hw / (avg * 1e-3) / 1e12:.1f} ({hw / (avg * 1e-3) / 78.6e12:.2f}) | "
                    f"{issue:.2f} | {rd / 1e9:.3f} | {wr / 1e9:.3f} |")
    with open(os.path.join(HERE, "traffic_model.json"), "w") as f:
        json.dump(model, f, indent=1)
    with open(out_md, "a") as f:
        f.write(f"\nCalibration: the 768 MiB copy gives FETCH_SIZE x{ff:.2f} and WRITE_SIZE "
                f"x{fw:.2f}.  Per dispatch (averages over the bench's dispatches of each "
                "kernel):\n\n")
        f.write("| kernel | calls | total ms | avg ms | VALU wave instr | FP64 add/mul/fma/trans "
                "wave instr | HW FP64 TFLOP/s (frac) | VALU issue | HBM read GB | HBM write GB "
                "|\n|---|---|---|---|---|---|---|---|---|---|\n")
        f.write("\n".join(rows) + "\n")
    print("wrote", os.path.join(HERE, "traffic_model.json"), "and", out_md)


def stats_main(trace, out_md):
    st = _stats(trace)
    tot = sum(ms for _, ms in st.values())
    with open(out_md, "a") as f:
        f.write("| kernel | calls | total ms | avg ms | % |\n|---|---|---|---|---|\n")
        for k, (calls, ms) in sorted(st.items(), key=lambda kv: -kv[1][1]):
            if ms < 0.5:
                continue
            f.write(f"| `{k[:90]}` | {calls} | {ms:.1f} | {ms / calls:.3f} | "
                    f"{100.0 * ms / tot:.2f} |\n")


if __name__ == "__main__":
    if sys.argv[1] == "copy":
        copy_main()
    elif sys.argv[1] == "stats":
        stats_main(*sys.argv[2:4])
    elif sys.argv[1] == "freeze":
        freeze_main(*sys.argv[2:9])
    else:
        raise SystemExit(__doc__)
