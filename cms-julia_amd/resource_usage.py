#!/usr/bin/env python3
"""Register, spill, scratch, LDS and occupancy figures of every gfx950 kernel
of the library, from the compiler's resource-usage remarks
(-Rpass-analysis=kernel-resource-usage), as a markdown table on stdout.

    python cms-julia_amd/resource_usage.py > profiles/kernel_resources_gfx950.md
"""
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SOURCES = ["pt_ca.hip", "pt_fit.hip", "pt_classify.hip", "pt_hist.hip", "pt_load_tracks.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "--cuda-device-only", "-Rpass-analysis=kernel-resource-usage", "-c", "-o", "/dev/null"]
KEYS = ("VGPRs", "AGPRs", "VGPRs Spill", "ScratchSize [bytes/lane]", "LDS Size [bytes/block]",
        "Occupancy [waves/SIMD]")


def demangle(names):
    try:
        out = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True).stdout
        return out.splitlines()
    except (OSError, subprocess.CalledProcessError):
        return names


def kernels(src):
    r = subprocess.run(["/opt/rocm/bin/hipcc", *FLAGS, str(PKG / "csrc" / src)], capture_output=True, text=True)
    if r.returncode != 0:
        sys.exit(f"compile of {src} failed:\n{r.stderr}")
    rows, cur = [], None
    for line in r.stderr.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = {"name": m.group(1)}
            rows.append(cur)
            continue
        m = re.search(r"remark:\s+([^:]+): (\S+) \[-Rpass", line)
        if m and cur is not None:
            cur[m.group(1).strip()] = m.group(2)
    return rows


def main():
    print("# Kernel resources on gfx950 (compiler remarks)\n")
    print("Command: `python cms-julia_amd/resource_usage.py` (hipcc " + " ".join(FLAGS[:-4]) +
          " -Rpass-analysis=kernel-resource-usage, each source of cms-julia_amd/csrc).\n")
    print("| source | kernel | VGPRs | AGPRs | VGPR spills | scratch B/lane | LDS B/block | waves/SIMD |")
    print("|---|---|---|---|---|---|---|---|")
    for src in SOURCES:
        rows = kernels(src)
        names = demangle([r["name"] for r in rows])
        for r, nm in zip(rows, names):
            nm = re.sub(r"\(.*\)$", "", nm).replace("pt::", "")
            print(f"| {src} | `{nm}` | " + " | ".join(r.get(k, "?") for k in KEYS) + " |")


if __name__ == "__main__":
    main()
