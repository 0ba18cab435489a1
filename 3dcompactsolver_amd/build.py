"""Build lib3dcompact_hip.so (gfx950) in-tree with hipcc.

Every .hip and .cpp file under csrc/ is compiled for gfx950 with hipcc into
an object in _build/, then linked into ../3dcompactsolver_amd/lib3dcompact_hip.so.
The product depends on nothing else (no torch, no oracle).

Every kernel must run without scratch memory: the .hip files are compiled
with the compiler's per-kernel resource report, and a kernel that spills
registers to scratch (ScratchSize above 0) fails the build, naming the
kernel.  A spill costs memory traffic on the hot path, and a build whose
spilling variant gave wrong results has been seen on this path before.
"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib3dcompact_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are not generated (they have given silent wrong
    # results on this hardware/compiler pair)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(HERE, "..", "include"),
]


RESOURCE_REPORT = "-Rpass-analysis=kernel-resource-usage"


def scratch_kernels(report):
    """[(kernel, bytes per lane)] of the kernels with scratch in a hipcc
    resource report (the remarks of RESOURCE_REPORT on stderr)"""
    out, name = [], None
    for line in report.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name and int(m.group(1)) > 0:
            out.append((name, int(m.group(1))))
    return out


def _compile(src):
    import time
    t0 = time.time()
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    else:
        cmd.append(RESOURCE_REPORT)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    spills = scratch_kernels(r.stderr)
    if spills:
        names = "\n".join(f"  {k}: {b} bytes/lane" for k, b in spills)
        raise RuntimeError(f"{os.path.basename(src)}: kernels spill to scratch:\n{names}")
    if os.environ.get("CS3D_BUILD_VERBOSE"):
        print(f"  {os.path.basename(src)}: {time.time() - t0:.1f}s", flush=True)
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
