"""Build libnls_amd.so for gfx950 in-tree with hipcc.

``python -m nls_amd.build`` (or ``build()``) compiles every ``csrc/*.hip``
into ``_build/`` in parallel and links ``nls_amd/libnls_amd.so``.

The build fails if any kernel uses private (scratch) memory, spills
registers or needs a dynamic stack, as the compiler's resource-usage
remarks report them: such code has no place in these kernels, and private
arrays are where two miscompiles came from (an undefined offset register in
a scratch store).  The per-kernel figures go to _build/resource_usage.txt.
"""
import concurrent.futures as cf
import os
import re
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libnls_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + INCLUDE, "-I" + CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


REMARKS = ["-Rpass-analysis=kernel-resource-usage"]
_REMARK = re.compile(r"remark:\s+(Function Name|ScratchSize \[bytes/lane\]|VGPRs Spill|SGPRs Spill|Dynamic Stack|"
                     r"VGPRs|LDS Size \[bytes/block\]|Occupancy \[waves/SIMD\]):\s*(\S+)")


def resource_usage(stderr):
    """Per-kernel figures from the compiler's kernel-resource-usage remarks:
    {kernel: {field: value}}."""
    out, cur = {}, None
    for m in _REMARK.finditer(stderr):
        key, val = m.group(1), m.group(2)
        if key == "Function Name":
            cur = out.setdefault(val, {})
        elif cur is not None:
            cur[key] = val
    return out


def check_resources(usage):
    """Kernels that use scratch memory, spill vector registers or need a
    dynamic stack.  (Scalar registers spill into lanes of vector registers,
    not to memory: that shows as SGPRs Spill with ScratchSize 0 and is
    allowed.)"""
    bad = []
    for name, f in usage.items():
        if (int(f.get("ScratchSize [bytes/lane]", "0")) > 0 or int(f.get("VGPRs Spill", "0")) > 0 or
                f.get("Dynamic Stack", "False") != "False"):
            bad.append((name, f))
    return bad


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + REMARKS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj, resource_usage(r.stderr)


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        res = list(ex.map(_compile, srcs))
    objs = [o for o, _ in res]
    usage = {}
    for _, u in res:
        usage.update(u)
    with open(os.path.join(BUILD, "resource_usage.txt"), "w") as f:
        for name in sorted(usage):
            f.write("%s %s\n" % (name, " ".join("%s=%s" % kv for kv in sorted(usage[name].items()))))
    bad = check_resources(usage)
    if bad:
        raise RuntimeError("kernels using scratch memory, spilling or with a dynamic stack:\n" +
                           "\n".join("  %s: %s" % (n, f) for n, f in bad))
    if not usage:
        raise RuntimeError("hipcc reported no kernel resource usage (remarks missing?)")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
