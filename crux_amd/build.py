"""Build libcrux_amd.so in-tree with hipcc for gfx950.

    python -m crux_amd.build

All HIP code is compiled without packed-FP32 instructions
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcrux_amd.so")
SOURCES = ["engine.hip", "kernels.hip", "model.cpp", "server.cpp"]
# kernels.hip is compiled once per traversal class (K, CM) plus once for the
# P-matrix kernel and the dispatch (K 0), so the classes build in parallel.
KERNEL_PARTS = [(0, 0)] + [(4, cm) for cm in (1, 2, 4, 8, 16)] + \
    [(k, cm) for k in (20, 22) for cm in (1, 2, 4, 8)]
ARCH = os.environ.get("CRUX_OFFLOAD_ARCH", "gfx950")


def hipcc():
    p = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def build(verbose=False):
    objs = []
    common = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"),
              "-I", CSRC, "-Wall", "-Wno-unused-result"]
    units = []
    for src in SOURCES:
        if src == "kernels.hip":
            units += [(src, "_%s.%d_%d.o" % (src, k, cm),
                       ["-DCRUX_PART_K=%d" % k, "-DCRUX_PART_CM=%d" % cm])
                      for k, cm in KERNEL_PARTS]
        else:
            units.append((src, "_" + src + ".o", []))
    jobs = max(1, min(len(units), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    running, done = [], []
    for src, oname, defs in units:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, oname)
        if src.endswith(".hip"):
            cmd = [hipcc(), "--offload-arch=" + ARCH, "-x", "hip", "-Xclang",
                   "-target-feature", "-Xclang", "-packed-fp32-ops",
                   # Matrix-core results straight into VGPRs: the 4-state
                   # kernel feeds them to vector instructions at once, and
                   # accumulation registers would cost a move per value.
                   "-mllvm", "-amdgpu-mfma-vgpr-form", "-c",
                   path, "-o", obj] + defs + common
        else:
            cmd = [hipcc(), "-c", path, "-o", obj] + common
        if verbose:
            print(" ".join(cmd))
        while len(running) >= jobs:
            # Output is read as each unit ends (a full pipe would stall it).
            p, name = running.pop(0)
            done.append((p, name, p.communicate()[0]))
        running.append((subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                         stderr=subprocess.STDOUT), oname))
        objs.append(obj)
    done += [(p, name, p.communicate()[0]) for p, name in running]
    failed = []
    for p, src, out in done:
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines()
                 if "packed-fp32-ops" not in l and "is not a recognized feature"
                 not in l and l.strip() != ""]
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        elif verbose and lines:
            sys.stderr.write("\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("hipcc failed for %s" % ", ".join(failed))
    tmp = LIB + ".tmp%d" % os.getpid()
    subprocess.check_call([hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC",
                           "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
