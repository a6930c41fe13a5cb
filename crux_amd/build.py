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
ARCH = os.environ.get("CRUX_OFFLOAD_ARCH", "gfx950")


def hipcc():
    p = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def build(verbose=False):
    objs = []
    common = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"),
              "-I", CSRC, "-Wall", "-Wno-unused-result"]
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, "_" + src + ".o")
        if src.endswith(".hip"):
            cmd = [hipcc(), "--offload-arch=" + ARCH, "-x", "hip", "-Xclang",
                   "-target-feature", "-Xclang", "-packed-fp32-ops",
                   # Matrix-core results straight into VGPRs: the 4-state
                   # kernel feeds them to vector instructions at once, and
                   # accumulation registers would cost a move per value.
                   "-mllvm", "-amdgpu-mfma-vgpr-form", "-c",
                   path, "-o", obj] + common
        else:
            cmd = [hipcc(), "-c", path, "-o", obj] + common
        if verbose:
            print(" ".join(cmd))
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                       stderr=subprocess.STDOUT), src))
        objs.append(obj)
    failed = []
    for p, src in procs:
        out, _ = p.communicate()
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
