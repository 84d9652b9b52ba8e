"""Build libmac_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m mac_amd.build      (or build() from Python)

Every .hip file under mac_amd/csrc is compiled for gfx950 and linked into
mac_amd/libmac_amd.so.  Packed-FP32 instructions are disabled for all device
code (-packed-fp32-ops): plain float code is compiled to scalar VALU ops.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libmac_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I" + INCLUDE,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False):
    """Compile every HIP source and link the shared library (always rebuilds)."""
    objs = []
    tmp = tempfile.mkdtemp(prefix="mac_amd_build_")
    procs = []
    for src in sources():
        obj = os.path.join(tmp, os.path.basename(src) + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write(f"--- {src}\n{text}\n")
        raise RuntimeError("HIP compilation failed: " + ", ".join(s for s, _ in failed))
    tmp_lib = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp_lib, LIB)
    for o in objs:
        os.remove(o)
    os.rmdir(tmp)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
