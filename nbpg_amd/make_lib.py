"""Compile the gfx950 HIP sources into nbpg_amd/libnbpg_amd.so (in-tree)."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnbpg_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include")]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def build(jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


TOOLS = {"Vamana": "vamana.cpp", "CompGT": "compgt.cpp", "PCACompress": "pcacompress.cpp"}
BIN = os.path.join(HERE, "bin")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_tools():
    """The host command-line tools (reference CLI flags and file formats) over
    libnbpg_amd.so: plain host C++ against the HIP runtime, found next to the
    library through the run path."""
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name, src in TOOLS.items():
        exe = os.path.join(BIN, name)
        cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
               "-I" + os.path.join(HERE, "..", "include"), "-I" + os.path.join(ROCM, "include"),
               os.path.join(HERE, "tools", src), "-o", exe + ".tmp",
               "-L" + HERE, "-lnbpg_amd", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
               "-Wl,-rpath,$ORIGIN/..:" + os.path.join(ROCM, "lib")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("tool %s failed:\n%s" % (name, r.stderr[-4000:]))
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build())
    print(build_tools())
    sys.exit(0)
