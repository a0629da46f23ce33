"""Build liblpma.so (HIP kernels for gfx950 + host C++ + C ABI) in-tree.

    python -m lpma_amd.build        # or lpma_amd.build.build()

Objects go to lpma_amd/_build/, the shared library to lpma_amd/liblpma.so.
hipcc cross-compiles for gfx950 without a GPU present.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "liblpma.so")
ARCH = os.environ.get("LPMA_OFFLOAD_ARCH", "gfx950")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# packed FP32 instructions are disabled for all device code
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas"]

SOURCES = [
    ("sort.hip", "hip"),
    ("update.hip", "hip"),
    ("views.hip", "hip"),
    ("analytics.hip", "hip"),
    ("gen_dev.hip", "hip"),
    ("lpma.cpp", "hip"),
    ("gen.cpp", "cxx"),
    ("server.cpp", "hip"),
]


def _cmd(src, kind):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(SRC, src)
    if kind == "hip":
        cmd = [_hipcc(), "-x", "hip", "--offload-arch=" + ARCH] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + COMMON + ["-c", path, "-o", obj]
    return cmd, obj


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    cmds = [_cmd(s, k) for s, k in SOURCES]

    def run(c):
        r = subprocess.run(c[0], capture_output=True, text=True)
        return c, r

    objs = []
    errors = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for (cmd, obj), r in ex.map(run, cmds):
            if verbose:
                print(" ".join(cmd))
            out = (r.stdout + r.stderr).strip()
            # the host pass does not know the packed-fp32 target feature
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if r.returncode != 0:
                errors.append("\n".join(lines))
            elif verbose and lines:
                print("\n".join(lines))
            objs.append(obj)
    if errors:
        raise RuntimeError("liblpma build failed:\n" + "\n\n".join(errors))
    tmp = OUT + ".tmp"
    link = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("liblpma link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
