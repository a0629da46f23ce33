"""Host code under the sanitizers (CPU only).

* The oracle (oracle/oracle.c, with its pthreaded merge and row offsets)
  under gcc's AddressSanitizer + UBSan and, separately, ThreadSanitizer:
  tests/san_oracle.c drives every oracle function over seeded random cases.
* The library's host code (the C ABI's argument checks, the device pool's
  host state, the generators and the edgelist reader) under clang's
  AddressSanitizer + UBSan: a copy of liblpma is built with only its host
  code instrumented (hipcc -Xarch_host -fsanitize=...; GPU code is never
  sanitized) into lpma_amd/_build/san/, and tests/san_host.c runs against it
  without a GPU.

Any sanitizer report makes the program exit nonzero
(-fno-sanitize-recover=all, halt_on_error).
"""
import glob
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SAN_DIR = os.path.join(REPO, "lpma_amd", "_build", "san")
CSRC = os.path.join(REPO, "lpma_amd", "csrc")
HIP_SOURCES = ["sort.hip", "update.hip", "views.hip", "analytics.hip", "gen_dev.hip", "lpma.cpp", "server.cpp"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")

HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
CLANG_SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
             "-shared-libasan"]


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, (" ".join(cmd), r.stdout[-2000:], r.stderr[-4000:])
    return r


def gcc_or_skip():
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    return cc


def test_oracle_asan_ubsan(tmp_path):
    cc = gcc_or_skip()
    exe = str(tmp_path / "san_oracle")
    run([cc, "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-pthread", os.path.join(HERE, "san_oracle.c"), "-o", exe, "-lm"])
    r = run([exe, "8"], env=ENV, timeout=300)
    assert "san oracle ok" in r.stdout


def test_oracle_tsan(tmp_path):
    cc = gcc_or_skip()
    exe = str(tmp_path / "san_oracle_tsan")
    run([cc, "-std=gnu11", "-O1", "-g", "-fsanitize=thread", "-pthread", os.path.join(HERE, "san_oracle.c"), "-o",
         exe, "-lm"])
    r = run([exe, "8"], env=ENV, timeout=300)
    assert "san oracle ok" in r.stdout


def _clang():
    for c in (os.path.join(ROCM, "llvm", "bin", "clang"), shutil.which("clang")):
        if c and os.path.exists(c):
            return c
    pytest.skip("clang not available")


def _asan_runtime_dirs():
    return sorted(set(os.path.dirname(p) for p in
                      glob.glob(os.path.join(ROCM, "llvm", "lib", "clang", "*", "lib", "*", "libclang_rt.asan*.so"))))


def _stale(out, sources):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(s) > t for s in sources)


@pytest.fixture(scope="module")
def san_lib():
    """liblpma with its host code instrumented (rebuilt when a source is newer)."""
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    clang = _clang()
    clangxx = clang + "++" if os.path.exists(clang + "++") else clang
    os.makedirs(SAN_DIR, exist_ok=True)
    out = os.path.join(SAN_DIR, "liblpma_san.so")
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + [os.path.join(REPO, "include", "lpma.h")]
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES] + [os.path.join(CSRC, "gen.cpp")]
    if _stale(out, srcs + headers + [os.path.abspath(__file__)]):
        common = ["-O1", "-g", "-std=c++17", "-fPIC", "-I" + os.path.join(REPO, "include")]

        def compile_one(src):
            obj = os.path.join(SAN_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
            if src.endswith("gen.cpp"):
                cmd = [clangxx] + common + CLANG_SAN + ["-c", src, "-o", obj]
            else:
                cmd = ([hipcc, "-x", "hip", "--offload-arch=gfx950"] + common + HOST_SAN +
                       ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-c", src, "-o", obj])
            run(cmd)
            return obj

        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
            objs = list(ex.map(compile_one, srcs))
        run([clangxx, "-shared"] + CLANG_SAN + ["-o", out] + objs +
            ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lpthread", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return out, clang


@pytest.mark.timeout(900)
def test_host_code_asan_ubsan(san_lib, tmp_path):
    lib, clang = san_lib
    exe = str(tmp_path / "san_host")
    run([clang, "-std=gnu11", "-O1", "-g"] + CLANG_SAN +
        ["-I" + os.path.join(REPO, "include"), os.path.join(HERE, "san_host.c"), "-o", exe, lib,
         "-Wl,-rpath," + os.path.dirname(lib)])
    env = dict(ENV)
    dirs = _asan_runtime_dirs()
    env["LD_LIBRARY_PATH"] = ":".join(dirs + [os.path.join(ROCM, "lib"), env.get("LD_LIBRARY_PATH", "")])
    r = run([exe, str(tmp_path)], env=env, timeout=300)
    assert "san host ok" in r.stdout
