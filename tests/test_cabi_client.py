"""A compiled C caller of include/lpma.h (tests/cabi_client.c).

CPU: the client compiles against the header (its static asserts pin the
struct sizes the ABI fixes), the header's struct layout equals the ctypes
mirror in lpma_amd field for field, and dlopen/dlsym resolve every entry
point the client binds.  GPU: the client runs an update stream through the
library on device 0 and checks every batch against its own set reference.
"""
import ctypes
import os
import shutil
import subprocess

import pytest

import lpma_amd as L

HERE = os.path.dirname(os.path.abspath(__file__))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def client(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    out = str(tmp_path_factory.mktemp("cabi") / "cabi_client")
    cmd = [cc, "-std=gnu11", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROCM, "include"),
           os.path.join(HERE, "cabi_client.c"), "-o", out, "-ldl", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_header_layout_matches_ctypes_mirror(client):
    r = subprocess.run([client, "layout"], capture_output=True, text=True, check=True)
    c = {}
    for line in r.stdout.splitlines():
        name, val = line.split()
        c[name] = int(val)
    mirror = {"lpma_config_t": L.Config, "lpma_view_t": L.View, "lpma_stats_t": L.Stats,
              "lpma_profile_t": L.Profile, "lpma_server_config_t": L.ServerConfig,
              "lpma_server_result_t": L.ServerResult, "lpma_server_stats_t": L.ServerStats}
    checked = 0
    for cname, st in mirror.items():
        assert c[cname] == ctypes.sizeof(st), cname
        fields = [f[0] for f in st._fields_]
        cfields = [k.split(".", 1)[1] for k in c if k.startswith(cname + ".")]
        assert cfields == fields, (cname, cfields, fields)
        for f in fields:
            assert c[f"{cname}.{f}"] == getattr(st, f).offset, (cname, f)
            checked += 1
    assert checked >= 70


def test_dlopen_resolves_entry_points(client):
    r = subprocess.run([client, "symbols", L.LIB_PATH], capture_output=True, text=True)
    assert r.returncode == 0 and "symbols ok" in r.stdout, r.stderr


@pytest.mark.gpu
def test_c_client_update_stream(client):
    r = subprocess.run([client, "stream", L.LIB_PATH], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "cabi stream ok" in r.stdout, (r.returncode, r.stderr)
