"""The C-ABI library loads on a CPU-only host and exports every function
include/*.h declares (no compute calls: those need a GPU)."""
import ctypes
import glob
import os
import re

from mac_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(mac_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    assert {"mac_create", "mac_project", "mac_solve_csr", "mac_upload_bricks"} <= declared()


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    # every declared entry point has a ctypes signature in the binding
    assert declared() <= set(L.SIGNATURES)


def test_no_gpu_is_an_error_not_a_fallback():
    import subprocess
    import sys
    code = ("import mac_amd.lib as L\n"
            "try:\n    L.Projection()\nexcept L.MacError as e:\n    print('ERR', e)\nelse:\n    print('OK')\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert "ERR" in out.stdout and "no HIP device" in out.stdout, out.stdout + out.stderr
