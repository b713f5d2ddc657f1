"""The C-ABI library loads (no GPU needed) and exports every function include/*.h declares."""
import ctypes
import glob
import os
import re

from mrtrix3_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef struct \w+ \{.*?\} \w+;", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(mr\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "mrcsd_run" in names and "mrcsd_create" in names and len(names) >= 12


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_strings():
    lib = _lib.load()
    assert b"gfx950" in lib.mrcsd_version()
    assert isinstance(lib.mrcsd_last_error(), bytes)


def test_null_arguments_are_errors_not_crashes():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.mrcsd_shared_create(None, 0, None, 0, None, ctypes.byref(h)) == -1
    assert b"null" in lib.mrcsd_last_error()
    assert lib.mrcsd_create(None, ctypes.byref(h)) == -1
    assert lib.mrcsd_run(None, None, 0, None, 0, None, None, None, None) == -1


def test_one_hip_runtime_whatever_the_import_order():
    # torch and ROCm each ship a libamdhip64 under the same soname; the package must never map both
    # (the second runtime to initialise finds no device, and torch pointers are foreign to it)
    import subprocess
    import sys
    code = ("import re\n"
            "from mrtrix3_amd import _lib\n"
            "_lib.load()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read()\n"
            "print(len(set(re.findall(r'\\S*/libamdhip64\\S*', maps))))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "1"
