"""The C-ABI boundary: both shared libraries load without a GPU and export every entry point that
include/fgmi.h and include/fgmi_host.h declare, and the Python bindings cover exactly those."""
import ctypes as C
import os
import re
import subprocess

import pytest

from feellgood_amd import native as nv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^[A-Za-z_][\w\s\*]*?\b(fgh?_\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in DECL.finditer(text):
        line = text[m.start():text.index("(", m.start())]
        if line.lstrip().startswith(("typedef", "#", "struct")):
            continue
        names.add(m.group(1))
    return names


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.parametrize("header,lib,table", [("fgmi.h", nv.HIP_LIB, nv.HIP_SYMBOLS),
                                              ("fgmi_host.h", nv.HOST_LIB, nv.HOST_SYMBOLS)])
def test_library_exports_every_declared_entry_point(header, lib, table):
    names = declared(header)
    assert len(names) >= 10, names
    assert os.path.exists(lib), f"{lib} not built"
    missing = names - exported(lib)
    assert not missing, f"{os.path.basename(lib)} lacks {sorted(missing)}"
    assert names == set(table), (sorted(names - set(table)), sorted(set(table) - names))


def test_libraries_load_without_a_gpu():
    host = nv.host_lib()
    assert host.fgh_last_error() is not None or True
    hip = nv.hip_lib()
    opt = nv.Options()
    hip.fg_default_options(C.byref(opt))
    assert opt.tol == 1e-6 and opt.max_iter == 700 and opt.fmm_order == 8
    # errors come back as status codes, never as exits
    assert hip.fg_destroy(None) == -1
    assert hip.fg_solve(None, 1e-14, None) == -1
    assert hip.fg_fmm_info(None, None) == -1
