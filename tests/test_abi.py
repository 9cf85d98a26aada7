"""CPU checks of the drop-in boundary: the HIP library builds for gfx950, loads,
and exports every function include/ptn_hip.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

import ptn_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            with open(os.path.join(ROOT, "include", fn)) as f:
                text = f.read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(ptn_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "ptn_cpd" in names and "ptn_mttkrp" in names and "ptn_nesterov_mnls" in names
    assert set(names) == set(ptn_hip.EXPORTED)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(ptn_hip.LIB_PATH), "build first: __graft_entry__.build()"
    lib = ctypes.CDLL(ptn_hip.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_options_defaults_match_reference_defaults():
    # DefaultValues, PARTENSOR_basic.hpp:184-198; inner cap NesterovMNLS.hpp:187
    o = ptn_hip.default_options()
    assert o.max_iter == 20 and o.threshold_error == 1e-3
    assert o.nesterov_delta_1 == 1e-2 and o.nesterov_delta_2 == 1e-2
    assert o.accel_coeff == 3 and o.accel_fail == 0 and o.acceleration == 1 and o.normalization == 1
    assert all(c == ptn_hip.NONNEGATIVITY for c in o.constraints)
    assert o.max_inner == 50


def test_struct_layouts_match_header(tmp_path):
    """sizeof and field offsets of every C ABI struct, as the host compiler lays them out
    from the header, equal the ctypes mirrors (catches a field added on one side only)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    mirrors = {"ptn_options": ptn_hip.Options, "ptn_result": ptn_hip.Result, "ptn_sweep_log": ptn_hip.SweepLog,
               "ptn_batch_log": ptn_hip.BatchLog, "ptn_batch_report": ptn_hip.BatchReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ptn_hip.h"', "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for cname, cls in mirrors.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got["%s.%s" % (cname, f[0])]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_invalid_handle_is_an_error_not_a_crash():
    lib = ptn_hip.lib()
    assert lib.ptn_abi_version() == 1
    assert lib.ptn_create(None, 0) != 0
