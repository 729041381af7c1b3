"""The reference-side shim of INTEGRATION.md compiles against the reference.

The shim is the C a maintainer adds to the reference tree to call this
library (src/fe/sw2/fe_sw2_hip.c in INTEGRATION.md).  This test extracts that
code block and runs `gcc -std=gnu99 -fsyntax-only -D_ADH_HIP` on it with the
reference's own headers and include/adh_sw2_hip.h.  The reference's
src/include/type.h ships with its struct includes commented out, so the test
writes, into a temporary directory only, a type.h with those includes
restored (plus the two super-model headers), an fnctn.h with fe_prototypes.h
enabled, and a global_header.h beside them so the quote-include search finds
these first.  Nothing of the reference is copied into this repository.
Skipped where the reference tree is absent.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")) or shutil.which("gcc") is None,
                                reason="reference tree or gcc absent")


def _shim_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## Shim"):]
    m = re.search(r"```c\n(.*?)```", sec, flags=re.S)
    return m.group(1)


def _include_shim(tmp):
    inc = os.path.join(REF, "include")
    t = open(os.path.join(inc, "type.h")).read()
    t = re.sub(r'^//(#include "[a-z0-9_]+\.h")', r"\1", t, flags=re.M)
    t += '\n#include "ssuperinterface.h"\n#include "ssuperModel.h"\n'
    open(os.path.join(tmp, "type.h"), "w").write(t)
    f = open(os.path.join(inc, "fnctn.h")).read().replace('//#include "fe_prototypes.h"', '#include "fe_prototypes.h"', 1)
    open(os.path.join(tmp, "fnctn.h"), "w").write(f)
    shutil.copy(os.path.join(inc, "global_header.h"), os.path.join(tmp, "global_header.h"))


def _gcc(tmp, src_path):
    incs = [tmp, os.path.join(ROOT, "include")] + [os.path.join(REF, d) for d in
                                                     ("include", "structs", "debug/include", "fe", "solver", "xdmf")]
    cmd = ["gcc", "-std=gnu99", "-fsyntax-only", "-D_ADH_HIP", "-Werror=implicit-function-declaration",
           "-Werror=incompatible-pointer-types", "-Werror=int-conversion"]
    cmd += [f"-I{d}" for d in incs] + [src_path]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_shim_compiles_against_the_reference_headers(tmp_path):
    tmp = str(tmp_path)
    _include_shim(tmp)
    src = os.path.join(tmp, "fe_sw2_hip.c")
    open(src, "w").write(_shim_source())
    r = _gcc(tmp, src)
    assert r.returncode == 0, r.stderr[-4000:]


def test_a_call_the_reference_lacks_fails_the_check(tmp_path):
    """The check is sharp: a shim that calls a function the reference does
    not declare (spv_realloc, which it never had) is rejected."""
    tmp = str(tmp_path)
    _include_shim(tmp)
    bad = _shim_source().replace("spv_alloc(row, sm->submodel[0].max_nsys_sq);", "spv_realloc(row, n);")
    assert "spv_realloc" in bad
    src = os.path.join(tmp, "fe_sw2_hip_bad.c")
    open(src, "w").write(bad)
    r = _gcc(tmp, src)
    assert r.returncode != 0 and "spv_realloc" in r.stderr
