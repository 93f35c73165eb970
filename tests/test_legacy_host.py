"""The legacy-symbol library (libhrrt_osem3d_legacy.so) without a GPU: it exports the
reference projector's C++ entry points with the reference's exact signatures
(hrrt_osem3d.h:197-211), include/hrrt_legacy.h declares the same, and every call fails with
a message, not a crash, before the library is set up."""
import os
import re
import subprocess
import sys

from conftest import ROOT
from legacy_util import DEMANGLED, LEGACY_PATH


def _exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", LEGACY_PATH], check=True, capture_output=True,
                         text=True).stdout
    return {line.split(" T ", 1)[1] for line in out.splitlines() if " T " in line}


def test_exports_the_reference_signatures():
    assert _exported() == set(DEMANGLED.values()) | {"hrrt_legacy_set_params", "hrrt_legacy_context",
                                                      "hrrt_legacy_set_norm_zero_mask"}


def _normalise(decl):
    name, args = re.match(r"\s*(?:int|void)\s+(\w+)\((.*)\)", decl).groups()
    types = []
    for a in filter(None, (x.strip() for x in args.split(","))):
        types.append(re.sub(r"\s+", "", re.sub(r"\w+$", "", a)) if re.search(r"\*|\s", a) else a)
    return f"{name}({', '.join(types)})"


def test_header_declares_the_same():
    text = open(os.path.join(ROOT, "include", "hrrt_legacy.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    cxx = text.split("#ifdef __cplusplus\n}", 1)[1]
    decls = [_normalise(d) for d in re.findall(r"^(?:int|void) \w+\([^;]*\);", cxx, flags=re.M)]
    assert sorted(decls) == sorted(DEMANGLED.values())
    assert "int hrrt_legacy_set_params(const hrrt_params *params);" in text


def test_calls_before_setup_fail_with_a_message():
    code = f"""
import sys
sys.path[:0] = [{ROOT!r}, {os.path.join(ROOT, 'tests')!r}]
from legacy_util import Legacy
L = Legacy()
assert L.dependencies(256, 288, 0) == 0
assert L.forward_proj3d_thread1(None, None, 0, 1, None, None) == 0
assert L.back_proj3d_thread1(None, None, 0, 1, None, None) == 0
assert L.rotate_image_buf(None, None, 1, None) == 0
assert L.L.hrrt_legacy_context() is None
assert L.L.hrrt_legacy_set_norm_zero_mask(None) == 0
L.free_dependencies()
print("ok")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
    assert "dependencies: hrrt_legacy_set_params() has not been called" in r.stderr
    assert "forward_proj3d_thread1: null argument" in r.stderr
