"""The drop-in boundary without a GPU: the C-ABI library loads, exports every function that
include/hussar_hip.h declares, the host structs have the header's exact layout, and the
default knobs are the reference's (path.h:153-169)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from hussar_amd import _lib, types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hussar_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hsr_\w+)\s*\(", src)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 24
    L = _lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding declares a signature for each of them
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version_and_error_string():
    L = _lib.lib()
    assert L.hsr_abi_version() == _lib.ABI_VERSION == 5
    assert isinstance(L.hsr_last_error(), bytes)


STRUCTS = {
    "hsr_rf_config": T.c_rf_config,
    "hsr_frame_config": T.c_frame_config,
    "hsr_antenna": T.c_antenna,
    "hsr_scene": T.c_scene,
    "hsr_knobs": T.c_knobs,
    "hsr_run_stats": T.c_run_stats,
    "hsr_device_info": T.c_device_info,
    "hsr_server_config": T.c_server_config,
    "hsr_request": T.c_request,
    "hsr_request_info": T.c_request_info,
    "hsr_server_stats": T.c_server_stats,
}


def test_struct_layouts_match_header():
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, ct in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", "-Wall", c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for cname, ct in STRUCTS.items():
        assert int(got[f"{cname} size"]) == C.sizeof(ct), cname
        for f, _ in ct._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(ct, f).offset, (cname, f)


def test_knobs_default_are_the_reference_defaults():
    k = T.c_knobs()
    _lib.lib().hsr_knobs_default(C.byref(k))
    # path.h:153-169
    assert (k.onlyIndirect, k.maxDepth, k.useGeometricalOptics, k.poDiffraction, k.doGuiding,
            k.clearBeforeIteration, k.correctPhase, k.filterGuiding, k.doFiltering, k.filteringSphere) == \
        (1, 10, 1, 0, 1, 1, 0, 1, 1, 1)
    assert (k.filteringMin, k.filteringMax, k.filteringRadius) == (600.0, 900.0, 160.0)
    assert k.primaryOrder == 0 and k.produceDebugImage == 0
    # the Python mirror carries the same defaults
    assert T.Knobs().to_c().maxDepth == 10 and T.Knobs().filteringRadius == 160.0


def test_invalid_arguments_are_rejected_without_a_device():
    L = _lib.lib()
    # null output handle and negative sizes fail before any device call
    assert L.hsr_hip_create(None, -1, None, 0, 0, None) == 1
    assert L.hsr_hip_run(None, None, None, 0, 0, 1.0, 0, None, None) == 1
    assert L.hsr_hip_configure_frame(None, None) == 1
    assert b"null" in L.hsr_last_error()
    assert L.hsr_server_create(None, None) == 1
    assert L.hsr_server_submit(None, None, None) == 1
    assert L.hsr_server_get_stats(None, None) == 1
    L.hsr_server_destroy(None)


def test_python_api_fails_loudly_without_the_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()
