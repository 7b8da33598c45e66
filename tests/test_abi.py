"""The C-ABI library (CPU, no GPU compute): it loads, exports every entry point
include/mkde.h declares, and the product never touches the oracle."""
import os
import re
import subprocess

import pytest

import mgpukde_tiled_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = M.load_library()
    declared = M.declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []
    # the Python mirror binds every declared entry point
    assert sorted(M.SIGNATURES) == declared


def test_exported_dynamic_symbols_are_c_names():
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(mkde_\w+)$", out, re.M))
    assert set(M.declared_symbols()) <= exported


def test_header_has_no_torch_or_gpu_runtime_types():
    txt = open(M.HEADER_PATH).read()
    code = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)  # declarations only, comments stripped
    for bad in ("torch", "cuda", "hipStream", "hipEvent", "hip_runtime", "at::", "Tensor", "c10"):
        assert bad not in code, bad


def test_product_does_not_reference_the_oracle():
    pkg = os.path.join(ROOT, "mgpukde_tiled_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower(), f
                assert "mkde_oracle" not in txt


def test_version_string():
    assert "gfx950" in M.version()


def test_job_footprint_bound_shape():
    # linear in points and cells; the leave-one-out options add their candidate arrays
    base = M.job_footprint(0, 0, 0)
    assert base > 0
    assert M.job_footprint(1000, 100, 100) < M.job_footprint(2000, 100, 100) < M.job_footprint(2000, 200, 200)
    for hopt, opt in ((1, False), (M.MKDE_H_SEARCH, False), (2, True)):
        assert M.job_footprint(10000, 64, 64, hopt, opt, 16) > M.job_footprint(10000, 64, 64, 0)
    assert M.job_footprint(10000, 64, 64, 2, False) == M.job_footprint(10000, 64, 64, 0)
    assert M.job_footprint(-1, 1, 1) == -1


def test_server_without_a_device_fails_loudly():
    try:
        srv = M.Server(slots=2)
    except M.MkdeError as e:
        assert e.code == 2 and "device" in str(e)   # MKDE_E_HIP, no silent CPU fallback
    else:
        srv.close()
        pytest.skip("a HIP device is present")
