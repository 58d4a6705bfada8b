"""The C-ABI unit's host code under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer (CPU, no device): tests/sanitize/plan_driver.cpp
runs xsh_plan - descriptor validation, lattice derivation, kernel choice and
every host-built table (propagators, twiddles, dense transform matrices and
their matrix-core operand copies, boundary patches and boundfun tables, k-space
filters, spectrum flags) - over 176 descriptors of every lattice family and
refusal path.  Any sanitizer finding aborts the driver."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "xspde_matlab_amd", "_build")
sys.path.insert(0, os.path.join(ROOT, "tests", "sanitize"))


def _stale(exe):
    if not os.path.exists(exe):
        return True
    srcs = [os.path.join(ROOT, "xspde_matlab_amd", "csrc", "xsh_capi.hip"),
            os.path.join(ROOT, "include", "xspde_hip.h"),
            os.path.join(ROOT, "tests", "sanitize", "plan_driver.cpp")]
    return os.path.getmtime(exe) < max(os.path.getmtime(s) for s in srcs)


def test_host_code_under_sanitizers():
    if not os.path.isdir(BUILD) or not any(f.endswith(".o") for f in os.listdir(BUILD)):
        pytest.skip("library objects not built (python -m xspde_matlab_amd.build)")
    import build_asan
    exe = build_asan.EXE
    if _stale(exe):
        build_asan.build()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-15:])
    assert r.returncode == 0, tail
    assert "176 cases, 0 failures" in r.stdout, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail


def test_plan_reports_the_handle_layout():
    """xsh_plan gives the sizes xsh_create's handle would report, with no device."""
    from xspde_matlab_amd import _lib, configs
    from xspde_matlab_amd.xsim import make_desc, preferences
    info = _lib.plan(make_desc(preferences(configs.cgl2d("fp32"))))
    assert (info.kernel_kind, info.threads, info.traj_per_block) == (0, 1024, 1)
    assert info.traj_steps == 6202368 and info.nchk == 2 and info.nsites == 16384
    assert info.obs_sites[0] == 1 and info.obs_sites[1] == 16384
    info = _lib.plan(make_desc(preferences(configs.cgl2d("fp64"))))
    assert info.kernel_kind == 1                       # streaming kernel: FP64 fields exceed the CU
    with pytest.raises(_lib.XshError, match="SDEs"):
        p = configs.shrink(configs.nls1d(), E=16, e2=1, points=3)
        p["transforms"] = [0, 1]
        _lib.plan(make_desc(preferences(p)))
