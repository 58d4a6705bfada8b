"""Recipe of the host-sanitizer driver (test infrastructure, CPU only).

Compiles the C-ABI unit xsh_capi.hip with AddressSanitizer and
UndefinedBehaviorSanitizer on its host code (device code as usual), links it
with the library's other objects from xspde_matlab_amd/_build and
tests/sanitize/plan_driver.cpp into _build/plan_driver_asan.

    python tests/sanitize/build_asan.py
"""
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BUILD = os.path.join(ROOT, "xspde_matlab_amd", "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
EXE = os.path.join(BUILD, "plan_driver_asan")
# host-only sanitizers: each -fsanitize= directly after -Xarch_host
SAN = shlex.split("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
                  "-Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer")
FLAGS = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-munsafe-fp-atomics", "-Wno-unused-result"]


def build():
    os.makedirs(BUILD, exist_ok=True)
    capi = os.path.join(BUILD, "xsh_capi_asan.o")
    drv = os.path.join(BUILD, "plan_driver_asan.o")
    src = os.path.join(ROOT, "xspde_matlab_amd", "csrc", "xsh_capi.hip")
    for cmd in ([HIPCC] + FLAGS + SAN + ["-c", src, "-o", capi],
                [HIPCC] + FLAGS + SAN + ["-c", os.path.join(HERE, "plan_driver.cpp"), "-o", drv]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    objs = [os.path.join(BUILD, f) for f in sorted(os.listdir(BUILD))
            if f.endswith(".o") and f not in ("xsh_capi.o", "xsh_capi_asan.o", "plan_driver_asan.o")]
    r = subprocess.run([HIPCC, "--offload-arch=gfx950"] + SAN + ["-o", EXE, drv, capi] + objs,
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    return EXE


if __name__ == "__main__":
    print(build())
    sys.exit(0)
