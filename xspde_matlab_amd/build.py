"""Build recipe of libxspde_hip.so (gfx950) - in-tree, no JIT cache.

    python -m xspde_matlab_amd.build            # compile + link
    python -m xspde_matlab_amd.build --resource # also print VGPR/LDS/scratch use
    python -m xspde_matlab_amd.build --changed  # recompile only units older than
                                                # their source or any header

Each translation unit is compiled by hipcc in parallel, then linked into
``xspde_matlab_amd/libxspde_hip.so``.  Packed FP32 instructions are
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libxspde_hip.so")
# longest translation units first (they run in parallel)
SOURCES = ["xsh_k_stream_mat.hip", "xsh_k_mat_pad.hip", "xsh_k_kn64.hip", "xsh_k_f64.hip", "xsh_k_kn.hip", "xsh_k_mat_f64.hip", "xsh_k_kn2d.hip", "xsh_k_f32_1d.hip", "xsh_k_mat_f32_1d.hip", "xsh_k_f32_var.hip",
           "xsh_k_mat_f32_1d256.hip", "xsh_k_mat_f32.hip", "xsh_k_f32_var2d.hip", "xsh_k_f32_2d.hip",
           "xsh_k_mat_f64_1d256.hip", "xsh_k_stream.hip", "xsh_capi.hip", "xsh_mlp.hip", "xsh_k_sde.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wno-unused-result"]


def _stale(src, obj):
    """The object is older than its source or than any header of the library."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src), os.path.join(HERE, "..", "include", "xspde_hip.h")]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".inc"))]
    return os.path.getmtime(obj) < max(os.path.getmtime(d) for d in deps)


def _compile(src, objdir, extra, changed_only=False):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    if changed_only and not _stale(src, obj):
        return obj, ""
    cmd = [HIPCC] + FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
    return obj, out


def build(resource=False, verbose=False, changed_only=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if resource else []
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda s: _compile(s, objdir, extra, changed_only and not resource), SOURCES))
    if resource or verbose:
        for _, out in results:
            if out:
                print(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(resource="--resource" in sys.argv, verbose="-v" in sys.argv, changed_only="--changed" in sys.argv))
