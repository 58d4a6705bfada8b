"""xsh_run_path_noalloc against xsh_run_path: the same path with the caller's
device buffers and the library's static scratch array instead of handle
allocations.  Each case runs two calls in a row, the second with a smaller
input, and compares both with xsh_run_path (final fields, every average):
register-resident FP32 and FP64 kernels, the FP64 128x128 streaming kernel
(field planes in the scratch, persistent grid capped by it) and k-space noise
planes.  Paths whose scratch would exceed XSH_SCRATCH_BYTES are refused before
any launch, and the calls leave the device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest
import torch

from xspde_matlab_amd import Problem, _lib
from xspde_matlab_amd.configs import cgl2d, nls1d, shrink

pytestmark = pytest.mark.gpu


def _vp(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _interleave(a):
    a = np.asarray(a, complex).reshape(-1)
    return np.ascontiguousarray(np.stack([a.real, a.imag], -1).reshape(-1))


def _noalloc(prob, j, nc, a0=None):
    h = prob.handle
    E, ns, avlen = h.desc.ensembles[0], h.info.nsites, h.info.avlen
    av = torch.full((2 * avlen,), float("nan"), dtype=torch.float64, device="cuda")
    aout = torch.zeros(2 * E * ns, dtype=torch.float64, device="cuda")
    a0d = torch.from_numpy(_interleave(a0)).cuda() if a0 is not None else None
    torch.cuda.synchronize()
    lib = _lib.load()
    _lib.check(lib.xsh_run_path_noalloc(h._h, 1, j, nc, _vp(a0d), _vp(aout), _vp(av), None))
    _lib.check(lib.xsh_synchronize(h._h, None))
    a = av.cpu().numpy()
    f = aout.cpu().numpy()
    return a[0::2] + 1j * a[1::2], f[0::2] + 1j * f[1::2]


def _check(p, j=1, nc=1, a0=None, field_exact=True):
    prob = Problem(p)
    av_n, a_n = _noalloc(prob, j, nc, a0)
    av_r, a_r = prob.handle.run_path(1, j, nc, a0=a0, want_final=True)
    kind = prob.info.kernel_kind
    prob.close()
    scale = np.abs(av_r).max()
    err_av = np.abs(av_n - av_r).max() / scale
    err_a = np.abs(a_n - a_r).max() / np.abs(a_r).max()
    print("%s kernel %d: averages %.2e, fields %.2e" % (p.get("name"), kind, err_av, err_a), flush=True)
    # the averages are sums of FP64 atomics in either order
    assert err_av < 1e-12
    if field_exact:
        assert np.array_equal(a_n, a_r)
    else:
        assert err_a < 1e-12


def test_register_kernels_two_calls():
    """FP32 64x64 with per-site averages, then a smaller FP64 1D lattice."""
    p = shrink(cgl2d("fp32", n=64), E=4, e2=2, points=4, steps=2)
    p["gpu"]["observe"] = ["abs2", "abs2_ave", "corr"]
    _check(p, j=2, nc=1)
    p = shrink(nls1d("fp64", nx=64), E=3, e2=1, points=3, steps=2)
    _check(p, j=1, nc=2)


def test_streaming_kernel_two_calls():
    """FP64 128x128: the field planes of the streaming kernel in the scratch."""
    p = shrink(cgl2d("fp64", n=128), E=3, e2=1, points=3, steps=1)
    p["gpu"]["observe"] = ["abs2_ave", "corr"]
    _check(p, nc=1)
    p = shrink(cgl2d("fp64", n=128), E=2, e2=1, points=2, steps=1)
    p["gpu"]["observe"] = ["abs2_ave"]
    _check(p, nc=2)


def test_given_initial_fields_and_knoise():
    """a0 from the caller's device buffer; k-space noise planes in the scratch."""
    p = shrink(nls1d("fp64", nx=128), E=20, e2=1, points=3, steps=2)
    p.update(noises=0, knoises=2)
    rng = np.random.default_rng(4)
    a0 = rng.standard_normal((20, 128)) + 1j * rng.standard_normal((20, 128))
    _check(p, a0=a0)
    p = shrink(nls1d("fp64", nx=64), E=5, e2=1, points=2, steps=1)
    p.update(noises=0, knoises=1)
    _check(p, nc=1)


def test_scratch_limit_and_no_allocation():
    big = shrink(cgl2d("fp32", n=128), E=2, e2=1, points=101, steps=1)
    big["gpu"]["observe"] = ["abs2", "a", "corr", "abs2k"]
    prob = Problem(big)
    av = torch.zeros(2 * prob.info.avlen, dtype=torch.float64, device="cuda")
    rc = _lib.load().xsh_run_path_noalloc(prob.handle._h, 1, 1, 1, None, None, _vp(av), None)
    msg = _lib.load().xsh_last_error().decode()
    prob.close()
    print("refused:", msg)
    assert rc == 1 and "scratch" in msg
    p = shrink(cgl2d("fp32", n=32), E=3, e2=1, points=3, steps=2)
    prob = Problem(p)
    _noalloc(prob, 1, 1)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        _noalloc(prob, 1, 1)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    prob.close()
    print("free device memory before %d after %d" % (free0, free1))
    assert free1 == free0
