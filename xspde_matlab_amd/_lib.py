"""ctypes binding of the C ABI declared in include/xspde_hip.h.

The library is loaded from the package directory (built in-tree by
``python -m xspde_matlab_amd.build``).  There is no fallback: if the shared
library is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxspde_hip.so")

XSH_MAX_OBS = 8
FP32, FP64 = 0, 1
RK4, MP = 0, 1
DERIV = {"zero": 0, "kubo": 1, "gp": 2}
INITIAL = {"const": 0, "sech": 1, "gauss": 2, "phase": 3, "coherent": 4}
OBSERVE = {"a": 0, "abs2": 1, "a_ave": 2, "abs2_ave": 3, "abs2_int": 4, "abs4_ave": 5, "corr": 6}
STATUS = {0: "XSH_OK", 1: "XSH_EINVAL", 2: "XSH_EUNSUPPORTED", 3: "XSH_ENOMEM", 4: "XSH_EDEVICE",
          5: "XSH_ESTATE"}

EXPORTS = ["xsh_version", "xsh_last_error", "xsh_create", "xsh_destroy", "xsh_query", "xsh_run_path",
           "xsh_run_ensemble", "xsh_synchronize", "xsh_read_averages", "xsh_probe_normals",
           "xsh_probe_propagate"]


class Desc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int), ("precision", C.c_int), ("dimensions", C.c_int),
        ("points", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("steps", C.c_int),
        ("checks", C.c_int), ("method", C.c_int), ("iterations", C.c_int),
        ("ensembles", C.c_int * 3), ("seed", C.c_int), ("noises", C.c_int),
        ("inrandoms", C.c_int), ("urandoms", C.c_int), ("boundaries", (C.c_int * 2) * 2),
        ("ranges", C.c_double * 3), ("origins", C.c_double * 3), ("deriv", C.c_int),
        ("deriv_par", C.c_double * 6), ("lin_l0", C.c_double * 2),
        ("lin_l1", (C.c_double * 2) * 2), ("lin_l2", (C.c_double * 2) * 2),
        ("has_linear", C.c_int), ("initial", C.c_int), ("init_par", C.c_double * 6),
        ("nobserve", C.c_int), ("observe", C.c_int * XSH_MAX_OBS), ("seq", C.c_int), ("inst", C.c_int),
    ]


class Info(C.Structure):
    _fields_ = [
        ("nsites", C.c_int), ("nchk", C.c_int), ("avlen", C.c_longlong),
        ("obs_offset", C.c_longlong * XSH_MAX_OBS), ("obs_sites", C.c_int * XSH_MAX_OBS),
        ("traj_steps", C.c_longlong), ("threads", C.c_int), ("traj_per_block", C.c_int),
    ]


class XshError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


_lib = None


def load():
    """Load libxspde_hip.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libxspde_hip.so is not built (run python -m xspde_matlab_amd.build)")
    lib = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    lib.xsh_version.restype = C.c_int
    lib.xsh_last_error.restype = C.c_char_p
    lib.xsh_create.argtypes = [C.POINTER(Desc), C.POINTER(C.c_void_p)]
    lib.xsh_destroy.argtypes = [C.c_void_p]
    lib.xsh_destroy.restype = None
    lib.xsh_query.argtypes = [C.c_void_p, C.POINTER(Info)]
    lib.xsh_run_path.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp, dp]
    lib.xsh_run_ensemble.argtypes = [C.c_void_p, C.c_int]
    lib.xsh_synchronize.argtypes = [C.c_void_p, dp]
    lib.xsh_read_averages.argtypes = [C.c_void_p, dp]
    lib.xsh_probe_normals.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
    lib.xsh_probe_propagate.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise XshError(rc, load().xsh_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Handle:
    """One problem on the GPU (owns device tables and buffers)."""

    def __init__(self, desc):
        lib = load()
        desc.struct_size = C.sizeof(Desc)
        h = C.c_void_p()
        check(lib.xsh_create(C.byref(desc), C.byref(h)))
        self._h = h
        self.desc = desc
        info = Info()
        check(lib.xsh_query(h, C.byref(info)))
        self.info = info

    def close(self):
        if getattr(self, "_h", None):
            load().xsh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_path(self, npar, j, nc, a0=None, want_final=False):
        E = self.desc.ensembles[0]
        ns = self.info.nsites
        av = np.zeros(2 * self.info.avlen)
        aout = np.zeros(2 * E * ns) if want_final else None
        a0c = None
        if a0 is not None:
            a0c = np.ascontiguousarray(np.stack([np.real(a0), np.imag(a0)], -1).ravel(), dtype=np.float64)
        check(load().xsh_run_path(self._h, npar, j, nc, _ptr(a0c), _ptr(aout), _ptr(av)))
        avc = av[0::2] + 1j * av[1::2]
        if want_final:
            return avc, aout[0::2] + 1j * aout[1::2]
        return avc

    def run_ensemble(self, npar):
        check(load().xsh_run_ensemble(self._h, npar))

    def synchronize(self):
        ms = C.c_double()
        check(load().xsh_synchronize(self._h, C.byref(ms)))
        return ms.value

    def read_averages(self):
        n = self.desc.ensembles[1] * self.info.nchk * self.info.avlen
        out = np.zeros(2 * n)
        check(load().xsh_read_averages(self._h, _ptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(self.desc.ensembles[1], self.info.nchk, self.info.avlen)

    def probe_normals(self, nsp, tag, draw, traj0, ntraj, count):
        out = np.zeros(ntraj * count)
        check(load().xsh_probe_normals(self._h, nsp, tag, draw, traj0, ntraj, count, _ptr(out)))
        return out.reshape(ntraj, count)

    def probe_propagate(self, nc, planes):
        planes = np.asarray(planes, dtype=complex)
        buf = np.ascontiguousarray(np.stack([planes.real, planes.imag], -1).ravel())
        check(load().xsh_probe_propagate(self._h, nc, planes.shape[0], _ptr(buf)))
        return (buf[0::2] + 1j * buf[1::2]).reshape(planes.shape)
