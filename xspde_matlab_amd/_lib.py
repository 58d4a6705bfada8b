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
XSH_MAX_MODES = 32
XSH_MAX_FIELDS = 4
FP32, FP64 = 0, 1
RK4, MP = 0, 1
DERIV = {"zero": 0, "kubo": 1, "gp": 2, "matrix": 3}
INITIAL = {"const": 0, "sech": 1, "gauss": 2, "phase": 3, "coherent": 4}
OBSERVE = {"a": 0, "abs2": 1, "a_ave": 2, "abs2_ave": 3, "abs2_int": 4, "abs4_ave": 5, "corr": 6, "lowk": 7, "abs2k": 8}
STATUS = {0: "XSH_OK", 1: "XSH_EINVAL", 2: "XSH_EUNSUPPORTED", 3: "XSH_ENOMEM", 4: "XSH_EDEVICE",
          5: "XSH_ESTATE"}

EXPORTS = ["xsh_version", "xsh_last_error", "xsh_create", "xsh_plan", "xsh_destroy", "xsh_query", "xsh_run_path",
           "xsh_run_path_noalloc",
           "xsh_run_ensemble", "xsh_synchronize", "xsh_read_averages", "xsh_read_trajectories", "xsh_read_raw",
           "xsh_nonfinite", "xsh_batch_nonfinite",
           "xsh_probe_normals", "xsh_probe_propagate", "xsh_mlp_create", "xsh_mlp_destroy",
           "xsh_mlp_last_error", "xsh_mlp_get_params", "xsh_mlp_forward", "xsh_mlp_loss_grad",
           "xsh_mlp_train", "xsh_mlp_set_translations", "xsh_mlp_gaussian_moments",
           "xsh_batch_create", "xsh_batch_destroy", "xsh_batch_run", "xsh_batch_synchronize",
           "xsh_batch_read_averages", "xsh_batch_query", "xsh_batch_plan"]
XSH_BATCH_AUTOTUNE = 1
XSH_BATCH_SERIAL = 2


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
        ("traj_out", C.c_int), ("nmodes", C.c_int), ("modes", (C.c_int * 2) * XSH_MAX_MODES),
        ("boundval", ((C.c_double * 2) * 2) * 2), ("raw", C.c_int),
        ("boundtab", C.POINTER(C.c_double)), ("boundlevels", C.c_int),
        ("knoises", C.c_int), ("krandoms", C.c_int),
        ("nfilter", C.POINTER(C.c_double)), ("rfilter", C.POINTER(C.c_double)),
        ("tspec", C.c_int * XSH_MAX_OBS),
        ("fields", C.c_int),
        ("fmat", ((C.c_double * 2) * XSH_MAX_FIELDS) * XSH_MAX_FIELDS),
        ("fnl", (C.c_double * 2) * XSH_MAX_FIELDS),
        ("fnoise", (C.c_double * 4) * XSH_MAX_FIELDS),
        ("fl0", (C.c_double * 2) * XSH_MAX_FIELDS),
        ("fc0", (C.c_double * 2) * XSH_MAX_FIELDS),
        ("obs_field", C.c_int * XSH_MAX_OBS),
    ]

    def set_filters(self, nfilter=None, rfilter=None):
        """Attach k-space filter tables [components][nx*ny] complex (FFT order)."""
        if nfilter is not None:
            self._nfilter = np.ascontiguousarray(nfilter, dtype=np.complex128)
            self.knoises = self._nfilter.shape[0]
            self.nfilter = self._nfilter.ctypes.data_as(C.POINTER(C.c_double))
        if rfilter is not None:
            self._rfilter = np.ascontiguousarray(rfilter, dtype=np.complex128)
            self.krandoms = self._rfilter.shape[0]
            self.rfilter = self._rfilter.ctypes.data_as(C.POINTER(C.c_double))

    def set_boundary_table(self, tab):
        """Attach a boundary-value table [levels][2][2][nb] complex (kept alive
        with the descriptor until xsh_create has copied it)."""
        tab = np.ascontiguousarray(tab, dtype=np.complex128)
        self._boundtab = tab
        self.boundtab = tab.ctypes.data_as(C.POINTER(C.c_double))
        self.boundlevels = tab.shape[0]


class Info(C.Structure):
    _fields_ = [
        ("nsites", C.c_int), ("nchk", C.c_int), ("avlen", C.c_longlong),
        ("obs_offset", C.c_longlong * XSH_MAX_OBS), ("obs_sites", C.c_int * XSH_MAX_OBS),
        ("traj_steps", C.c_longlong), ("threads", C.c_int), ("traj_per_block", C.c_int),
        ("kernel_kind", C.c_int), ("ensembles", C.c_int * 3), ("points", C.c_int),
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
    lib.xsh_plan.argtypes = [C.POINTER(Desc), C.POINTER(Info)]
    lib.xsh_run_path.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp, dp]
    lib.xsh_run_path_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    lib.xsh_run_ensemble.argtypes = [C.c_void_p, C.c_int]
    lib.xsh_synchronize.argtypes = [C.c_void_p, dp]
    lib.xsh_read_averages.argtypes = [C.c_void_p, dp]
    lib.xsh_probe_normals.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
    lib.xsh_probe_propagate.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
    fp = C.POINTER(C.c_float)
    lib.xsh_read_trajectories.argtypes = [C.c_void_p, dp, fp]
    lib.xsh_read_raw.argtypes = [C.c_void_p, dp]
    lib.xsh_nonfinite.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    lib.xsh_batch_nonfinite.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    lib.xsh_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, fp, C.POINTER(C.c_void_p)]
    lib.xsh_mlp_destroy.argtypes = [C.c_void_p]
    lib.xsh_mlp_destroy.restype = None
    lib.xsh_mlp_last_error.restype = C.c_char_p
    lib.xsh_mlp_get_params.argtypes = [C.c_void_p, fp]
    lib.xsh_mlp_forward.argtypes = [C.c_void_p, C.c_longlong, fp, fp]
    lib.xsh_mlp_loss_grad.argtypes = [C.c_void_p, C.c_int, fp, fp, dp, fp]
    lib.xsh_mlp_train.argtypes = [C.c_void_p, C.c_int, fp, fp, C.c_int, C.c_int, C.POINTER(C.c_int),
                                  C.c_float, dp]
    lib.xsh_mlp_gaussian_moments.argtypes = [C.c_void_p, C.c_longlong, C.c_int, dp, dp]
    lib.xsh_mlp_set_translations.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int]
    lib.xsh_batch_create.argtypes = [C.c_int, C.POINTER(Desc), C.c_int, C.POINTER(C.c_void_p)]
    lib.xsh_batch_destroy.argtypes = [C.c_void_p]
    lib.xsh_batch_destroy.restype = None
    lib.xsh_batch_run.argtypes = [C.c_void_p, C.c_int]
    lib.xsh_batch_synchronize.argtypes = [C.c_void_p, dp]
    lib.xsh_batch_read_averages.argtypes = [C.c_void_p, C.c_int, dp]
    lib.xsh_batch_query.argtypes = [C.c_void_p, C.c_int, C.POINTER(Info)]
    lib.xsh_batch_plan.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise XshError(rc, load().xsh_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def plan(desc):
    """xsh_plan: the Info xsh_create would produce, with no device (raises
    XshError with XSH_EUNSUPPORTED for projects that stay on the CPU path)."""
    desc.struct_size = C.sizeof(Desc)
    info = Info()
    check(load().xsh_plan(C.byref(desc), C.byref(info)))
    return info


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
            if np.size(a0) != E * ns:
                raise ValueError("a0 holds %d values; the path needs ensembles(1) x sites = %d" % (np.size(a0), E * ns))
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

    def nonfinite(self):
        """True if a field value at a stored point of the last run was NaN or infinite."""
        f = C.c_int()
        check(load().xsh_nonfinite(self._h, C.byref(f)))
        return bool(f.value)

    def read_averages(self):
        n = self.desc.ensembles[1] * self.info.nchk * self.info.avlen
        out = np.zeros(2 * n)
        check(load().xsh_read_averages(self._h, _ptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(self.desc.ensembles[1], self.info.nchk, self.info.avlen)

    def read_trajectories(self):
        """Per-trajectory lattice means [e2, nchk, E, points, nobs] and noise modes [e2, nchk, E, 2*nmodes]."""
        d = self.desc
        e2, nchk, E = d.ensembles[1], self.info.nchk, d.ensembles[0]
        obs = np.zeros(e2 * nchk * E * d.points * d.nobserve) if d.traj_out else None
        feat = np.zeros(e2 * nchk * E * 2 * d.nmodes, dtype=np.float32) if d.nmodes else None
        fptr = feat.ctypes.data_as(C.POINTER(C.c_float)) if feat is not None else None
        check(load().xsh_read_trajectories(self._h, _ptr(obs), fptr))
        if obs is not None:
            obs = obs.reshape(e2, nchk, E, d.points, d.nobserve)
        if feat is not None:
            feat = feat.reshape(e2, nchk, E, 2 * d.nmodes)
        return obs, feat

    def read_raw(self):
        """Raw fields [e2, nchk, points, E, nx*ny] complex of the last run_ensemble."""
        d = self.desc
        e2, nchk, E = d.ensembles[1], self.info.nchk, d.ensembles[0]
        ns = self.info.nsites
        out = np.zeros(2 * e2 * nchk * d.points * E * ns)
        check(load().xsh_read_raw(self._h, _ptr(out)))
        # per path p.d.raw = [1, points, nx, ny, E] column-major: C order [E, site, points]
        r = (out[0::2] + 1j * out[1::2]).reshape(e2, nchk, E, ns, d.points)
        return np.ascontiguousarray(np.transpose(r, (0, 1, 4, 2, 3)))

    def probe_normals(self, nsp, tag, draw, traj0, ntraj, count):
        out = np.zeros(ntraj * count)
        check(load().xsh_probe_normals(self._h, nsp, tag, draw, traj0, ntraj, count, _ptr(out)))
        return out.reshape(ntraj, count)

    def probe_propagate(self, nc, planes):
        planes = np.asarray(planes, dtype=complex)
        buf = np.ascontiguousarray(np.stack([planes.real, planes.imag], -1).ravel())
        check(load().xsh_probe_propagate(self._h, nc, planes.shape[0], _ptr(buf)))
        return (buf[0::2] + 1j * buf[1::2]).reshape(planes.shape)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mcheck(rc):
    if rc != 0:
        raise XshError(rc, load().xsh_mlp_last_error().decode())


class MLP:
    """The control-variate network on the GPU (xsh_mlp_* entry points)."""

    def __init__(self, theta0, nin=64, nh=256, batch_max=65536):
        lib = load()
        th = np.ascontiguousarray(theta0, dtype=np.float32)
        h = C.c_void_p()
        mcheck(lib.xsh_mlp_create(nin, nh, batch_max, _fptr(th), C.byref(h)))
        self._h, self.nin, self.nh, self.P = h, nin, nh, th.size

    def close(self):
        if getattr(self, "_h", None):
            load().xsh_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.P, dtype=np.float32)
        mcheck(load().xsh_mlp_get_params(self._h, _fptr(out)))
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[0], dtype=np.float32)
        mcheck(load().xsh_mlp_forward(self._h, x.shape[0], _fptr(x), _fptr(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.P, dtype=np.float32)
        loss = C.c_double()
        mcheck(load().xsh_mlp_loss_grad(self._h, x.shape[0], _fptr(x), _fptr(t), C.byref(loss), _fptr(g)))
        return loss.value, g

    def train(self, x, t, epochs, batch, lr, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n = x.shape[0]
        rng = np.random.default_rng(seed)
        order = np.ascontiguousarray(np.concatenate([rng.permutation(n) for _ in range(epochs)]).astype(np.int32))
        losses = np.zeros(epochs)
        mcheck(load().xsh_mlp_train(self._h, n, _fptr(x), _fptr(t), epochs, batch,
                                    order.ctypes.data_as(C.POINTER(C.c_int)), lr, _ptr(losses)))
        return losses

    def set_translations(self, modes, nx, ny, seed=0):
        """Train with random lattice translations of the (kx, ky) feature modes (None: off)."""
        if modes is None:
            mcheck(load().xsh_mlp_set_translations(self._h, 0, None, 1, 1, 0))
            return
        k = np.ascontiguousarray(np.asarray(modes, dtype=np.int32).reshape(-1))
        mcheck(load().xsh_mlp_set_translations(self._h, len(k) // 2, k.ctypes.data_as(C.POINTER(C.c_int)),
                                               nx, ny, seed))

    def gaussian_moments(self, n, seed):
        s, s2 = C.c_double(), C.c_double()
        mcheck(load().xsh_mlp_gaussian_moments(self._h, n, seed, C.byref(s), C.byref(s2)))
        return s.value, s2.value


class Batch:
    """Many instances in one graph launch (xsh_batch_*)."""

    def __init__(self, descs, autotune=False, serial=False):
        lib = load()
        for d in descs:
            d.struct_size = C.sizeof(Desc)
        arr = (Desc * len(descs))(*descs)
        b = C.c_void_p()
        flags = (XSH_BATCH_AUTOTUNE if autotune else 0) | (XSH_BATCH_SERIAL if serial else 0)
        check(lib.xsh_batch_create(len(descs), arr, flags, C.byref(b)))
        self._b, self.descs = b, list(descs)
        self.infos = []
        for i in range(len(descs)):
            info = Info()
            check(lib.xsh_batch_query(self._b, i, C.byref(info)))
            self.infos.append(info)

    def close(self):
        if getattr(self, "_b", None):
            load().xsh_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, npar=1):
        check(load().xsh_batch_run(self._b, npar))

    def synchronize(self):
        ms = C.c_double()
        check(load().xsh_batch_synchronize(self._b, C.byref(ms)))
        return ms.value

    def nonfinite(self, i):
        """Instance i's divergence flag for the last run."""
        f = C.c_int()
        check(load().xsh_batch_nonfinite(self._b, i, C.byref(f)))
        return bool(f.value)

    def read_averages(self, i):
        d, info = self.descs[i], self.infos[i]
        n = d.ensembles[1] * info.nchk * info.avlen
        out = np.zeros(2 * n)
        check(load().xsh_batch_read_averages(self._b, i, _ptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(d.ensembles[1], info.nchk, info.avlen)

    def plan(self):
        n = C.c_int()
        check(load().xsh_batch_plan(self._b, C.byref(n), None))
        buf = (C.c_int * (8 * n.value))()
        check(load().xsh_batch_plan(self._b, C.byref(n), buf))
        keys = ("precision", "nx", "ny", "instances", "threads", "traj_per_block", "radix", "grid")
        return [dict(zip(keys, buf[8 * g:8 * g + 8])) for g in range(n.value)]
