"""CPU-side checks of the C ABI: the library loads, exports every declared
entry point, and rejects invalid descriptors before touching a device
(the analogue of the reference's error() aborts, xtransform.m:51)."""
import ctypes as C

import numpy as np
import os
import re

import pytest

from xspde_matlab_amd import _lib, configs
from xspde_matlab_amd.xsim import GpuFormError, make_desc, preferences

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "xspde_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(xsh_\w+)\(", hdr, re.M)))


def test_header_matches_binding():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_all_symbols():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.xsh_version() == 1


def _create(desc):
    lib = _lib.load()
    desc.struct_size = C.sizeof(_lib.Desc)
    h = C.c_void_p()
    rc = lib.xsh_create(C.byref(desc), C.byref(h))
    return rc, lib.xsh_last_error().decode()


def test_invalid_boundary_rejected():
    q = preferences(configs.shrink(configs.nls1d(), E=16, e2=1, points=3))
    d = make_desc(q)
    d.boundaries[0][0] = 2
    rc, msg = _create(d)
    assert rc == 1 and "Invalid boundary" in msg


def test_unsupported_forms_rejected():
    q = preferences(configs.shrink(configs.nls1d(), E=16, e2=1, points=3))
    d = make_desc(q)
    d.deriv = 9
    rc, msg = _create(d)
    assert rc == 2 and "drift" in msg
    d = make_desc(preferences(configs.shrink(configs.cgl2d(n=32), E=2, e2=1, points=3)))
    d.nx = 129   # odd 2D lattice above the padded 128 x 64 kernels
    rc, msg = _create(d)
    assert rc == 2 and "odd" in msg
    d = make_desc(q)
    d.struct_size = 4
    d2 = d
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.xsh_create(C.byref(d2), C.byref(h)) == 1


def test_bad_arity_rejected():
    q = preferences(configs.shrink(configs.kubo(), E=16, e2=1, points=3))
    d = make_desc(q)
    d.noises = 0
    rc, msg = _create(d)
    assert rc == 1 and "noises" in msg
    d = make_desc(q)
    d.checks = 2
    assert _create(d)[0] == 1


def test_user_handles_stay_on_cpu_path():
    p = configs.kubo()
    p["deriv"] = lambda a, w, p: 1j * w * a
    with pytest.raises(GpuFormError):
        preferences(p)
    p = configs.kubo()
    p["method"] = "Euler"
    with pytest.raises(GpuFormError):
        preferences(p)
    p = configs.kubo()
    p["gpu"] = dict(p["gpu"], deriv="unknown")
    with pytest.raises(GpuFormError):
        make_desc(preferences(p))


def test_refined_lattice_and_defaults():
    # xpreferences.m:137 points(2:nd) = 1+(points-1)*steps(2:nd); :101-105 default method
    q = preferences(dict(dimensions=2, points=[11, 33], steps=[2, 2], ensembles=[4, 2],
                         gpu=dict(deriv="zero")))
    assert q["points"] == [11, 65]
    assert q["method"] == "MP"
    q = preferences(dict(dimensions=2, gpu=dict(deriv="zero")))
    assert q["points"] == [51, 35] and q["method"] == "RK4" and q["checks"] == [1, 0]


def test_batch_instances_generate_valid_descriptors():
    """The uneven batch workload: 64 instances over every lattice family, with
    the drawn parameters inside their stated ranges and descriptors that pass
    the host-side form checks."""
    from xspde_matlab_amd.configs import BATCH_1D, BATCH_2D, BATCH_E, batch_instances
    from xspde_matlab_amd.xsim import make_desc, preferences
    ps = batch_instances(count=64, seed=0)
    shapes = {tuple(p["points"][1:]) for p in ps}
    assert shapes == {(n,) for n in BATCH_1D} | {(n, n) for n in BATCH_2D}
    for p in ps:
        g = p["gpu"]
        gamma, gnl = -g["u"], -g["g"].imag
        assert 0.01 <= gamma <= 0.5 and 0.1 <= gnl <= 2.0
        assert p["ensembles"][0] * p["ensembles"][1] in BATCH_E
        assert (p["points"][0] - 1) * p["steps"][0] == 500
        d = make_desc(preferences(p))
        assert d.dimensions == len(p["points"]) and d.method == 0


def test_boundary_table_levels_and_times():
    """p.boundfun tabulated per half fine step: level m holds the values at
    origins(1) + m dtr/2 (dtr the fine step), 2 n (1 + checks) + 1 levels for n
    method calls per path; values that do not change in time give one level."""
    from xspde_matlab_amd.xsim import boundary_table
    p = configs.shrink(configs.nls1d(), E=16, e2=1, points=3, steps=2)
    p["boundaries"] = [[1, -1]]
    p["origins"] = [0.5, -10.0]
    p["boundfun"] = lambda a, c, d, p: (np.exp(1j * p.t), p.t ** 2)
    q = preferences(p)
    tab = boundary_table(q)
    calls = 3 * 2 - 1
    assert tab.shape == (2 * calls * 2 + 1, 2, 2, 1)
    dtr = q["dt"] / (2 * 2)
    t = 0.5 + np.arange(tab.shape[0]) * dtr / 2
    assert np.allclose(tab[:, 0, 0, 0], np.exp(1j * t)) and np.allclose(tab[:, 0, 1, 0], t ** 2)
    assert not tab[:, 1].any()
    p["boundfun"] = lambda a, c, d, p: (0.25, -1j)
    tab = boundary_table(preferences(p))
    assert tab.shape == (1, 2, 2, 1) and tab[0, 0, 0, 0] == 0.25 and tab[0, 0, 1, 0] == -1j
    d = make_desc(preferences(p))
    assert d.boundlevels == 1
    d.boundlevels = 7
    rc, msg = _create(d)
    assert rc == 1 and "boundlevels" in msg


def test_boundary_table_lines_2d():
    """2D: the x ends' values run along y and the y ends' along x, whether given
    as vectors, lattice-shaped arrays or scalars."""
    from xspde_matlab_amd.xsim import boundary_table
    p = configs.shrink(configs.cgl2d(n=32), E=2, e2=1, points=2, steps=1)
    p["boundaries"] = [[-1, -1], [1, 1]]
    p["boundfun"] = lambda a, c, d, p: ((np.cos(p.y), 2.0) if d == 2 else (np.arange(32) * p.t, np.sin(p.x)))
    q = preferences(p)
    tab = boundary_table(q)
    y = q["origins"][2] + np.arange(32) * q["ranges"][2] / 31
    x = q["origins"][1] + np.arange(32) * q["ranges"][1] / 31
    assert np.allclose(tab[3, 0, 0], np.cos(y)) and np.allclose(tab[3, 0, 1], 2.0)
    assert np.allclose(tab[3, 1, 1], np.sin(x))
    t3 = q["origins"][0] + 3 * q["dt"] / 2 / 2
    assert np.allclose(tab[3, 1, 0], np.arange(32) * t3)


def test_filter_table_site_order_and_linearity():
    """k-space filters (p.nfilter, p.rfilter) become multiplier tables over the
    FFT-order k sites kx + nx*ky; a filter that is not a site-wise
    multiplication has no table and the project stays on the CPU path."""
    from xspde_matlab_amd.xsim import filter_table
    p = configs.shrink(configs.cgl2d(n=32), E=2, e2=1, points=2, steps=1)
    p.update(noises=0, knoises=2, nfilter=lambda w, p: w * (p.kx + 2.0 * p.ky + 0.5j))
    q = preferences(p)
    F = filter_table(q, "nfilter", 2)
    dk = 2 * np.pi / (32 * q["ranges"][1] / 31)
    k = np.fft.ifftshift((np.arange(32) - 16) * dk)
    kx, ky = np.meshgrid(k, k, indexing="xy")          # [ky, kx] -> site kx + 32 ky
    assert F.shape == (2, 1024)
    assert np.allclose(F[0], (kx + 2.0 * ky + 0.5j).reshape(-1)) and np.allclose(F[1], F[0])
    p["nfilter"] = lambda w, p: w ** 2
    with pytest.raises(GpuFormError):
        filter_table(preferences(p), "nfilter", 2)
    assert filter_table(preferences(dict(p, nfilter=None)), "nfilter", 1).shape == (1, 1024)


def test_time_spectra_flags():
    """p.transforms{o}(1) reaches the descriptor as tspec[o]; spectra run on SDEs
    only, and a flag other than 0/1 is invalid (rejected before any device call)."""
    p = configs.shrink(configs.nls1d(), E=16, e2=1, points=3)
    p["transforms"] = [0, [1, 0]]
    d = make_desc(preferences(p))
    assert list(d.tspec)[:2] == [0, 1]
    rc, msg = _create(d)
    assert rc == 2 and "SDEs" in msg
    q = preferences(configs.shrink(configs.kubo(), E=16, e2=1, points=3))
    d = make_desc(q)
    d.tspec[0] = 2
    rc, msg = _create(d)
    assert rc == 1 and "tspec" in msg


def test_run_path_rejects_a_wrongly_sized_a0():
    """The initial fields handed to xsh_run_path hold ensembles(1) x sites
    complex values; the binding refuses any other size before the call (a
    handle shell from xsh_plan's Info: no device is touched)."""
    p = dict(dimensions=2, points=[5, 16], ranges=[1.0, 8.0], ensembles=[3, 1, 1], noises=0,
             gpu=dict(deriv="zero", linear=dict(l2=0.5j), initial="gauss", observe=["abs2"]))
    d = make_desc(preferences(p))
    h = _lib.Handle.__new__(_lib.Handle)
    h.desc, h.info, h._h = d, _lib.plan(d), None
    assert h.info.nsites == 16
    with pytest.raises(ValueError, match="ensembles\\(1\\) x sites = 48"):
        h.run_path(1, 1, 1, a0=np.zeros((1, 16), complex))
