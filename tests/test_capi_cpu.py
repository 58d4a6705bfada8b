"""CPU-side checks of the C ABI: the library loads, exports every declared
entry point, and rejects invalid descriptors before touching a device
(the analogue of the reference's error() aborts, xtransform.m:225-226)."""
import ctypes as C
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
    d = make_desc(q)
    d.nx = 201   # odd lattice
    rc, msg = _create(d)
    assert rc == 2
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
