"""Splitting a project with several field components (p.fields > 1) into
one-component descriptors, without a device: boundary rows and values per
field, the per-field boundary function, noise shares, observable placement,
and the device-free plan of every component (xsh_plan)."""
import numpy as np
import pytest

from xspde_matlab_amd import _lib
from xspde_matlab_amd.xsim import GpuFormError, field_problems, make_desc, preferences


def _four_fields(**kw):
    p = dict(dimensions=2, fields=4, noises=4, points=[11, 161], ranges=[1.0, 10.0],
             boundaries=[[[1, 1], [-1, -1], [1, -1], [-1, 1]]],
             boundval=[[[1.0, 2.0], [0.5, -0.5], [0.0, 1.0], [2.0, 0.0]]],
             gpu=dict(deriv="gp", g=1j, linear=dict(l2=0.5j), initial="const", observe=["a", "abs2", "a_ave"],
                      ofield=[2, 4, 2], fields=[{}, dict(g=2j), {}, dict(s=0.1)], precision="fp64"))
    p.update(kw)
    return p


def test_rows_values_noises_and_observables_per_field():
    p = _four_fields()
    subs, where = field_problems(p, preferences(p))
    assert [s["boundaries"][0] for s in subs] == [[1, 1], [-1, -1], [1, -1], [-1, 1]]
    assert [s["boundval"][0] for s in subs] == [[1.0, 2.0], [0.5, -0.5], [0.0, 1.0], [2.0, 0.0]]
    assert [s["noises"] for s in subs] == [1, 1, 1, 1]
    assert [s["gpu"]["inst"] for s in subs] == [0, 1, 2, 3]
    assert subs[1]["gpu"]["g"] == 2j and subs[0]["gpu"]["g"] == 1j and subs[3]["gpu"]["s"] == 0.1
    # observables 1 and 3 on field 2, observable 2 on field 4; fields 1 and 3 carry a placeholder
    assert subs[1]["gpu"]["observe"] == ["a", "a_ave"] and subs[3]["gpu"]["observe"] == ["abs2"]
    assert subs[0]["gpu"]["observe"] == ["a_ave"]
    assert where == [(1, 0), (3, 0), (1, 1)]
    for s in subs:
        d = make_desc(preferences(s))
        info = _lib.plan(d)
        assert info.nsites == 161


def test_per_field_boundary_function_is_tabulated_per_field():
    def ends(a, c, d, p):
        return [(p.t, 0.0), (1.0, -p.t), (2.0, 3.0), (0.0, 0.0)]
    p = _four_fields(boundfun=ends, boundval=None)
    subs, _ = field_problems(p, preferences(p))
    from xspde_matlab_amd.xsim import boundary_table
    tabs = [boundary_table(preferences(s)) for s in subs]
    q = preferences(subs[0])
    dtr = q["dt"] / (q["steps"][0] * 2)
    m = 3
    assert np.allclose(tabs[0][m, 0, 0], m * dtr / 2) and np.allclose(tabs[0][m, 0, 1], 0)
    assert np.allclose(tabs[1][m, 0, 1], -m * dtr / 2)
    assert tabs[2].shape[0] == 1 and np.allclose(tabs[2][0, 0, :, 0], [2.0, 3.0])


def test_refusals():
    with pytest.raises(GpuFormError, match="noises"):
        p = _four_fields(noises=2)
        field_problems(p, preferences(p))
    with pytest.raises(GpuFormError, match="ofield"):
        p = _four_fields()
        p["gpu"] = dict(p["gpu"], ofield=[1, 5, 1])
        field_problems(p, preferences(p))
    with pytest.raises(GpuFormError, match="one entry per field"):
        p = _four_fields()
        p["gpu"] = dict(p["gpu"], fields=[{}, {}])
        field_problems(p, preferences(p))
    with pytest.raises(GpuFormError, match="one field component"):
        make_desc(preferences(_four_fields()))
