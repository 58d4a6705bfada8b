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
    assert [s["noises"] for s in subs] == [0, 0, 0, 1]      # only field 4's form reads noise (s = 0.1)
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
        p["gpu"] = dict(p["gpu"], s=0.1)                 # a noise term that reads the cells
        field_problems(p, preferences(p))
    # cells no form reads may be any number (Peregrine.m: noises = 1, four fields)
    p = _four_fields(noises=1)
    p["gpu"] = dict(p["gpu"], fields=[{}, dict(g=2j), {}, {}])
    subs, _ = field_problems(p, preferences(p))
    assert [s["noises"] for s in subs] == [0, 0, 0, 0]
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


def _coupled(**kw):
    p = dict(dimensions=1, fields=3, noises=2, points=[6], ranges=[1.0], steps=[2],
             gpu=dict(deriv="matrix", M=np.eye(3) * 0.1j, S=[[1, 0], [0, 1], [0.5, 0.5]], initial="const",
                      c0=[1, 2, 3], observe=["a", "abs2"], ofield=[3, 1], precision="fp64"))
    p.update(kw)
    return p


def test_coupled_components_descriptor_and_plan():
    """Coupled SDE components: one descriptor (fields, M, g, S, l0 and c0 per
    component, the component of each observable), planned without a device:
    a0 / a_out hold the components of a trajectory (nsites = fields), each
    observable one value per time point."""
    d = make_desc(preferences(_coupled()))
    assert d.fields == 3 and d.deriv == _lib.DERIV["matrix"]
    assert [d.obs_field[o] for o in range(2)] == [2, 0]
    assert d.fnoise[2][0] == 0.5 and d.fmat[1][1][1] == 0.1 and d.fc0[2][0] == 3.0
    info = _lib.plan(d)
    assert info.nsites == 3 and list(info.obs_sites)[:2] == [1, 1] and info.avlen == 12


def _set(d, **kw):
    for k, v in kw.items():
        if k == "obs_field0":
            d.obs_field[0] = v
        elif k == "tspec0":
            d.tspec[0] = v
        else:
            setattr(d, k, v)
    return d


@pytest.mark.parametrize("change,status", [
    (dict(fields=1), "EINVAL"),                          # the matrix form with one component
    (dict(fields=5), "EUNSUPPORTED"),                    # more components than the kernels hold
    (dict(dimensions=2, nx=32), "EUNSUPPORTED"),          # coupled lattice components
    (dict(noises=6), "EUNSUPPORTED"),                    # more than 4 noise components
    (dict(obs_field0=3), "EINVAL"),                      # observable of a component that is not there
    (dict(tspec0=1), "EUNSUPPORTED"),                    # time-domain spectra of coupled components
    (dict(deriv=_lib.DERIV["gp"]), "EINVAL"),            # fields > 1 without the matrix form
])
def test_coupled_components_refusals(change, status):
    """The C ABI refuses coupled descriptors it cannot integrate (no device needed)."""
    d = _set(make_desc(preferences(_coupled())), **change)
    if "dimensions" in change:
        d.ranges[1] = 5.0
    with pytest.raises(_lib.XshError, match=status):
        _lib.plan(d)
    with pytest.raises(GpuFormError, match="4 noise components"):
        make_desc(preferences(_coupled(noises=6)))


def test_oracle_sho_closed_form():
    """Examples/SDE_1D/SHO.m's deterministic project in the oracle (the matrix
    form, MP, order 2): the mean of a_1 is cos(t) within the extrapolated step
    error, that of a_2 is sin(t), and a wrong sign of the coupling fails."""
    from oracle import forms, xspde as X
    p = dict(dimensions=1, fields=2, noises=0, points=[51], ranges=[10.0], steps=[20], checks=1, order=2,
             ensembles=[3, 1, 1], method="MP",
             gpu=dict(deriv="matrix", M=[[0, -1], [1, 0]], initial="const", c0=[1.0, 0.0], observe=["a", "a"],
                      ofield=[1, 2]),
             compare=[lambda p: np.cos(p.t), lambda p: np.sin(p.t)])
    err = X.xsim(forms.apply_forms(p))[0]
    assert err[0] < 1e-4                                # MP, dt = 0.025, order 2: 5.6e-6
    bad = dict(p, gpu=dict(p["gpu"], M=[[0, 1], [-1, 0]]))
    assert X.xsim(forms.apply_forms(bad))[0][0] > 1e-2


class _FakeHandle:
    """Stands in for a GPU handle: fixed per-path averages, no device."""

    def __init__(self, av_all):
        self.av_all = av_all

    def run_ensemble(self, npar):
        pass

    def synchronize(self):
        return 0.0

    def nonfinite(self):
        return False

    def read_averages(self):
        return self.av_all


def test_fields_problem_assembles_observables_in_project_order():
    """FieldsProblem.xensemble on fake per-component handles: observable o is
    read from its component's handle at its place there, and the D, D^2 sums
    equal those of ensemble_moments on the same averages taken directly."""
    import importlib
    import types
    XS = importlib.import_module("xspde_matlab_amd.xsim")
    p = _four_fields()
    q = preferences(p)
    subs, where = field_problems(p, q)
    rng = np.random.default_rng(3)
    e2, nchk, npts = q["ensembles"][1], 1 + q["checks"][0], q["points"][0]
    parts = []
    for sk in subs:
        qk = preferences(sk)
        nobs = len(sk["gpu"]["observe"])
        ns = [161 if o in ("a", "abs2") else 1 for o in sk["gpu"]["observe"]]
        offs = np.concatenate([[0], np.cumsum([npts * n for n in ns])])
        info = types.SimpleNamespace(obs_offset=list(offs[:-1]), obs_sites=ns, avlen=int(offs[-1]))
        av_all = rng.standard_normal((e2, nchk, info.avlen)) + 1j * rng.standard_normal((e2, nchk, info.avlen))
        parts.append(types.SimpleNamespace(q=qk, info=info, handle=_FakeHandle(av_all)))
    fp = XS.FieldsProblem.__new__(XS.FieldsProblem)
    fp.q, fp.where, fp.parts, fp.info = q, where, parts, parts[0].info
    fp.outputs = [lambda av, p, _n=n: av[_n] for n in range(len(where))]
    data = fp.xensemble(1)
    assert len(data) == 3
    for o, (k, i) in enumerate(where):
        part = parts[k]
        direct = XS.ensemble_moments(part.q, part.info, [lambda av, p, _i=i: av[_i]], part.handle.av_all)
        assert np.allclose(data[o], direct[0])
