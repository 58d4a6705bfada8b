"""Oracle counterparts of the workloads in xspde_matlab_amd.configs."""
from xspde_matlab_amd.configs import cgl2d, kubo, nls1d, shrink  # noqa: F401


def oracle_problem(p):
    """The oracle's version of the same project (callables from the gpu forms)."""
    from oracle import forms
    return forms.apply_forms(p)
