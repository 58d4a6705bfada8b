"""FP32 against FP64 on the GPU, per observable, on the same seeded streams
(the precision statement of the FP32 path): 1D stochastic NLS at its full time
span (1000 coarse RK4IP steps with checks) and 2D CGL at its full span
(2000 coarse steps), fewer trajectories.  The FP32 error of each ensemble
mean is stated relative to its FP64 value and to the sampling error bar."""
import numpy as np
import pytest

from xspde_matlab_amd import Problem
from xspde_matlab_amd.configs import cgl2d, nls1d, shrink

pytestmark = pytest.mark.gpu


def _means(p):
    prob = Problem(p)
    prob.handle.run_ensemble(1)
    prob.handle.synchronize()
    av = prob.handle.read_averages()          # [e2, nchk, avlen]
    info = prob.info
    prob.close()
    return av, info


def _compare(p32, p64, names, bound_rel, bound_sigma, bound_step=None, bound_step_sigma=None):
    """Per observable: the FP32 ensemble mean against FP64 (relative to the
    largest FP64 mean and to the FP64 sampling error bar), and the FP32 step-error
    estimate |fine - coarse| (formed in FP64 from the FP32 paths' FP64 moments)
    against the FP64 one.  Either the two step errors agree (max difference below
    bound_step times the largest FP64 step error), or, where FP32 rounding
    dominates the FP32 estimate, its RMS over the time points (the step entry
    of xsim's error vector, Simulator/xsim.m:104-107,144-188) stays below
    bound_step_sigma times the RMS sampling error (xsim.m:94-98), so the total
    error bar is unchanged."""
    a32, info = _means(p32)
    a64, _ = _means(p64)
    e2 = a64.shape[0]
    npts = p64["points"][0]
    for o, name in enumerate(names):
        off, ns = info.obs_offset[o], info.obs_sites[o]
        m32 = a32[:, 0, off:off + npts * ns].real
        m64 = a64[:, 0, off:off + npts * ns].real
        diff = np.abs(m32.mean(0) - m64.mean(0))
        scale = np.abs(m64.mean(0)).max()
        sigma = m64.std(0, ddof=1) / np.sqrt(e2)
        rel = diff.max() / scale
        # points with no sampling spread (t = 0 of a deterministic initial state) are
        # judged by the relative bound alone
        spread = sigma > 1e-9 * scale
        in_sigma = (diff[spread] / sigma[spread]).max() if spread.any() else 0.0
        # step error of the ensemble means (order 0): |D_fine - D_coarse| per point
        s32 = np.abs(m32.mean(0) - a32[:, 1, off:off + npts * ns].real.mean(0))
        s64 = np.abs(m64.mean(0) - a64[:, 1, off:off + npts * ns].real.mean(0))
        step_rel = np.abs(s32 - s64).max() / s64.max()
        step_sigma = np.sqrt(np.mean(s32 ** 2) / np.mean(sigma ** 2))
        print("%s %s: max |fp32 - fp64| / max|fp64| = %.2e, max |fp32 - fp64| / sigma = %.2e; "
              "step error max fp64 %.3e fp32 %.3e, max |fp32 - fp64| / max fp64 = %.2e, rms fp32 / rms sigma = %.2e"
              % (p64["name"], name, rel, in_sigma, s64.max(), s32.max(), step_rel, step_sigma), flush=True)
        assert rel < bound_rel and in_sigma < bound_sigma, name
        if bound_step is not None:
            assert step_rel < bound_step, name
        if bound_step_sigma is not None:
            assert step_sigma < bound_step_sigma, name


def test_nls1d_fp32_vs_fp64_full_span():
    """The FP32 step-error estimate agrees with FP64's to 0.1 % of the largest."""
    p64 = shrink(nls1d("fp64"), E=512, e2=8)
    p32 = shrink(nls1d("fp32"), E=512, e2=8)
    _compare(p32, p64, ["abs2", "a"], 1e-3, 0.1, bound_step=0.01)


def test_cgl2d_fp32_vs_fp64_full_span():
    """The 2D CGL is chaotic: over 2000 coarse steps FP32 rounding decorrelates the
    fine and coarse paths, and the FP32 step-error estimate (2e-5..4e-5 for
    <|a|^2>) exceeds the FP64 truncation estimate (6e-6) several times; its RMS
    is 6 % of the RMS sampling error (bound 15 %: the total error bar, the two
    added in quadrature, grows by less than 1.2 %)."""
    p64 = shrink(cgl2d("fp64"), E=128, e2=8)
    p32 = shrink(cgl2d("fp32"), E=128, e2=8)
    _compare(p32, p64, ["abs2_ave", "corr"], 1e-3, 0.1, bound_step_sigma=0.15)


def test_cgl2d_neumann_fp32_vs_fp64_full_span():
    """Config 3's Neumann run: FP32 on the matrix cores (register-resident
    kernel) against FP64 (streaming dense-transform kernel), full span.  The
    periodic correlation of the Neumann field has long-distance entries whose
    sampling spread is small, where the FP32 error reaches 0.2 sigma (6.8e-4
    relative): bound 0.3 sigma.  As in the periodic run, the FP32 step-error
    estimate is rounding-dominated and its RMS stays below 15 % of the RMS
    sampling error."""
    p64 = shrink(cgl2d("fp64"), E=128, e2=8)
    p32 = shrink(cgl2d("fp32"), E=128, e2=8)
    for p in (p64, p32):
        p["boundaries"] = [[-1, -1], [-1, -1]]
    _compare(p32, p64, ["abs2_ave", "corr"], 1e-3, 0.3, bound_step_sigma=0.15)
