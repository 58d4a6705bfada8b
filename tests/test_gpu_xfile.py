"""xsim with p.file and p.rawdata on the GPU: the saved '.mat' file read back
with xread holds the run's error vector, its statistics in the
[lines, points, nx, ny, 3] layout, and every path's raw fields as p.d.raw,
which match the oracle's fields at the stored points (xsim.m:207-210,
xpath.m:73-78)."""
import numpy as np
import pytest

from oracle import xspde as X
from xspde_matlab_amd.xfile import xread
from xspde_matlab_amd.xsim import xsim

from _problems import cgl2d, oracle_problem, shrink
from test_gpu_parity import _oracle_snapshots


@pytest.mark.gpu
def test_xsim_file_with_raw(tmp_path):
    f = str(tmp_path / "cgl.mat")
    p = shrink(cgl2d("fp64", n=32), E=3, e2=2, points=3, steps=2)
    p["rawdata"] = 1
    p["file"] = f
    err, outd, q, raw = xsim(p)
    inp, data, rawf = xread(f, verbose=False)
    assert inp[0]["version"] == "xSPDE4.21" and list(inp[0]["points"]) == [3, 32, 32]
    npts, nx, ny = 3, 32, 32
    assert len(data) == 1 and len(data[0]) == len(outd)
    for o, d in enumerate(data[0]):
        if d.ndim == 5:
            assert d.shape == (1, npts, nx, ny, 3)
            # back to the C-order D column of the output av[o] = [1, points, sites]
            flat = np.transpose(d, (0, 1, 3, 2, 4)).reshape(-1, 3)
        else:
            assert d.shape == (1, npts, 1, 3)
            flat = d.reshape(-1, 3)
        np.testing.assert_allclose(flat, outd[o], rtol=0, atol=0)
    assert len(rawf) == 2 and len(rawf[0]) == 2
    qo = X.xpreferences(oracle_problem(p))
    for nc in (1, 2):
        for j in (1, 2):
            got = rawf[nc - 1][j - 1]
            assert got.shape == (1, npts, nx, ny, 3)
            np.testing.assert_array_equal(got, raw[nc - 1][j - 1])
            for t, a in enumerate(_oracle_snapshots(qo, j, nc)):
                ref = np.asarray(a[0])                        # [nx, ny, E]
                e = np.abs(got[0, t] - ref).max() / np.abs(ref).max()
                assert e < 1e-10, (nc, j, t, e)
    print("xsim file: error, %d outputs, 4 raw paths match the oracle" % len(outd))
