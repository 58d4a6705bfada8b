"""The '.mat' data file of xsim.m:207-210 and its reader xread.m:20-33 (CPU).

The MATLAB-visible structure (variable names, classes, sizes) is checked with
scipy.io.whosmat, the values by a save/load round trip, and the array layout
of per-site outputs against the column-major definition of p.d.av
(xpreferences.m:340).  Whether MATLAB's own load/xread accept the file is
parity unpinned: MATLAB is not part of this image."""
import types

import numpy as np
import pytest
from scipy.io import whosmat

from xspde_matlab_amd.xfile import xread, xsave
from xspde_matlab_amd.xsim import matlab_data, matlab_raw, split_averages


def _inp():
    return {"name": "test", "dimensions": 3, "points": [3, 4, 2], "ranges": [1.0, 2.0, 2.0],
            "ensembles": [5, 2, 1], "checks": [1, 0, 0], "seed": 7, "rawdata": 1, "file": "x.mat",
            "boundaries": [[-1, -1], [0, 0]], "gpu": {"deriv": "gp", "g": -1 + 0.5j, "observe": ["a", "abs2_ave"]},
            "output": [None, lambda av, p: av[1]], "drift": lambda a, w, p: a, "_private": 3}


def test_file_structure_and_round_trip(tmp_path):
    f = str(tmp_path / "run.mat")
    rng = np.random.default_rng(3)
    error = rng.random(6)
    data = [rng.random((1, 3, 4, 2, 3)), rng.random((1, 3, 3))]
    raw = [[rng.random((1, 3, 4, 2, 5)) + 1j * rng.random((1, 3, 4, 2, 5)) for j in range(2)] for nc in range(2)]
    xsave(f, error, data, _inp(), raw)
    w = {name: (shape, cls) for name, shape, cls in whosmat(f)}
    assert w["error"] == ((1, 6), "double")
    assert w["data"] == ((1, 1), "cell")
    assert w["input"] == ((1, 1), "cell")
    assert w["raw"] == ((1, 2, 2), "cell")
    inp, d2, r2 = xread(f, verbose=False)
    q = inp[0]
    assert q["version"] == "xSPDE4.21" and q["name"] == "test" and q["seed"] == 7
    assert list(q["points"]) == [3, 4, 2]
    assert np.array_equal(np.asarray(q["boundaries"]), [[-1, -1], [0, 0]])
    assert q["gpu"]["deriv"] == "gp" and q["gpu"]["g"] == -1 + 0.5j
    assert q["gpu"]["observe"] == ["a", "abs2_ave"]
    assert "drift" not in q and "_private" not in q        # callables and private keys are not written
    assert len(d2) == 1 and len(d2[0]) == 2
    for a, b in zip(data, d2[0]):
        assert a.shape == b.shape and np.array_equal(a, b)
    for nc in range(2):
        for j in range(2):
            assert r2[nc][j].shape == (1, 3, 4, 2, 5) and np.array_equal(r2[nc][j], raw[nc][j])


def test_file_without_raw_and_single_path(tmp_path):
    f = str(tmp_path / "one.mat")
    q = _inp()
    q["ensembles"], q["checks"] = [5, 1, 1], [0, 0, 0]
    xsave(f, np.zeros(6), [np.ones((1, 3, 3))], q, None)
    w = {name: (shape, cls) for name, shape, cls in whosmat(f)}
    assert w["raw"][1] == "cell"
    inp, data, raw = xread(f, verbose=False)
    assert raw == [[None]]
    assert np.array_equal(data[0][0], np.ones((1, 3, 3)))


def test_file_names(tmp_path):
    with pytest.raises(ValueError):
        xsave(str(tmp_path / "run.h5"), np.zeros(6), [], _inp())
    with pytest.raises(ValueError):
        xread(str(tmp_path / "run"))
    with pytest.raises(FileNotFoundError):
        xread(str(tmp_path / "missing.mat"))


def test_version_printed(tmp_path, capsys):
    f = str(tmp_path / "v.mat")
    xsave(f, np.zeros(6), [np.zeros((1, 2, 3))], _inp())
    xread(f)
    assert "generated by xSPDE4.21" in capsys.readouterr().out


def test_per_site_output_layout():
    """data{n}(1, t, x, y, c) of a per-site observable is the statistic c of
    the average at (t, x, y) of p.d.av{n}, whose flat block is column-major
    [lines, points, nx, ny] (time fastest, then x, then y)."""
    npts, nx, ny = 3, 4, 2
    ns = nx * ny
    q = {"points": [npts, nx, ny], "dimensions": 3, "gpu": {"observe": ["abs2", "abs2_ave"]}}
    info = types.SimpleNamespace(obs_offset=[0, npts * ns], obs_sites=[ns, 1], avlen=npts * ns + npts)
    outputs = [lambda av, p: av[0], lambda av, p: av[1]]
    flat = np.arange(info.avlen, dtype=float)
    # the D column of each output is the C-order flatten of output(av), as ensemble_moments builds it
    outd = []
    for o, out in enumerate(outputs):
        v = np.real(np.asarray(out(split_averages(q, info, flat), q))).reshape(-1)
        outd.append(np.stack([v, 10 * v, 100 * v], axis=1))
    arr = matlab_data(q, split_averages(q, info, np.zeros(info.avlen, dtype=complex)), outputs, outd)
    assert arr[0].shape == (1, npts, nx, ny, 3) and arr[1].shape == (1, npts, 1, 3)
    block = flat[:npts * ns].reshape(1, npts, nx, ny, order="F")
    for t in range(npts):
        for x in range(nx):
            for y in range(ny):
                want = flat[t + npts * (x + nx * y)]
                assert block[0, t, x, y] == want
                assert np.array_equal(arr[0][0, t, x, y], [want, 10 * want, 100 * want])
    assert np.array_equal(arr[1][0, :, 0, 0], flat[npts * ns:])


def test_raw_layout():
    """p.d.raw = [1, points, nx, ny, E] from read_raw's [points, E, sites]."""
    npts, nx, ny, E = 3, 4, 2, 5
    r = np.arange(npts * E * nx * ny).reshape(npts, E, nx * ny) * (1 + 1j)
    m = matlab_raw({"points": [npts, nx, ny], "dimensions": 3}, r)
    assert m.shape == (1, npts, nx, ny, E)
    for t in range(npts):
        for x in range(nx):
            for y in range(ny):
                assert np.array_equal(m[0, t, x, y], r[t, :, x + nx * y])
    assert matlab_raw({"points": [npts], "dimensions": 1}, r[:, :, :1]).shape == (1, npts, E)
    assert matlab_raw({"points": [npts, 8], "dimensions": 2}, r).shape == (1, npts, 8, E)
