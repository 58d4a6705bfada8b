"""xSPDE's MATLAB data files: the save at the end of ``xsim`` and ``xread``.

``xsave`` writes what Simulator/xsim.m:207-210 saves when ``p.file`` is not
blank -- the variables ``error``, ``data``, ``input`` and ``raw`` -- as a MAT
file (level 5, zlib-compressed, readable by MATLAB's ``load``).  ``xread``
reads such a file back as Graphics/DATA/xread.m:20-33 does for a '.mat' name:
it returns ``(input, data, raw)`` and prints the version that wrote it.

MATLAB shapes (what ``load`` shows):
  error        1 x n double (xsim.m:41, 177-185)
  data         1 x 1 cell holding the 1 x nout cell of sequence 1
               (``data = cell(seq)``, xsim.m:36); output n is an array of
               size [lines, points, space..., 3] with the mean, the step
               error and the sampling error along the last dimension
               (xsim.m:150-157)
  input        1 x 1 cell holding the parameter struct (xsim.m:177);
               function handles and other callables are not written
  raw          seq x nchk x ne cell (xsim.m:42, 56); with p.rawdata each
               entry is the path's complex fields [1, points, space..., E]
               (p.d.raw, xpreferences.m:148), else empty

Only '.mat' names are accepted, as xreadname (xread.m:137-160) does for the
MATLAB branch; the HDF5 branch of xread needs a writer the reference does
not ship with xsim and is not provided.
"""
import os

import numpy as np

VERSION = "xSPDE4.21"   # p.version default, Simulator/xpreferences.m:19


def _check_name(filename):
    if os.path.splitext(filename)[1] != ".mat":
        raise ValueError("Filename %s must end with .mat" % filename)


def _to_matlab(v):
    """A parameter value as scipy.io writes it, or None to leave it out."""
    if v is None or callable(v):
        return None
    if isinstance(v, dict):
        s = {}
        for k, x in v.items():
            if not isinstance(k, str) or not k[:1].isalpha():
                continue
            y = _to_matlab(x)
            if y is not None:
                s[k] = y
        return s
    if isinstance(v, str):
        return v
    if isinstance(v, (bool, np.bool_)):
        return float(v)
    if isinstance(v, (int, float, complex, np.number)):
        return v
    if isinstance(v, np.ndarray):
        return v
    if isinstance(v, (list, tuple)):
        try:
            arr = np.array(v)                     # a rectangular numeric list: a matrix
            if arr.dtype.kind in "biufc":
                return np.atleast_2d(arr) if arr.ndim == 1 else arr
        except (ValueError, TypeError):
            pass
        items = [_to_matlab(x) for x in v]
        cell = np.empty((1, len(items)), dtype=object)
        for i, x in enumerate(items):
            cell[0, i] = np.zeros((0, 0)) if x is None else x
        return cell
    return None


def _cell(shape):
    c = np.empty(shape, dtype=object)
    for idx in np.ndindex(*shape):
        c[idx] = np.zeros((0, 0))
    return c


def xsave(filename, error, data, inp, raw=None):
    """save(filename, 'error', 'data', 'input', 'raw') of xsim.m:207-210.

    error: the error vector; data: per output the array of MATLAB size
    [lines, points, space..., 3]; inp: the parameter dictionary; raw: nested
    list raw[nc][j] of per-path arrays (or None)."""
    from scipy.io import savemat
    _check_name(filename)
    outs = _cell((1, len(data)))
    for n, d in enumerate(data):
        outs[0, n] = np.asarray(d)
    dcell = _cell((1, 1))
    dcell[0, 0] = outs
    q = dict(inp)
    q.setdefault("version", VERSION)
    icell = _cell((1, 1))
    icell[0, 0] = _to_matlab(q)
    if raw:
        nchk, ne = len(raw), len(raw[0])
        rcell = _cell((1, nchk, ne))
        for nc in range(nchk):
            for j in range(ne):
                if raw[nc][j] is not None:
                    rcell[0, nc, j] = np.asarray(raw[nc][j])
    else:
        ens = list(q.get("ensembles", [1, 1, 1])) + [1, 1, 1]
        rcell = _cell((1, 1 + int(q.get("checks", [0])[0]), int(ens[1]) * int(ens[2])))
    savemat(filename, {"error": np.atleast_2d(np.asarray(error, dtype=float)), "data": dcell,
                       "input": icell, "raw": rcell}, format="5", do_compression=True, oned_as="row")


def _from_struct(s):
    """A loaded MATLAB struct (mat_struct) as a dictionary."""
    out = {}
    for k in s._fieldnames:
        v = getattr(s, k)
        out[k] = _from_value(v)
    return out


def _from_value(v):
    from scipy.io.matlab import mat_struct
    if isinstance(v, mat_struct):
        return _from_struct(v)
    if isinstance(v, np.ndarray):
        if v.dtype == object:
            items = v.reshape(-1, order="F")
            if items.size == 1 and isinstance(items[0], mat_struct):    # a 1 x 1 struct
                return _from_struct(items[0])
            return [_from_value(x) for x in items]
        if v.dtype.kind == "U":
            return "".join(v.reshape(-1).tolist())
        if v.size == 1:
            return v.reshape(-1)[0].item()
        if v.ndim == 2 and v.shape[0] == 1:
            return v[0]
    return v


def xread(filename, verbose=True):
    """[input, data, raw] = xread('filename.mat') (xread.m:20-33): input is the
    list of parameter dictionaries, data the list (per sequence) of per-output
    arrays, raw the nested list raw[nc][j] (None where a path has no fields)."""
    from scipy.io import loadmat
    _check_name(filename)
    if not os.path.exists(filename):
        raise FileNotFoundError("File %s doesn't exist" % filename)
    m = loadmat(filename, squeeze_me=False, struct_as_record=False, mat_dtype=False)
    inp = [_from_value(s) for s in m["input"].reshape(-1, order="F")]
    try:
        v = inp[0]["version"]
        if verbose:
            print("Reading Matlab file %s generated by %s" % (filename, v))
    except (KeyError, IndexError, TypeError):
        if verbose:
            print("Warning in xread: file %s is not an xGRAPH file" % filename)
    data = []
    for seq in m["data"].reshape(-1, order="F")[:len(inp)]:
        data.append([np.asarray(x) for x in seq.reshape(-1, order="F")])
    rc = m["raw"]
    if rc.ndim == 2:                       # MATLAB drops a trailing singleton ne
        rc = rc[:, :, None]
    raw = [[None if rc[0, nc, j].size == 0 else np.asarray(rc[0, nc, j]) for j in range(rc.shape[2])]
           for nc in range(rc.shape[1])]
    return inp, data, raw
