"""Counter-based Gaussian stream of the oracle (test infrastructure only).

The reference draws its noise with MATLAB ``rng(nsp,'Threefry'); randn``
(Simulator/xensemble.m:26-32, xnoise.m:31, xrandom.m:21).  That stream is
closed, so the GPU path and this oracle share a stream of their own:

* Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
  SC'11), the published round function and constants.  Pinned by the
  Random123 known-answer vectors in ``tests/test_oracle_rng.py``.
* Box-Muller: (x0, x1) -> sqrt(-2 ln u0) * (cos 2 pi u1, sin 2 pi u1) with
  u = (x + 1/2) 2^-32 in (0, 1).  Pinned by moment/KS tests.

Stream layout (identical in ``xspde_matlab_amd/csrc/rng.hpp``):

    key     = (nsp, tag | seq << 8)      nsp = seed + npar + (j-1)*e3,
                                          xensemble.m:26
    counter = (group, draw, traj, inst)   group = L // 4, lane = L % 4

where ``L = site * comps + comp`` is the normal's index inside one draw of
one trajectory (site = ix + nx*iy, x fastest like MATLAB's column-major
layout), ``draw`` counts noise draws along the path (fine and coarse runs
see the same draws, xnoise.m:33-35), ``traj`` is the trajectory index in
its sub-ensemble and ``inst`` an instance id for batched problems.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

TAG_NOISE = 0
TAG_GAUSS_INIT = 1
TAG_UNIFORM_INIT = 2
TAG_KNOISE = 3          # k-space filtered noises (xnoise.m:39-50)
TAG_KRAND_INIT = 4      # k-space filtered initial randoms (xrandom.m:24-29)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32 with 10 rounds on broadcastable uint32-valued arrays.

    Returns four uint64 arrays holding the 32-bit outputs.
    """
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & MASK for c in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    k0 = np.asarray(k0, dtype=np.uint64) & MASK
    k1 = np.asarray(k1, dtype=np.uint64) & MASK
    for r in range(10):
        p0 = c0 * M0
        p1 = c2 * M1
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
        if r < 9:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
    return c0, c1, c2, c3


def uniform_open01(x):
    """uint32 -> float64 in (0,1): (x + 1/2) * 2^-32 (exact in float64)."""
    return (x.astype(np.float64) + 0.5) * 2.0 ** -32


def box_muller(x0, x1):
    u0 = uniform_open01(x0)
    u1 = uniform_open01(x1)
    r = np.sqrt(-2.0 * np.log(u0))
    th = 2.0 * np.pi * u1
    return r * np.cos(th), r * np.sin(th)


def key1(tag, seq=0):
    return (int(tag) & 0xFF) | ((int(seq) & 0xFFFFFF) << 8)


def normals(nsp, tag, draw, traj, count, seq=0, inst=0):
    """Standard normals N[traj, L] for L in [0, count) of one draw.

    ``traj`` is an int array of trajectory indices.  Shape (len(traj), count).
    """
    traj = np.asarray(traj, dtype=np.uint64).reshape(-1, 1)
    ngroups = (count + 3) // 4
    g = np.arange(ngroups, dtype=np.uint64).reshape(1, -1)
    x0, x1, x2, x3 = philox4x32_10(g, np.uint64(draw), traj, np.uint64(inst),
                                    np.uint64(nsp), np.uint64(key1(tag, seq)))
    n0, n1 = box_muller(x0, x1)
    n2, n3 = box_muller(x2, x3)
    out = np.stack([n0, n1, n2, n3], axis=-1).reshape(traj.shape[0], 4 * ngroups)
    return out[:, :count]


def uniforms(nsp, tag, draw, traj, count, seq=0, inst=0):
    """Uniforms in (0,1) U[traj, L], same counter layout as ``normals``."""
    traj = np.asarray(traj, dtype=np.uint64).reshape(-1, 1)
    ngroups = (count + 3) // 4
    g = np.arange(ngroups, dtype=np.uint64).reshape(1, -1)
    xs = philox4x32_10(g, np.uint64(draw), traj, np.uint64(inst),
                       np.uint64(nsp), np.uint64(key1(tag, seq)))
    out = np.stack([uniform_open01(x) for x in xs], axis=-1).reshape(traj.shape[0], 4 * ngroups)
    return out[:, :count]
