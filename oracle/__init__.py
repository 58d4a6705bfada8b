"""FP64 CPU oracle for the xSPDE ensemble integration path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It restates, in numpy/scipy float64, the reference algorithms of
peterddrummond/xspde_matlab (xSPDE 4.21):

* ``philox``     - the counter-based random stream shared with the GPU
                   (the reference uses MATLAB ``randn``/Threefry,
                   Simulator/xensemble.m:28-32, which cannot be reproduced;
                   parity at that boundary is therefore statistical only).
* ``transforms`` - xdct1/2/3, xdst1/2/3, fft/ifft and xtransform
                   (Applications/SPDE/xdct1.m .. xdst3.m, xtransform.m).
* ``xspde``      - xpreferences/spdpreferences/xgrid (lattice and scales),
                   xpropfactor/xprop, RK4 and MP interaction-picture steps,
                   xnoise/xrandom, xpath, xdata, xensemble, xsim statistics.

Pinning status of every function is listed in each module's header.
"""
