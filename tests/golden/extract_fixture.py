"""Extract the reference's golden NSVEparticles checkpoint into .npz fixtures.

Source: ``bfps/test/B32p1e4_checkpoint_0.h5`` (read by ``h5mini``), the data
file pinned by ``bfps/test/test_bfps_NSVEparticles.py:35-50``.  It holds a
32^3 single-precision NSVEparticles run with 10^4 tracers (neighbours 1,
smoothness 1, 4 Adams-Bashforth steps) at iterations 0, 32 and 64.

Output: ``tests/golden/B32p1e4.npz`` with arrays

* ``vort_<it>``   complex64 ``[ky=32][kz=32][kx=17][3]`` (transposed layout)
* ``state_<it>``  float64   ``[10000][3]`` tracer positions (sorted by index)
* ``rhs_<it>``    float64   ``[4][10000][3]`` Adams-Bashforth history

Run: ``python tests/golden/extract_fixture.py`` (needs the reference tree).
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import h5mini  # noqa: E402

SRC = "/root/reference/bfps/test/B32p1e4_checkpoint_0.h5"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "B32p1e4.npz")


def main():
    f = h5mini.File(SRC)
    out = {}
    for it in (0, 32, 64):
        v = f.get("vorticity/complex/%d" % it).read()
        out["vort_%d" % it] = (v["r"] + 1j * v["i"]).astype(np.complex64)
        out["state_%d" % it] = f.get("tracers0/state/%d" % it).read()
        out["rhs_%d" % it] = f.get("tracers0/rhs/%d" % it).read()
    np.savez_compressed(DST, **out)
    print("wrote", DST, os.path.getsize(DST), "bytes")


if __name__ == "__main__":
    main()
