"""Writes tests/golden/lbgk_fixtures.npz: self-generated regression pins.

ASL ships no LBGK test vectors, so these fixtures come from this project's
CPU oracle (oracle/, the C++ restatement of ASL's algorithm) and pin its
results -- and the GPU's, which must equal them bit for bit -- against later
changes.  Cases (ghost-inclusive arrays, interior values only are compared):

* cavity8_q{15,19,27}: 8^3 lid-driven cavity, FP64, nu = 1/30, lid u = (0.05, 0, 0)
  on ZE executed first, no-slip on the other faces; f after 10 steps, rho and
  u after 100 steps.
* channel_q19_f32: 16x8x8 channel past a normalised level-set sphere (r = 2.5
  at (5, 4, 4)), FP32, nu = 0.02, velocity inlet X0 u = (0.04, 0, 0), pressure
  outlet XE, no-slip walls and sphere; rho and u after 50 steps.

Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from asl_amd import geometry  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbgk_fixtures.npz")


def cavity(q, steps):
    O = oracle.OracleLattice((8, 8, 8), q, fp64=True, nu=1.0 / 30.0)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, 32, u=(0.05, 0.0, 0.0)), O.add_slice_bc(O.NOSLIP, 31)]
    for b in ids:
        O.execute_bc(b)
    O.run(steps, ids)
    return O


def channel():
    size = (16, 8, 8)
    body = geometry.normalize(geometry.sphere(size, 2.5, (5.0, 4.0, 4.0)))
    O = oracle.OracleLattice(size, 19, fp64=False, nu=0.02)
    O.init_f(1.0, (0.04, 0.0, 0.0))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, 1, u=(0.04, 0.0, 0.0)), O.add_slice_bc(O.CONSTANT_PRESSURE, 2, p=1.0),
           O.add_slice_bc(O.NOSLIP, 4 | 8 | 16 | 32), O.add_map_bc(O.NOSLIP_MAP, body)]
    for b in ids:
        O.execute_bc(b)
    O.run(50, ids)
    return O, body


def main():
    data = {}
    for q in (15, 19, 27):
        data["cavity8_q%d_f10" % q] = cavity(q, 10).f()
        O = cavity(q, 100)
        data["cavity8_q%d_rho100" % q] = O.rho()
        data["cavity8_q%d_u100" % q] = O.velocity()
    O, body = channel()
    data["channel_q19_f32_map"] = body
    data["channel_q19_f32_rho50"] = O.rho().astype(np.float32)
    data["channel_q19_f32_u50"] = O.velocity().astype(np.float32)
    np.savez_compressed(OUT, **data)
    print("wrote %s (%d arrays, %d bytes)" % (OUT, len(data), os.path.getsize(OUT)))


if __name__ == "__main__":
    main()
