"""Writes topo64_oracle_impedances.npz: the FP64 oracle's impedances and
tippers for the topography workload's generator at 64x64x48 cells (the
128x128x96 full-tensor block model at half resolution with the seeded
undulating ground carved into it, up to ~950 m below the datum), 0.1 Hz, both
polarisations, 4 EXCMG levels, 20x20 stations, solved with COCR to an
interior residual of 1e-10.  The station fields are taken on each column's
ground surface (the topography branch of post_process, surface planes read off
the carved conductivity).  The GPU test compares against these values without
running the oracle on the GPU host.

    python tests/golden/make_topo64_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from sfd_excmg_amd import synth  # noqa: E402

AIR_MAX = 1e-6   # S/m: cells at or below this conductivity are air


def model():
    return synth.make_aniso_topography_128(nx=64, nz_earth=40, nair=8, nstations=20)


def main():
    m = model()
    t0 = time.time()
    ex, st = O.fwd_solver(m.a, m.b, m.c, m.nair, m.level, 1e-10, m.sigma, m.freqs,
                          O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior"))
    print("oracle solve", time.time() - t0, "s; stats", st, flush=True)
    surf = O.surface_planes(m.a, m.b, m.c, m.sigma, AIR_MAX)
    res, tip = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2, surf=surf)
    out = os.path.join(HERE, "topo64_oracle_impedances.npz")
    np.savez(out, zxy=res[:, 0] + 1j * res[:, 1], zyx=res[:, 2] + 1j * res[:, 3],
             tx=tip[:, 0] + 1j * tip[:, 1], ty=tip[:, 2] + 1j * tip[:, 3], surf=surf.astype(np.int32),
             freqs=np.asarray(m.freqs, float), recv=np.asarray(m.recv, float), sigma_sum=float(np.sum(m.sigma)))
    print(f"wrote {out}: {len(res)} stations, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
