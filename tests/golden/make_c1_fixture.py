"""Writes c1_oracle_impedances.npz: the FP64 oracle's impedances for the
32x32x32 block half-space (3 EXCMG levels, 1 Hz, both polarisations, 10x10
stations), solved to an interior residual of 1e-11 with COCR.  The GPU test
compares its own solve against these values without running the oracle on
the GPU host (the oracle takes minutes there).

    python tests/golden/make_c1_fixture.py
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


def main():
    m = synth.make_block_halfspace_32()
    t0 = time.time()
    ex, st = O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs,
                          O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior"))
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2)
    zxy, zyx = res[:, 0] + 1j * res[:, 1], res[:, 2] + 1j * res[:, 3]
    out = os.path.join(HERE, "c1_oracle_impedances.npz")
    np.savez(out, zxy=zxy, zyx=zyx, freqs=np.asarray(m.freqs, float), recv=np.asarray(m.recv, float),
             sigma_sum=float(np.sum(m.sigma)))
    print(f"wrote {out}: {len(zxy)} stations, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
