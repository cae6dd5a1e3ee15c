"""Writes c2_oracle_impedances.npz: the FP64 oracle's impedances for the
benchmark's full-tensor anisotropic block model (128x128x96 cells, 16 seeded
blocks with rotated tensors, 0.1 Hz, both polarisations, 4 EXCMG levels,
40x40 stations), solved with COCR to an interior residual of 1e-10 (tens of
minutes on one core).  The GPU test compares against these values without
running the oracle on the GPU host.

    python tests/golden/make_c2_fixture.py
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


def model():
    return synth.make_aniso_blocks_128()


def main():
    m = model()
    t0 = time.time()
    ex, st = O.fwd_solver(m.a, m.b, m.c, m.nair, m.level, 1e-10, m.sigma, m.freqs,
                          O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior"))
    print("oracle solve", time.time() - t0, "s; stats", st, flush=True)
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2)
    out = os.path.join(HERE, "c2_oracle_impedances.npz")
    np.savez(out, zxy=res[:, 0] + 1j * res[:, 1], zyx=res[:, 2] + 1j * res[:, 3],
             freqs=np.asarray(m.freqs, float), recv=np.asarray(m.recv, float), sigma_sum=float(np.sum(m.sigma)))
    print(f"wrote {out}: {len(res)} stations, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
