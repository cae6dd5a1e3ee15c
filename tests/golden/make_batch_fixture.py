"""Writes batch_oracle_impedances.npz: the FP64 oracle's impedances for the two
smallest 4-level models of the batch workload's 8 isotropic models of uneven
sizes (seed 4): 48x48x48 and 64x64x48 (the 3-level 56^3 model's second coarse
level, 28^3, is beyond SciPy's direct solver here), 8 air layers, seeded
random blocks in [1,1000] ohm-m over 100 ohm-m, 1 Hz, 10x10 stations, COCR
to an interior residual of 1e-10 (iteration cap raised to 5000 per level, every
level checked for convergence).  The GPU test runs all 8 models through the
batched entry point and compares these two with the oracle.

    python tests/golden/make_batch_fixture.py
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

PICK = (0, 2)   # 48^3 and 64x64x48 in synth.UNEVEN_ISO_SHAPES


def main():
    t0 = time.time()
    ms = synth.make_uneven_iso_models()
    out = {}
    for q in PICK:
        m = ms[q]
        # the reference caps EXCMG levels at 1000 iterations (fdfe_main.f90:1414); the
        # 64x64x48 model's fine level needs more, so the cap is raised and every
        # level's convergence is checked
        ex, rep = O.fwd_solver(m.a, m.b, m.c, m.nair, m.level, 1e-10, m.sigma, m.freqs,
                               O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior", maxit_excmg=5000))
        for r in rep[0]:
            assert all(s.status == 0 for s in r.stats), (q, r)
        res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2)
        out[f"zxy_{q}"] = res[:, 0] + 1j * res[:, 1]
        out[f"zyx_{q}"] = res[:, 2] + 1j * res[:, 3]
        out[f"sigma_sum_{q}"] = float(np.sum(m.sigma))
        print(f"model {q} {m.nx}x{m.ny}x{m.nz}: {time.time() - t0:.1f} s", flush=True)
    path = os.path.join(HERE, "batch_oracle_impedances.npz")
    np.savez(path, pick=np.array(PICK), **out)
    print(f"wrote {path}, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
