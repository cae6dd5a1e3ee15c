"""Writes c5_48_oracle_gradient.npz: the FP64 oracle's misfit and its
derivative with respect to log sigma on 3 cells, for the differentiation
workload's generator at 48x48x32 cells (log10 rho a seeded Gaussian random
field, mean 2, sd 0.5, correlation length 8 cells; 8 air layers; 10x10
stations; frequencies 0.1 and 10 Hz).  The observed data are the oracle's
responses of the same generator with the next seed (out_type 2), as in the
full-size workload.  Each derivative is a central difference of the oracle's
forward solve (EXCMG, COCR with the reference's lexicographic SSOR and a
divergence correction every 50 iterations, interior residual 1e-11) with step
1e-3 in log sigma: brute force, independent of any adjoint.

    python tests/golden/make_c5_48_fixture.py
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

FREQS = np.array([0.1, 10.0])
STEP = 1e-3
OPTS = O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior", maxit_excmg=5000)


def model(seed=5):
    m = synth.make_random_field_96(seed=seed, nx=48, ny=48, nz_earth=24, nair=8, nstations=10)
    m.freqs = FREQS.copy()
    return m


def cells(m):
    """3 seeded cells in the core, in the top third of the earth."""
    rng = np.random.default_rng(7)
    nx, ny = m.nx, m.ny
    out = []
    while len(out) < 3:
        i, j = (int(v) for v in rng.integers(nx // 4, 3 * nx // 4, 2))
        k = int(rng.integers(m.nair, m.nair + (m.nz - m.nair) // 3))
        c = (k * ny + j) * nx + i
        if c not in out:
            out.append(c)
    return out


def solve(m, sigma):
    ex, rep = O.fwd_solver(m.a, m.b, m.c, m.nair, m.level, 1e-11, sigma, m.freqs, OPTS)
    for fr in rep:
        for r in fr:
            assert all(s.status == 0 for s in r.stats), r
    return ex


def main():
    t0 = time.time()
    m = model()
    mo = model(6)
    obs, _ = O.post_process(mo.a, mo.b, mo.c, mo.nair, mo.freqs, mo.recv, solve(mo, mo.sigma), out_type=2)
    print(f"observed data {time.time() - t0:.1f} s", flush=True)
    phi = O.misfit(m.a, m.b, m.c, m.nair, m.freqs, m.recv, solve(m, m.sigma), obs)
    cl = cells(m)
    grad = []
    for cell in cl:
        sp_, sm_ = m.sigma.copy(), m.sigma.copy()
        sp_[cell] *= np.exp(STEP)
        sm_[cell] *= np.exp(-STEP)
        fp = O.misfit(m.a, m.b, m.c, m.nair, m.freqs, m.recv, solve(m, sp_), obs)
        fm = O.misfit(m.a, m.b, m.c, m.nair, m.freqs, m.recv, solve(m, sm_), obs)
        grad.append((fp - fm) / (2 * STEP))
        print(f"cell {cell}: {grad[-1]:.6e} ({time.time() - t0:.1f} s)", flush=True)
    out = os.path.join(HERE, "c5_48_oracle_gradient.npz")
    np.savez(out, misfit=phi, cells=np.array(cl), grad=np.array(grad), obs=obs, freqs=FREQS,
             sigma_sum=float(np.sum(m.sigma)), step=STEP)
    print(f"wrote {out}: misfit {phi:.6e}, {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
