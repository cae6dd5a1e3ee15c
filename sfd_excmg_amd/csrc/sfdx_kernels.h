// sfdx_kernels.h -- launcher declarations and the device-resident Krylov scalars.
#pragma once
#include "sfdx_internal.h"

namespace sfdx {

// Device scalar slots (doubles).  Reductions land in SL_TMP.. and a scalar
// step (executed by the last thread of the second-stage reduction) turns them
// into the Krylov coefficients, so no host round trip is needed per step.
enum Slot {
    SL_RHO = 0,      // complex
    SL_ALPHA = 2,    // complex
    SL_BETA = 4,     // complex
    SL_OMEGA = 6,    // complex
    SL_TMP = 8,      // 8 doubles of reduction output
    SL_RN = 16,      // |r_unscaled|^2
    SL_PRHO = 18,    // PCG rho (real)
    SL_PALPHA = 20,  // complex
    SL_PBETA = 22,   // complex
    SL_PR = 24,      // PCG |r|^2
    SL_T = 26,       // |t|^2 of the DC right-hand side
    SL_BREAK = 28,   // breakdown flag
    SL_ETA = 30,     // complex (GPBiCG)
    SL_COUNT = 32
};

enum ScalarOp {
    OP_NONE = 0,
    OP_RN = 1,          // RN = tmp[0]
    OP_COCR_RHO = 2,    // RHO = tmp[0..1]
    OP_COCR_ALPHA = 3,  // ALPHA = RHO / tmp
    OP_COCR_BETA = 4,   // BETA = tmp / RHO ; RHO = tmp
    OP_BICG_RHO = 5,    // BETA = (tmp/RHO)(ALPHA/OMEGA); RHO = tmp
    OP_BICG_ALPHA = 6,  // ALPHA = RHO / tmp
    OP_BICG_OMEGA = 7,  // OMEGA = tmp[0..1] / tmp[2]
    OP_PCG_INIT = 8,    // PRHO = tmp[0]
    OP_PCG_ALPHA = 9,   // PALPHA = PRHO / tmp
    OP_PCG_BETA = 10,   // PBETA = tmp[0]/PRHO ; PRHO = tmp[0] ; PR = tmp[1]
    OP_T = 11,          // T = tmp[0]
    OP_BICG_INIT = 12,  // RHO = ALPHA = OMEGA = 1
    OP_GP_INIT = 13,    // RHO = tmp[0..1]; BETA = 0
    OP_GP_OMEGA1 = 14,  // first GPBiCG step: OMEGA = (at,t)/(at,at); ETA = 0
    OP_GP_OMEGA = 15,   // OMEGA, ETA from the 2x2 minimisation (tmp[0..7])
    OP_GP_BETA = 16     // RN = tmp[0]; BETA = (ALPHA/OMEGA) tmp[1..2]/RHO; RHO = tmp[1..2]
};

__device__ __forceinline__ void cdiv(double ar, double ai, double br, double bi, double &cr, double &ci)
{
    double den = br * br + bi * bi;
    cr = (ar * br + ai * bi) / den;
    ci = (ai * br - ar * bi) / den;
}

__device__ __forceinline__ void scalar_step(double *s, int op)
{
    double *t = s + SL_TMP;
    switch (op) {
    case OP_RN: s[SL_RN] = t[0]; break;
    case OP_COCR_RHO:
        s[SL_RHO] = t[0]; s[SL_RHO + 1] = t[1];
        s[SL_BETA] = 0; s[SL_BETA + 1] = 0;  // p = z at a fresh start (direction folded one step behind)
        s[SL_BREAK] = 0;  // fresh start of the recurrences
        break;
    case OP_COCR_ALPHA:
        cdiv(s[SL_RHO], s[SL_RHO + 1], t[0], t[1], s[SL_ALPHA], s[SL_ALPHA + 1]);
        if (!isfinite(s[SL_ALPHA]) || !isfinite(s[SL_ALPHA + 1])) s[SL_BREAK] = 1;
        break;
    case OP_COCR_BETA:
        cdiv(t[0], t[1], s[SL_RHO], s[SL_RHO + 1], s[SL_BETA], s[SL_BETA + 1]);
        s[SL_RHO] = t[0]; s[SL_RHO + 1] = t[1];
        if (!isfinite(s[SL_BETA]) || !isfinite(s[SL_BETA + 1])) s[SL_BREAK] = 1;
        break;
    case OP_BICG_RHO: {
        double qr, qi, ar, ai;
        cdiv(t[0], t[1], s[SL_RHO], s[SL_RHO + 1], qr, qi);
        cdiv(s[SL_ALPHA], s[SL_ALPHA + 1], s[SL_OMEGA], s[SL_OMEGA + 1], ar, ai);
        s[SL_BETA] = qr * ar - qi * ai;
        s[SL_BETA + 1] = qr * ai + qi * ar;
        s[SL_RHO] = t[0]; s[SL_RHO + 1] = t[1];
        if (t[0] == 0 && t[1] == 0) s[SL_BREAK] = 1;
        break;
    }
    case OP_BICG_ALPHA:
        cdiv(s[SL_RHO], s[SL_RHO + 1], t[0], t[1], s[SL_ALPHA], s[SL_ALPHA + 1]);
        if (t[0] == 0 && t[1] == 0) s[SL_BREAK] = 1;
        break;
    case OP_BICG_OMEGA:
        s[SL_OMEGA] = t[0] / t[2];
        s[SL_OMEGA + 1] = t[1] / t[2];
        if (t[2] == 0) s[SL_BREAK] = 1;
        break;
    case OP_PCG_INIT: s[SL_PRHO] = t[0]; break;
    case OP_PCG_ALPHA:
        cdiv(s[SL_PRHO], 0.0, t[0], t[1], s[SL_PALPHA], s[SL_PALPHA + 1]);
        break;
    case OP_PCG_BETA:
        s[SL_PBETA] = t[0] / s[SL_PRHO];
        s[SL_PBETA + 1] = 0;
        s[SL_PRHO] = t[0];
        s[SL_PR] = t[1];
        break;
    case OP_T: s[SL_T] = t[0]; break;
    case OP_GP_INIT:
        s[SL_RHO] = t[0]; s[SL_RHO + 1] = t[1];
        s[SL_BETA] = 0; s[SL_BETA + 1] = 0;
        s[SL_BREAK] = 0;
        break;
    case OP_GP_OMEGA1:
        // tmp: [0] (y,y) [1] (at,at) [2..3] (at,t) [4..5] (y,t) [6..7] (at,y), conj first argument
        s[SL_OMEGA] = t[2] / t[1];
        s[SL_OMEGA + 1] = t[3] / t[1];
        s[SL_ETA] = 0; s[SL_ETA + 1] = 0;
        if (t[1] == 0) s[SL_BREAK] = 1;
        break;
    case OP_GP_OMEGA: {
        // den = (at,at)(y,y) - (y,at)(at,y) = (at,at)(y,y) - |(at,y)|^2  (real)
        const double den = t[1] * t[0] - (t[6] * t[6] + t[7] * t[7]);
        // omega = ((y,y)(at,t) - (y,t)(at,y)) / den
        s[SL_OMEGA] = (t[0] * t[2] - (t[4] * t[6] - t[5] * t[7])) / den;
        s[SL_OMEGA + 1] = (t[0] * t[3] - (t[4] * t[7] + t[5] * t[6])) / den;
        // eta = ((at,at)(y,t) - (y,at)(at,t)) / den with (y,at) = conj((at,y))
        s[SL_ETA] = (t[1] * t[4] - (t[6] * t[2] + t[7] * t[3])) / den;
        s[SL_ETA + 1] = (t[1] * t[5] - (t[6] * t[3] - t[7] * t[2])) / den;
        if (den == 0) s[SL_BREAK] = 1;
        break;
    }
    case OP_GP_BETA: {
        s[SL_RN] = t[0];
        double qr, qi, ar, ai;
        cdiv(t[1], t[2], s[SL_RHO], s[SL_RHO + 1], qr, qi);
        cdiv(s[SL_ALPHA], s[SL_ALPHA + 1], s[SL_OMEGA], s[SL_OMEGA + 1], ar, ai);
        s[SL_BETA] = qr * ar - qi * ai;
        s[SL_BETA + 1] = qr * ai + qi * ar;
        s[SL_RHO] = t[1]; s[SL_RHO + 1] = t[2];
        if (t[1] == 0 && t[2] == 0) s[SL_BREAK] = 1;
        break;
    }
    case OP_BICG_INIT:
        s[SL_RHO] = 1; s[SL_RHO + 1] = 0;
        s[SL_ALPHA] = 1; s[SL_ALPHA + 1] = 0;
        s[SL_OMEGA] = 1; s[SL_OMEGA + 1] = 0;
        s[SL_BREAK] = 0;
        break;
    default: break;
    }
}

// Prolongation weight tables on the device (per dimension).
struct ProlongDev {
    int *nb[3] = {nullptr, nullptr, nullptr};
    double *nw[3] = {nullptr, nullptr, nullptr};
    int *eb[3] = {nullptr, nullptr, nullptr};
    double *ew[3] = {nullptr, nullptr, nullptr};
};

void reduce_final(RedBuf &rb, int64_t nblk, int K, int dst, int op, hipStream_t st);

template <typename R>
int64_t launch_spmv(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode, const cx<R> *w,
                    RedBuf &rb, hipStream_t st);  // returns the number of partial sums written (dot modes 1-3)
// tile autotuner: median of `reps` timed launches per variant, returns the
// fastest (host-synchronising; not under stream capture)
template <typename R>
int tune_spmv(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, RedBuf &rb, hipStream_t st, int reps,
              float *ms_out);
// k-marching stencil with an explicit chunk of node planes per block (kc <= 0:
// the default) and tile variant (0: 32x8 with prefetch, 1: 32x8 without, 2:
// 32x4 with prefetch, 3: 32x4 without)
template <typename R>
int64_t launch_spmv_march(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode,
                          const cx<R> *w, RedBuf &rb, hipStream_t st, int kc, int variant);
// one-row-per-thread gather form of the same operator
template <typename R>
int64_t launch_spmv_gather(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode,
                           const cx<R> *w, RedBuf &rb, hipStream_t st);
int spmv_chunk(const Dims &d, int ty, int64_t slots);
template <typename R>
void launch_mass(const DevGrid<R> &g, const cx<R> *x, cx<R> *w, hipStream_t st);
template <typename R>
void launch_ssor(const DevGrid<R> &g, R wmu, const cx<R> *in, cx<R> *out, R omega, hipStream_t st);
template <typename R>
void launch_vec(int mode, const DevGrid<R> &g, RedBuf &rb, cx<R> *v0, cx<R> *v1, cx<R> *v2, cx<R> *v3, cx<R> *v4,
                cx<R> *v5, hipStream_t st);
template <typename R>
void launch_dot(int mode, int64_t n, const cx<R> *u, const cx<R> *v, RedBuf &rb, hipStream_t st, int64_t &nblk);
template <typename R>
void launch_copy(int64_t n, const cx<R> *s, cx<R> *t, hipStream_t st);
// bandwidth-tuned streaming copy of `bytes` (a multiple of 16), units = 4 or 8
// 16-byte loads in flight per thread (measurement only)
void launch_stream_copy(size_t bytes, const void *s, void *t, int units, hipStream_t st);
template <typename RS, typename RD>
void launch_convert(int64_t n, const cx<RS> *s, cx<RD> *t, hipStream_t st);
void launch_add_lo(int64_t n, zd *x, const zf *d, hipStream_t st);
template <typename R>
void launch_axpby(int64_t n, cx<R> al, const cx<R> *u, cx<R> be, cx<R> *v, hipStream_t st);
template <typename R>
void launch_dirichlet(const Dims &d, const cx<R> *edc, int pol, cx<R> *x, int zero_inner, hipStream_t st,
                      int full = 0);  // full: edc holds the full-tensor boundary data (4 nz + 4)
template <typename R>
void launch_node_div(const DevGrid<R> &g, const cx<R> *w, cx<R> *t, RedBuf &rb, hipStream_t st, int64_t &nblk);
template <typename R>
void launch_pbuild(const Dims &d, const double *a, const double *b, const double *c, const double *sig, R *P,
                   hipStream_t st);
template <typename R>
void launch_papply(const Dims &d, const R *P, const cx<R> *x, cx<R> *y, int dot, RedBuf &rb, hipStream_t st,
                   int64_t &nblk);
template <typename R>
void launch_pcg(int mode, int64_t np, const R *Pdiag, RedBuf &rb, cx<R> *u, cx<R> *r, cx<R> *p, const cx<R> *v,
                hipStream_t st, int64_t &nblk);
template <typename R>
void launch_grad_update(const DevGrid<R> &g, const cx<R> *phi, cx<R> *x, hipStream_t st);
template <typename R>
void launch_prolong(const Dims &df, const Dims &dc, int m, const ProlongDev &T, const cx<R> *uc, cx<R> *uf,
                    hipStream_t st);
template <typename R>
void launch_hybrid(const DevGrid<R> &g, R wmu, const R *P, const cx<R> *in, cx<R> *out, cx<R> *tnode, cx<R> *phi,
                   R omega, hipStream_t st, int nodal = 0, const cx<R> *dotw = nullptr, double *partials = nullptr,
                   bool store = true);  // store false: out = edge SSOR output, only the dot of the sum (dotw required)
// COCR update (k_vec mode 6) with q = zs + s G phi formed from launch_hybrid's pieces (store false)
template <typename R>
void launch_cocr_upd_grad(const DevGrid<R> &g, RedBuf &rb, cx<R> *x, cx<R> *pv, const cx<R> *Ap, cx<R> *r, cx<R> *z,
                          const cx<R> *zs, const cx<R> *phi, R wmu, hipStream_t st);
// sigma of a coarsened grid from the fine sigma, all on the device: coarse
// cell (I, J, K) is the volume-weighted mean of the fine cells
// [sx[I], sx[I+1]) x [sy[J], sy[J+1]) x [sz[K], sz[K+1]); ac, bc, cc are the
// coarse widths (the sums of the fine widths of each range)
void launch_mesh_coarsen(const Dims &df, const Dims &dc, const int *sx, const int *sy, const int *sz, const double *a,
                         const double *b, const double *c, const double *ac, const double *bc, const double *cc,
                         const double *sig, double *sigc, hipStream_t st);
// sigma (6 x nc, component-major) in the cell-plane march's pair layout, and in
// FP32 with the diagonal components raised to floor (0: unchanged)
void launch_sigma_pairs(int64_t nc, const double *sig, double *sigp, hipStream_t st);
void launch_sigma_f32(int64_t nc, double floor, const double *sig, float *out, hipStream_t st);
// y = A_s (z + (i/wmu) G psi), with the partial sums of u^T y (u = z + (i/wmu) G psi);
// returns the number of partial sums
int64_t launch_split_product(const DevGrid<float> &g, float wmu, const zf *z, const zf *psi, zf *y, RedBuf &rb,
                             hipStream_t st);
// COCR update with complex64 split directions: node part (p_psi = z_psi + beta
// p_psi, z_psi -= alpha q_psi), then edge part (p_e = z_e + beta p_e; x +=
// alpha (p_e + s G p_psi); r -= alpha Ap; z_e -= alpha q_e; |r/ve|^2 partials)
void launch_cocr_split_upd(const DevGrid<double> &g, RedBuf &rb, zd *x, zd *r, const zf *Ap, zf *ze, const zf *qe,
                           zf *pe, zf *ppsi, zf *zpsi, const zf *qpsi, double wmu, hipStream_t st);
void launch_psym32(int64_t np, const double *P, float *Ps32, hipStream_t st);
void launch_hybrid_edge32(const DevGrid<double> &g64, const DevGrid<float> &g32, double wmu, const float *Ps32,
                          const zd *in, zd *out, zf *in32, zf *z32, zd *tnode, zd *phi, float omega,
                          hipStream_t st, const zd *dotw = nullptr, double *partials = nullptr,
                          bool convert_in = true,  // convert_in false: in32 already holds in
                          bool store = true);      // store false: only the dot (dotw required)
// COCR update (k_vec mode 6) with q formed from z32 and phi of launch_hybrid_edge32
void launch_cocr_upd_hyb(const DevGrid<double> &g, RedBuf &rb, zd *x, zd *pv, const zd *Ap, zd *r, zd *z,
                         const zf *z32, const zd *phi, double wmu, hipStream_t st);
// the same with the nodal part in complex64 (the complex64 V-cycle's)
void launch_cocr_upd_hyb(const DevGrid<double> &g, RedBuf &rb, zd *x, zd *pv, const zd *Ap, zd *r, zd *z,
                         const zf *z32, const zf *phi, double wmu, hipStream_t st);
// Ap = Az + beta Ap (copy: Ap = Az); ap32, when set, receives a complex64 copy
template <typename R>
void launch_cocr_ap(int64_t n, RedBuf &rb, cx<R> *Ap, const cx<R> *Az, zf *ap32, bool copy, hipStream_t st);
void launch_hybrid_mixed(const DevGrid<double> &g64, const DevGrid<float> &g32, double wmu, const float *P32,
                         const zd *in, zd *out, zf *in32, zf *z32, zf *tnode, zf *phi, float omega, hipStream_t st,
                         int nodal = 0, const zd *dotw = nullptr, double *partials = nullptr);
// One hybrid Gauss-Seidel step: out = GS_e(in), phi = GS_n(G^T in); with
// add_grad the nodal part is folded in (out += (i/wmu) G phi).
template <typename R>
void launch_hybrid_gs(const DevGrid<R> &g, R wmu, const R *P, const cx<R> *in, cx<R> *out, cx<R> *tnode, cx<R> *phi,
                      int dir, hipStream_t st, bool add_grad = true,
                      int nodal = 0);  // nodal 0: 8-colour GS on P in the sweep's direction; 1: Jacobi polynomial
// t = r - A (z + (i/wmu) G psi) on interior edges (0 on boundary edges), formed
// as r - K z + i wmu M (z + (i/wmu) G psi) so that the (possibly huge)
// gradient part never meets the curl-curl term (K G = 0).
template <typename R>
void launch_resid_split(const DevGrid<R> &g, R wmu, const cx<R> *z, const cx<R> *psi, const cx<R> *r, cx<R> *t,
                        hipStream_t st);
// Transfers between a grid and a coarser one whose cells each merge one or two
// consecutive fine cells along every axis a: st[a][c] is the fine node index
// of coarse node c (nc + 1 entries), lo[a][q] the coarse node at or below fine
// node q and wl[a][q] the linear weight of that coarse node (exactly 1 where
// fine node q is a coarse node, below 1 elsewhere), nf + 1 entries each.
// Standard 2:1 coarsening is st[c] = 2c.
struct MGTables {
    const double *wl[3];
    const int *lo[3];
    const int *st[3];
};
template <typename R>
void launch_prolong_node_add(const Dims &df, const Dims &dc, const MGTables &T, const cx<R> *pc, cx<R> *pf,
                             hipStream_t st);
// z += s G phi; with dotw the partial sums of dotw^T (z + s G phi), and with
// store false z is left unchanged (only the dot)
template <typename R>
void launch_grad_add(const DevGrid<R> &g, cx<R> s, const cx<R> *phi, cx<R> *z, hipStream_t st,
                     const cx<R> *dotw = nullptr, double *partials = nullptr, bool store = true);
// z64 = widen(z32) + s G phi, phi in complex64; dotw / store as launch_grad_add
void launch_hyb_combine(const DevGrid<double> &g64, zd s, const zf *phi, const zf *z32, zd *z64, hipStream_t st,
                        const zd *dotw = nullptr, double *partials = nullptr, bool store = true);
template <typename R>
void launch_restrict(const Dims &df, const Dims &dc, const MGTables &T, const cx<R> *rf, cx<R> *rc, hipStream_t st);
template <typename R>
void launch_prolong_add(const Dims &df, const Dims &dc, const MGTables &T, const cx<R> *zc, cx<R> *zf,
                        hipStream_t st);
// GPBiCG work vectors (pgpbicg, src/solvers.f90:137-315)
template <typename R>
struct GPVecs {
    cx<R> *p, *rp, *u, *w, *at, *ap, *y, *t, *r, *q, *tp, *tmp, *z, *x, *r0;
};
// mode 0: w = at + beta ap; p = rp + beta (p - u)
// mode 1: y = t - r - alpha w + alpha ap; t = r - alpha ap
// mode 2: q = tp - rp + beta u; tp = rp - alpha tmp
// mode 3: partial sums of (y,y) (at,at) (at,t) (y,t) (at,y) (conjugated first argument)
// mode 4: u = omega tmp + eta q; z = omega rp + eta z - alpha u; x += alpha p + z;
//         r = t - eta y - omega at; partial sums |r/ve|^2 (interior) and (r0,r)
template <typename R>
void launch_gpbicg(int mode, const DevGrid<R> &g, RedBuf &rb, const GPVecs<R> &v, hipStream_t st);

// adjoint gradient accumulation over cells (see k_cell_grad) and boundary zeroing
void launch_cell_grad(const DevGrid<double> &g, double wmu, const zd *e, const zd *lam, double *grad, hipStream_t st);
void launch_zero_boundary(const Dims &d, zd *v, hipStream_t st);
// surface E and H at the cell centres of the surface node plane (nair, or surf[i + nx*j]
// per column when surf is not null), 5 arrays of nx*ny
void launch_surface_fields(const Dims &d, int nair, const int *surf, const double *a, const double *b,
                           const double *c, double om, const zd *x, zd *out, hipStream_t st);

template <typename R>
void launch_extrap(int64_t n, R c1, const cx<R> *u1, R c2, cx<R> *w, hipStream_t st);
// out[i] = 1 / w[i], i < n
void launch_reciprocal(int n, const double *w, double *out, hipStream_t st);

}  // namespace sfdx
