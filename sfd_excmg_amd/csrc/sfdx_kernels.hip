// sfdx_kernels.hip -- gfx950 kernels of the MT forward solve.
//
//  K1 k_spmv_march  matrix-free 13-point edge stencil y = alpha*A_s x, marching
//                   up node planes with LDS-staged planes, optional fused dot
//                   products (replaces zamux on the CSR of fd_assemb + dcfast,
//                   sfd_excmg/src/div_correc.f90:431-449), and with a node
//                   potential ring the split residual of the multigrid
//                   V-cycle; k_spmv is the one-row-per-thread form (comparison)
//  K2 k_ssor        one colour sweep of multicolour SSOR (replaces the
//                   lexicographic ssorfac + lusol, src/div_correc.f90:171-220,
//                   src/solvers.f90:721-762)
//  K3 k_vec, k_cocr_ap, k_gpbicg, k_dot
//                   fused Krylov vector updates with wavefront/block partial
//                   reductions and a deterministic second-stage reduction
//  K4 k_prolong     EXCMG block prolongation with separable Lagrange weights
//                   (src/excmg.f90:357-772)
//  K5 k_mass, k_node_div, k_pbuild, k_psym, k_papply, k_grad_update
//                   device divergence correction (src/div_correc.f90:11-147)
//  K6 k_dirichlet   boundary data from the 1-D field (src/fdfe_main.f90:470-494)
//  K7 k_gradT, k_pjacobi, k_pssor, k_grad_add, k_hyb_combine
//                   nodal half of the hybrid preconditioner
//  K8 k_restrict, k_prolong_add, k_prolong_node_add
//                   multigrid transfers
//  K9 k_cell_grad, k_surface_fields, k_zero_boundary
//                   adjoint gradient and device post-process
//
// Launch geometry: edge kernels use one grid row (blockIdx.y) per
// (orientation, k) slice and 256 threads over the contiguous (i,j) plane of that
// slice, so consecutive lanes touch consecutive edges and the (o,i,j,k) decode
// needs one 32-bit division.  Reductions: each block writes its partial sums
// (wavefront shuffle, then LDS across the 4 wavefronts) and k_reduce_final sums
// the partials in a fixed order, so results are run-to-run deterministic.
#include "sfdx_kernels.h"

#include <stdexcept>
#include <cstdint>

namespace sfdx {

HipError::HipError(hipError_t c, const char *expr, const char *file, int line) : code(c)
{
    what = std::string(hipGetErrorString(c)) + " at " + file + ":" + std::to_string(line) + " (" + expr + ")";
}

static constexpr int TPB = 256;

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8 XCDs
// (linear block id mod 8), and each XCD has its own L2.  The launchers pad every
// 2-D grid to a multiple of 8 blocks, and the kernels renumber their block so
// that XCD x runs the contiguous tile range [x n/8, (x+1) n/8).  Tiles are
// ordered plane by plane (and, for edges, orientation within a plane), so the
// k-1/k+1 planes and the other orientations a stencil row reads are fetched
// into the same L2 instead of into several.
struct Tile {
    int x, y;
};
__device__ __forceinline__ Tile tile_id()
{
    const unsigned gx = gridDim.x, n = gx * gridDim.y;
    const unsigned lin = blockIdx.y * gx + blockIdx.x;
    const unsigned t = (lin & 7u) * (n >> 3) + (lin >> 3);
    Tile T;
    T.y = (int)(t / gx);
    T.x = (int)(t - (unsigned)T.y * gx);
    return T;
}
// smallest gy' >= gy with gx * gy' a multiple of 8
static unsigned pad_y(unsigned gx, unsigned gy)
{
    while ((gx * gy) % 8u) gy++;
    return gy;
}
static constexpr int kNoTile = 1 << 30;  // thread index of a padding tile: fails every range test

// ------------------------------------------------------------------ slices
struct Slice {
    int o, k, nI, nJ;
    int64_t base;
};

__device__ __forceinline__ Slice edge_slice(const Dims &d, int s)
{
    // slice s = 3k + o: the three orientations of plane k are adjacent tiles
    Slice q;
    q.k = s / 3;
    q.o = s - 3 * q.k;
    if (q.k > d.nz || (q.o == 2 && q.k == d.nz)) {  // padding tile
        q.o = 0; q.k = 0; q.nI = 0; q.nJ = 0; q.base = 0;
        return q;
    }
    if (q.o == 0) {
        q.nI = d.nx; q.nJ = d.ny + 1;
        q.base = (int64_t)q.k * q.nI * q.nJ;
    } else if (q.o == 1) {
        q.nI = d.nx + 1; q.nJ = d.ny;
        q.base = d.XN + (int64_t)q.k * q.nI * q.nJ;
    } else {
        q.nI = d.nx + 1; q.nJ = d.ny + 1;
        q.base = d.XN + d.YN + (int64_t)q.k * q.nI * q.nJ;
    }
    return q;
}

SliceLaunch edge_launch(const Dims &d)
{
    SliceLaunch L;
    int64_t plane = (int64_t)(d.nx + 1) * (d.ny + 1);
    L.nslices = 3 * d.nz + 2;
    const unsigned gx = (unsigned)((plane + TPB - 1) / TPB);
    L.grid = dim3(gx, pad_y(gx, (unsigned)L.nslices), 1);
    L.block = dim3(TPB, 1, 1);
    return L;
}

SliceLaunch node_launch(const Dims &d)
{
    SliceLaunch L;
    int64_t plane = (int64_t)(d.nx + 1) * (d.ny + 1);
    L.nslices = d.nz + 1;
    const unsigned gx = (unsigned)((plane + TPB - 1) / TPB);
    L.grid = dim3(gx, pad_y(gx, (unsigned)L.nslices), 1);
    L.block = dim3(TPB, 1, 1);
    return L;
}

// Compact colour mapping: thread p of slice k -> (i,j) of colour (o,par).
SliceLaunch colour_launch(const Dims &d, int colour)
{
    int o = colour >> 1;
    int W = (d.nx + 2) / 2;
    int64_t plane;
    int ns;
    if (o == 0) { plane = (int64_t)((d.ny + 2) / 2) * d.nx; ns = d.nz + 1; }
    else if (o == 1) { plane = (int64_t)d.ny * W; ns = d.nz + 1; }
    else { plane = (int64_t)(d.ny + 1) * W; ns = d.nz; }
    SliceLaunch L;
    L.nslices = ns;
    const unsigned gx = (unsigned)((plane + TPB - 1) / TPB);
    L.grid = dim3(gx, pad_y(gx, (unsigned)ns), 1);
    L.block = dim3(TPB, 1, 1);
    return L;
}

template <int O>
__device__ __forceinline__ bool colour_decode(const Dims &d, int par, int k, int p, int &i, int &j)
{
    if (O == 0) {
        int j0 = (par - k) & 1;
        int r = p / d.nx;
        i = p - r * d.nx;
        j = j0 + 2 * r;
        return j <= d.ny;
    } else if (O == 1) {
        int W = (d.nx + 2) / 2;
        j = p / W;
        int q = p - j * W;
        i = ((par - k) & 1) + 2 * q;
        return j < d.ny && i <= d.nx;
    } else {
        int W = (d.nx + 2) / 2;
        j = p / W;
        int q = p - j * W;
        i = ((par - j) & 1) + 2 * q;
        return j <= d.ny && i <= d.nx;
    }
}

// ------------------------------------------------------------------ reductions
template <int K, int NTHR = TPB>
__device__ __forceinline__ void block_reduce_store(double (&v)[K], double *partials)
{
    __shared__ double sm[K][NTHR / 64];
#pragma unroll
    for (int q = 0; q < K; q++) {
        double t = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        v[q] = t;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < K; q++) sm[q][wid] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
        const int64_t nblk = (int64_t)gridDim.x * gridDim.y;
#pragma unroll
        for (int q = 0; q < K; q++) {
            double t = 0;
            for (int w = 0; w < NTHR / 64; w++) t += sm[q][w];
            partials[q * nblk + blk] = t;  // component-major: the final pass reads contiguously
        }
    }
}

// Sum partials[K][nblk] (component-major) into slots[dst..dst+K) then apply the
// scalar step `op` (single thread).  1024 threads, fixed summation order.
__global__ __launch_bounds__(1024) void k_reduce_final(const double *__restrict__ partials, int64_t nblk, int K,
                                                       double *slots, int dst, int op)
{
    __shared__ double sm[8][16];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // four independent loads in flight per component, summed in increasing b
    const int64_t step = blockDim.x;
    int64_t b = threadIdx.x;
    for (; b + 3 * step < nblk; b += 4 * step)
        for (int q = 0; q < K; q++) {
            const double *pq = partials + q * nblk;
            const double v0 = pq[b], v1 = pq[b + step], v2 = pq[b + 2 * step], v3 = pq[b + 3 * step];
            acc[q] += v0;
            acc[q] += v1;
            acc[q] += v2;
            acc[q] += v3;
        }
    for (; b < nblk; b += step)
        for (int q = 0; q < K; q++) acc[q] += partials[q * nblk + b];
    for (int q = 0; q < K; q++) {
        double t = acc[q];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        acc[q] = t;
    }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < K; q++) sm[q][wid] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < K; q++) {
            double t = 0;
            for (int w = 0; w < (int)(blockDim.x / 64); w++) t += sm[q][w];
            slots[dst + q] = t;
        }
        scalar_step(slots, op);
    }
}

// ------------------------------------------------------------------ K1 stencil
template <typename R>
__device__ __forceinline__ cx<R> coef_of(R kg, R km, R wmu) { return cx<R>(kg, -wmu * km); }

template <typename R>
__device__ __forceinline__ cx<R> row_apply(const Row<R> &r, R wmu, const cx<R> *__restrict__ x, cx<R> xe)
{
    cx<R> acc = coef_of(r.kgd, r.kmd, wmu) * xe;
#pragma unroll
    for (int s = 0; s < 12; s++) acc = fma_c(acc, coef_of(r.kg[s], r.km[s], wmu), x[r.nb[s]]);
    return acc;
}

template <typename R, int DOT>
__global__ __launch_bounds__(TPB) void k_spmv(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                              const R *__restrict__ c, const R *__restrict__ ra,
                                              const R *__restrict__ rb, const R *__restrict__ rc,
                                              const R *__restrict__ sig, R wmu,
                                              const cx<R> *__restrict__ x, cx<R> *__restrict__ y, R alpha,
                                              const cx<R> *__restrict__ w, double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        cx<R> out(0);
        if (edge_interior(d, S.o, i, j, S.k)) {
            Row<R> r;
            row_coeffs<R, true>(d, a, b, c, sig, S.o, i, j, S.k, r, ra, rb, rc);
            out = alpha * row_apply(r, wmu, x, x[e]);
            if (DOT == 4) out = w[e] + out;
        }
        y[e] = out;
        if (DOT == 1) {  // x^T y
            cx<R> xv = x[e];
            acc[0] = (double)xv.re * out.re - (double)xv.im * out.im;
            acc[1] = (double)xv.re * out.im + (double)xv.im * out.re;
        } else if (DOT == 2) {  // conj(w)^T y
            cx<R> wv = w[e];
            acc[0] = (double)wv.re * out.re + (double)wv.im * out.im;
            acc[1] = (double)wv.re * out.im - (double)wv.im * out.re;
        } else if (DOT == 3) {  // conj(y)^T w and |y|^2
            cx<R> wv = w[e];
            acc[0] = (double)out.re * wv.re + (double)out.im * wv.im;
            acc[1] = (double)out.re * wv.im - (double)out.im * wv.re;
            acc[2] = (double)out.re * out.re + (double)out.im * out.im;
        }
    }
    if (DOT >= 1 && DOT <= 3) block_reduce_store<4>(acc, partials);
}

// ------------------------------------------------------------------ K1 k-marching stencil
// The operator of k_spmv, organised so that each value is fetched from memory
// about once per block instead of once per row that reads it.  A block owns a
// TX x TY tile of node columns (i,j) and marches up a chunk of node planes; at
// step k every column of the tile produces its rows Ex(i,j,k), Ey(i,j,k) and
// Ez(i,j,k).  Those rows read Ex/Ey at planes k and k+1, Ez at k-1 and k, and
// the cell tensors of the cell planes k-1 and k.  These planes live in LDS
// rings of two planes over the tile plus a one-column halo: at each step one
// thread per footprint column loads Ex/Ey(k+1), Ez(k) and sigma(k) (9 vector
// loads for 3 rows, against 23 per row when every row gathers its 13
// neighbours and 12 cell values itself) into the ring slots that held the
// planes no longer needed.  Each thread forms the rows of one tile column and
// loads one or two footprint columns (the tile plus its halo), so every
// wavefront both loads and computes.
namespace mt {
template <int TY>
struct Geo {
    static constexpr int TX = 32;
    static constexpr int PX = TX + 2, PY = TY + 2;  // edge footprint: i0-1..i0+TX, j0-1..j0+TY
    static constexpr int SX = TX + 1, SY = TY + 1;  // cell footprint: i0-1..i0+TX-1, j0-1..j0+TY-1
    static constexpr int NP = PX * PY, NS = SX * SY;
    static constexpr int NT = TX * TY;  // threads: one per tile column; each loads footprint columns t, t + NT
    static_assert(NP <= 2 * NT, "a thread loads at most two footprint columns");
    // LDS words of type R: Ex, Ey, Ez rings (2 complex planes each), sigma ring
    // (2 planes x 6 components), widths a, 1/a over i0-1..i0+TX-1 and b, 1/b
    // over j0-1..j0+TY-1
    static constexpr int LDS_R = 3 * 2 * 2 * NP + 2 * 6 * NS + 2 * (TX + 1) + 2 * (TY + 1);
};
}  // namespace mt

// K - i*wmu*M from the real-coefficient accumulators K and M, times alpha
template <typename R>
__device__ __forceinline__ cx<R> km_combine(cx<R> K, cx<R> M, R wmu, R alpha)
{
    return cx<R>(alpha * (K.re + wmu * M.im), alpha * (K.im - wmu * M.re));
}

// the value of row e that memit stores, with the fused dot partial sums
template <typename R, int DOT>
__device__ __forceinline__ cx<R> mvalue(const cx<R> *__restrict__ w, int64_t e, cx<R> out, bool interior, cx<R> xe,
                                        double (&acc)[4])
{
    if ((DOT == 4 || DOT == 5) && interior) out = w[e] + out;
    if (DOT == 1 || DOT == 6) {
        acc[0] += (double)xe.re * out.re - (double)xe.im * out.im;
        acc[1] += (double)xe.re * out.im + (double)xe.im * out.re;
    } else if ((DOT == 2 || DOT == 3) && interior) {
        const cx<R> wv = w[e];
        if (DOT == 2) {
            acc[0] += (double)wv.re * out.re + (double)wv.im * out.im;
            acc[1] += (double)wv.re * out.im - (double)wv.im * out.re;
        } else {
            acc[0] += (double)out.re * wv.re + (double)out.im * wv.im;
            acc[1] += (double)out.re * wv.im - (double)out.im * wv.re;
            acc[2] += (double)out.re * out.re + (double)out.im * out.im;
        }
    }
    return out;
}

template <typename R, int DOT>
__device__ __forceinline__ void memit(cx<R> *__restrict__ y, const cx<R> *__restrict__ w, int64_t e, cx<R> out,
                                      bool interior, cx<R> xe, double (&acc)[4])
{
    y[e] = mvalue<R, DOT>(w, e, out, interior, xe, acc);
}

// one footprint column of the k-marching stencil: where it lives, which of
// its edges and cell exist, and the values entering at the current step
// With PS the column also carries the node potential psi of the split
// residual: psi(k+1) enters at step k, scaled by i/wmu, into a ring of three
// node planes (plane p in slot p mod 3); the node of the column has the index
// of its Ez edge less XN + YN.
template <typename R, bool PS = false>
struct MLoad {
    int lf, ls;                   // LDS index in the edge / cell footprint
    bool on, xv, yv, zv, sv, cv;  // in the footprint; Ex, Ey, Ez exist; in the cell footprint; cell exists
    int64_t ex, ey, ez, cc;       // indices at k = 0
    cx<R> inX, inY, inZ;          // Ex/Ey(k+1), Ez(k)
    R iS[6];                      // sigma(k)
    cx<R> inP;                    // (i/wmu) psi(k+1)

    template <int PX, int SX, int SY>
    __device__ __forceinline__ void init(const Dims &d, int f, int np, int i0, int j0)
    {
        on = f < np;
        const int lj = f / PX, li = f - lj * PX;
        const int gi = i0 - 1 + li, gj = j0 - 1 + lj;
        const bool iin = gi >= 0 && gi <= d.nx, jin = gj >= 0 && gj <= d.ny;
        xv = on && iin && jin && gi < d.nx;
        yv = on && iin && jin && gj < d.ny;
        zv = on && iin && jin;
        sv = on && li < SX && lj < SY;
        cv = sv && gi >= 0 && gi < d.nx && gj >= 0 && gj < d.ny;
        lf = lj * PX + li;
        ls = lj * SX + li;
        ex = (int64_t)gj * d.nx + gi;
        ey = d.XN + (int64_t)gj * (d.nx + 1) + gi;
        ez = d.XN + d.YN + (int64_t)gj * (d.nx + 1) + gi;
        cc = (int64_t)gj * d.nx + gi;
    }
    // i/wmu times psi at the column's node of plane k (0 outside the grid)
    __device__ __forceinline__ cx<R> node_psi(const Dims &d, const cx<R> *__restrict__ psi, int k, int64_t ZP,
                                              R rw) const
    {
        if (!zv) return cx<R>(0);
        const cx<R> v = psi[ez - d.XN - d.YN + (int64_t)k * ZP];
        return cx<R>(-v.im * rw, v.re * rw);
    }
    __device__ __forceinline__ void fetch(const Dims &d, const cx<R> *__restrict__ x, const R *__restrict__ sig,
                                          int k, int64_t XP, int64_t YP, int64_t ZP, int64_t CP,
                                          const cx<R> *__restrict__ psi = nullptr, R rw = 0)
    {
        inX = cx<R>(0); inY = cx<R>(0); inZ = cx<R>(0);
        if (PS) inP = cx<R>(0);
#pragma unroll
        for (int q = 0; q < 6; q++) iS[q] = 0;
        if (k < d.nz) {
            if (xv) inX = x[ex + (int64_t)(k + 1) * XP];
            if (yv) inY = x[ey + (int64_t)(k + 1) * YP];
            if (zv) inZ = x[ez + (int64_t)k * ZP];
            if (PS) inP = node_psi(d, psi, k + 1, ZP, rw);
            if (cv) {
                const int64_t q = cc + (int64_t)k * CP;
#pragma unroll
                for (int u = 0; u < 6; u++) iS[u] = sig[u * d.NC + q];
            }
        }
    }
    // before the first step kb: Ex/Ey(kb), Ez(kb-1), sigma(kb-1) into their ring slots
    // (with PS also psi(kb) and psi(kb-1) into slots kb mod 3 and (kb+2) mod 3 of PN)
    __device__ __forceinline__ void prime(const Dims &d, const cx<R> *__restrict__ x, const R *__restrict__ sig,
                                          int kb, int64_t XP, int64_t YP, int64_t ZP, int64_t CP, cx<R> *EX,
                                          cx<R> *EY, cx<R> *EZ, R *SG, int np, int ns,
                                          const cx<R> *__restrict__ psi = nullptr, R rw = 0, cx<R> *PN = nullptr)
    {
        const bool lo = kb >= 1;
        const int s0 = kb & 1, s1 = s0 ^ 1;
        if (on) {
            EX[s0 * np + lf] = xv ? x[ex + (int64_t)kb * XP] : cx<R>(0);
            EY[s0 * np + lf] = yv ? x[ey + (int64_t)kb * YP] : cx<R>(0);
            EZ[s1 * np + lf] = (zv && lo) ? x[ez + (int64_t)(kb - 1) * ZP] : cx<R>(0);
            if (PS) {
                PN[(kb % 3) * np + lf] = node_psi(d, psi, kb, ZP, rw);
                PN[((kb + 2) % 3) * np + lf] = lo ? node_psi(d, psi, kb - 1, ZP, rw) : cx<R>(0);
            }
        }
        if (sv) {
            const int64_t q = cc + (int64_t)(kb - 1) * CP;
#pragma unroll
            for (int u = 0; u < 6; u++) SG[(s1 * 6 + u) * ns + ls] = (cv && lo) ? sig[u * d.NC + q] : (R)0;
        }
    }
    // at step k: Ex/Ey(k+1) into slot (k+1)&1, Ez(k) and sigma(k) into slot k&1
    // (with PS also psi(k+1) into slot pn = (k+1) mod 3 of PN)
    __device__ __forceinline__ void store(int k, cx<R> *EX, cx<R> *EY, cx<R> *EZ, R *SG, int np, int ns,
                                          cx<R> *PN = nullptr, int pn = 0) const
    {
        const int sk = k & 1, sn = sk ^ 1;
        if (on) {
            EX[sn * np + lf] = inX;
            EY[sn * np + lf] = inY;
            EZ[sk * np + lf] = inZ;
            if (PS) PN[pn * np + lf] = inP;
        }
        if (sv) {
#pragma unroll
            for (int u = 0; u < 6; u++) SG[(sk * 6 + u) * ns + ls] = iS[u];
        }
    }
};

template <typename R, int DOT, int TY, bool PF>
__global__ __launch_bounds__(mt::Geo<TY>::NT) void k_spmv_march(
    Dims d, const R *__restrict__ a, const R *__restrict__ b, const R *__restrict__ c, const R *__restrict__ ra,
    const R *__restrict__ rb, const R *__restrict__ rc, const R *__restrict__ sig, R wmu,
    const cx<R> *__restrict__ x, cx<R> *__restrict__ y, R alpha, const cx<R> *__restrict__ w,
    double *__restrict__ partials, int ntx, int kc, const cx<R> *__restrict__ psi)
{
    using G = mt::Geo<TY>;
    constexpr int TX = G::TX, PX = G::PX, SX = G::SX, SY = G::SY, NP = G::NP, NS = G::NS;
    constexpr bool PS = DOT == 5 || DOT == 6;
    __shared__ __align__(16) R lds[G::LDS_R + (PS ? 6 * NP : 0)];
    cx<R> *const EX = reinterpret_cast<cx<R> *>(lds);  // [2][NP], slot k&1 holds plane k
    cx<R> *const EY = EX + 2 * NP;
    cx<R> *const EZ = EY + 2 * NP;
    R *const SG = lds + 12 * NP;                        // [2][6][NS], components xx yy zz xy xz yz
    R *const TA = SG + 12 * NS;                         // a, then 1/a
    R *const TB = TA + 2 * (TX + 1);                    // b, then 1/b
    cx<R> *const PN = reinterpret_cast<cx<R> *>(lds + G::LDS_R);  // PS: [3][NP] (i/wmu) psi rings
    const R rw = (R)1 / wmu;
    const Tile TL = tile_id();
    const int tj = TL.x / ntx, ti = TL.x - tj * ntx;
    const int i0 = ti * TX, j0 = tj * TY;
    const int kb = TL.y * kc, ke = min(kb + kc, d.nz + 1);
    const int t = threadIdx.x;
    const int64_t XP = (int64_t)(d.ny + 1) * d.nx, YP = (int64_t)d.ny * (d.nx + 1);
    const int64_t ZP = (int64_t)(d.ny + 1) * (d.nx + 1), CP = (int64_t)d.nx * d.ny, NC = d.NC;
    // loader role: footprint columns t and t + NT
    constexpr int NT = G::NT;
    MLoad<R, PS> L0, L1;
    L0.template init<PX, SX, SY>(d, t, NP, i0, j0);
    L1.template init<PX, SX, SY>(d, t + NT, NP, i0, j0);
    const bool two = t + NT < NP;
    // compute role: tile column (i, j)
    const int cy = t / TX, cxl = t - cy * TX;
    const int i = i0 + cxl, j = j0 + cy;
    const bool own = i < d.nx && j < d.ny;
    const int fo = (cy + 1) * PX + cxl + 1, so = (cy + 1) * SX + cxl + 1;
    const int64_t ex0 = (int64_t)j * d.nx + i, ey0 = d.XN + (int64_t)j * (d.nx + 1) + i;
    const int64_t ez0 = d.XN + d.YN + (int64_t)j * (d.nx + 1) + i;
    cx<R> exP, eyP;  // own Ex, Ey at k-1
    if (kb < ke) {
        // prologue: the first planes into their ring slots, width tables
        const bool lo = kb >= 1;
        L0.prime(d, x, sig, kb, XP, YP, ZP, CP, EX, EY, EZ, SG, NP, NS, psi, rw, PN);
        if (two) L1.prime(d, x, sig, kb, XP, YP, ZP, CP, EX, EY, EZ, SG, NP, NS, psi, rw, PN);
        if (t <= TX) {
            const int ii = i0 - 1 + t;
            const bool ok = ii >= 0 && ii < d.nx;
            TA[t] = ok ? a[ii] : (R)0;
            TA[TX + 1 + t] = ok ? ra[ii] : (R)0;
        }
        if (t <= TY) {
            const int jj = j0 - 1 + t;
            const bool ok = jj >= 0 && jj < d.ny;
            TB[t] = ok ? b[jj] : (R)0;
            TB[TY + 1 + t] = ok ? rb[jj] : (R)0;
        }
        if (own && lo) {
            exP = x[ex0 + (int64_t)(kb - 1) * XP];
            eyP = x[ey0 + (int64_t)(kb - 1) * YP];
        }
        L0.fetch(d, x, sig, kb, XP, YP, ZP, CP, psi, rw);
        if (two) L1.fetch(d, x, sig, kb, XP, YP, ZP, CP, psi, rw);
    }
    const R hE = (R)0.5, q4 = (R)0.25, q8 = (R)0.125;
    double acc[4] = {0, 0, 0, 0};
    for (int k = kb; k < ke; k++) {
        const int sk = k & 1, sn = sk ^ 1;  // ring slots of plane k and of planes k+1 / k-1
        __syncthreads();                    // the previous step's reads of the slots refilled here are done
        const int pk = k % 3, pm = pk == 0 ? 2 : pk - 1, pp = pk == 2 ? 0 : pk + 1;  // psi slots of k, k-1, k+1
        L0.store(k, EX, EY, EZ, SG, NP, NS, PN, pp);
        if (two) L1.store(k, EX, EY, EZ, SG, NP, NS, PN, pp);
        // with PF the next step's loads are in flight while this step's rows are formed
        if (PF && k + 1 < ke) {
            L0.fetch(d, x, sig, k + 1, XP, YP, ZP, CP, psi, rw);
            if (two) L1.fetch(d, x, sig, k + 1, XP, YP, ZP, CP, psi, rw);
        }
        __syncthreads();
        if (own) {
            const cx<R> *EXK = EX + sk * NP, *EXN = EX + sn * NP, *EYK = EY + sk * NP, *EYN = EY + sn * NP;
            const cx<R> *EZP = EZ + sn * NP, *EZK = EZ + sk * NP;
            const R *SP = SG + sn * 6 * NS, *SK = SG + sk * 6 * NS;  // component u at u*NS
            // PS: the mass terms act on u = x + (i/wmu) G psi, the stiffness on x
            // alone (K G psi = 0): w - A (x + (i/wmu) G psi) = w - K x + i wmu M u
            const cx<R> *PK = PN + pk * NP, *PM = PN + pm * NP, *PP = PN + pp * NP;
            const bool zk = k >= 1 && k < d.nz;  // both cell planes k-1, k exist
            const R cm = zk ? c[k - 1] : (R)0, cp = zk ? c[k] : (R)0;
            const R ai = TA[cxl + 1], am = TA[cxl], rai = TA[TX + 2 + cxl], ram = TA[TX + 1 + cxl];
            const R bj = TB[cy + 1], bm = TB[cy], rbj = TB[TY + 2 + cy], rbm = TB[TY + 1 + cy];
            const cx<R> exK = EXK[fo], exN = EXN[fo], eyK = EYK[fo], eyN = EYN[fo];
            const cx<R> ezP = EZP[fo], ezK = EZK[fo];
            auto S2 = [&](int u, int o) { return cm * SP[u * NS + o] + cp * SK[u * NS + o]; };
            // ---- Ex(i,j,k)
            {
                cx<R> out(0), xo = exK;  // xo: the own value the fused dot takes (mode 6: u)
                const bool in = j >= 1 && zk;
                if (in) {
                    const R rcm = rc[k - 1], rcp = rc[k];
                    const R hb = hE * ai * (bm + bj), hc = hE * ai * (cm + cp);
                    const R wkm = hb * rcm, wkp = hb * rcp, wjm = hc * rbm, wjp = hc * rbj;
                    const R hy = hE * (cm + cp), hz = hE * (bm + bj);
                    const R kmd = ai * (bm * S2(0, so - SX) + bj * S2(0, so)) * q4;
                    const R my1 = ai * bm * S2(3, so - SX) * q8, my2 = ai * bj * S2(3, so) * q8;
                    const R mz1 = ai * cm * (bm * SP[4 * NS + so - SX] + bj * SP[4 * NS + so]) * q8;
                    const R mz2 = ai * cp * (bm * SK[4 * NS + so - SX] + bj * SK[4 * NS + so]) * q8;
                    const cx<R> ey00 = EYK[fo - PX], ey10 = EYK[fo - PX + 1], ey11 = EYK[fo + 1];
                    const cx<R> ez1 = EZP[fo + 1], ez2 = EZK[fo + 1];
                    const cx<R> K = (wkm + wkp + wjm + wjp) * exK - wkm * exP - wkp * exN - wjm * EXK[fo - PX]
                                    - wjp * EXK[fo + PX] + hy * (ey00 - ey10 - eyK + ey11)
                                    + hz * (ezP - ez1 - ezK + ez2);
                    cx<R> uK = exK, u00 = ey00, u10 = ey10, uyK = eyK, u11 = ey11;
                    cx<R> uzP = ezP, uz1 = ez1, uzK = ezK, uz2 = ez2;
                    if (PS) {
                        const cx<R> p0 = PK[fo], p1 = PK[fo + 1];
                        uK = uK + (p1 - p0) * rai;
                        u00 = u00 + (p0 - PK[fo - PX]) * rbm;
                        u10 = u10 + (p1 - PK[fo - PX + 1]) * rbm;
                        uyK = uyK + (PK[fo + PX] - p0) * rbj;
                        u11 = u11 + (PK[fo + PX + 1] - p1) * rbj;
                        uzP = uzP + (p0 - PM[fo]) * rcm;
                        uz1 = uz1 + (p1 - PM[fo + 1]) * rcm;
                        uzK = uzK + (PP[fo] - p0) * rcp;
                        uz2 = uz2 + (PP[fo + 1] - p1) * rcp;
                    }
                    const cx<R> M = kmd * uK + my1 * (u00 + u10) + my2 * (uyK + u11) + mz1 * (uzP + uz1)
                                    + mz2 * (uzK + uz2);
                    out = km_combine(K, M, wmu, alpha);
                    if (DOT == 6) xo = uK;
                }
                memit<R, DOT>(y, w, ex0 + (int64_t)k * XP, out, in, xo, acc);
            }
            // rows one after another: the scheduler may not hoist the next row's
            // LDS reads above this point, which bounds the registers live at once
            __builtin_amdgcn_sched_barrier(0);
            // ---- Ey(i,j,k)
            {
                cx<R> out(0), xo = eyK;
                const bool in = i >= 1 && zk;
                if (in) {
                    const R rcm = rc[k - 1], rcp = rc[k];
                    const R ha = hE * bj * (am + ai), hc = hE * bj * (cm + cp);
                    const R wkm = ha * rcm, wkp = ha * rcp, wim = hc * ram, wip = hc * rai;
                    const R hx = hE * (cm + cp), hz = hE * (am + ai);
                    const R kmd = bj * (am * S2(1, so - 1) + ai * S2(1, so)) * q4;
                    const R mx1 = bj * am * S2(3, so - 1) * q8, mx2 = bj * ai * S2(3, so) * q8;
                    const R mz1 = bj * cm * (am * SP[5 * NS + so - 1] + ai * SP[5 * NS + so]) * q8;
                    const R mz2 = bj * cp * (am * SK[5 * NS + so - 1] + ai * SK[5 * NS + so]) * q8;
                    const cx<R> x0 = EXK[fo - 1], x2 = EXK[fo + PX - 1], x3 = EXK[fo + PX];
                    const cx<R> z1 = EZP[fo + PX], z3 = EZK[fo + PX];
                    const cx<R> K = (wkm + wkp + wim + wip) * eyK - wkm * eyP - wkp * eyN - wim * EYK[fo - 1]
                                    - wip * EYK[fo + 1] + hx * (x0 - exK - x2 + x3) + hz * (ezP - z1 - ezK + z3);
                    cx<R> uK = eyK, u0 = x0, u2 = x2, uxK = exK, u3 = x3, uzP = ezP, uz1 = z1, uzK = ezK, uz3 = z3;
                    if (PS) {
                        const cx<R> p0 = PK[fo], p2 = PK[fo + PX];
                        uK = uK + (p2 - p0) * rbj;
                        u0 = u0 + (p0 - PK[fo - 1]) * ram;
                        u2 = u2 + (p2 - PK[fo + PX - 1]) * ram;
                        uxK = uxK + (PK[fo + 1] - p0) * rai;
                        u3 = u3 + (PK[fo + PX + 1] - p2) * rai;
                        uzP = uzP + (p0 - PM[fo]) * rcm;
                        uz1 = uz1 + (p2 - PM[fo + PX]) * rcm;
                        uzK = uzK + (PP[fo] - p0) * rcp;
                        uz3 = uz3 + (PP[fo + PX] - p2) * rcp;
                    }
                    const cx<R> M = kmd * uK + mx1 * (u0 + u2) + mx2 * (uxK + u3) + mz1 * (uzP + uz1)
                                    + mz2 * (uzK + uz3);
                    out = km_combine(K, M, wmu, alpha);
                    if (DOT == 6) xo = uK;
                }
                memit<R, DOT>(y, w, ey0 + (int64_t)k * YP, out, in, xo, acc);
            }
            __builtin_amdgcn_sched_barrier(0);
            // ---- Ez(i,j,k), between node planes k and k+1
            if (k < d.nz) {
                cx<R> out(0), xo = ezK;
                const bool in = i >= 1 && j >= 1;
                if (in) {
                    const R ck = c[k];
                    const R ha = hE * ck * (am + ai), hb = hE * ck * (bm + bj);
                    const R wjm = ha * rbm, wjp = ha * rbj, wim = hb * ram, wip = hb * rai;
                    const R hx = hE * (bm + bj), hy = hE * (am + ai);
                    const R *zz = SK + 2 * NS, *xz = SK + 4 * NS, *yz = SK + 5 * NS;
                    const int o00 = so - SX - 1, o01 = so - 1, o10 = so - SX;
                    const R kmd = ck * (am * (bm * zz[o00] + bj * zz[o01]) + ai * (bm * zz[o10] + bj * zz[so])) * q4;
                    const R mx1 = ck * am * (bm * xz[o00] + bj * xz[o01]) * q8;
                    const R mx2 = ck * ai * (bm * xz[o10] + bj * xz[so]) * q8;
                    const R my1 = ck * bm * (am * yz[o00] + ai * yz[o10]) * q8;
                    const R my2 = ck * bj * (am * yz[o01] + ai * yz[so]) * q8;
                    const cx<R> exm = EXK[fo - 1], exmN = EXN[fo - 1];
                    const cx<R> eym = EYK[fo - PX], eymN = EYN[fo - PX];
                    const cx<R> K = (wjm + wjp + wim + wip) * ezK - wjm * EZK[fo - PX] - wjp * EZK[fo + PX]
                                    - wim * EZK[fo - 1] - wip * EZK[fo + 1] + hx * (exm - exK - exmN + exN)
                                    + hy * (eym - eyK - eymN + eyN);
                    cx<R> uK = ezK, uxm = exm, uxK = exK, uxmN = exmN, uxN = exN;
                    cx<R> uym = eym, uyK = eyK, uymN = eymN, uyN = eyN;
                    if (PS) {
                        const R rck = rc[k];
                        const cx<R> p0 = PK[fo], q0 = PP[fo];
                        uK = uK + (q0 - p0) * rck;
                        uxm = uxm + (p0 - PK[fo - 1]) * ram;
                        uxK = uxK + (PK[fo + 1] - p0) * rai;
                        uxmN = uxmN + (q0 - PP[fo - 1]) * ram;
                        uxN = uxN + (PP[fo + 1] - q0) * rai;
                        uym = uym + (p0 - PK[fo - PX]) * rbm;
                        uyK = uyK + (PK[fo + PX] - p0) * rbj;
                        uymN = uymN + (q0 - PP[fo - PX]) * rbm;
                        uyN = uyN + (PP[fo + PX] - q0) * rbj;
                    }
                    const cx<R> M = kmd * uK + mx1 * (uxm + uxmN) + mx2 * (uxK + uxN) + my1 * (uym + uymN)
                                    + my2 * (uyK + uyN);
                    out = km_combine(K, M, wmu, alpha);
                    if (DOT == 6) xo = uK;
                }
                memit<R, DOT>(y, w, ez0 + (int64_t)k * ZP, out, in, xo, acc);
            }
            // boundary edges of the last column i = nx and the last row j = ny
            const cx<R> zero(0);
            if (i == d.nx - 1) {
                y[ey0 + 1 + (int64_t)k * YP] = zero;
                if (k < d.nz) y[ez0 + 1 + (int64_t)k * ZP] = zero;
            }
            if (j == d.ny - 1) {
                y[ex0 + d.nx + (int64_t)k * XP] = zero;
                if (k < d.nz) {
                    y[ez0 + (d.nx + 1) + (int64_t)k * ZP] = zero;
                    if (i == d.nx - 1) y[ez0 + (d.nx + 2) + (int64_t)k * ZP] = zero;
                }
            }
            exP = exK;
            eyP = eyK;
        }
        if (!PF && k + 1 < ke) {
            L0.fetch(d, x, sig, k + 1, XP, YP, ZP, CP, psi, rw);
            if (two) L1.fetch(d, x, sig, k + 1, XP, YP, ZP, CP, psi, rw);
        }
    }
    if ((DOT >= 1 && DOT <= 3) || DOT == 6) block_reduce_store<4, G::NT>(acc, partials);
}

// ------------------------------------------------------------------ K1b
// Cell-plane march of the same complex128 operator.  A block owns a 32x8 tile
// of node columns and walks up the cell planes c of its chunk of node planes.
// Cell plane c lies between node planes c and c+1: it couples the Ex/Ey edges
// of both planes and the Ez edges between them.  Step c completes the rows of
// node plane c (their part from cell plane c-1 was formed at step c-1 and is
// carried in registers), forms the Ez rows of plane c and the part of the
// Ex/Ey rows of plane c+1 that comes from cell plane c.  It reads Ex/Ey of
// planes c and c+1, Ez of plane c and the cell tensors of plane c only, so the
// LDS rings hold 3 planes of Ex/Ey (one in flight) and 2 of Ez and sigma.
// Buffer loads write the planes straight into LDS one step ahead of their use
// (no registers stage them, one barrier per step); lanes outside the grid load
// zeros (offset beyond the buffer).  sigma comes from a pair layout of three
// 16-byte pieces per cell, (xx,yy), (zz,xy), (xz,yz), so that every load moves
// 16 bytes per lane.  LDS: 79.9 KB, two blocks per CU.
namespace cm {
constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr int PX = TX + 2, PY = TY + 2, NP = PX * PY;  // edge footprint: i0-1..i0+TX, j0-1..j0+TY
constexpr int SX = TX + 1, SY = TY + 1, NS = SX * SY;  // cell footprint: i0-1..i0+TX-1, j0-1..j0+TY-1
constexpr int NPP = 384, NSP = 320;  // footprints padded to whole wavefronts of loads (a wave fills 64 slots)
static_assert(NP <= NPP && NPP % 64 == 0 && NS <= NSP && NSP % 64 == 0, "footprint padding");
static_assert(NPP <= 2 * NT && NSP <= 2 * NT, "two load rounds per footprint");
// LDS in 16-byte units: Ex ring [3][NPP], Ey ring [3][NPP], Ez ring [2][NPP], sigma ring [2][3][NSP]
constexpr int EXo = 0, EYo = 3 * NPP, EZo = 6 * NPP, SGo = 8 * NPP, N16 = 8 * NPP + 6 * NSP;
constexpr unsigned OOB = 0x80000000u;  // a byte offset past every buffer: the load returns zeros
}  // namespace cm

typedef __attribute__((address_space(3))) void lds_void_t;

// 64 lanes x 16 bytes from a raw buffer into LDS at `dst` (wave-uniform; lane l
// writes dst + 16 l); offset = voff + soff bytes
__device__ __forceinline__ void lds_load16(__amdgpu_buffer_rsrc_t rs, void *dst, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void_t *)dst, 16, voff, soff, 0, 0);
}

template <int DOT>
__global__ __launch_bounds__(cm::NT) void k_spmv_cell(
    Dims d, const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
    const double *__restrict__ ra, const double *__restrict__ rb, const double *__restrict__ rc,
    const double *__restrict__ sigp, double wmu, const zd *__restrict__ x, zd *__restrict__ y, double alpha,
    const zd *__restrict__ w, double *__restrict__ partials, int ntx, int kc)
{
    using namespace cm;
    __shared__ __align__(16) double lds[2 * N16];
    zd *const EX = reinterpret_cast<zd *>(lds) + EXo;
    zd *const EY = reinterpret_cast<zd *>(lds) + EYo;
    zd *const EZ = reinterpret_cast<zd *>(lds) + EZo;
    double2 *const SG = reinterpret_cast<double2 *>(lds) + SGo;
    const Tile TL = tile_id();
    const int tj = TL.x / ntx, ti = TL.x - tj * ntx;
    const int i0 = ti * TX, j0 = tj * TY;
    const int kb = TL.y * kc, ke = min(kb + kc, d.nz + 1);
    const int t = threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const unsigned XP16 = (unsigned)((d.ny + 1) * d.nx) * 16u, YP16 = (unsigned)(d.ny * (d.nx + 1)) * 16u;
    const unsigned ZP16 = (unsigned)((d.ny + 1) * (d.nx + 1)) * 16u, CP16 = (unsigned)(d.nx * d.ny) * 16u;
    const unsigned NC16 = (unsigned)d.NC * 16u;
    const __amdgpu_buffer_rsrc_t rsx = __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)(d.NL * 16), 0x00020000);
    const __amdgpu_buffer_rsrc_t rss = __builtin_amdgcn_make_buffer_rsrc((void *)sigp, (short)0, (int)(3 * NC16), 0x00020000);
    // loader role: footprint slots t and NT + t (the second only where a wave's 64 slots start inside the footprint)
    unsigned ox[2], oy[2], oz[2], os[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int f = q * NT + t;
        {
            const int lj = f / PX, li = f - lj * PX, gi = i0 - 1 + li, gj = j0 - 1 + lj;
            const bool on = f < NP && gi >= 0 && gi <= d.nx && gj >= 0 && gj <= d.ny;
            ox[q] = on && gi < d.nx ? (unsigned)(gj * d.nx + gi) * 16u : OOB;
            oy[q] = on && gj < d.ny ? (unsigned)(d.XN + gj * (d.nx + 1) + gi) * 16u : OOB;
            oz[q] = on ? (unsigned)(d.XN + d.YN + gj * (d.nx + 1) + gi) * 16u : OOB;
        }
        {
            const int lj = f / SX, li = f - lj * SX, gi = i0 - 1 + li, gj = j0 - 1 + lj;
            const bool on = f < NS && gi >= 0 && gi < d.nx && gj >= 0 && gj < d.ny;
            os[q] = on ? (unsigned)(gj * d.nx + gi) * 16u : OOB;
        }
    }
    const bool e2 = NT + 64 * wv < NP, s2 = NT + 64 * wv < NS;  // wave-uniform
    // Ex, Ey of node plane p into ring slot p mod 3
    auto load_xy = [&](int p) {
        zd *ex = EX + (p % 3) * NPP + 64 * wv, *ey = EY + (p % 3) * NPP + 64 * wv;
        lds_load16(rsx, ex, ox[0], (unsigned)p * XP16);
        lds_load16(rsx, ey, oy[0], (unsigned)p * YP16);
        if (e2) {
            lds_load16(rsx, ex + NT, ox[1], (unsigned)p * XP16);
            lds_load16(rsx, ey + NT, oy[1], (unsigned)p * YP16);
        }
    };
    // Ez of node plane p and the cell tensors of cell plane p into slot p mod 2
    auto load_zs = [&](int p) {
        zd *ez = EZ + (p & 1) * NPP + 64 * wv;
        double2 *sg = SG + (p & 1) * 3 * NSP + 64 * wv;
        lds_load16(rsx, ez, oz[0], (unsigned)p * ZP16);
        if (e2) lds_load16(rsx, ez + NT, oz[1], (unsigned)p * ZP16);
#pragma unroll
        for (int u = 0; u < 3; u++) {
            lds_load16(rss, sg + u * NSP, os[0], (unsigned)p * CP16 + u * NC16);
            if (s2) lds_load16(rss, sg + u * NSP + NT, os[1], (unsigned)p * CP16 + u * NC16);
        }
    };
    // compute role: tile column (i, j)
    const int cy = t / TX, cxl = t - cy * TX;
    const int i = i0 + cxl, j = j0 + cy;
    const bool own = i < d.nx && j < d.ny;
    const int fo = (cy + 1) * PX + cxl + 1, so = (cy + 1) * SX + cxl + 1;
    const int64_t XP = (int64_t)(d.ny + 1) * d.nx, YP = (int64_t)d.ny * (d.nx + 1), ZP = (int64_t)(d.ny + 1) * (d.nx + 1);
    const int64_t ex0 = (int64_t)j * d.nx + i, ey0 = d.XN + (int64_t)j * (d.nx + 1) + i;
    const int64_t ez0 = d.XN + d.YN + (int64_t)j * (d.nx + 1) + i;
    double ai = 0, am = 0, bj = 0, bm = 0, rai = 0, ram = 0, rbj = 0, rbm = 0;
    if (own) {
        ai = a[i]; rai = ra[i]; bj = b[j]; rbj = rb[j];
        if (i >= 1) { am = a[i - 1]; ram = ra[i - 1]; }
        if (j >= 1) { bm = b[j - 1]; rbm = rb[j - 1]; }
    }
    double acc[4] = {0, 0, 0, 0};
    if (kb < ke) {
        // zero the rings (the padding slots and any slot a load does not fill stay 0)
        for (int u = t; u < N16; u += NT) reinterpret_cast<zd *>(lds)[u] = zd(0);
        __syncthreads();
        const int cs = kb > 0 ? kb - 1 : 0;  // first cell plane: the one under the chunk's first node plane
        load_xy(cs);
        if (cs + 1 <= d.nz) load_xy(cs + 1);
        if (cs < d.nz) load_zs(cs);
        const double hE = 0.5, q4 = 0.25, q8 = 0.125;
        zd pEx(0), pEy(0);  // part of the Ex, Ey rows of node plane c from cell plane c-1
        // The rows of a step are stored at the start of the next one, after its
        // loads are issued: the wait at the top of a step (for the loads) then
        // does not also wait for the stores issued just before it.
        zd sEx(0), sEy(0), sEz(0);
        bool pX = false, pY = false, pZ = false;
        int pc = 0;
        auto flush = [&]() {
            if (pX) y[ex0 + (int64_t)pc * XP] = sEx;
            if (pY) y[ey0 + (int64_t)pc * YP] = sEy;
            if (pZ) y[ez0 + (int64_t)pc * ZP] = sEz;
            pX = pY = pZ = false;
        };
        for (int cp = cs; cp < ke; cp++) {
            __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of step cp are in LDS
            __syncthreads();                // every wave's; the reads of step cp-1 are done
            if (cp + 1 < ke) {
                if (cp + 2 <= d.nz) load_xy(cp + 2);
                if (cp + 1 < d.nz) load_zs(cp + 1);
            }
            flush();
            if (!own) continue;
            pc = cp;
            const zd *X0 = EX + (cp % 3) * NPP, *X1 = EX + ((cp + 1) % 3) * NPP;
            const zd *Y0 = EY + (cp % 3) * NPP, *Y1 = EY + ((cp + 1) % 3) * NPP;
            const zd *Z = EZ + (cp & 1) * NPP;
            const double2 *S0 = SG + (cp & 1) * 3 * NSP, *S1 = S0 + NSP, *S2 = S1 + NSP;
            const bool row = cp >= kb;       // the rows of plane cp are this block's
            const bool nxt = cp + 1 < ke;    // so are those of plane cp + 1
            const bool cell = cp < d.nz;     // cell plane cp exists
            const bool inc = cp >= 1 && cp <= d.nz - 1, inn = cp + 1 <= d.nz - 1;  // interior in k
            const double cc = cell ? c[cp] : 0.0, rcc = cell ? rc[cp] : 0.0;
            // ---- Ex rows: plane cp (above part, completes the row) and plane cp+1 (below part)
            {
                const bool up = row && inc && j >= 1, dn = nxt && inn && j >= 1;
                zd out = pEx;
                if (cell && (up || dn)) {
                    const double2 p1 = S0[so - SX], p2 = S0[so], q1 = S1[so - SX], q2 = S1[so], r1 = S2[so - SX], r2 = S2[so];
                    const double acc_ = ai * cc;
                    const double mxx = acc_ * (bm * p1.x + bj * p2.x) * q4;   // xx of cells (i,j-1), (i,j)
                    const double my1 = acc_ * bm * q1.y * q8, my2 = acc_ * bj * q2.y * q8;  // xy
                    const double mxz = acc_ * (bm * r1.x + bj * r2.x) * q8;  // xz
                    const double wk = hE * ai * (bm + bj) * rcc, hm = hE * acc_ * rbm, hp = hE * acc_ * rbj;
                    const double hc = hE * cc, hz = hE * (bm + bj);
                    const zd z0 = Z[fo], z1 = Z[fo + 1];
                    const zd mz = mxz * (z0 + z1);
                    if (up) {
                        const zd e = X0[fo], en = X1[fo], em = X0[fo - PX], ep = X0[fo + PX];
                        const zd y00 = Y0[fo - PX], y10 = Y0[fo - PX + 1], yk = Y0[fo], y11 = Y0[fo + 1];
                        const zd K = wk * (e - en) + hm * (e - em) + hp * (e - ep) + hc * (y00 - y10 - yk + y11)
                                     + hz * (z1 - z0);
                        const zd M = mxx * e + my1 * (y00 + y10) + my2 * (yk + y11) + mz;
                        out = out + km_combine(K, M, wmu, alpha);
                    }
                    if (dn) {
                        const zd e = X1[fo], eb = X0[fo], em = X1[fo - PX], ep = X1[fo + PX];
                        const zd y00 = Y1[fo - PX], y10 = Y1[fo - PX + 1], yk = Y1[fo], y11 = Y1[fo + 1];
                        const zd K = wk * (e - eb) + hm * (e - em) + hp * (e - ep) + hc * (y00 - y10 - yk + y11)
                                     + hz * (z0 - z1);
                        const zd M = mxx * e + my1 * (y00 + y10) + my2 * (yk + y11) + mz;
                        pEx = km_combine(K, M, wmu, alpha);
                    }
                }
                if (row) {
                    sEx = mvalue<double, DOT>(w, ex0 + (int64_t)cp * XP, up ? out : zd(0), up, X0[fo], acc);
                    pX = true;
                }
            }
            // ---- Ey rows
            {
                const bool up = row && inc && i >= 1, dn = nxt && inn && i >= 1;
                zd out = pEy;
                if (cell && (up || dn)) {
                    const double2 p1 = S0[so - 1], p2 = S0[so], q1 = S1[so - 1], q2 = S1[so], r1 = S2[so - 1], r2 = S2[so];
                    const double bcc = bj * cc;
                    const double myy = bcc * (am * p1.y + ai * p2.y) * q4;  // yy of cells (i-1,j), (i,j)
                    const double mx1 = bcc * am * q1.y * q8, mx2 = bcc * ai * q2.y * q8;  // xy
                    const double myz = bcc * (am * r1.y + ai * r2.y) * q8;  // yz
                    const double wk = hE * bj * (am + ai) * rcc, hm = hE * bcc * ram, hp = hE * bcc * rai;
                    const double hc = hE * cc, hz = hE * (am + ai);
                    const zd z0 = Z[fo], z3 = Z[fo + PX];
                    const zd mz = myz * (z0 + z3);
                    if (up) {
                        const zd e = Y0[fo], en = Y1[fo], em = Y0[fo - 1], ep = Y0[fo + 1];
                        const zd x0 = X0[fo - 1], xk = X0[fo], x2 = X0[fo + PX - 1], x3 = X0[fo + PX];
                        const zd K = wk * (e - en) + hm * (e - em) + hp * (e - ep) + hc * (x0 - xk - x2 + x3)
                                     + hz * (z3 - z0);
                        const zd M = myy * e + mx1 * (x0 + x2) + mx2 * (xk + x3) + mz;
                        out = out + km_combine(K, M, wmu, alpha);
                    }
                    if (dn) {
                        const zd e = Y1[fo], eb = Y0[fo], em = Y1[fo - 1], ep = Y1[fo + 1];
                        const zd x0 = X1[fo - 1], xk = X1[fo], x2 = X1[fo + PX - 1], x3 = X1[fo + PX];
                        const zd K = wk * (e - eb) + hm * (e - em) + hp * (e - ep) + hc * (x0 - xk - x2 + x3)
                                     + hz * (z0 - z3);
                        const zd M = myy * e + mx1 * (x0 + x2) + mx2 * (xk + x3) + mz;
                        pEy = km_combine(K, M, wmu, alpha);
                    }
                }
                if (row) {
                    sEy = mvalue<double, DOT>(w, ey0 + (int64_t)cp * YP, up ? out : zd(0), up, Y0[fo], acc);
                    pY = true;
                }
            }
            // ---- Ez row of plane cp (cell plane cp only)
            if (row && cell) {
                zd out(0);
                const bool in = i >= 1 && j >= 1;
                if (in) {
                    const double ha = hE * cc * (am + ai), hb = hE * cc * (bm + bj);
                    const double wjm = ha * rbm, wjp = ha * rbj, wim = hb * ram, wip = hb * rai;
                    const double hx = hE * (bm + bj), hy = hE * (am + ai);
                    const int o00 = so - SX - 1, o01 = so - 1, o10 = so - SX;
                    const double2 u00 = S1[o00], u01 = S1[o01], u10 = S1[o10], u11 = S1[so];  // .x = zz
                    const double2 v00 = S2[o00], v01 = S2[o01], v10 = S2[o10], v11 = S2[so];  // xz, yz
                    const double kmd = cc * (am * (bm * u00.x + bj * u01.x) + ai * (bm * u10.x + bj * u11.x)) * q4;
                    const double mx1 = cc * am * (bm * v00.x + bj * v01.x) * q8;
                    const double mx2 = cc * ai * (bm * v10.x + bj * v11.x) * q8;
                    const double my1 = cc * bm * (am * v00.y + ai * v10.y) * q8;
                    const double my2 = cc * bj * (am * v01.y + ai * v11.y) * q8;
                    const zd ez = Z[fo];
                    const zd exm = X0[fo - 1], exk = X0[fo], exmN = X1[fo - 1], exN = X1[fo];
                    const zd eym = Y0[fo - PX], eyk = Y0[fo], eymN = Y1[fo - PX], eyN = Y1[fo];
                    const zd K = (wjm + wjp + wim + wip) * ez - wjm * Z[fo - PX] - wjp * Z[fo + PX] - wim * Z[fo - 1]
                                 - wip * Z[fo + 1] + hx * (exm - exk - exmN + exN) + hy * (eym - eyk - eymN + eyN);
                    const zd M = kmd * ez + mx1 * (exm + exmN) + mx2 * (exk + exN) + my1 * (eym + eymN)
                                 + my2 * (eyk + eyN);
                    out = km_combine(K, M, wmu, alpha);
                }
                sEz = mvalue<double, DOT>(w, ez0 + (int64_t)cp * ZP, out, in, Z[fo], acc);
                pZ = true;
            }
            if (row) {
                // boundary edges of the last column i = nx and the last row j = ny
                const zd zero(0);
                if (i == d.nx - 1) {
                    y[ey0 + 1 + (int64_t)cp * YP] = zero;
                    if (cell) y[ez0 + 1 + (int64_t)cp * ZP] = zero;
                }
                if (j == d.ny - 1) {
                    y[ex0 + d.nx + (int64_t)cp * XP] = zero;
                    if (cell) {
                        y[ez0 + (d.nx + 1) + (int64_t)cp * ZP] = zero;
                        if (i == d.nx - 1) y[ez0 + (d.nx + 2) + (int64_t)cp * ZP] = zero;
                    }
                }
            }
        }
        flush();
    }
    if (DOT >= 1 && DOT <= 3) block_reduce_store<4, cm::NT>(acc, partials);
}

// w = M_sigma x on interior edges
template <typename R>
__global__ __launch_bounds__(TPB) void k_mass(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                              const R *__restrict__ c, const R *__restrict__ ra,
                                              const R *__restrict__ rb, const R *__restrict__ rc,
                                              const R *__restrict__ sig,
                                              const cx<R> *__restrict__ x, cx<R> *__restrict__ w)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int j = p / S.nI, i = p - j * S.nI;
    const int64_t e = S.base + p;
    cx<R> out(0);
    if (edge_interior(d, S.o, i, j, S.k)) {
        Row<R> r;
        row_coeffs<R, true>(d, a, b, c, sig, S.o, i, j, S.k, r, ra, rb, rc);
        out = r.kmd * x[e];
#pragma unroll
        for (int s = 4; s < 12; s++) out = out + r.km[s] * x[r.nb[s]];
    }
    w[e] = out;
}

// ------------------------------------------------------------------ K2 SSOR
// Forward (DIR=0): out_e = omega*(in_e - sum_{lower colours} A_en out_n)/D_e
// Backward (DIR=1): out_e = omega*(out_e*D_e*(2-omega)/omega - sum_{upper} A_en out_n)/D_e
template <typename R, int O, int PAR, int DIR, bool XM>
__global__ __launch_bounds__(TPB) void k_ssor(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                              const R *__restrict__ c, const R *__restrict__ ra,
                                              const R *__restrict__ rb, const R *__restrict__ rc,
                                              const R *__restrict__ sig, R wmu,
                                              const cx<R> *__restrict__ in, cx<R> *__restrict__ out, R omega)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k < (O == 2 ? d.nz : d.nz + 1)) ? TL.x * TPB + threadIdx.x : kNoTile;
    int i, j;
    if (!colour_decode<O>(d, PAR, k, p, i, j)) return;
    const int64_t e = (O == 0) ? ex_id(d, i, j, k) : (O == 1) ? ey_id(d, i, j, k) : ez_id(d, i, j, k);
    if (!edge_interior(d, O, i, j, k)) {
        if (DIR != 1) out[e] = cx<R>(0);
        return;
    }
    Row<R> r;
    row_coeffs<R, true, XM>(d, a, b, c, sig, O, i, j, k, r, ra, rb, rc);
    // group orientations: O=0 -> (1,2); O=1 -> (0,2); O=2 -> (0,1)
    constexpr int OA = (O == 0) ? 1 : 0;
    constexpr int OB = (O == 2) ? 1 : 2;
    constexpr bool SAME = (DIR == 0) ? (PAR == 1) : (PAR == 0);
    constexpr bool GA = (DIR == 0) ? (OA < O) : (OA > O);
    constexpr bool GB = (DIR == 0) ? (OB < O) : (OB > O);
    cx<R> s(0);
    if (SAME) {
#pragma unroll
        for (int q = 0; q < 4; q++) s = fma_c(s, cx<R>(r.kg[q], (R)0), out[r.nb[q]]);
    }
    if (GA) {
#pragma unroll
        for (int q = 4; q < 8; q++)
            s = XM ? fma_c(s, coef_of(r.kg[q], r.km[q], wmu), out[r.nb[q]]) : fma_c(s, cx<R>(r.kg[q], (R)0), out[r.nb[q]]);
    }
    if (GB) {
#pragma unroll
        for (int q = 8; q < 12; q++)
            s = XM ? fma_c(s, coef_of(r.kg[q], r.km[q], wmu), out[r.nb[q]]) : fma_c(s, cx<R>(r.kg[q], (R)0), out[r.nb[q]]);
    }
    const cx<R> D = coef_of(r.kgd, r.kmd, wmu);
    const R rden = (R)1 / (D.re * D.re + D.im * D.im);
    const cx<R> Dinv(D.re * rden, -D.im * rden);
    if (DIR == 0) {
        out[e] = omega * ((in[e] - s) * Dinv);
    } else if (DIR == 1) {
        const cx<R> wv = ((R)(2 - omega) / omega) * (out[e] * D);
        out[e] = omega * ((wv - s) * Dinv);
    } else {  // backward Gauss-Seidel: (D + U) out = in
        out[e] = (in[e] - s) * Dinv;
    }
}

// ------------------------------------------------------------------ K3 vectors
// Generic fused vector update over all edges with an unscaled-residual norm:
//  mode 0 (COCR, 9 vectors; kernel micro-benchmark): x += alpha p; r -= alpha Ap;
//                         z -= alpha q; acc |r/ve|^2
//  mode 1 (BiCGStab s):   s = r - alpha v
//  mode 2 (BiCGStab end): x += alpha y + omega z; r = s - omega t; acc |r/ve|^2, conj(rh)^T r
//  mode 4 (BiCGStab p):   p = r + beta (p - omega v)
//  mode 5 (norm only):    acc |r/ve|^2 (r = v0)
//  mode 6 (COCR):         p = z + beta p; x += alpha p; r -= alpha Ap; z -= alpha q;
//                         acc |r/ve|^2 (the direction is kept one step behind)
template <typename R, int MODE>
__global__ __launch_bounds__(TPB) void k_vec(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                             const R *__restrict__ c, const double *__restrict__ slots,
                                             cx<R> *v0, cx<R> *v1, cx<R> *v2, cx<R> *v3, cx<R> *v4, cx<R> *v5,
                                             double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const bool inner = edge_interior(d, S.o, i, j, S.k);
        if (MODE == 0) {
            // after a breakdown (flag set by the scalar step) the update is a no-op,
            // so an iteration queued before the host saw the flag leaves x intact
            const bool brk = slots[SL_BREAK] != 0;
            const cx<R> al(brk ? (R)0 : (R)slots[SL_ALPHA], brk ? (R)0 : (R)slots[SL_ALPHA + 1]);
            cx<R> ap = v2[e], q = v4[e];
            v0[e] = fma_c(v0[e], al, v1[e]);
            cx<R> rr = v3[e] - al * ap;
            v3[e] = rr;
            v5[e] = v5[e] - al * q;
            if (inner) {
                double ve = (double)dual_volume(a, b, c, S.o, i, j, S.k);
                acc[0] = ((double)rr.re * rr.re + (double)rr.im * rr.im) / (ve * ve);
            }
        } else if (MODE == 1) {
            const cx<R> al((R)slots[SL_ALPHA], (R)slots[SL_ALPHA + 1]);
            v0[e] = v1[e] - al * v2[e];
        } else if (MODE == 2) {
            // v0 = x, v1 = y, v2 = z, v3 = r (out), v4 = s, v5 = t ; rhat via partial in acc[1..2]
            const cx<R> al((R)slots[SL_ALPHA], (R)slots[SL_ALPHA + 1]);
            const cx<R> om((R)slots[SL_OMEGA], (R)slots[SL_OMEGA + 1]);
            v0[e] = v0[e] + al * v1[e] + om * v2[e];
            cx<R> rr = v4[e] - om * v5[e];
            v3[e] = rr;
            if (inner) {
                double ve = (double)dual_volume(a, b, c, S.o, i, j, S.k);
                acc[0] = ((double)rr.re * rr.re + (double)rr.im * rr.im) / (ve * ve);
            }
        } else if (MODE == 4) {
            const cx<R> be((R)slots[SL_BETA], (R)slots[SL_BETA + 1]);
            const cx<R> om((R)slots[SL_OMEGA], (R)slots[SL_OMEGA + 1]);
            v0[e] = v1[e] + be * (v0[e] - om * v2[e]);
        } else if (MODE == 5) {
            if (inner) {
                cx<R> rr = v0[e];
                double ve = (double)dual_volume(a, b, c, S.o, i, j, S.k);
                acc[0] = ((double)rr.re * rr.re + (double)rr.im * rr.im) / (ve * ve);
            }
        } else if (MODE == 6) {
            // COCR update with the direction folded in (p is kept one step behind):
            // v0 = x, v1 = p, v2 = Ap, v3 = r, v4 = q, v5 = z
            //   p = z + beta p; x += alpha p; r -= alpha Ap; z -= alpha q
            // after a breakdown nothing is written, so x, p, r and z stay finite
            const bool brk = slots[SL_BREAK] != 0;
            cx<R> rr = v3[e];
            if (!brk) {
                const cx<R> al((R)slots[SL_ALPHA], (R)slots[SL_ALPHA + 1]);
                const cx<R> be((R)slots[SL_BETA], (R)slots[SL_BETA + 1]);
                const cx<R> zv = v5[e];
                const cx<R> pn = fma_c(zv, be, v1[e]);
                v1[e] = pn;
                v0[e] = fma_c(v0[e], al, pn);
                rr = rr - al * v2[e];
                v3[e] = rr;
                v5[e] = zv - al * v4[e];
            }
            if (inner) {
                double ve = (double)dual_volume(a, b, c, S.o, i, j, S.k);
                acc[0] = ((double)rr.re * rr.re + (double)rr.im * rr.im) / (ve * ve);
            }
        }
    }
    if (MODE == 0 || MODE == 2 || MODE == 5 || MODE == 6) block_reduce_store<4>(acc, partials);
}

// COCR direction of A: Ap = Az + beta Ap (copy: Ap = Az), with an optional
// complex64 copy of the result for a complex64 preconditioner (which then needs
// no conversion pass of its own)
template <typename R>
__global__ __launch_bounds__(TPB) void k_cocr_ap(int64_t n, const double *__restrict__ slots, cx<R> *__restrict__ Ap,
                                                 const cx<R> *__restrict__ Az, zf *__restrict__ ap32, int copy)
{
    const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (e >= n) return;
    cx<R> v = Az[e];
    if (!copy) {
        const cx<R> be((R)slots[SL_BETA], (R)slots[SL_BETA + 1]);
        v = fma_c(v, be, Ap[e]);
    }
    Ap[e] = v;
    if (ap32) ap32[e] = zf((float)v.re, (float)v.im);
}

// GPBiCG fused vector kernels (modes documented at launch_gpbicg)
template <typename R, int MODE>
__global__ __launch_bounds__(TPB) void k_gpbicg(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                                const R *__restrict__ c, const double *__restrict__ slots,
                                                GPVecs<R> v, double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const cx<R> al((R)slots[SL_ALPHA], (R)slots[SL_ALPHA + 1]);
        const cx<R> be((R)slots[SL_BETA], (R)slots[SL_BETA + 1]);
        const cx<R> om((R)slots[SL_OMEGA], (R)slots[SL_OMEGA + 1]);
        const cx<R> et((R)slots[SL_ETA], (R)slots[SL_ETA + 1]);
        if (MODE == 0) {
            v.w[e] = v.at[e] + be * v.ap[e];
            v.p[e] = v.rp[e] + be * (v.p[e] - v.u[e]);
        } else if (MODE == 1) {
            const cx<R> r = v.r[e], ap = v.ap[e];
            v.y[e] = v.t[e] - r - al * v.w[e] + al * ap;
            v.t[e] = r - al * ap;
        } else if (MODE == 2) {
            const cx<R> rp = v.rp[e];
            v.q[e] = v.tp[e] - rp + be * v.u[e];
            v.tp[e] = rp - al * v.tmp[e];
        } else if (MODE == 3) {
            const cx<R> y = v.y[e], at = v.at[e], t = v.t[e];
            acc[0] = (double)y.re * y.re + (double)y.im * y.im;
            acc[1] = (double)at.re * at.re + (double)at.im * at.im;
            acc[2] = (double)at.re * t.re + (double)at.im * t.im;   // conj(at) t
            acc[3] = (double)at.re * t.im - (double)at.im * t.re;
            acc[4] = (double)y.re * t.re + (double)y.im * t.im;     // conj(y) t
            acc[5] = (double)y.re * t.im - (double)y.im * t.re;
            acc[6] = (double)at.re * y.re + (double)at.im * y.im;   // conj(at) y
            acc[7] = (double)at.re * y.im - (double)at.im * y.re;
        } else {
            const cx<R> u = om * v.tmp[e] + et * v.q[e];
            v.u[e] = u;
            const cx<R> z = om * v.rp[e] + et * v.z[e] - al * u;
            v.z[e] = z;
            v.x[e] = v.x[e] + al * v.p[e] + z;
            const cx<R> r = v.t[e] - et * v.y[e] - om * v.at[e];
            v.r[e] = r;
            if (edge_interior(d, S.o, i, j, S.k)) {
                const double ve = (double)dual_volume(a, b, c, S.o, i, j, S.k);
                acc[0] = ((double)r.re * r.re + (double)r.im * r.im) / (ve * ve);
            }
            const cx<R> r0 = v.r0[e];
            acc[1] = (double)r0.re * r.re + (double)r0.im * r.im;
            acc[2] = (double)r0.re * r.im - (double)r0.im * r.re;
        }
    }
    if (MODE == 3) block_reduce_store<8>(acc, partials);
    if (MODE == 4) {
        double a3[3] = {acc[0], acc[1], acc[2]};
        block_reduce_store<3>(a3, partials);
    }
}

// g_c += Re(i wmu lam^T M_c e): derivative of the misfit with respect to the log
// of a scale factor on cell c's conductivity tensor.  M_c is the cell's
// corner-lumped contribution to the row-scaled edge mass: at each of its 8
// corners the x/y/z edges meeting there couple through (V_c/8) sigma_c, so
// that the operator's mass term is sum_c M_c and dA_s/dlog(s_c) = -i wmu M_c.
// One thread per cell; the 12 edge values of e and lam are read once.
__global__ __launch_bounds__(TPB) void k_cell_grad(Dims d, const double *__restrict__ a, const double *__restrict__ b,
                                                   const double *__restrict__ c, const double *__restrict__ sig,
                                                   double wmu, const zd *__restrict__ e, const zd *__restrict__ lam,
                                                   double *__restrict__ g)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k < d.nz) ? TL.x * TPB + threadIdx.x : kNoTile;
    if (p >= d.nx * d.ny) return;
    const int j = p / d.nx, i = p - j * d.nx;
    const int64_t cell = (int64_t)k * d.nx * d.ny + p;
    const double sxx = sig[cell], syy = sig[d.NC + cell], szz = sig[2 * d.NC + cell];
    const double sxy = sig[3 * d.NC + cell], sxz = sig[4 * d.NC + cell], syz = sig[5 * d.NC + cell];
    double sre = 0, sim = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int ci = q & 1, cj = (q >> 1) & 1, ck = (q >> 2) & 1;
        const int64_t kx = ex_id(d, i, j + cj, k + ck), ky = ey_id(d, i + ci, j, k + ck), kz = ez_id(d, i + ci, j + cj, k);
        const zd ex = e[kx], ey = e[ky], ez = e[kz];
        const zd lx = lam[kx], ly = lam[ky], lz = lam[kz];
        const zd sx = sxx * ex + sxy * ey + sxz * ez;
        const zd sy = sxy * ex + syy * ey + syz * ez;
        const zd sz = sxz * ex + syz * ey + szz * ez;
        const zd t = lx * sx + ly * sy + lz * sz;
        sre += t.re;
        sim += t.im;
    }
    const double V = a[i] * b[j] * c[k];
    g[cell] += -wmu * (V / 8.0) * sim;  // Re(i wmu (V/8) sum)
}

// v_e = 0 on boundary edges
__global__ __launch_bounds__(TPB) void k_zero_boundary(Dims d, zd *__restrict__ v)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int j = p / S.nI, i = p - j * S.nI;
    if (!edge_interior(d, S.o, i, j, S.k)) v[S.base + p] = zd(0);
}

// Surface fields at the cell centres of the surface node plane (post_process,
// src/fdfe_main.f90:549-640): E averaged from the two bounding edges, H = curl
// E / (i omega mu) with a one-sided z difference over the cell below the
// surface.  The surface is node plane nair (the flat datum) or, with
// topography, surf[i + nx*j] per column (0 <= surf < nz, checked by the host).
// One thread per surface cell; out[q*nx*ny + i + nx*j] for q = ex, ey, hx, hy, hz.
__global__ __launch_bounds__(TPB) void k_surface_fields(Dims d, int nair, const int *__restrict__ surf,
                                                        const double *__restrict__ a, const double *__restrict__ b,
                                                        const double *__restrict__ c, double om,
                                                        const zd *__restrict__ x, zd *__restrict__ out)
{
    const int p = blockIdx.x * TPB + threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    if (p >= nx * ny) return;
    const int j = p / nx, i = p - j * nx;
    const int m = surf ? surf[p] : nair;
    const int64_t n1 = d.XN, n2 = d.YN;
    const int64_t idx = (int64_t)m * nx * (ny + 1) + (int64_t)j * nx + i;
    const int64_t idx2 = n1 + (int64_t)m * ny * (nx + 1) + (int64_t)j * (nx + 1) + i;
    const int64_t idx3 = n1 + n2 + (int64_t)m * (nx + 1) * (ny + 1) + (int64_t)j * (nx + 1) + i;
    const int64_t P = (int64_t)nx * (ny + 1), Q = (int64_t)ny * (nx + 1);
    const double cz = c[m], ai = a[i], bj = b[j];
    const zd ex = (x[idx] + x[idx + nx]) * 0.5;
    const zd ey = (x[idx2] + x[idx2 + 1]) * 0.5;
    const zd exz = ((x[idx + P] + x[idx + P + nx]) * 0.5 - ex) * (1.0 / cz);
    const zd ezx = ((x[idx3 + 1] + x[idx3 + nx + 2]) * 0.5 - (x[idx3 + nx + 1] + x[idx3]) * 0.5) * (1.0 / ai);
    const zd eyz = ((x[idx2 + Q] + x[idx2 + Q + 1]) * 0.5 - ey) * (1.0 / cz);
    const zd ezy = ((x[idx3 + nx + 1] + x[idx3 + nx + 2]) * 0.5 - (x[idx3] + x[idx3 + 1]) * 0.5) * (1.0 / bj);
    const zd exy = (x[idx + nx] - x[idx]) * (1.0 / bj);
    const zd eyx = (x[idx2 + 1] - x[idx2]) * (1.0 / ai);
    // 1/(i om) = -i/om
    auto div_kap = [&](zd v) { return zd(v.im / om, -v.re / om); };
    const int64_t s = p, ns = (int64_t)nx * ny;
    out[s] = ex;
    out[ns + s] = ey;
    out[2 * ns + s] = div_kap(ezy - eyz);
    out[3 * ns + s] = div_kap(exz - ezx);
    out[4 * ns + s] = div_kap(exy - eyx);
}

// generic dots over all edges: mode 0: u^T v (unconjugated); 1: conj(u)^T v
template <typename R, int MODE>
__global__ __launch_bounds__(TPB) void k_dot(int64_t n, const cx<R> *__restrict__ u, const cx<R> *__restrict__ v,
                                             double *__restrict__ partials)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    double acc[2] = {0, 0};
    if (t < n) {
        cx<R> x = u[t], y = v[t];
        if (MODE == 0) {
            acc[0] = (double)x.re * y.re - (double)x.im * y.im;
            acc[1] = (double)x.re * y.im + (double)x.im * y.re;
        } else {
            acc[0] = (double)x.re * y.re + (double)x.im * y.im;
            acc[1] = (double)x.re * y.im - (double)x.im * y.re;
        }
    }
    // reuse the 2-D block index layout of block_reduce_store (gridDim.y == 1)
    block_reduce_store<2>(acc, partials);
}

template <typename R>
__global__ __launch_bounds__(TPB) void k_copy(int64_t n, const cx<R> *__restrict__ s, cx<R> *__restrict__ t)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q < n) t[q] = s[q];
}

// Streaming copy tuned for bandwidth (measurement only): 16-byte units, U loads
// in flight per thread before the U non-temporal stores; the block's U x TPB
// units are contiguous and each of the U load instructions is coalesced.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <int U>
__global__ __launch_bounds__(TPB) void k_stream_copy(int64_t n16, const u32x4 *__restrict__ s, u32x4 *__restrict__ t)
{
    const int64_t base = (int64_t)blockIdx.x * (TPB * U) + threadIdx.x;
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t q = base + (int64_t)u * TPB;
        if (q < n16) v[u] = s[q];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t q = base + (int64_t)u * TPB;
        if (q < n16) __builtin_nontemporal_store(v[u], &t[q]);
    }
}

template <typename RS, typename RD>
__global__ __launch_bounds__(TPB) void k_convert(int64_t n, const cx<RS> *__restrict__ s, cx<RD> *__restrict__ t)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q < n) t[q] = cx<RD>(s[q]);
}

// x_hi += (hi) d_lo  (iterative refinement update, interior only: d is 0 on the boundary)
__global__ __launch_bounds__(TPB) void k_add_lo(int64_t n, zd *__restrict__ x, const zf *__restrict__ dlo)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q < n) {
        zf v = dlo[q];
        x[q] = x[q] + zd((double)v.re, (double)v.im);
    }
}

template <typename R>
__global__ __launch_bounds__(TPB) void k_axpby(int64_t n, cx<R> al, const cx<R> *__restrict__ u, cx<R> be,
                                               cx<R> *__restrict__ v)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q < n) v[q] = al * u[q] + be * v[q];
}

// ------------------------------------------------------------------ K6 Dirichlet
// x_e = g_e on boundary edges; interior edges untouched (or zeroed if zero_inner).
template <typename R>
__global__ __launch_bounds__(TPB) void k_dirichlet(Dims d, const cx<R> *__restrict__ edc, int pol,
                                                   cx<R> *__restrict__ x, int zero_inner, int full)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int j = p / S.nI, i = p - j * S.nI;
    const int64_t e = S.base + p;
    if (edge_interior(d, S.o, i, j, S.k)) {
        if (zero_inner) x[e] = cx<R>(0);
        return;
    }
    cx<R> v(0);
    dirichlet_value(d, edc, pol, S.o, i, j, S.k, v, full != 0);
    x[e] = v;
}

// ------------------------------------------------------------------ K5 divergence correction
// t_n = (G^T w)_n = sum over incident edges of +-w_e/l_e on interior nodes, 0 on
// the boundary; acc |t|^2.
template <typename R>
__global__ __launch_bounds__(TPB) void k_node_div(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                                  const R *__restrict__ c, const cx<R> *__restrict__ w,
                                                  cx<R> *__restrict__ t, double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k <= d.nz) ? TL.x * TPB + threadIdx.x : kNoTile;
    const int nI = d.nx + 1, nJ = d.ny + 1;
    double acc[2] = {0, 0};
    if (p < nI * nJ) {
        const int j = p / nI, i = p - j * nI;
        const int64_t n = (int64_t)k * nI * nJ + p;
        cx<R> out(0);
        if (!node_boundary(d, i, j, k)) {
            // (G^T w)_n: +w/l on edges ending at n, -w/l on edges starting at n
            out = w[ex_id(d, i - 1, j, k)] * ((R)1 / a[i - 1]) - w[ex_id(d, i, j, k)] * ((R)1 / a[i]);
            out = out + w[ey_id(d, i, j - 1, k)] * ((R)1 / b[j - 1]) - w[ey_id(d, i, j, k)] * ((R)1 / b[j]);
            out = out + w[ez_id(d, i, j, k - 1)] * ((R)1 / c[k - 1]) - w[ez_id(d, i, j, k)] * ((R)1 / c[k]);
        }
        t[n] = out;
        acc[0] = (double)out.re * out.re + (double)out.im * out.im;
    }
    block_reduce_store<2>(acc, partials);
}

// P = G^T M_sigma G, 19-point, assembled per node from the 8 surrounding cells
// (corner quadrature of int sigma grad phi . grad psi).  Boundary rows: identity.
// Slot s <-> offset (dx,dy,dz) in P19_OFF.
__constant__ int c_p19[19][3] = {{0, 0, 0},  {-1, 0, 0}, {1, 0, 0},   {0, -1, 0}, {0, 1, 0},  {0, 0, -1}, {0, 0, 1},
                                 {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                                 {1, 0, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};

template <typename R>
__global__ __launch_bounds__(TPB) void k_pbuild(Dims d, const double *__restrict__ a, const double *__restrict__ b,
                                                const double *__restrict__ c, const double *__restrict__ sig,
                                                R *__restrict__ P)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k <= d.nz) ? TL.x * TPB + threadIdx.x : kNoTile;
    const int nI = d.nx + 1, nJ = d.ny + 1;
    if (p >= nI * nJ) return;
    const int j = p / nI, i = p - j * nI;
    const int64_t n = (int64_t)k * nI * nJ + p;
    double band[27];
    for (int s = 0; s < 27; s++) band[s] = 0;
    if (node_boundary(d, i, j, k)) {
        band[13] = 1.0;
    } else {
        for (int oz = 0; oz < 2; oz++)
            for (int oy = 0; oy < 2; oy++)
                for (int ox = 0; ox < 2; ox++) {
                    const int ci = i - 1 + ox, cj = j - 1 + oy, ck = k - 1 + oz;
                    const int64_t ce = cell_id(d, ci, cj, ck);
                    const double h[3] = {a[ci], b[cj], c[ck]};
                    const double V8 = h[0] * h[1] * h[2] / 8;
                    double s3[3][3];
                    s3[0][0] = sig[ce]; s3[1][1] = sig[d.NC + ce]; s3[2][2] = sig[2 * d.NC + ce];
                    s3[0][1] = s3[1][0] = sig[3 * d.NC + ce];
                    s3[0][2] = s3[2][0] = sig[4 * d.NC + ce];
                    s3[1][2] = s3[2][1] = sig[5 * d.NC + ce];
                    const int l[3] = {1 - ox, 1 - oy, 1 - oz};  // node position inside the cell
                    for (int cz = 0; cz < 2; cz++)
                        for (int cy = 0; cy < 2; cy++)
                            for (int cxn = 0; cxn < 2; cxn++) {
                                const int cc[3] = {cxn, cy, cz};
                                for (int d1 = 0; d1 < 3; d1++) {
                                    bool on = true;
                                    for (int dd = 0; dd < 3; dd++)
                                        if (dd != d1 && l[dd] != cc[dd]) on = false;
                                    if (!on) continue;
                                    const double g1 = (l[d1] == 1 ? 1.0 : -1.0) / h[d1];
                                    for (int d2 = 0; d2 < 3; d2++) {
                                        const double wv = V8 * s3[d1][d2] * g1 / h[d2];
                                        for (int eend = 0; eend < 2; eend++) {
                                            int m[3] = {cc[0], cc[1], cc[2]};
                                            m[d2] = eend;
                                            const int sx = m[0] - l[0] + 1, sy = m[1] - l[1] + 1, sz = m[2] - l[2] + 1;
                                            band[sz * 9 + sy * 3 + sx] += (eend ? wv : -wv);
                                        }
                                    }
                                }
                            }
                }
    }
    for (int s = 0; s < 19; s++) {
        const int sx = c_p19[s][0] + 1, sy = c_p19[s][1] + 1, sz = c_p19[s][2] + 1;
        P[(int64_t)s * d.NP + n] = (R)band[sz * 9 + sy * 3 + sx];
    }
}

// Symmetric form of P for the Jacobi smoother, stored after the 19 slots:
// slot 19 = 1/diag(P), slots 20..28 = the 9 couplings to the upper neighbours
// (positive linear offset).  The coupling of node n to a lower neighbour
// m = n - o is the upper coupling of m, P_nm = P_mn, so one pass reads 10
// values per node from memory instead of 19 (boundary rows, whose couplings
// are dropped this way, only ever meet a zero input).
__constant__ int c_pup[9] = {2, 4, 6, 9, 10, 13, 14, 17, 18};

template <typename R>
__global__ __launch_bounds__(TPB) void k_psym(int64_t np, R *__restrict__ P)
{
    const int64_t n = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (n >= np) return;
    P[19 * np + n] = (R)1 / P[n];
#pragma unroll
    for (int q = 0; q < 9; q++) P[(20 + q) * np + n] = P[(int64_t)c_pup[q] * np + n];
}

// y = P x (+ acc conj(x)^T y when DOT)
template <typename R, int DOT>
__global__ __launch_bounds__(TPB) void k_papply(Dims d, const R *__restrict__ P, const cx<R> *__restrict__ x,
                                                cx<R> *__restrict__ y, double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k <= d.nz) ? TL.x * TPB + threadIdx.x : kNoTile;
    const int nI = d.nx + 1, nJ = d.ny + 1;
    double acc[2] = {0, 0};
    if (p < nI * nJ) {
        const int j = p / nI, i = p - j * nI;
        const int64_t n = (int64_t)k * nI * nJ + p;
        cx<R> out(0);
        if (node_boundary(d, i, j, k)) {
            out = x[n];
        } else {
            const int64_t sy = nI, sz = (int64_t)nI * nJ;
#pragma unroll
            for (int s = 0; s < 19; s++) {
                const int64_t off = c_p19[s][0] + c_p19[s][1] * sy + c_p19[s][2] * sz;
                out = out + P[(int64_t)s * d.NP + n] * x[n + off];
            }
        }
        y[n] = out;
        if (DOT) {
            cx<R> xv = x[n];
            acc[0] = (double)xv.re * out.re + (double)xv.im * out.im;
            acc[1] = (double)xv.re * out.im - (double)xv.im * out.re;
        }
    }
    if (DOT) block_reduce_store<2>(acc, partials);
}

// PCG vector steps (Jacobi, P diagonal in slot 0):
//  mode 0: u += alpha p; r -= alpha v; acc sum |r|^2/mj, |r|^2
//  mode 1: p = r/mj + beta p
//  mode 2: init: u = 0; r = t; p = r/mj; acc sum |r|^2/mj
template <typename R, int MODE>
__global__ __launch_bounds__(TPB) void k_pcg(int64_t np, const R *__restrict__ Pdiag, const double *__restrict__ slots,
                                             cx<R> *u, cx<R> *r, cx<R> *p, const cx<R> *v, double *__restrict__ partials)
{
    const int64_t n = (int64_t)blockIdx.x * TPB + threadIdx.x;
    double acc[2] = {0, 0};
    if (n < np) {
        const R mj = Pdiag[n];
        if (MODE == 0) {
            const cx<R> al((R)slots[SL_PALPHA], (R)slots[SL_PALPHA + 1]);
            u[n] = fma_c(u[n], al, p[n]);
            cx<R> rr = r[n] - al * v[n];
            r[n] = rr;
            double a2 = (double)rr.re * rr.re + (double)rr.im * rr.im;
            acc[0] = a2 / mj;
            acc[1] = a2;
        } else if (MODE == 1) {
            const cx<R> be((R)slots[SL_PBETA], (R)slots[SL_PBETA + 1]);
            p[n] = fma_c(((R)1 / mj) * r[n], be, p[n]);
        } else {
            cx<R> rr = v[n];
            u[n] = cx<R>(0);
            r[n] = rr;
            p[n] = ((R)1 / mj) * rr;
            acc[0] = ((double)rr.re * rr.re + (double)rr.im * rr.im) / mj;
        }
    }
    if (MODE != 1) block_reduce_store<2>(acc, partials);
}

// x_e -= (phi(end) - phi(start))/l_e for edges not having both ends on the boundary
template <typename R>
__global__ __launch_bounds__(TPB) void k_grad_update(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                                     const R *__restrict__ c, const cx<R> *__restrict__ phi,
                                                     cx<R> *__restrict__ x)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int j = p / S.nI, i = p - j * S.nI;
    const int64_t e = S.base + p;
    int64_t n0, n1;
    const R l = edge_len_nodes(d, a, b, c, S.o, i, j, S.k, n0, n1);
    int i1 = i, j1 = j, k1 = S.k;
    if (S.o == 0) i1++; else if (S.o == 1) j1++; else k1++;
    if (node_boundary(d, i, j, S.k) && node_boundary(d, i1, j1, k1)) return;
    x[e] = x[e] - (phi[n1] - phi[n0]) * ((R)1 / l);
}

// ------------------------------------------------------------------ K7 auxiliary nodal space
// Hybrid (Hiptmair-type) preconditioner pieces.  The gradient fields G phi are
// the near-null space of A_s where sigma is tiny (the air); their Galerkin
// operator is G^T A_s G = -kappa P exactly (K G = 0), so the preconditioner
// adds G [(-1/kappa) SSOR_P^-1] G^T to the edge SSOR.
// Nodal colours (i&1) + 2(j&1) + 4(k&1): the 19-point P never couples two
// nodes of one colour.
template <typename R, int COL, int DIR>
__global__ __launch_bounds__(TPB) void k_pssor(Dims d, const R *__restrict__ P, const cx<R> *__restrict__ t,
                                               cx<R> *__restrict__ phi)
{
    constexpr int ci = COL & 1, cj = (COL >> 1) & 1, ck = (COL >> 2) & 1;
    const Tile TL = tile_id();
    const int k = ck + 2 * TL.y;
    const int W = (d.nx + 2) / 2;
    const int p = TL.x * TPB + threadIdx.x;
    const int r = p / W, q = p - r * W;
    const int i = ci + 2 * q, j = cj + 2 * r;
    if (k > d.nz || i > d.nx || j > d.ny) return;
    const int nI = d.nx + 1, nJ = d.ny + 1;
    const int64_t n = ((int64_t)k * nJ + j) * nI + i;
    if (node_boundary(d, i, j, k)) {
        if (DIR != 1) phi[n] = cx<R>(0);
        return;
    }
    const int64_t sy = nI, sz = (int64_t)nI * nJ;
    cx<R> s(0);
#pragma unroll
    for (int q2 = 1; q2 < 19; q2++) {
        const int dx = c_p19[q2][0], dy = c_p19[q2][1], dz = c_p19[q2][2];
        const int cm = COL ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
        const bool take = (DIR == 0) ? (cm < COL) : (cm > COL);  // DIR 1, 2: upper colours
        if (take) s = s + P[(int64_t)q2 * d.NP + n] * phi[n + dx + dy * sy + dz * sz];
    }
    const R dinv = (R)1 / P[n];
    if (DIR == 0 || DIR == 2) phi[n] = (t[n] - s) * dinv;
    else phi[n] = phi[n] - s * dinv;
}

// t = G^T r on interior nodes (0 on the boundary), no reduction; with SCALE
// the result is multiplied by dinv = 1/diag(P) (the input of k_pjacobi)
template <typename R, bool SCALE = false, typename RP = R>
__global__ __launch_bounds__(TPB) void k_gradT(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                               const R *__restrict__ c, const cx<R> *__restrict__ w,
                                               cx<R> *__restrict__ t, const RP *__restrict__ dinv)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k <= d.nz) ? TL.x * TPB + threadIdx.x : kNoTile;
    const int nI = d.nx + 1, nJ = d.ny + 1;
    if (p >= nI * nJ) return;
    const int j = p / nI, i = p - j * nI;
    const int64_t n = (int64_t)k * nI * nJ + p;
    cx<R> out(0);
    if (!node_boundary(d, i, j, k)) {
        out = w[ex_id(d, i - 1, j, k)] * ((R)1 / a[i - 1]) - w[ex_id(d, i, j, k)] * ((R)1 / a[i]);
        out = out + w[ey_id(d, i, j - 1, k)] * ((R)1 / b[j - 1]) - w[ey_id(d, i, j, k)] * ((R)1 / b[j]);
        out = out + w[ez_id(d, i, j, k - 1)] * ((R)1 / c[k - 1]) - w[ez_id(d, i, j, k)] * ((R)1 / c[k]);
        if (SCALE) out = out * (R)dinv[n];
    }
    t[n] = out;
}

// z_e += s * (G phi)_e; with DOT also the partial sums of w^T z (unconjugated)
// z_e + s (G phi)_e: the hybrid preconditioner's output on edge e, z the edge
// SSOR output.  Shared by k_grad_add and k_cocr_upd_grad.
template <typename R>
__device__ __forceinline__ cx<R> grad_add_value(const Dims &d, const R *__restrict__ a, const R *__restrict__ b,
                                                const R *__restrict__ c, cx<R> s, const cx<R> *__restrict__ phi,
                                                cx<R> z, const Slice &S, int i, int j)
{
    int64_t n0, n1;
    const R l = edge_len_nodes(d, a, b, c, S.o, i, j, S.k, n0, n1);
    const cx<R> g = (phi[n1] - phi[n0]) * ((R)1 / l);
    return fma_c(z, s, g);
}

// z += s G phi (STORE), with the unconjugated dot w^T (z + s G phi) (DOT)
template <typename R, bool DOT, bool STORE = true>
__global__ __launch_bounds__(TPB) void k_grad_add(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                                  const R *__restrict__ c, cx<R> s, const cx<R> *__restrict__ phi,
                                                  cx<R> *__restrict__ z, const cx<R> *__restrict__ w,
                                                  double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[2] = {0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const cx<R> zn = grad_add_value<R>(d, a, b, c, s, phi, z[e], S, i, j);
        if (STORE) z[e] = zn;
        if (DOT) {
            const cx<R> wv = w[e];
            acc[0] = (double)wv.re * zn.re - (double)wv.im * zn.im;
            acc[1] = (double)wv.re * zn.im + (double)wv.im * zn.re;
        }
    }
    if (DOT) block_reduce_store<2>(acc, partials);
}

// COCR update of k_vec mode 6 with q = zs + s G phi formed here from the edge
// SSOR output zs and the nodal correction phi of launch_hybrid (not stored).
template <typename R>
__global__ __launch_bounds__(TPB) void k_cocr_upd_grad(Dims d, const R *__restrict__ a, const R *__restrict__ b,
                                                       const R *__restrict__ c, const double *__restrict__ slots,
                                                       cx<R> *x, cx<R> *pv, const cx<R> *__restrict__ Ap, cx<R> *r,
                                                       cx<R> *z, const cx<R> *__restrict__ zs,
                                                       const cx<R> *__restrict__ phi, cx<R> s,
                                                       double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const bool inner = edge_interior(d, S.o, i, j, S.k);
        const bool brk = slots[SL_BREAK] != 0;
        cx<R> rr = r[e];
        if (!brk) {
            const cx<R> al((R)slots[SL_ALPHA], (R)slots[SL_ALPHA + 1]);
            const cx<R> be((R)slots[SL_BETA], (R)slots[SL_BETA + 1]);
            const cx<R> zv = z[e];
            const cx<R> pn = fma_c(zv, be, pv[e]);
            pv[e] = pn;
            x[e] = fma_c(x[e], al, pn);
            rr = rr - al * Ap[e];
            r[e] = rr;
            z[e] = zv - al * grad_add_value<R>(d, a, b, c, s, phi, zs[e], S, i, j);
        }
        if (inner) {
            double ve = (double)dual_volume(a, b, c, S.o, i, j, S.k);
            acc[0] = ((double)rr.re * rr.re + (double)rr.im * rr.im) / (ve * ve);
        }
    }
    block_reduce_store<4>(acc, partials);
}

template <typename R>
void launch_cocr_upd_grad(const DevGrid<R> &g, RedBuf &rb, cx<R> *x, cx<R> *pv, const cx<R> *Ap, cx<R> *r, cx<R> *z,
                          const cx<R> *zs, const cx<R> *phi, R wmu, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
    hipLaunchKernelGGL((k_cocr_upd_grad<R>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, rb.slots, x, pv, Ap, r, z, zs,
                       phi, cx<R>((R)0, (R)1 / wmu), rb.partials);
}

// Mixed-precision recombination of the hybrid preconditioner:
// z64_e = z32_e + s * (G phi)_e (phi in complex64 or complex128), with G
// applied in FP64 so that later FP64 operator applications see the gradient
// part exactly; the complex64 SSOR output z32 is widened here, in the same pass.
// The hybrid preconditioner's output on edge e of slice S: widened complex64
// edge SSOR value plus s times the gradient of the nodal correction.  Shared by
// k_hyb_combine and k_cocr_upd_hyb, so both form the same value.
template <typename RP>
__device__ __forceinline__ zd hyb_value(const Dims &d, const double *__restrict__ a, const double *__restrict__ b,
                                        const double *__restrict__ c, zd s, const cx<RP> *__restrict__ phi,
                                        const zf *__restrict__ z32, const Slice &S, int i, int j, int64_t e)
{
    int64_t n0, n1;
    const double l = edge_len_nodes(d, a, b, c, S.o, i, j, S.k, n0, n1);
    const cx<RP> f1 = phi[n1], f0 = phi[n0];
    const zf zz = z32[e];
    const zd g(((double)f1.re - (double)f0.re) / l, ((double)f1.im - (double)f0.im) / l);
    return fma_c(zd((double)zz.re, (double)zz.im), s, g);
}

// z64 = widen(z32) + s G phi (STORE), with the unconjugated dot w^T z64 (DOT)
template <typename RP, bool DOT, bool STORE = true>
__global__ __launch_bounds__(TPB) void k_hyb_combine(Dims d, const double *__restrict__ a, const double *__restrict__ b,
                                                     const double *__restrict__ c, zd s, const cx<RP> *__restrict__ phi,
                                                     const zf *__restrict__ z32, zd *__restrict__ z64,
                                                     const zd *__restrict__ w, double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[2] = {0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const zd zn = hyb_value<RP>(d, a, b, c, s, phi, z32, S, i, j, e);
        if (STORE) z64[e] = zn;
        if (DOT) {
            const zd wv = w[e];
            acc[0] = wv.re * zn.re - wv.im * zn.im;
            acc[1] = wv.re * zn.im + wv.im * zn.re;
        }
    }
    if (DOT) block_reduce_store<2>(acc, partials);
}

// COCR update of k_vec mode 6 with q = M Ap formed from the complex64-edge
// preconditioner's pieces (z32, phi: the hybrid preconditioner's, or the
// complex64 V-cycle's with phi in complex64) instead of read from a stored
// vector: p = z + beta p; x += alpha p; r -= alpha Ap; z -= alpha q; |r/ve|^2.
template <typename RP>
__global__ __launch_bounds__(TPB) void k_cocr_upd_hyb(Dims d, const double *__restrict__ a,
                                                      const double *__restrict__ b, const double *__restrict__ c,
                                                      const double *__restrict__ slots, zd *x, zd *pv,
                                                      const zd *__restrict__ Ap, zd *r, zd *z, const zf *__restrict__ z32,
                                                      const cx<RP> *__restrict__ phi, zd s, double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const bool inner = edge_interior(d, S.o, i, j, S.k);
        const bool brk = slots[SL_BREAK] != 0;
        zd rr = r[e];
        if (!brk) {
            const zd al(slots[SL_ALPHA], slots[SL_ALPHA + 1]);
            const zd be(slots[SL_BETA], slots[SL_BETA + 1]);
            const zd zv = z[e];
            const zd pn = fma_c(zv, be, pv[e]);
            pv[e] = pn;
            x[e] = fma_c(x[e], al, pn);
            rr = rr - al * Ap[e];
            r[e] = rr;
            z[e] = zv - al * hyb_value<RP>(d, a, b, c, s, phi, z32, S, i, j, e);
        }
        if (inner) {
            double ve = dual_volume(a, b, c, S.o, i, j, S.k);
            acc[0] = (rr.re * rr.re + rr.im * rr.im) / (ve * ve);
        }
    }
    block_reduce_store<4>(acc, partials);
}

void launch_cocr_upd_hyb(const DevGrid<double> &g, RedBuf &rb, zd *x, zd *pv, const zd *Ap, zd *r, zd *z,
                         const zf *z32, const zd *phi, double wmu, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
    hipLaunchKernelGGL(k_cocr_upd_hyb<double>, L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, rb.slots, x, pv, Ap, r, z,
                       z32, phi, zd(0.0, 1.0 / wmu), rb.partials);
}

void launch_cocr_upd_hyb(const DevGrid<double> &g, RedBuf &rb, zd *x, zd *pv, const zd *Ap, zd *r, zd *z,
                         const zf *z32, const zf *phi, double wmu, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
    hipLaunchKernelGGL(k_cocr_upd_hyb<float>, L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, rb.slots, x, pv, Ap, r, z,
                       z32, phi, zd(0.0, 1.0 / wmu), rb.partials);
}

// COCR with complex64 directions in split form (edge part and node values:
// v = v_e + s G v_psi, s = i/wmu), complex128 solution and residual.  The node
// part runs first: p_psi = z_psi + beta p_psi (the direction kept one step
// behind), z_psi -= alpha q_psi.
__global__ __launch_bounds__(TPB) void k_cocr_split_node(int64_t np, const double *__restrict__ slots,
                                                         zf *__restrict__ ppsi, zf *__restrict__ zpsi,
                                                         const zf *__restrict__ qpsi)
{
    const int64_t n = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (n >= np || slots[SL_BREAK] != 0) return;
    const zd al(slots[SL_ALPHA], slots[SL_ALPHA + 1]), be(slots[SL_BETA], slots[SL_BETA + 1]);
    const zf zv = zpsi[n], pv = ppsi[n], qv = qpsi[n];
    const zd pn = fma_c(zd(zv.re, zv.im), be, zd(pv.re, pv.im));
    const zd zn = zd(zv.re, zv.im) - al * zd(qv.re, qv.im);
    ppsi[n] = zf((float)pn.re, (float)pn.im);
    zpsi[n] = zf((float)zn.re, (float)zn.im);
}

// The edge part: p_e = z_e + beta p_e; x += alpha (p_e + s G p_psi) (in FP64,
// from the stored complex64 direction and the already updated p_psi);
// r -= alpha Ap; z_e -= alpha q_e; partial sums of |r/ve|^2 on interior edges.
__global__ __launch_bounds__(TPB) void k_cocr_split_upd(Dims d, const double *__restrict__ a,
                                                        const double *__restrict__ b, const double *__restrict__ c,
                                                        const double *__restrict__ slots, zd *__restrict__ x,
                                                        zd *__restrict__ r, const zf *__restrict__ Ap,
                                                        zf *__restrict__ ze, const zf *__restrict__ qe,
                                                        zf *__restrict__ pe, const zf *__restrict__ ppsi, zd s,
                                                        double *__restrict__ partials)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(d, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (p < S.nI * S.nJ) {
        const int j = p / S.nI, i = p - j * S.nI;
        const int64_t e = S.base + p;
        const bool inner = edge_interior(d, S.o, i, j, S.k);
        zd rr = r[e];
        if (slots[SL_BREAK] == 0) {
            const zd al(slots[SL_ALPHA], slots[SL_ALPHA + 1]), be(slots[SL_BETA], slots[SL_BETA + 1]);
            const zf zv = ze[e], pv = pe[e], av = Ap[e], qv = qe[e];
            const zd pn = fma_c(zd(zv.re, zv.im), be, zd(pv.re, pv.im));
            const zf pn32((float)pn.re, (float)pn.im);
            pe[e] = pn32;
            int64_t n0, n1;
            const double l = edge_len_nodes(d, a, b, c, S.o, i, j, S.k, n0, n1);
            const zf f1 = ppsi[n1], f0 = ppsi[n0];
            const zd g(((double)f1.re - (double)f0.re) / l, ((double)f1.im - (double)f0.im) / l);
            const zd pf = fma_c(zd(pn32.re, pn32.im), s, g);
            x[e] = fma_c(x[e], al, pf);
            rr = rr - al * zd(av.re, av.im);
            r[e] = rr;
            const zd zn = zd(zv.re, zv.im) - al * zd(qv.re, qv.im);
            ze[e] = zf((float)zn.re, (float)zn.im);
        }
        if (inner) {
            const double ve = dual_volume(a, b, c, S.o, i, j, S.k);
            acc[0] = (rr.re * rr.re + rr.im * rr.im) / (ve * ve);
        }
    }
    block_reduce_store<4>(acc, partials);
}

void launch_cocr_split_upd(const DevGrid<double> &g, RedBuf &rb, zd *x, zd *r, const zf *Ap, zf *ze, const zf *qe,
                           zf *pe, zf *ppsi, zf *zpsi, const zf *qpsi, double wmu, hipStream_t st)
{
    const int64_t np = g.d.NP;
    hipLaunchKernelGGL(k_cocr_split_node, dim3((unsigned)((np + TPB - 1) / TPB)), dim3(TPB), 0, st, np, rb.slots,
                       ppsi, zpsi, qpsi);
    SliceLaunch L = edge_launch(g.d);
    hipLaunchKernelGGL(k_cocr_split_upd, L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, rb.slots, x, r, Ap, ze, qe, pe,
                       (const zf *)ppsi, zd(0.0, 1.0 / wmu), rb.partials);
}

// ------------------------------------------------------------------ K8 multigrid transfers
// Edge prolongation between nested grids (each coarse cell merges one or two
// fine cells along every axis, MGTables): the tangential field is constant
// along a coarse edge and linear across it, so a fine edge takes the coarse
// edge containing it (along) with linear weights over the one or two
// bracketing coarse nodes in each transverse direction.  Restriction is the
// exact transpose.

// fine node q of axis a lies strictly inside a coarse cell (two bracketing
// coarse nodes) rather than on a coarse node: its weight is below 1 (a fine
// node on a coarse node has weight exactly 1, one inside lies a whole fine
// width past the lower coarse node)
__device__ __forceinline__ bool node_between(const MGTables &T, int a, int q) { return T.wl[a][q] != 1.0; }

// Layout of one edge orientation seen along the edge: edge (a along o; q1, q2
// across, in the dimensions t1 < t2 other than o) has index
// base + a*sa + q1*s1 + q2*s2, and it is interior when 0 < q1 < n1 and
// 0 < q2 < n2.
struct EdgeAxes {
    int64_t base, sa, s1, s2;
    int n1, n2;
};

__device__ __forceinline__ EdgeAxes edge_axes(const Dims &d, int o)
{
    EdgeAxes A;
    if (o == 0) {
        A.base = 0; A.sa = 1; A.s1 = d.nx; A.s2 = (int64_t)d.nx * (d.ny + 1); A.n1 = d.ny; A.n2 = d.nz;
    } else if (o == 1) {
        A.base = d.XN; A.sa = d.nx + 1; A.s1 = 1; A.s2 = (int64_t)d.ny * (d.nx + 1); A.n1 = d.nx; A.n2 = d.nz;
    } else {
        A.base = d.XN + d.YN; A.sa = (int64_t)(d.nx + 1) * (d.ny + 1); A.s1 = 1; A.s2 = d.nx + 1; A.n1 = d.nx;
        A.n2 = d.ny;
    }
    return A;
}

// zf_e += (P zc)_e on interior fine edges: the coarse edge containing e along
// its direction, linear weights over the (one or two) bracketing coarse nodes
// in each transverse direction; terms in the order (lo,lo), (lo,hi), (hi,lo),
// (hi,hi), zero weights and boundary coarse edges skipped
template <typename R>
__global__ __launch_bounds__(TPB) void k_prolong_add(Dims df, Dims dc, MGTables T, const cx<R> *__restrict__ zc,
                                                     cx<R> *__restrict__ zf)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(df, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int j = p / S.nI, i = p - j * S.nI, k = S.k;
    if (!edge_interior(df, S.o, i, j, k)) return;
    const int64_t e = S.base + p;
    const int o = S.o, t1 = (o == 0) ? 1 : 0, t2 = (o == 2) ? 1 : 2;
    const int ia = (o == 0) ? i : (o == 1) ? j : k;
    const int q1 = (o == 0) ? j : i, q2 = (o == 2) ? j : k;
    const EdgeAxes C = edge_axes(dc, o);
    const int l1 = T.lo[t1][q1], l2 = T.lo[t2][q2];
    const bool h1 = node_between(T, t1, q1), h2 = node_between(T, t2, q2);
    const double w1 = h1 ? T.wl[t1][q1] : 1.0, w2 = h2 ? T.wl[t2][q2] : 1.0;
    const int64_t c00 = C.base + (int64_t)T.lo[o][ia] * C.sa + (int64_t)l1 * C.s1 + (int64_t)l2 * C.s2;
    double sre = 0, sim = 0;
    auto term = [&](int c1, int c2, double w, int64_t ce) {
        if (w == 0 || c1 <= 0 || c1 >= C.n1 || c2 <= 0 || c2 >= C.n2) return;
        const cx<R> v = zc[ce];
        sre += w * (double)v.re;
        sim += w * (double)v.im;
    };
    term(l1, l2, w1 * w2, c00);
    if (h2) term(l1, l2 + 1, w1 * (1.0 - w2), c00 + C.s2);
    if (h1) term(l1 + 1, l2, (1.0 - w1) * w2, c00 + C.s1);
    if (h1 && h2) term(l1 + 1, l2 + 1, (1.0 - w1) * (1.0 - w2), c00 + C.s1 + C.s2);
    zf[e] = zf[e] + cx<R>((R)sre, (R)sim);
}

// rc = P^T rf on interior coarse edges (0 on the coarse boundary): the 1 or 2
// fine edges along the coarse edge times the fine nodes across it (the coarse
// node's own and, where they lie inside a coarse cell, its two neighbours),
// summed transverse offset 1 outer, offset 2, then along (interior fine edges)
template <typename R>
__global__ __launch_bounds__(TPB) void k_restrict(Dims df, Dims dc, MGTables T, const cx<R> *__restrict__ rf,
                                                  cx<R> *__restrict__ rc)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(dc, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int J = p / S.nI, I = p - J * S.nI, K = S.k;
    const int64_t ec = S.base + p;
    if (!edge_interior(dc, S.o, I, J, K)) { rc[ec] = cx<R>(0); return; }
    const int o = S.o, t1 = (o == 0) ? 1 : 0, t2 = (o == 2) ? 1 : 2;
    const int A = (o == 0) ? I : (o == 1) ? J : K;
    const int C1 = (o == 0) ? J : I, C2 = (o == 2) ? J : K;
    const EdgeAxes F = edge_axes(df, o);
    const int a0 = T.st[o][A], na = T.st[o][A + 1] - a0;
    const int64_t fa0 = F.base + (int64_t)a0 * F.sa;
    double sre = 0, sim = 0;
#pragma unroll
    for (int f1 = -1; f1 <= 1; f1++) {
        const int q1 = T.st[t1][C1] + f1;
        if (q1 <= 0 || q1 >= F.n1 || (f1 != 0 && !node_between(T, t1, q1))) continue;
        const double w1 = f1 == 0 ? 1.0 : (f1 < 0 ? 1.0 - T.wl[t1][q1] : T.wl[t1][q1]);
#pragma unroll
        for (int f2 = -1; f2 <= 1; f2++) {
            const int q2 = T.st[t2][C2] + f2;
            if (q2 <= 0 || q2 >= F.n2 || (f2 != 0 && !node_between(T, t2, q2))) continue;
            const double w2 = f2 == 0 ? 1.0 : (f2 < 0 ? 1.0 - T.wl[t2][q2] : T.wl[t2][q2]);
            const int64_t fe = fa0 + (int64_t)q1 * F.s1 + (int64_t)q2 * F.s2;
#pragma unroll
            for (int fa = 0; fa < 2; fa++) {
                if (fa >= na) break;
                const cx<R> v = rf[fe + fa * F.sa];
                sre += w1 * w2 * (double)v.re;
                sim += w1 * w2 * (double)v.im;
            }
        }
    }
    rc[ec] = cx<R>((R)sre, (R)sim);
}

// psi_f += P_n psi_c on interior fine nodes: trilinear nodal interpolation with
// the same node weights as the edge transfers, so that P_e G_c = G_f P_n;
// terms in the order of the corner bits (x, y, z), zero weights and boundary
// coarse nodes skipped
template <typename R>
__global__ __launch_bounds__(TPB) void k_prolong_node_add(Dims df, Dims dc, MGTables T, const cx<R> *__restrict__ pc,
                                                          cx<R> *__restrict__ pf)
{
    const Tile TL = tile_id();
    const int k = TL.y;
    const int p = (k <= df.nz) ? TL.x * TPB + threadIdx.x : kNoTile;
    const int nI = df.nx + 1, nJ = df.ny + 1;
    if (p >= nI * nJ) return;
    const int j = p / nI, i = p - j * nI;
    if (node_boundary(df, i, j, k)) return;
    const int q[3] = {i, j, k};
    int lo[3];
    bool h[3];
    double w[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = T.lo[a][q[a]];
        h[a] = node_between(T, a, q[a]);
        w[a] = h[a] ? T.wl[a][q[a]] : 1.0;
    }
    const int cI = dc.nx + 1, cJ = dc.ny + 1;
    const int64_t c0 = ((int64_t)lo[2] * cJ + lo[1]) * cI + lo[0];
    double sre = 0, sim = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bx = b & 1, by = b & 2, bz = b & 4;
        if ((bx && !h[0]) || (by && !h[1]) || (bz && !h[2])) continue;  // weight 1 - 1 = 0
        const double wq = (bx ? 1.0 - w[0] : w[0]) * (by ? 1.0 - w[1] : w[1]) * (bz ? 1.0 - w[2] : w[2]);
        const int ci = lo[0] + bx, cj = lo[1] + by, ck = lo[2] + bz;
        if (wq == 0 || node_boundary(dc, ci, cj, ck)) continue;
        const cx<R> v = pc[c0 + (bx ? 1 : 0) + (by ? (int64_t)cI : 0) + (bz ? (int64_t)cI * cJ : 0)];
        sre += wq * (double)v.re;
        sim += wq * (double)v.im;
    }
    const int64_t n = (int64_t)k * nI * nJ + p;
    pf[n] = pf[n] + cx<R>((R)sre, (R)sim);
}

// ------------------------------------------------------------------ coarsening
// sigma of a coarsening: each coarse cell is the volume-weighted mean of its
// fine cells (mesh_coarsen, src/fdfe_main.f90:18-84, with the fine index taken
// inside the block, for 2^k coarsenings; the multigrid grids also merge one or
// two cells per axis).  One coarse cell per thread; the products and sums come
// in the order of the host restatement (z outer, x inner, a*b*c, then the
// division by the coarse volume) and are not contracted into fused
// multiply-adds, so both give the same bits.
__global__ __launch_bounds__(TPB) void k_mesh_coarsen(Dims df, Dims dc, const int *__restrict__ sx,
                                                      const int *__restrict__ sy, const int *__restrict__ sz,
                                                      const double *__restrict__ a, const double *__restrict__ b,
                                                      const double *__restrict__ c, const double *__restrict__ ac,
                                                      const double *__restrict__ bc, const double *__restrict__ cc,
                                                      const double *__restrict__ sig, double *__restrict__ sigc)
{
#pragma clang fp contract(off)
    const int64_t ic = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (ic >= dc.NC) return;
    const int I = (int)(ic % dc.nx), J = (int)((ic / dc.nx) % dc.ny), K = (int)(ic / ((int64_t)dc.nx * dc.ny));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int fk = sz[K]; fk < sz[K + 1]; fk++)
        for (int fj = sy[J]; fj < sy[J + 1]; fj++)
            for (int fi = sx[I]; fi < sx[I + 1]; fi++) {
                const int64_t ie = ((int64_t)fk * df.ny + fj) * df.nx + fi;
                const double v = a[fi] * b[fj] * c[fk];
#pragma unroll
                for (int q = 0; q < 6; q++) acc[q] += sig[q * df.NC + ie] * v;
            }
    const double V = ac[I] * bc[J] * cc[K];
#pragma unroll
    for (int q = 0; q < 6; q++) sigc[q * dc.NC + ic] = acc[q] / V;
}

// Device copies of a level's sigma: the (xx,yy), (zz,xy), (xz,yz) pair layout
// of the complex128 cell-plane march, and FP32 with the diagonal raised to
// `floor` (0: unchanged).  One cell (pairs) or one value (FP32) per thread.
__global__ __launch_bounds__(TPB) void k_sigma_pairs(int64_t nc, const double *__restrict__ sig,
                                                     double *__restrict__ sigp)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= nc) return;
#pragma unroll
    for (int u = 0; u < 3; u++) {
        sigp[(u * nc + q) * 2] = sig[2 * u * nc + q];
        sigp[(u * nc + q) * 2 + 1] = sig[(2 * u + 1) * nc + q];
    }
}

__global__ __launch_bounds__(TPB) void k_sigma_f32(int64_t nc, double floor, const double *__restrict__ sig,
                                                   float *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= 6 * nc) return;
    float v = (float)sig[q];
    if (floor > 0 && q < 3 * nc) v = fmaxf(v, (float)floor);
    out[q] = v;
}

void launch_sigma_pairs(int64_t nc, const double *sig, double *sigp, hipStream_t st)
{
    hipLaunchKernelGGL(k_sigma_pairs, dim3((unsigned)((nc + TPB - 1) / TPB)), dim3(TPB), 0, st, nc, sig, sigp);
}

void launch_sigma_f32(int64_t nc, double floor, const double *sig, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sigma_f32, dim3((unsigned)((6 * nc + TPB - 1) / TPB)), dim3(TPB), 0, st, nc, floor, sig, out);
}

void launch_mesh_coarsen(const Dims &df, const Dims &dc, const int *sx, const int *sy, const int *sz, const double *a,
                         const double *b, const double *c, const double *ac, const double *bc, const double *cc,
                         const double *sig, double *sigc, hipStream_t st)
{
    const unsigned nb = (unsigned)((dc.NC + TPB - 1) / TPB);
    hipLaunchKernelGGL(k_mesh_coarsen, dim3(nb), dim3(TPB), 0, st, df, dc, sx, sy, sz, a, b, c, ac, bc, cc, sig,
                       sigc);
}

// ------------------------------------------------------------------ K4 prolongation
// Fine edge value = sum over the block's coarse edges of separable Lagrange
// weights: along-edge table (m weights) x two transverse node tables (m+1).
struct ProlongTables {
    // per dimension: node tables [n+1] (cbase, 5 weights) and edge tables [n] (cbase, 4 weights)
    const int *nb[3];
    const double *nw[3];
    const int *eb[3];
    const double *ew[3];
};

template <typename R>
__global__ __launch_bounds__(TPB) void k_prolong(Dims df, Dims dc, int m, ProlongTables T,
                                                 const cx<R> *__restrict__ uc, cx<R> *__restrict__ uf)
{
    const Tile TL = tile_id();
    const Slice S = edge_slice(df, TL.y);
    const int p = TL.x * TPB + threadIdx.x;
    if (p >= S.nI * S.nJ) return;
    const int j = p / S.nI, i = p - j * S.nI, k = S.k;
    const int64_t e = S.base + p;
    const int idx[3] = {i, j, k};
    // along dim = o; transverse dims t1 < t2
    const int o = S.o;
    const int t1 = (o == 0) ? 1 : 0, t2 = (o == 2) ? 1 : 2;
    const int bl = T.eb[o][idx[o]], b1 = T.nb[t1][idx[t1]], b2 = T.nb[t2][idx[t2]];
    const double *wl = T.ew[o] + (int64_t)idx[o] * 4;
    const double *w1 = T.nw[t1] + (int64_t)idx[t1] * 5;
    const double *w2 = T.nw[t2] + (int64_t)idx[t2] * 5;
    double sre = 0, sim = 0;
    for (int kk = 0; kk < m; kk++)
        for (int jj = 0; jj <= m; jj++) {
            const double wkj = wl[kk] * w2[jj];
            for (int ii = 0; ii <= m; ii++) {
                int ci[3];
                ci[o] = bl + kk;
                ci[t1] = b1 + ii;
                ci[t2] = b2 + jj;
                const int64_t ce = (o == 0) ? ex_id(dc, ci[0], ci[1], ci[2])
                                 : (o == 1) ? ey_id(dc, ci[0], ci[1], ci[2])
                                            : ez_id(dc, ci[0], ci[1], ci[2]);
                const double wgt = wkj * w1[ii];
                const cx<R> v = uc[ce];
                sre += wgt * (double)v.re;
                sim += wgt * (double)v.im;
            }
        }
    uf[e] = cx<R>((R)sre, (R)sim);
}

// w = c1 * u1 - c2 * w   (Richardson extrapolation on level i-1, src/excmg.f90:60-63)
template <typename R>
__global__ __launch_bounds__(TPB) void k_extrap(int64_t n, R c1, const cx<R> *__restrict__ u1, R c2,
                                                cx<R> *__restrict__ w)
{
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q < n) w[q] = c1 * u1[q] - c2 * w[q];
}

// ==================================================================== launchers
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + TPB - 1) / TPB), 1, 1); }

void reduce_final(RedBuf &rb, int64_t nblk, int K, int dst, int op, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, st, rb.partials, nblk, K, rb.slots, dst, op);
}

// Chunk of node planes per block of the k-marching stencil.  A block re-loads
// one plane at the start of its chunk, so long chunks are cheaper, but the grid
// should fill the device in whole rounds: with `slots` blocks resident at once,
// the time goes as rounds * (kc + 1).  kc is taken in [1, 16] minimising that
// (the longer chunk on ties), so grids too small to fill the device get short
// chunks and more blocks; slots <= 0 falls back to ~1000 blocks.
int spmv_chunk(const Dims &d, int ty, int64_t slots)
{
    const int64_t ntile = (int64_t)((d.nx + 31) / 32) * ((d.ny + ty - 1) / ty);
    const int np = d.nz + 1;
    if (slots <= 0) {
        int kc = 16;
        while (kc > 4 && ntile * ((np + kc - 1) / kc) < 1024) kc /= 2;
        return kc;
    }
    int best = 16;
    int64_t best_cost = INT64_MAX;
    for (int kc = 16; kc >= 1; kc--) {
        const int64_t nblk = ntile * ((np + kc - 1) / kc);
        const int64_t cost = ((nblk + slots - 1) / slots) * (kc + 1);
        if (cost < best_cost) { best_cost = cost; best = kc; }
    }
    return best;
}

// The cell-plane march (complex128): its buffer loads take 32-bit byte
// offsets, so the edge vector and the sigma pairs must stay below 2 GiB
// (about 130 M edges); the modes are the plain product, its dots and the
// residual form (0-4).
static bool cell_march_fits(const DevGrid<double> &g, int dot_mode)
{
    return g.sigp && dot_mode >= 0 && dot_mode <= 4 && g.d.NL * 16 < ((int64_t)1 << 31) &&
           g.d.NC * 48 < ((int64_t)1 << 31);
}

static int64_t cell_go(const DevGrid<double> &g, double wmu, const zd *x, zd *y, double alpha, int dot_mode,
                       const zd *w, RedBuf &rb, hipStream_t st, int kc)
{
    const Dims &d = g.d;
    if (kc <= 0) {
        static const int64_t slots = []() -> int64_t {
            int per_cu = 0, dev = 0, ncu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)&k_spmv_cell<0>, cm::NT, 0) !=
                    hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 0;
            return (int64_t)per_cu * ncu;
        }();
        kc = spmv_chunk(d, cm::TY, slots);
    }
    const int ntx = (d.nx + cm::TX - 1) / cm::TX, nty = (d.ny + cm::TY - 1) / cm::TY;
    const unsigned gx = (unsigned)(ntx * nty), nch = (unsigned)((d.nz + kc) / kc);
    const dim3 grid(gx, pad_y(gx, nch), 1), block(cm::NT, 1, 1);
    const int64_t nblk = (int64_t)grid.x * grid.y;
    if (dot_mode >= 1 && dot_mode <= 3 && nblk > rb.max_blocks)
        throw std::runtime_error("reduction buffer smaller than the stencil grid");
#define SFDX_CELL(M)                                                                                         \
    hipLaunchKernelGGL((k_spmv_cell<M>), grid, block, 0, st, d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sigp, wmu, x, \
                       y, alpha, w, rb.partials, ntx, kc)
    switch (dot_mode) {
    case 0: SFDX_CELL(0); break;
    case 1: SFDX_CELL(1); break;
    case 2: SFDX_CELL(2); break;
    case 3: SFDX_CELL(3); break;
    default: SFDX_CELL(4); break;
    }
#undef SFDX_CELL
    return nblk;
}

template <typename R, int TY, bool PF>
static int64_t march_go(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode,
                        const cx<R> *w, RedBuf &rb, hipStream_t st, int kc, const cx<R> *psi = nullptr)
{
    using G = mt::Geo<TY>;
    const Dims &d = g.d;
    if (kc <= 0) {
        // blocks of this instance resident on the whole device, queried once
        // (a magic static: safe when several host threads launch, as the batch does)
        static const int64_t slots = []() -> int64_t {
            int per_cu = 0, dev = 0, ncu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)&k_spmv_march<R, 0, TY, PF>, G::NT, 0) !=
                    hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 0;
            return (int64_t)per_cu * ncu;
        }();
        kc = spmv_chunk(d, TY, slots);
    }
    const int ntx = (d.nx + G::TX - 1) / G::TX, nty = (d.ny + TY - 1) / TY;
    const unsigned gx = (unsigned)(ntx * nty), nch = (unsigned)((d.nz + kc) / kc);
    const dim3 grid(gx, pad_y(gx, nch), 1), block(G::NT, 1, 1);
    const int64_t nblk = (int64_t)grid.x * grid.y;
    if (((dot_mode >= 1 && dot_mode <= 3) || dot_mode == 6) && nblk > rb.max_blocks)
        throw std::runtime_error("reduction buffer smaller than the stencil grid");
#define SFDX_MARCH(M)                                                                                      \
    hipLaunchKernelGGL((k_spmv_march<R, M, TY, PF>), grid, block, 0, st, d, g.a, g.b, g.c, g.ra, g.rb, g.rc, \
                       g.sig, wmu, x, y, alpha, w, rb.partials, ntx, kc, psi)
    switch (dot_mode) {
    case 0: SFDX_MARCH(0); break;
    case 1: SFDX_MARCH(1); break;
    case 2: SFDX_MARCH(2); break;
    case 3: SFDX_MARCH(3); break;
    case 4: SFDX_MARCH(4); break;
    case 5: SFDX_MARCH(5); break;
    case 6:
        if constexpr (sizeof(R) == 4) {
            SFDX_MARCH(6);
            break;
        }
        throw std::runtime_error("launch_spmv_march: mode 6 is complex64 only");
    default: throw std::runtime_error("launch_spmv_march: bad mode");
    }
#undef SFDX_MARCH
    return nblk;
}

// variant: 0 = tile 32x8 (256 threads) with prefetch, 1 = 32x8 without,
// 2 = 32x4 (128 threads) with prefetch, 3 = 32x4 without
template <typename R>
int64_t launch_spmv_march(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode,
                          const cx<R> *w, RedBuf &rb, hipStream_t st, int kc, int variant)
{
    switch (variant) {
    case 0:
        if constexpr (sizeof(R) == 8)
            if (cell_march_fits(g, dot_mode)) return cell_go(g, wmu, x, y, alpha, dot_mode, w, rb, st, kc);
        return march_go<R, 8, true>(g, wmu, x, y, alpha, dot_mode, w, rb, st, kc);
    case 1: return march_go<R, 8, false>(g, wmu, x, y, alpha, dot_mode, w, rb, st, kc);
    case 2: return march_go<R, 4, true>(g, wmu, x, y, alpha, dot_mode, w, rb, st, kc);
    case 3: return march_go<R, 4, false>(g, wmu, x, y, alpha, dot_mode, w, rb, st, kc);
    default: return march_go<R, 8, true>(g, wmu, x, y, alpha, dot_mode, w, rb, st, kc);
    }
}

// One row per thread (13-neighbour gather); kept for comparison with the
// k-marching form.
template <typename R>
int64_t launch_spmv_gather(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode,
                           const cx<R> *w, RedBuf &rb, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
    switch (dot_mode) {
    case 0: hipLaunchKernelGGL((k_spmv<R, 0>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sig, wmu, x, y, alpha, w, rb.partials); break;
    case 1: hipLaunchKernelGGL((k_spmv<R, 1>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sig, wmu, x, y, alpha, w, rb.partials); break;
    case 2: hipLaunchKernelGGL((k_spmv<R, 2>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sig, wmu, x, y, alpha, w, rb.partials); break;
    case 3: hipLaunchKernelGGL((k_spmv<R, 3>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sig, wmu, x, y, alpha, w, rb.partials); break;
    default: hipLaunchKernelGGL((k_spmv<R, 4>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sig, wmu, x, y, alpha, w, rb.partials); break;
    }
    return L.nblocks();
}

template <typename R>
int64_t launch_spmv(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, R alpha, int dot_mode, const cx<R> *w,
                    RedBuf &rb, hipStream_t st)
{
    // the level's kernel: 0 unless set by tune_spmv or the caller's option; among
    // the k-marching tiles the sweep of sfdx_bench_kernels on 128x128x96 found
    // 32x8 columns with prefetch fastest (0.058 ms in complex128, 32x4: 0.059;
    // 0.035 ms in complex64, 32x4: 0.037)
    return launch_spmv_march<R>(g, wmu, x, y, alpha, dot_mode, w, rb, st, 0, g.spmv_variant);
}

// Kernel autotuner: time the five SpMV variants on this grid (as the Krylov
// solvers launch them, with the fused dot, chunk from the occupancy rule) and
// return the fastest.  One untimed launch per variant, then the
// median of `reps` timed ones (HIP events on st).  Host-synchronising: call
// it at set-up, never inside a stream capture.  x must hold NL finite values,
// y is overwritten.
template <typename R>
int tune_spmv(const DevGrid<R> &g, R wmu, const cx<R> *x, cx<R> *y, RedBuf &rb, hipStream_t st, int reps,
              float *ms_out)
{
    hipEvent_t e0, e1;
    SFDX_CHECK(hipEventCreate(&e0));
    SFDX_CHECK(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 0;
    std::vector<float> t(reps > 0 ? reps : 1);
    for (int v = 0; v < 5; v++) {
        launch_spmv_march<R>(g, wmu, x, y, (R)1, 1, nullptr, rb, st, 0, v);
        for (auto &ms : t) {
            SFDX_CHECK(hipEventRecord(e0, st));
            launch_spmv_march<R>(g, wmu, x, y, (R)1, 1, nullptr, rb, st, 0, v);
            SFDX_CHECK(hipEventRecord(e1, st));
            SFDX_CHECK(hipEventSynchronize(e1));
            SFDX_CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (ms_out) ms_out[v] = med;
        if (v == 0 || med < best_ms) { best = v; best_ms = med; }
    }
    SFDX_CHECK(hipEventDestroy(e0));
    SFDX_CHECK(hipEventDestroy(e1));
    return best;
}

// t = r - A (z + (i/wmu) G psi) on interior edges, 0 on boundary edges: the
// k-marching stencil with the node potential in a third LDS ring, forming
// r - K z + i wmu M (z + (i/wmu) G psi) so that the curl part never acts on
// G psi (K G = 0; in resistive cells psi is orders of magnitude larger than z
// and K G psi would be rounding noise of that size)
template <typename R>
void launch_resid_split(const DevGrid<R> &g, R wmu, const cx<R> *z, const cx<R> *psi, const cx<R> *r, cx<R> *t,
                        hipStream_t st)
{
    RedBuf none;
    switch (g.spmv_variant) {
    case 1: march_go<R, 8, false>(g, wmu, z, t, (R)-1, 5, r, none, st, 0, psi); break;
    case 2: march_go<R, 4, true>(g, wmu, z, t, (R)-1, 5, r, none, st, 0, psi); break;
    case 3: march_go<R, 4, false>(g, wmu, z, t, (R)-1, 5, r, none, st, 0, psi); break;
    default: march_go<R, 8, true>(g, wmu, z, t, (R)-1, 5, r, none, st, 0, psi); break;
    }
}

// y = A_s (z + (i/wmu) G psi) on interior edges (0 on boundary edges) with the
// partial sums of u^T y, u = z + (i/wmu) G psi (unconjugated): the operator on
// the split directions of the complex64 COCR, the curl part acting on z alone
int64_t launch_split_product(const DevGrid<float> &g, float wmu, const zf *z, const zf *psi, zf *y, RedBuf &rb,
                             hipStream_t st)
{
    return march_go<float, 8, true>(g, wmu, z, y, 1.0f, 6, nullptr, rb, st, 0, psi);
}

template <typename R>
void launch_mass(const DevGrid<R> &g, const cx<R> *x, cx<R> *w, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
    hipLaunchKernelGGL((k_mass<R>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc, g.sig, x, w);
}

template <typename R, int O, int PAR, int DIR>
static void ssor_one(const DevGrid<R> &g, R wmu, const cx<R> *in, cx<R> *out, R omega, hipStream_t st)
{
    SliceLaunch L = colour_launch(g.d, 2 * O + PAR);
    if (g.ssor_xmass)
        hipLaunchKernelGGL((k_ssor<R, O, PAR, DIR, true>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb, g.rc,
                           g.sig, wmu, in, out, omega);
    else
        hipLaunchKernelGGL((k_ssor<R, O, PAR, DIR, false>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, g.ra, g.rb,
                           g.rc, g.sig, wmu, in, out, omega);
}

template <typename R>
void launch_ssor(const DevGrid<R> &g, R wmu, const cx<R> *in, cx<R> *out, R omega, hipStream_t st)
{
    ssor_one<R, 0, 0, 0>(g, wmu, in, out, omega, st);
    ssor_one<R, 0, 1, 0>(g, wmu, in, out, omega, st);
    ssor_one<R, 1, 0, 0>(g, wmu, in, out, omega, st);
    ssor_one<R, 1, 1, 0>(g, wmu, in, out, omega, st);
    ssor_one<R, 2, 0, 0>(g, wmu, in, out, omega, st);
    ssor_one<R, 2, 1, 0>(g, wmu, in, out, omega, st);
    // the backward step of the last colour has no upper neighbours: it only
    // scales by (2 - omega), which is the identity for omega = 1
    if (omega != (R)1) ssor_one<R, 2, 1, 1>(g, wmu, in, out, omega, st);
    ssor_one<R, 2, 0, 1>(g, wmu, in, out, omega, st);
    ssor_one<R, 1, 1, 1>(g, wmu, in, out, omega, st);
    ssor_one<R, 1, 0, 1>(g, wmu, in, out, omega, st);
    ssor_one<R, 0, 1, 1>(g, wmu, in, out, omega, st);
    ssor_one<R, 0, 0, 1>(g, wmu, in, out, omega, st);
}

template <typename R>
void launch_cocr_ap(int64_t n, RedBuf &rb, cx<R> *Ap, const cx<R> *Az, zf *ap32, bool copy, hipStream_t st)
{
    hipLaunchKernelGGL((k_cocr_ap<R>), dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, rb.slots, Ap, Az,
                       ap32, copy ? 1 : 0);
}

template <typename R>
void launch_vec(int mode, const DevGrid<R> &g, RedBuf &rb, cx<R> *v0, cx<R> *v1, cx<R> *v2, cx<R> *v3, cx<R> *v4,
                cx<R> *v5, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
#define SFDX_VEC(M) hipLaunchKernelGGL((k_vec<R, M>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, rb.slots, v0, v1, v2, v3, v4, v5, rb.partials)
    switch (mode) {
    case 0: SFDX_VEC(0); break;
    case 1: SFDX_VEC(1); break;
    case 2: SFDX_VEC(2); break;
    case 4: SFDX_VEC(4); break;
    case 5: SFDX_VEC(5); break;
    case 6: SFDX_VEC(6); break;
    default: throw std::runtime_error("launch_vec: bad mode");
    }
#undef SFDX_VEC
}

template <typename R>
void launch_dot(int mode, int64_t n, const cx<R> *u, const cx<R> *v, RedBuf &rb, hipStream_t st, int64_t &nblk)
{
    dim3 gr = grid1(n);
    nblk = gr.x;
    if (mode == 0) hipLaunchKernelGGL((k_dot<R, 0>), gr, dim3(TPB), 0, st, n, u, v, rb.partials);
    else hipLaunchKernelGGL((k_dot<R, 1>), gr, dim3(TPB), 0, st, n, u, v, rb.partials);
}

template <typename R>
void launch_copy(int64_t n, const cx<R> *s, cx<R> *t, hipStream_t st)
{
    hipLaunchKernelGGL((k_copy<R>), grid1(n), dim3(TPB), 0, st, n, s, t);
}

void launch_stream_copy(size_t bytes, const void *s, void *t, int units, hipStream_t st)
{
    if (bytes % 16) throw std::runtime_error("launch_stream_copy: size must be a multiple of 16 bytes");
    const int64_t n16 = (int64_t)(bytes / 16);
    if (units == 8) {
        const dim3 g((unsigned)((n16 + TPB * 8 - 1) / (TPB * 8)));
        hipLaunchKernelGGL((k_stream_copy<8>), g, dim3(TPB), 0, st, n16, (const u32x4 *)s, (u32x4 *)t);
    } else {
        const dim3 g((unsigned)((n16 + TPB * 4 - 1) / (TPB * 4)));
        hipLaunchKernelGGL((k_stream_copy<4>), g, dim3(TPB), 0, st, n16, (const u32x4 *)s, (u32x4 *)t);
    }
}

template <typename RS, typename RD>
void launch_convert(int64_t n, const cx<RS> *s, cx<RD> *t, hipStream_t st)
{
    hipLaunchKernelGGL((k_convert<RS, RD>), grid1(n), dim3(TPB), 0, st, n, s, t);
}

void launch_add_lo(int64_t n, zd *x, const zf *d, hipStream_t st)
{
    hipLaunchKernelGGL(k_add_lo, grid1(n), dim3(TPB), 0, st, n, x, d);
}

template <typename R>
void launch_axpby(int64_t n, cx<R> al, const cx<R> *u, cx<R> be, cx<R> *v, hipStream_t st)
{
    hipLaunchKernelGGL((k_axpby<R>), grid1(n), dim3(TPB), 0, st, n, al, u, be, v);
}

template <typename R>
void launch_dirichlet(const Dims &d, const cx<R> *edc, int pol, cx<R> *x, int zero_inner, hipStream_t st, int full)
{
    SliceLaunch L = edge_launch(d);
    hipLaunchKernelGGL((k_dirichlet<R>), L.grid, L.block, 0, st, d, edc, pol, x, zero_inner, full);
}

template <typename R>
void launch_node_div(const DevGrid<R> &g, const cx<R> *w, cx<R> *t, RedBuf &rb, hipStream_t st, int64_t &nblk)
{
    SliceLaunch L = node_launch(g.d);
    nblk = L.nblocks();
    hipLaunchKernelGGL((k_node_div<R>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, w, t, rb.partials);
}

// P holds 29 * NP values: the 19 slots of k_pbuild, then the symmetric form
// (1/diag, 9 upper couplings) of k_psym
template <typename R>
void launch_pbuild(const Dims &d, const double *a, const double *b, const double *c, const double *sig, R *P,
                   hipStream_t st)
{
    SliceLaunch L = node_launch(d);
    hipLaunchKernelGGL((k_pbuild<R>), L.grid, L.block, 0, st, d, a, b, c, sig, P);
    hipLaunchKernelGGL((k_psym<R>), dim3((unsigned)((d.NP + TPB - 1) / TPB)), dim3(TPB), 0, st, d.NP, P);
}

template <typename R>
void launch_papply(const Dims &d, const R *P, const cx<R> *x, cx<R> *y, int dot, RedBuf &rb, hipStream_t st,
                   int64_t &nblk)
{
    SliceLaunch L = node_launch(d);
    nblk = L.nblocks();
    if (dot) hipLaunchKernelGGL((k_papply<R, 1>), L.grid, L.block, 0, st, d, P, x, y, rb.partials);
    else hipLaunchKernelGGL((k_papply<R, 0>), L.grid, L.block, 0, st, d, P, x, y, rb.partials);
}

template <typename R>
void launch_pcg(int mode, int64_t np, const R *Pdiag, RedBuf &rb, cx<R> *u, cx<R> *r, cx<R> *p, const cx<R> *v,
                hipStream_t st, int64_t &nblk)
{
    dim3 gr = grid1(np);
    nblk = gr.x;
    if (mode == 0) hipLaunchKernelGGL((k_pcg<R, 0>), gr, dim3(TPB), 0, st, np, Pdiag, rb.slots, u, r, p, v, rb.partials);
    else if (mode == 1) hipLaunchKernelGGL((k_pcg<R, 1>), gr, dim3(TPB), 0, st, np, Pdiag, rb.slots, u, r, p, v, rb.partials);
    else hipLaunchKernelGGL((k_pcg<R, 2>), gr, dim3(TPB), 0, st, np, Pdiag, rb.slots, u, r, p, v, rb.partials);
}

template <typename R>
void launch_grad_update(const DevGrid<R> &g, const cx<R> *phi, cx<R> *x, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
    hipLaunchKernelGGL((k_grad_update<R>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, phi, x);
}

template <typename R>
void launch_prolong(const Dims &df, const Dims &dc, int m, const ProlongDev &T, const cx<R> *uc, cx<R> *uf,
                    hipStream_t st)
{
    SliceLaunch L = edge_launch(df);
    ProlongTables t;
    for (int q = 0; q < 3; q++) { t.nb[q] = T.nb[q]; t.nw[q] = T.nw[q]; t.eb[q] = T.eb[q]; t.ew[q] = T.ew[q]; }
    hipLaunchKernelGGL((k_prolong<R>), L.grid, L.block, 0, st, df, dc, m, t, uc, uf);
}

// out[i] = 1 / w[i] (reciprocal cell widths of one axis, FP64, correctly rounded
// as the host's 1.0 / w)
__global__ __launch_bounds__(TPB) void k_reciprocal(int n, const double *__restrict__ w, double *__restrict__ out)
{
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = 1.0 / w[i];
}

void launch_reciprocal(int n, const double *w, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reciprocal, grid1(n), dim3(TPB), 0, st, n, w, out);
}

template <typename R>
void launch_extrap(int64_t n, R c1, const cx<R> *u1, R c2, cx<R> *w, hipStream_t st)
{
    hipLaunchKernelGGL((k_extrap<R>), grid1(n), dim3(TPB), 0, st, n, c1, u1, c2, w);
}

template <typename R, int COL, int DIR>
static void pssor_one(const Dims &d, const R *P, const cx<R> *t, cx<R> *phi, hipStream_t st)
{
    const int W = (d.nx + 2) / 2, H = (d.ny + 2) / 2;
    const unsigned gx = (unsigned)(((int64_t)W * H + TPB - 1) / TPB);
    dim3 gr(gx, pad_y(gx, (unsigned)((d.nz + 2) / 2)), 1);
    hipLaunchKernelGGL((k_pssor<R, COL, DIR>), gr, dim3(TPB), 0, st, d, P, t, phi);
}

// The symmetric slots (1/diag and the 9 upper couplings) rounded to FP32, for a
// complex128 nodal smoother that reads the operator in half the bytes (the
// operator is data: rounding it keeps the smoother a fixed linear operator)
__global__ __launch_bounds__(TPB) void k_psym32(int64_t np, const double *__restrict__ P, float *__restrict__ Ps32)
{
    const int64_t n = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (n >= np) return;
#pragma unroll
    for (int q = 0; q < 10; q++) Ps32[q * np + n] = (float)P[(19 + q) * np + n];
}

void launch_psym32(int64_t np, const double *P, float *Ps32, hipStream_t st)
{
    hipLaunchKernelGGL(k_psym32, dim3((unsigned)((np + TPB - 1) / TPB)), dim3(TPB), 0, st, np, P, Ps32);
}

// Two damped Jacobi steps on P from phi = 0, as one symmetric polynomial:
//   phi = omega D^-1 ((2 - omega) t - omega sum_{m != n} P_nm D^-1_m t_m)
//       = omega (2 - omega) u_n - omega^2 D^-1_n sum_{m != n} P_nm u_m,  u = D^-1 t,
// i.e. [2 omega D^-1 - omega^2 D^-1 P D^-1] t: one pass over P with every node
// in flight, instead of the 16 dependent colour sweeps of the nodal SSOR.
// Input u = D^-1 t (k_gradT with SCALE), Ps = the symmetric slots 19..28.
// One block per 32x8 tile of nodes of one plane; u of planes k-1, k, k+1 over
// the tile and a one-node halo is staged in LDS, so each node issues 4 vector
// loads of u (instead of 19 gathers) plus its own 10 and its 9 lower
// neighbours' coupling values.  The 9 upper couplings, in the order of the
// symmetric slots (P slots 2, 4, 6, 9, 10, 13, 14, 17, 18): offsets
// (1,0,0) (0,1,0) (0,0,1) (-1,1,0) (1,1,0) (-1,0,1) (1,0,1) (0,-1,1) (0,1,1).
namespace pj {
constexpr int TX = 32, TY = 8, PX = TX + 2, PY = TY + 2, NPL = PX * PY;
constexpr int DX[9] = {1, 0, 0, -1, 1, -1, 1, 0, 0};
constexpr int DY[9] = {0, 1, 0, 1, 1, 0, 0, -1, 1};
constexpr int DZ[9] = {0, 0, 1, 0, 0, 1, 1, 1, 1};
}  // namespace pj

template <typename R, typename RP = R>
__global__ __launch_bounds__(pj::TX * pj::TY) void k_pjacobi(Dims d, const RP *__restrict__ Ps,
                                                           const cx<R> *__restrict__ u, cx<R> *__restrict__ phi,
                                                           R omega, int ntx)
{
    using namespace pj;
    __shared__ R ure[3 * NPL], uim[3 * NPL];  // planes k-1, k, k+1
    const Tile TL = tile_id();
    const int k = TL.y;
    if (k > d.nz) return;  // padding tile (uniform over the block)
    const int tj = TL.x / ntx, ti = TL.x - tj * ntx;
    const int i0 = ti * TX, j0 = tj * TY;
    const int nI = d.nx + 1, nJ = d.ny + 1;
    const int64_t sy = nI, sz = (int64_t)nI * nJ, np = d.NP;
    for (int f = threadIdx.x; f < 3 * NPL; f += TX * TY) {
        const int pl = f / NPL, r = f - pl * NPL, lj = r / PX, li = r - lj * PX;
        const int gi = i0 - 1 + li, gj = j0 - 1 + lj, gk = k - 1 + pl;
        cx<R> v(0);
        if (gi >= 0 && gi < nI && gj >= 0 && gj < nJ && gk >= 0 && gk <= d.nz) v = u[gk * sz + (int64_t)gj * sy + gi];
        ure[f] = v.re;
        uim[f] = v.im;
    }
    __syncthreads();
    const int li = threadIdx.x % TX, lj = threadIdx.x / TX;
    const int i = i0 + li, j = j0 + lj;
    if (i >= nI || j >= nJ) return;
    const int64_t n = k * sz + (int64_t)j * sy + i;
    if (node_boundary(d, i, j, k)) {
        phi[n] = cx<R>(0);
        return;
    }
    const int c = NPL + (lj + 1) * PX + li + 1;  // own node in the k plane of LDS
    cx<R> s(0);
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const int64_t o = DX[q] + DY[q] * sy + DZ[q] * sz;
        const int lo = DX[q] + DY[q] * PX + DZ[q] * NPL;
        const RP *Pu = Ps + (1 + q) * np;
        const cx<R> up(ure[c + lo], uim[c + lo]), dn(ure[c - lo], uim[c - lo]);
        s = s + (R)Pu[n] * up + (R)Pu[n - o] * dn;  // u = 0 on boundary nodes
    }
    const cx<R> un(ure[c], uim[c]);
    phi[n] = (omega * ((R)2 - omega)) * un - (omega * omega * (R)Ps[n]) * s;
}

template <typename R, typename RP>
static void launch_pjacobi(const Dims &d, const RP *Ps, const cx<R> *u, cx<R> *phi, R omega, hipStream_t st)
{
    const int ntx = (d.nx + 1 + pj::TX - 1) / pj::TX, nty = (d.ny + 1 + pj::TY - 1) / pj::TY;
    const unsigned gx = (unsigned)(ntx * nty);
    hipLaunchKernelGGL((k_pjacobi<R, RP>), dim3(gx, pad_y(gx, (unsigned)(d.nz + 1))), dim3(pj::TX * pj::TY), 0, st, d, Ps,
                       u, phi, omega, ntx);
}

// phi = N_P (G^T in): nodal 1 the Jacobi polynomial (on D^-1 G^T in), else the
// 8-colour nodal SSOR; tnode is the nodal work vector
template <typename R>
static void nodal_part(const DevGrid<R> &g, const R *P, const cx<R> *in, cx<R> *tnode, cx<R> *phi, int nodal,
                       hipStream_t st)
{
    const Dims &d = g.d;
    SliceLaunch Ln = node_launch(d);
    if (nodal == 1) {
        const R *Ps = P + 19 * d.NP;
        hipLaunchKernelGGL((k_gradT<R, true>), Ln.grid, Ln.block, 0, st, d, g.a, g.b, g.c, in, tnode, Ps);
        launch_pjacobi<R, R>(d, Ps, tnode, phi, (R)0.8, st);
        return;
    }
    hipLaunchKernelGGL((k_gradT<R, false>), Ln.grid, Ln.block, 0, st, d, g.a, g.b, g.c, in, tnode, (const R *)nullptr);
    const cx<R> *t = tnode;
    pssor_one<R, 0, 0>(d, P, t, phi, st);
    pssor_one<R, 1, 0>(d, P, t, phi, st);
    pssor_one<R, 2, 0>(d, P, t, phi, st);
    pssor_one<R, 3, 0>(d, P, t, phi, st);
    pssor_one<R, 4, 0>(d, P, t, phi, st);
    pssor_one<R, 5, 0>(d, P, t, phi, st);
    pssor_one<R, 6, 0>(d, P, t, phi, st);
    pssor_one<R, 7, 0>(d, P, t, phi, st);
    pssor_one<R, 7, 1>(d, P, t, phi, st);
    pssor_one<R, 6, 1>(d, P, t, phi, st);
    pssor_one<R, 5, 1>(d, P, t, phi, st);
    pssor_one<R, 4, 1>(d, P, t, phi, st);
    pssor_one<R, 3, 1>(d, P, t, phi, st);
    pssor_one<R, 2, 1>(d, P, t, phi, st);
    pssor_one<R, 1, 1>(d, P, t, phi, st);
    pssor_one<R, 0, 1>(d, P, t, phi, st);
}

// out = SSOR_e(in) + s * G N_P(G^T in), s = -1/kappa, N_P the nodal smoother
// (nodal 0: 8-colour SSOR on P; 1: two-step Jacobi polynomial)
template <typename R>
void launch_hybrid(const DevGrid<R> &g, R wmu, const R *P, const cx<R> *in, cx<R> *out, cx<R> *tnode,
                   cx<R> *phi, R omega, hipStream_t st, int nodal, const cx<R> *dotw, double *partials, bool store)
{
    launch_ssor<R>(g, wmu, in, out, omega, st);
    nodal_part<R>(g, P, in, tnode, phi, nodal, st);
    // s = -1/kappa = -1/(i wmu) = i/wmu
    const cx<R> s((R)0, (R)1 / wmu);
    SliceLaunch Le = edge_launch(g.d);
    if (!store) {
        // only the dot: out keeps the edge SSOR output, the caller adds s G phi itself
        if (!dotw) throw std::runtime_error("launch_hybrid: nothing to do without store or dot");
        hipLaunchKernelGGL((k_grad_add<R, true, false>), Le.grid, Le.block, 0, st, g.d, g.a, g.b, g.c, s, phi, out, dotw,
                           partials);
    } else if (dotw)
        hipLaunchKernelGGL((k_grad_add<R, true>), Le.grid, Le.block, 0, st, g.d, g.a, g.b, g.c, s, phi, out, dotw,
                           partials);
    else
        hipLaunchKernelGGL((k_grad_add<R, false>), Le.grid, Le.block, 0, st, g.d, g.a, g.b, g.c, s, phi, out,
                           (const cx<R> *)nullptr, (double *)nullptr);
}

// Hybrid preconditioner with the edge SSOR in complex64 and the nodal part
// (one Jacobi pass over P, where the air's large potentials live) in
// complex128: out = SSOR32(in32) + s G N_P(G^T in), in32 = complex64 copy of
// in; P's symmetric slots are read in FP32 (Ps32, k_psym32); the complex64
// SSOR output is widened inside the recombination pass (optional fused
// unconjugated dot with dotw).
void launch_hybrid_edge32(const DevGrid<double> &g64, const DevGrid<float> &g32, double wmu, const float *Ps32,
                          const zd *in, zd *out, zf *in32, zf *z32, zd *tnode, zd *phi, float omega,
                          hipStream_t st, const zd *dotw, double *partials, bool convert_in, bool store)
{
    if (convert_in) launch_convert<double, float>(g64.d.NL, in, in32, st);
    launch_ssor<float>(g32, (float)wmu, in32, z32, omega, st);
    // nodal part in complex128 on the FP32 symmetric operator
    SliceLaunch Ln = node_launch(g64.d);
    hipLaunchKernelGGL((k_gradT<double, true, float>), Ln.grid, Ln.block, 0, st, g64.d, g64.a, g64.b, g64.c, in, tnode,
                       Ps32);
    launch_pjacobi<double, float>(g64.d, Ps32, tnode, phi, 0.8, st);
    const zd s(0.0, 1.0 / wmu);
    SliceLaunch Le = edge_launch(g64.d);
    if (!store) {
        // only the dot: the caller forms the output itself from z32 and phi
        if (!dotw) throw std::runtime_error("launch_hybrid_edge32: nothing to do without store or dot");
        hipLaunchKernelGGL((k_hyb_combine<double, true, false>), Le.grid, Le.block, 0, st, g64.d, g64.a, g64.b, g64.c, s,
                           phi, z32, out, dotw, partials);
    } else if (dotw)
        hipLaunchKernelGGL((k_hyb_combine<double, true>), Le.grid, Le.block, 0, st, g64.d, g64.a, g64.b, g64.c, s, phi,
                           z32, out, dotw, partials);
    else
        hipLaunchKernelGGL((k_hyb_combine<double, false>), Le.grid, Le.block, 0, st, g64.d, g64.a, g64.b, g64.c, s,
                           phi, z32, out, (const zd *)nullptr, (double *)nullptr);
}

// out64 = SSOR32(in) + s * G SSOR_P32(G^T in) with the FP32 pieces on a
// complex64 copy of the input and the recombination in FP64.
void launch_hybrid_mixed(const DevGrid<double> &g64, const DevGrid<float> &g32, double wmu, const float *P32,
                         const zd *in, zd *out, zf *in32, zf *z32, zf *tnode, zf *phi, float omega, hipStream_t st,
                         int nodal, const zd *dotw, double *partials)
{
    const int64_t nl = g64.d.NL;
    launch_convert<double, float>(nl, in, in32, st);
    launch_ssor<float>(g32, (float)wmu, in32, z32, omega, st);
    nodal_part<float>(g32, P32, in32, tnode, phi, nodal, st);
    const zd s(0.0, 1.0 / wmu);
    SliceLaunch Le = edge_launch(g64.d);
    if (dotw)
        hipLaunchKernelGGL((k_hyb_combine<float, true>), Le.grid, Le.block, 0, st, g64.d, g64.a, g64.b, g64.c, s, phi, z32, out,
                           dotw, partials);
    else
        hipLaunchKernelGGL((k_hyb_combine<float, false>), Le.grid, Le.block, 0, st, g64.d, g64.a, g64.b, g64.c, s, phi, z32, out,
                           (const zd *)nullptr, (double *)nullptr);
}

// One Gauss-Seidel smoothing step of the hybrid smoother:
//   out = GS_e(in) + s * G GS_n(G^T in),  forward (dir 0) or backward (dir 1)
// sweeps, so that the backward step is the transpose of the forward one.
template <typename R>
void launch_hybrid_gs(const DevGrid<R> &g, R wmu, const R *P, const cx<R> *in, cx<R> *out, cx<R> *tnode, cx<R> *phi,
                      int dir, hipStream_t st, bool add_grad, int nodal)
{
    if (dir == 0) {
        ssor_one<R, 0, 0, 0>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 0, 1, 0>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 1, 0, 0>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 1, 1, 0>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 2, 0, 0>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 2, 1, 0>(g, wmu, in, out, (R)1, st);
    } else {
        ssor_one<R, 2, 1, 2>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 2, 0, 2>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 1, 1, 2>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 1, 0, 2>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 0, 1, 2>(g, wmu, in, out, (R)1, st);
        ssor_one<R, 0, 0, 2>(g, wmu, in, out, (R)1, st);
    }
    SliceLaunch Ln = node_launch(g.d);
    if (nodal == 1) {
        // the two-step Jacobi polynomial on the symmetric form of P: one pass, the
        // same symmetric operator before and after the coarse correction
        const R *Ps = P + 19 * g.d.NP;
        hipLaunchKernelGGL((k_gradT<R, true>), Ln.grid, Ln.block, 0, st, g.d, g.a, g.b, g.c, in, tnode, Ps);
        launch_pjacobi<R, R>(g.d, Ps, tnode, phi, (R)0.8, st);
    } else {
    hipLaunchKernelGGL((k_gradT<R, false>), Ln.grid, Ln.block, 0, st, g.d, g.a, g.b, g.c, in, tnode,
                       (const R *)nullptr);
    if (dir == 0) {
        pssor_one<R, 0, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 1, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 2, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 3, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 4, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 5, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 6, 0>(g.d, P, tnode, phi, st);
        pssor_one<R, 7, 0>(g.d, P, tnode, phi, st);
    } else {
        pssor_one<R, 7, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 6, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 5, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 4, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 3, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 2, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 1, 2>(g.d, P, tnode, phi, st);
        pssor_one<R, 0, 2>(g.d, P, tnode, phi, st);
    }
    }
    if (!add_grad) return;  // caller keeps the nodal part phi (split form)
    const cx<R> s((R)0, (R)1 / wmu);
    SliceLaunch Le = edge_launch(g.d);
    hipLaunchKernelGGL((k_grad_add<R, false>), Le.grid, Le.block, 0, st, g.d, g.a, g.b, g.c, s, phi, out,
                       (const cx<R> *)nullptr, (double *)nullptr);
}

template <typename R>
void launch_restrict(const Dims &df, const Dims &dc, const MGTables &T, const cx<R> *rf, cx<R> *rc, hipStream_t st)
{
    SliceLaunch L = edge_launch(dc);
    hipLaunchKernelGGL((k_restrict<R>), L.grid, L.block, 0, st, df, dc, T, rf, rc);
}

void launch_cell_grad(const DevGrid<double> &g, double wmu, const zd *e, const zd *lam, double *grad, hipStream_t st)
{
    const int64_t plane = (int64_t)g.d.nx * g.d.ny;
    const unsigned gx = (unsigned)((plane + TPB - 1) / TPB);
    hipLaunchKernelGGL(k_cell_grad, dim3(gx, pad_y(gx, (unsigned)g.d.nz), 1), dim3(TPB), 0, st, g.d, g.a, g.b, g.c,
                       g.sig, wmu, e, lam, grad);
}

void launch_surface_fields(const Dims &d, int nair, const int *surf, const double *a, const double *b,
                           const double *c, double om, const zd *x, zd *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_surface_fields, grid1((int64_t)d.nx * d.ny), dim3(TPB), 0, st, d, nair, surf, a, b, c, om, x,
                       out);
}

void launch_zero_boundary(const Dims &d, zd *v, hipStream_t st)
{
    SliceLaunch L = edge_launch(d);
    hipLaunchKernelGGL(k_zero_boundary, L.grid, L.block, 0, st, d, v);
}

template <typename R>
void launch_gpbicg(int mode, const DevGrid<R> &g, RedBuf &rb, const GPVecs<R> &v, hipStream_t st)
{
    SliceLaunch L = edge_launch(g.d);
#define SFDX_GP(M) hipLaunchKernelGGL((k_gpbicg<R, M>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, rb.slots, v, rb.partials)
    switch (mode) {
    case 0: SFDX_GP(0); break;
    case 1: SFDX_GP(1); break;
    case 2: SFDX_GP(2); break;
    case 3: SFDX_GP(3); break;
    default: SFDX_GP(4); break;
    }
#undef SFDX_GP
}

template <typename R>
void launch_prolong_node_add(const Dims &df, const Dims &dc, const MGTables &T, const cx<R> *pc, cx<R> *pf,
                             hipStream_t st)
{
    SliceLaunch L = node_launch(df);
    hipLaunchKernelGGL((k_prolong_node_add<R>), L.grid, L.block, 0, st, df, dc, T, pc, pf);
}

template <typename R>
void launch_grad_add(const DevGrid<R> &g, cx<R> s, const cx<R> *phi, cx<R> *z, hipStream_t st, const cx<R> *dotw,
                     double *partials, bool store)
{
    SliceLaunch L = edge_launch(g.d);
    if (dotw && !store)
        hipLaunchKernelGGL((k_grad_add<R, true, false>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, s, phi, z, dotw,
                           partials);
    else if (dotw)
        hipLaunchKernelGGL((k_grad_add<R, true>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, s, phi, z, dotw, partials);
    else
        hipLaunchKernelGGL((k_grad_add<R, false>), L.grid, L.block, 0, st, g.d, g.a, g.b, g.c, s, phi, z,
                           (const cx<R> *)nullptr, (double *)nullptr);
}

void launch_hyb_combine(const DevGrid<double> &g64, zd s, const zf *phi, const zf *z32, zd *z64, hipStream_t st,
                        const zd *dotw, double *partials, bool store)
{
    SliceLaunch L = edge_launch(g64.d);
    if (dotw && !store)
        hipLaunchKernelGGL((k_hyb_combine<float, true, false>), L.grid, L.block, 0, st, g64.d, g64.a, g64.b, g64.c, s,
                           phi, z32, z64, dotw, partials);
    else if (dotw)
        hipLaunchKernelGGL((k_hyb_combine<float, true>), L.grid, L.block, 0, st, g64.d, g64.a, g64.b, g64.c, s, phi,
                           z32, z64, dotw, partials);
    else
        hipLaunchKernelGGL((k_hyb_combine<float, false>), L.grid, L.block, 0, st, g64.d, g64.a, g64.b, g64.c, s, phi,
                           z32, z64, (const zd *)nullptr, (double *)nullptr);
}

template <typename R>
void launch_prolong_add(const Dims &df, const Dims &dc, const MGTables &T, const cx<R> *zc, cx<R> *zf,
                        hipStream_t st)
{
    SliceLaunch L = edge_launch(df);
    hipLaunchKernelGGL((k_prolong_add<R>), L.grid, L.block, 0, st, df, dc, T, zc, zf);
}

// explicit instantiations
#define SFDX_INST(R)                                                                                              \
    template int tune_spmv<R>(const DevGrid<R> &, R, const cx<R> *, cx<R> *, RedBuf &, hipStream_t, int, float *);  \
    template int64_t launch_spmv<R>(const DevGrid<R> &, R, const cx<R> *, cx<R> *, R, int, const cx<R> *,         \
                                    RedBuf &, hipStream_t);                                                        \
    template int64_t launch_spmv_march<R>(const DevGrid<R> &, R, const cx<R> *, cx<R> *, R, int, const cx<R> *,   \
                                          RedBuf &, hipStream_t, int, int);                                        \
    template int64_t launch_spmv_gather<R>(const DevGrid<R> &, R, const cx<R> *, cx<R> *, R, int, const cx<R> *,  \
                                           RedBuf &, hipStream_t);                                                 \
    template void launch_mass<R>(const DevGrid<R> &, const cx<R> *, cx<R> *, hipStream_t);                        \
    template void launch_ssor<R>(const DevGrid<R> &, R, const cx<R> *, cx<R> *, R, hipStream_t);                 \
    template void launch_cocr_ap<R>(int64_t, RedBuf &, cx<R> *, const cx<R> *, zf *, bool, hipStream_t);          \
    template void launch_vec<R>(int, const DevGrid<R> &, RedBuf &, cx<R> *, cx<R> *, cx<R> *, cx<R> *, cx<R> *,   \
                                cx<R> *, hipStream_t);                                                            \
    template void launch_dot<R>(int, int64_t, const cx<R> *, const cx<R> *, RedBuf &, hipStream_t, int64_t &);    \
    template void launch_copy<R>(int64_t, const cx<R> *, cx<R> *, hipStream_t);                                   \
    template void launch_axpby<R>(int64_t, cx<R>, const cx<R> *, cx<R>, cx<R> *, hipStream_t);                    \
    template void launch_dirichlet<R>(const Dims &, const cx<R> *, int, cx<R> *, int, hipStream_t, int);         \
    template void launch_node_div<R>(const DevGrid<R> &, const cx<R> *, cx<R> *, RedBuf &, hipStream_t,          \
                                     int64_t &);                                                                  \
    template void launch_pbuild<R>(const Dims &, const double *, const double *, const double *, const double *, \
                                   R *, hipStream_t);                                                             \
    template void launch_papply<R>(const Dims &, const R *, const cx<R> *, cx<R> *, int, RedBuf &, hipStream_t,   \
                                   int64_t &);                                                                    \
    template void launch_pcg<R>(int, int64_t, const R *, RedBuf &, cx<R> *, cx<R> *, cx<R> *, const cx<R> *,      \
                                hipStream_t, int64_t &);                                                          \
    template void launch_grad_update<R>(const DevGrid<R> &, const cx<R> *, cx<R> *, hipStream_t);                 \
    template void launch_prolong<R>(const Dims &, const Dims &, int, const ProlongDev &, const cx<R> *, cx<R> *,   \
                                    hipStream_t);                                                                 \
    template void launch_extrap<R>(int64_t, R, const cx<R> *, R, cx<R> *, hipStream_t);                        \
    template void launch_hybrid<R>(const DevGrid<R> &, R, const R *, const cx<R> *, cx<R> *, cx<R> *, cx<R> *, R, \
                                   hipStream_t, int, const cx<R> *, double *, bool);                             \
    template void launch_cocr_upd_grad<R>(const DevGrid<R> &, RedBuf &, cx<R> *, cx<R> *, const cx<R> *, cx<R> *,  \
                                          cx<R> *, const cx<R> *, const cx<R> *, R, hipStream_t);                 \
    template void launch_hybrid_gs<R>(const DevGrid<R> &, R, const R *, const cx<R> *, cx<R> *, cx<R> *, cx<R> *,  \
                                      int, hipStream_t, bool, int);                                                    \
    template void launch_resid_split<R>(const DevGrid<R> &, R, const cx<R> *, const cx<R> *, const cx<R> *, cx<R> *, \
                                        hipStream_t);                                           \
    template void launch_prolong_node_add<R>(const Dims &, const Dims &, const MGTables &, const cx<R> *,             \
                                             cx<R> *, hipStream_t);                                               \
    template void launch_grad_add<R>(const DevGrid<R> &, cx<R>, const cx<R> *, cx<R> *, hipStream_t, const cx<R> *, \
                                     double *, bool);            \
    template void launch_gpbicg<R>(int, const DevGrid<R> &, RedBuf &, const GPVecs<R> &, hipStream_t);           \
    template void launch_restrict<R>(const Dims &, const Dims &, const MGTables &, const cx<R> *, cx<R> *,            \
                                     hipStream_t);                                                                \
    template void launch_prolong_add<R>(const Dims &, const Dims &, const MGTables &, const cx<R> *,                  \
                                        cx<R> *, hipStream_t);

SFDX_INST(double)
SFDX_INST(float)
template void launch_convert<double, float>(int64_t, const zd *, zf *, hipStream_t);
template void launch_convert<float, double>(int64_t, const zf *, zd *, hipStream_t);

}  // namespace sfdx
