// sfdx_api.hip -- host orchestration of the forward solve and the C ABI.
//
// fwd_solver (sfd_excmg/src/fdfe_main.f90:1565-1594) loops over frequencies;
// for each one doublemode_acc (:1258-1563) solves the two coarsest levels
// directly (u1u2, :1228-1256; here one MKL PARDISO factorisation serves both
// polarisations), then for levels 3..L builds the EXCMG initial guess on the
// device (extint_v), imposes the Dirichlet data of the level's 1-D background
// and runs the preconditioned Krylov solver on the device with the level
// tolerance eps*10^(i-L) (:1312-1331).  Without EXCMG, doublemode_nor (:892-1063)
// solves the fine level from x = 0.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "../../include/sfdx.h"
#include "sfdx_host.h"
#include "sfdx_kernels.h"

using namespace sfdx;
typedef std::complex<double> zc;

static thread_local std::string g_err;

// Scratch of the allocation-free entry point sfdx_apply_operator_noalloc: the
// reciprocal cell widths of the three axes, kNoallocMaxDim doubles each.  One
// copy per GPU lives until the process ends.  On the host this name is only a
// stand-in for the symbol: its device address is looked up once per device
// (noalloc_scratch_init, from sfdx_create) and only that address is used.
static constexpr int kNoallocMaxDim = 4096;
// tightest relative residual asked of the device solves of EXCMG levels 1-2
static constexpr double kCoarseTolFloor = 1e-12;
static __device__ __align__(256) double scratch[3 * kNoallocMaxDim];
static constexpr int kMaxDevices = 64;
static double *g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

namespace {

struct Pool {
    std::map<std::string, std::pair<void *, size_t>> bufs;
    void *get(const std::string &tag, size_t bytes)
    {
        auto it = bufs.find(tag);
        if (it != bufs.end()) {
            if (it->second.second >= bytes) return it->second.first;
            SFDX_CHECK(hipFree(it->second.first));
            bufs.erase(it);
        }
        void *p = nullptr;
        SFDX_CHECK(hipMalloc(&p, bytes ? bytes : 16));
        bufs[tag] = {p, bytes};
        return p;
    }
    void release()
    {
        for (auto &kv : bufs) (void)hipFree(kv.second.first);
        bufs.clear();
    }
};

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double s() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

}  // namespace

struct sfdx_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    RedBuf rb;
    Pool pool;
    double *hslots = nullptr;  // pinned
    double *hlag = nullptr;    // pinned, 2 x SL_COUNT: lagged convergence test
    hipEvent_t lag_ev[2] = {nullptr, nullptr};
    std::vector<sfdx_ctx *> workers;  // per-stream contexts of sfdx_solve_batch
    std::map<std::array<int, 4>, int> tiles;  // autotuned SpMV tile per (nx, ny, nz, sizeof R)
};

namespace {

template <typename T>
T *pool_vec(sfdx_ctx *ctx, const std::string &tag, size_t n)
{
    return (T *)ctx->pool.get(tag, n * sizeof(T));
}

void ensure_partials(sfdx_ctx *ctx, int64_t nblk)
{
    if (nblk > ctx->rb.max_blocks) {
        if (ctx->rb.partials) SFDX_CHECK(hipFree(ctx->rb.partials));
        SFDX_CHECK(hipMalloc(&ctx->rb.partials, (size_t)nblk * 8 * sizeof(double)));
        ctx->rb.max_blocks = nblk;
    }
}

// SpMV kernel of a level from the option spmv_tile: fixed (0..4), or (-1) timed
// once per grid shape and precision by tune_spmv and cached in the context.
// Needs the reduction buffer sized for the level (ensure_partials).
template <typename R>
void choose_tile(sfdx_ctx *ctx, DevGrid<R> &g, const std::string &tag, int opt, int verbose)
{
    if (opt >= 0) {
        g.spmv_variant = opt;
        return;
    }
    const Dims &d = g.d;
    const std::array<int, 4> key{d.nx, d.ny, d.nz, (int)sizeof(R)};
    auto it = ctx->tiles.find(key);
    if (it == ctx->tiles.end()) {
        cx<R> *x = pool_vec<cx<R>>(ctx, tag + "/tune_x", d.NL), *y = pool_vec<cx<R>>(ctx, tag + "/tune_y", d.NL);
        SFDX_CHECK(hipMemsetAsync(x, 0, d.NL * sizeof(cx<R>), ctx->st));
        float ms[5];
        const int v = tune_spmv<R>(g, (R)1, x, y, ctx->rb, ctx->st, 5, ms);
        if (verbose)
            fprintf(stderr, "  spmv kernel %dx%dx%d %s: %.4f %.4f %.4f %.4f %.4f ms -> variant %d\n", d.nx, d.ny,
                    d.nz, sizeof(R) == 8 ? "c128" : "c64", ms[0], ms[1], ms[2], ms[3], ms[4], v);
        it = ctx->tiles.emplace(key, v).first;
    }
    g.spmv_variant = it->second;
}

void read_slots(sfdx_ctx *ctx)
{
    SFDX_CHECK(hipMemcpyAsync(ctx->hslots, ctx->rb.slots, SL_COUNT * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    SFDX_CHECK(hipStreamSynchronize(ctx->st));
}

// asynchronous copy of the scalar slots into pinned buffer `slot` (0 or 1)
void post_slots(sfdx_ctx *ctx, int slot)
{
    SFDX_CHECK(hipMemcpyAsync(ctx->hlag + slot * SL_COUNT, ctx->rb.slots, SL_COUNT * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->st));
    SFDX_CHECK(hipEventRecord(ctx->lag_ev[slot], ctx->st));
}

const double *wait_slots(sfdx_ctx *ctx, int slot)
{
    SFDX_CHECK(hipEventSynchronize(ctx->lag_ev[slot]));
    return ctx->hlag + slot * SL_COUNT;
}

// Device copy of one grid level in precision R (plus FP64 copies for P assembly).
template <typename R>
struct LevelDev {
    DevGrid<R> g;
    double *a64 = nullptr, *b64 = nullptr, *c64 = nullptr, *sig64 = nullptr;
    R *P = nullptr;  // 29 * NP: nodal operator G^T M G (19 slots) and its symmetric form
    float *Ps32 = nullptr;  // 10 * NP: the symmetric form in FP32 (complex64-edge hybrid)
};

template <typename T>
T *upload_vec(sfdx_ctx *ctx, const std::string &tag, const std::vector<T> &v)
{
    T *d = pool_vec<T>(ctx, tag, v.size());
    SFDX_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->st));
    return d;
}

// Device copy of one level in precision R from its widths (host) and its FP64
// sigma already on the device (sig64, 6 x ncell component-major, which stays
// the level's FP64 copy): sigma in R (FP32: the diagonal raised to
// sigma_floor, 0 = unchanged) or, in FP64, sig64 itself plus the pair layout of
// the cell-plane march, both formed on the device.
template <typename R>
LevelDev<R> level_from_device(sfdx_ctx *ctx, const HostLevel &L, double *sig64, const std::string &tag,
                              double sigma_floor = 0)
{
    LevelDev<R> D;
    D.g.d = make_dims(L.nx, L.ny, L.nz);
    const int64_t nc = (int64_t)L.nx * L.ny * L.nz;
    std::vector<R> a(L.a.begin(), L.a.end()), b(L.b.begin(), L.b.end()), c(L.c.begin(), L.c.end());
    const std::string p = tag + (sizeof(R) == 8 ? "/d" : "/f");
    D.g.a = upload_vec(ctx, p + "a", a);
    D.g.b = upload_vec(ctx, p + "b", b);
    D.g.c = upload_vec(ctx, p + "c", c);
    if constexpr (sizeof(R) == 8) {
        if (sigma_floor > 0) throw std::runtime_error("sigma_floor applies to FP32 levels only");
        D.g.sig = sig64;
        double *dp = pool_vec<double>(ctx, p + "sp", 6 * nc);
        launch_sigma_pairs(nc, sig64, dp, ctx->st);
        D.g.sigp = dp;
    } else {
        float *sf = pool_vec<float>(ctx, p + "s", 6 * nc);
        launch_sigma_f32(nc, sigma_floor, sig64, sf, ctx->st);
        D.g.sig = sf;
    }
    // reciprocal widths (computed in FP64, stored in R) for the stencil rows
    std::vector<R> ra(L.a.size()), rb(L.b.size()), rc(L.c.size());
    for (size_t q = 0; q < ra.size(); q++) ra[q] = (R)(1.0 / L.a[q]);
    for (size_t q = 0; q < rb.size(); q++) rb[q] = (R)(1.0 / L.b[q]);
    for (size_t q = 0; q < rc.size(); q++) rc[q] = (R)(1.0 / L.c[q]);
    D.g.ra = upload_vec(ctx, p + "ra", ra);
    D.g.rb = upload_vec(ctx, p + "rb", rb);
    D.g.rc = upload_vec(ctx, p + "rc", rc);
    D.a64 = upload_vec(ctx, tag + "/a64", L.a);
    D.b64 = upload_vec(ctx, tag + "/b64", L.b);
    D.c64 = upload_vec(ctx, tag + "/c64", L.c);
    D.sig64 = sig64;
    SFDX_CHECK(hipStreamSynchronize(ctx->st));  // the host vectors above end here
    return D;
}

// Device copy of one grid level in precision R (plus FP64 copies for P
// assembly) from a host level.
template <typename R>
LevelDev<R> upload_level(sfdx_ctx *ctx, const HostLevel &L, const std::string &tag, double sigma_floor = 0)
{
    const size_t n = 6 * L.ncell();
    double *s64 = pool_vec<double>(ctx, tag + "/s64", n);
    SFDX_CHECK(hipMemcpyAsync(s64, L.S(), n * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    return level_from_device<R>(ctx, L, s64, tag, sigma_floor);
}

template <typename R>
void ensure_P(sfdx_ctx *ctx, LevelDev<R> &D, const std::string &tag)
{
    if (D.P) return;
    D.P = pool_vec<R>(ctx, tag + (sizeof(R) == 8 ? "/P64" : "/P32"), (size_t)29 * D.g.d.NP);
    launch_pbuild<R>(D.g.d, D.a64, D.b64, D.c64, D.sig64, D.P, ctx->st);
}

// Fine node index of each coarse node of a 2^k coarsening of n cells.
std::vector<int> uniform_starts(int n, int k)
{
    std::vector<int> st((size_t)(n >> k) + 1);
    for (size_t c = 0; c < st.size(); c++) st[c] = (int)(c << k);
    return st;
}

using Starts = std::array<std::vector<int>, 3>;

// The grid whose cell c along axis a merges the fine cells [st[a][c],
// st[a][c+1]): widths summed on the host, sigma averaged on the device from
// the fine grid's resident FP64 copy (k_mesh_coarsen; for a 2^k coarsening,
// mesh_coarsen of src/fdfe_main.f90:18-84, the same bits as the host coarsen)
// and copied back for the host-side users (1-D background, direct solves)
// unless host_sig is false; *dev_sig (when given) receives the device sigma.
HostLevel coarsen_by_starts(sfdx_ctx *ctx, const HostLevel &F, const LevelDev<double> &DF, const Starts &st,
                            const std::string &tag, double **dev_sig = nullptr, bool host_sig = true)
{
    HostLevel C;
    const std::vector<double> *wf[3] = {&F.a, &F.b, &F.c};
    std::vector<double> *wc[3] = {&C.a, &C.b, &C.c};
    for (int a = 0; a < 3; a++) {
        wc[a]->assign(st[a].size() - 1, 0.0);
        for (size_t c = 0; c + 1 < st[a].size(); c++)
            for (int q = st[a][c]; q < st[a][c + 1]; q++) (*wc[a])[c] += (*wf[a])[q];
    }
    C.nx = (int)C.a.size(); C.ny = (int)C.b.size(); C.nz = (int)C.c.size();
    C.nair = 0;  // coarse cells wholly in the air
    while (C.nair < C.nz && st[2][C.nair + 1] <= F.nair) C.nair++;
    const Dims df = make_dims(F.nx, F.ny, F.nz), dc = make_dims(C.nx, C.ny, C.nz);
    const int *sx = upload_vec(ctx, tag + "/sx", st[0]), *sy = upload_vec(ctx, tag + "/sy", st[1]),
              *sz = upload_vec(ctx, tag + "/sz", st[2]);
    const double *ac = upload_vec(ctx, tag + "/ca", C.a), *bc = upload_vec(ctx, tag + "/cb", C.b),
                 *cc = upload_vec(ctx, tag + "/cc", C.c);
    double *sc = pool_vec<double>(ctx, tag + "/cs", 6 * dc.NC);
    launch_mesh_coarsen(df, dc, sx, sy, sz, DF.a64, DF.b64, DF.c64, ac, bc, cc, DF.sig64, sc, ctx->st);
    if (host_sig) {
        C.sig.assign(6 * dc.NC, 0.0);
        SFDX_CHECK(hipMemcpyAsync(C.sig.data(), sc, 6 * dc.NC * 8, hipMemcpyDeviceToHost, ctx->st));
    }
    SFDX_CHECK(hipStreamSynchronize(ctx->st));
    if (dev_sig) *dev_sig = sc;
    return C;
}

// The grid of a 2^k coarsening (mesh_coarsen, src/fdfe_main.f90:18-84).
HostLevel coarsen_device(sfdx_ctx *ctx, const HostLevel &F, const LevelDev<double> &DF, int k, const std::string &tag,
                         double **dev_sig = nullptr)
{
    const Starts st = {uniform_starts(F.nx, k), uniform_starts(F.ny, k), uniform_starts(F.nz, k)};
    return coarsen_by_starts(ctx, F, DF, st, tag, dev_sig);
}

// Coarse node starts of one axis of the merge-ratio coarsening: two
// consecutive cells (scanning from index 0) are merged when both are at most
// thr wide and node `keep` (the air-earth interface; -1: none) does not lie
// between them; every other cell stays as it is.
std::vector<int> merge_starts(const std::vector<double> &w, double thr, int keep)
{
    std::vector<int> st{0};
    const int n = (int)w.size();
    for (int i = 0; i < n;) {
        const bool pair = i + 1 < n && w[i] <= thr && w[i + 1] <= thr && i + 1 != keep;
        i += pair ? 2 : 1;
        st.push_back(i);
    }
    return st;
}

// Helper context of the second polarisation (its own stream, reduction
// buffers, pinned slots and work-vector pool), created on first use and kept
// as the first of the context's worker contexts (which the batch also uses).
sfdx_ctx *pol_helper(sfdx_ctx *ctx)
{
    if (ctx->workers.empty()) {
        sfdx_ctx *w = nullptr;
        if (sfdx_create(ctx->device, &w) != SFDX_OK) throw std::runtime_error("helper context: " + g_err);
        ctx->workers.push_back(w);
    }
    sfdx_ctx *w = ctx->workers[0];
    ensure_partials(w, ctx->rb.max_blocks);
    return w;
}

// Stagnation window of the restarted COCR with a complex64 preconditioner:
// iterations without a new best residual after which the recurrences are
// taken to have stalled.  Two decades at the reduction rate reached since the
// last restart (it_best - it_restart iterations took the residual from rel0 to
// best), between 20 and 400: a V-cycle that gains a decade every ~10
// iterations is restarted after ~20 idle ones, the one-level preconditioner
// (a decade per ~150 iterations) only after 300-400, as its residual wanders
// that long between new bests.
inline double stall_window(double rel0, double best, int it_restart, int it_best)
{
    if (!(best < rel0)) return 400;
    return std::min(400.0, std::max(20.0, 2.0 * (it_best - it_restart) / std::log10(rel0 / best)));
}

struct SolveInfo {
    int iters = 0, dc_calls = 0, pcg_iters = 0, status = 0, restarts = 0, fallback = 0;
    int stalled = 0;   // COCR stopped on unproductive restarts (not on maxit)
    int switched = 0;  // the solve was continued by GPBiCG after such a stop
    double relres = 0;
    double bytes = 0;
};

// ---------------------------------------------------------------- divergence correction
// x <- x - G phi with P phi = G^T M_sigma x on interior nodes (div_correc,
// src/div_correc.f90:11-147; Jacobi-PCG as pcg, :343-402, itmax 2000).
template <typename R>
int div_correct(sfdx_ctx *ctx, LevelDev<R> &D, const std::string &tag, cx<R> *x, double tol)
{
    const Dims &d = D.g.d;
    ensure_P(ctx, D, tag);
    const std::string p = tag + (sizeof(R) == 8 ? "/dc64" : "/dc32");
    cx<R> *w = pool_vec<cx<R>>(ctx, p + "w", d.NL);
    cx<R> *t = pool_vec<cx<R>>(ctx, p + "t", d.NP);
    cx<R> *u = pool_vec<cx<R>>(ctx, p + "u", d.NP);
    cx<R> *r = pool_vec<cx<R>>(ctx, p + "r", d.NP);
    cx<R> *pp = pool_vec<cx<R>>(ctx, p + "p", d.NP);
    cx<R> *v = pool_vec<cx<R>>(ctx, p + "v", d.NP);
    int64_t nb;
    launch_mass<R>(D.g, x, w, ctx->st);
    launch_node_div<R>(D.g, w, t, ctx->rb, ctx->st, nb);
    reduce_final(ctx->rb, nb, 1, SL_TMP, OP_T, ctx->st);
    launch_pcg<R>(2, d.NP, D.P, ctx->rb, u, r, pp, t, ctx->st, nb);
    reduce_final(ctx->rb, nb, 1, SL_TMP, OP_PCG_INIT, ctx->st);
    read_slots(ctx);
    const double tn = std::sqrt(ctx->hslots[SL_T]);
    if (!(tn > 0)) return 0;
    int it = 0;
    for (; it < 2000; it++) {
        launch_papply<R>(d, D.P, pp, v, 1, ctx->rb, ctx->st, nb);
        reduce_final(ctx->rb, nb, 2, SL_TMP, OP_PCG_ALPHA, ctx->st);
        launch_pcg<R>(0, d.NP, D.P, ctx->rb, u, r, pp, v, ctx->st, nb);
        reduce_final(ctx->rb, nb, 2, SL_TMP, OP_PCG_BETA, ctx->st);
        launch_pcg<R>(1, d.NP, D.P, ctx->rb, u, r, pp, v, ctx->st, nb);
        read_slots(ctx);
        const double err = std::sqrt(ctx->hslots[SL_PR]) / tn;
        if (!(err == err)) break;
        if (err < tol) { it++; break; }
    }
    launch_grad_update<R>(D.g, u, x, ctx->st);
    return it;
}

// ---------------------------------------------------------------- multigrid
// Geometric V-cycle over the EXCMG grid hierarchy (levels 1..L, each a 2x
// coarsening of the next), used as the preconditioner of the Krylov solve:
// forward hybrid Gauss-Seidel pre-smoothing, residual restriction (P^T),
// recursive coarse correction, prolongation, backward post-smoothing (the
// transpose of the pre-smoother, so the cycle is symmetric), and a fixed number
// of smoothing pairs on the coarsest level.  Coarse operators are the same
// stencil on the volume-averaged conductivity of the coarse grid.
//
// Split form: every level's correction is held as z + (i/wmu) G psi with an
// edge part z and a nodal part psi.  In resistive cells (air) psi is many
// orders of magnitude larger than z; applying the curl-curl stencil to G psi
// would cancel to rounding noise of that size (K G = 0), so residuals use
// A G psi = -i wmu M G psi instead, and the edge/nodal prolongations commute
// (P_e G_c = G_f P_n) so psi is carried through the hierarchy as nodes.
template <typename R>
struct MGHier {
    std::vector<LevelDev<R> *> lev;  // index = level (1-based), lev[0] unused
    std::vector<MGTables> tr;        // tr[l]: transfer tables of level l vs l-1
    int coarse_pairs = 4;
    int nodal = 0;  // nodal part of the smoother: 0 8-colour Gauss-Seidel on P, 1 Jacobi polynomial
    std::string name = "mg";  // prefix of the work-vector and operator pool tags
};

// The level data the solvers build on first use (the nodal operators P of the
// level and of every multigrid grid, the FP32 copy of the symmetric form),
// built here so that two polarisations solved at the same time only read it.
void prebuild_level(sfdx_ctx *ctx, LevelDev<double> &D64, LevelDev<float> *D32, MGHier<double> *h64,
                    MGHier<float> *h32, const std::string &tag, bool mixed)
{
    ensure_P(ctx, D64, tag);
    if (mixed && !D64.Ps32) {
        D64.Ps32 = pool_vec<float>(ctx, tag + "/Ps32", (size_t)10 * D64.g.d.NP);
        launch_psym32(D64.g.d.NP, D64.P, D64.Ps32, ctx->st);
    }
    if (D32) ensure_P(ctx, *D32, tag);
    for (size_t l = 1; h64 && l < h64->lev.size(); l++)
        if (h64->lev[l]) ensure_P(ctx, *h64->lev[l], h64->name + std::to_string(l) + "d");
    for (size_t l = 1; h32 && l < h32->lev.size(); l++)
        if (h32->lev[l]) ensure_P(ctx, *h32->lev[l], h32->name + std::to_string(l) + "f");
    SFDX_CHECK(hipStreamSynchronize(ctx->st));
}

// Transfer tables between grid F and the coarser grid C whose node starts are
// st (MGTables): the coarse node at or below each fine node and its linear
// weight from the node positions, uploaded under `tag`.
MGTables mg_tables(sfdx_ctx *ctx, const HostLevel &F, const HostLevel &C, const Starts &st, const std::string &tag)
{
    MGTables T;
    const std::vector<double> *wf[3] = {&F.a, &F.b, &F.c}, *wc[3] = {&C.a, &C.b, &C.c};
    std::vector<int> lo[3];
    std::vector<double> wl[3];
    for (int a = 0; a < 3; a++) {
        const size_t nf = wf[a]->size(), nc = wc[a]->size();
        std::vector<double> xf(nf + 1, 0.0), xc(nc + 1, 0.0);
        for (size_t i = 0; i < nf; i++) xf[i + 1] = xf[i] + (*wf[a])[i];
        for (size_t i = 0; i < nc; i++) xc[i + 1] = xc[i] + (*wc[a])[i];
        lo[a].assign(nf + 1, (int)nc);
        wl[a].assign(nf + 1, 1.0);
        for (size_t c = 0; c < nc; c++)
            for (int q = st[a][c]; q < st[a][c + 1]; q++) {
                lo[a][q] = (int)c;
                if (q != st[a][c]) {
                    wl[a][q] = (xc[c + 1] - xf[q]) / (xc[c + 1] - xc[c]);
                    // the kernels tell a node inside a coarse cell by its weight below 1
                    if (!(wl[a][q] > 0 && wl[a][q] < 1)) throw std::runtime_error("degenerate multigrid cell widths");
                }
            }
        const std::string t = tag + "_" + std::to_string(a);
        T.wl[a] = upload_vec(ctx, t + "w", wl[a]);
        T.lo[a] = upload_vec(ctx, t + "l", lo[a]);
        T.st[a] = upload_vec(ctx, t + "s", st[a]);
    }
    SFDX_CHECK(hipStreamSynchronize(ctx->st));  // the host tables above end here
    return T;
}

// (z, psi) ~= A^-1 r on level l (split form, see above)
template <typename R>
void mg_vcycle(sfdx_ctx *ctx, MGHier<R> &H, int l, R wmu, const cx<R> *r, cx<R> *z, cx<R> *psi)
{
    LevelDev<R> &D = *H.lev[l];
    const Dims &d = D.g.d;
    const std::string tg = H.name + std::to_string(l) + (sizeof(R) == 8 ? "d" : "f");
    cx<R> *t = pool_vec<cx<R>>(ctx, tg + "t", d.NL), *w = pool_vec<cx<R>>(ctx, tg + "w", d.NL);
    cx<R> *tn = pool_vec<cx<R>>(ctx, tg + "tn", d.NP), *ph = pool_vec<cx<R>>(ctx, tg + "ph", d.NP);
    ensure_P(ctx, D, tg);
    hipStream_t st = ctx->st;
    // pre-smoothing from zero: z = GS_e r, psi = GS_n G^T r
    launch_hybrid_gs<R>(D.g, wmu, D.P, r, z, tn, psi, 0, st, false, H.nodal);
    auto smooth_add = [&](int dir) {
        launch_resid_split<R>(D.g, wmu, z, psi, r, t, st);
        launch_hybrid_gs<R>(D.g, wmu, D.P, t, w, tn, ph, dir, st, false, H.nodal);
        launch_axpby<R>(d.NL, cx<R>(1), w, cx<R>(1), z, st);
        launch_axpby<R>(d.NP, cx<R>(1), ph, cx<R>(1), psi, st);
    };
    if (l == 1) {
        for (int q = 0; q < H.coarse_pairs; q++) {
            if (q > 0) smooth_add(0);
            smooth_add(1);
        }
        return;
    }
    LevelDev<R> &C = *H.lev[l - 1];
    const std::string tc = H.name + std::to_string(l - 1) + (sizeof(R) == 8 ? "d" : "f");
    cx<R> *rc = pool_vec<cx<R>>(ctx, tc + "r", C.g.d.NL), *zc = pool_vec<cx<R>>(ctx, tc + "z", C.g.d.NL);
    cx<R> *pc = pool_vec<cx<R>>(ctx, tc + "p", C.g.d.NP);
    launch_resid_split<R>(D.g, wmu, z, psi, r, t, st);
    launch_restrict<R>(d, C.g.d, H.tr[l], t, rc, st);
    mg_vcycle<R>(ctx, H, l - 1, wmu, rc, zc, pc);
    launch_prolong_add<R>(d, C.g.d, H.tr[l], zc, z, st);
    launch_prolong_node_add<R>(d, C.g.d, H.tr[l], pc, psi, st);
    smooth_add(1);
}

// One captured HIP graph of a fixed launch sequence on a stream (kernel nodes
// only: no memset or host-synchronising call is issued inside the capture).
struct StreamGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    template <typename F>
    void capture(hipStream_t st, F &&enqueue)
    {
        SFDX_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue();
        SFDX_CHECK(hipStreamEndCapture(st, &g));
        SFDX_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    }
    void launch(hipStream_t st) { SFDX_CHECK(hipGraphLaunch(exec, st)); }
    void reset()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (g) (void)hipGraphDestroy(g);
        exec = nullptr;
        g = nullptr;
    }
    ~StreamGraph() { reset(); }
};

// ---------------------------------------------------------------- Krylov
struct KOpts {
    int corr_intv = 0;
    double corr_tol = 1e-3;
    double omega = 1.0;
    int verbose = 0;
    int dc_restart = 0;
    int dc_at_start = 1;
    int precond = SFDX_PRECOND_HYBRID;
    LevelDev<float> *lo = nullptr;  // FP64 Krylov with an FP32 preconditioner when set
    MGHier<double> *mg64 = nullptr;  // multigrid hierarchy (precond == MG)
    MGHier<float> *mg32 = nullptr;
    int mg_level = 0;                // level being solved
    double restart_drop = 0;         // COCR: restart from the true residual after this drop
    int use_graph = 1;               // replay the steady-state iteration as a HIP graph
    int nodal = 0;                   // hybrid preconditioner: 0 nodal SSOR, 1 two-step Jacobi
};

// out = M^-1 in with the chosen preconditioner
template <typename R>
// With dotw the partial sums of dotw^T out are produced by the last kernel when
// the preconditioner allows it (hybrid); returns whether it did.
bool apply_prec(sfdx_ctx *ctx, LevelDev<R> &D, const std::string &tag, R wmu, const cx<R> *in, cx<R> *out,
                const KOpts &ko, const cx<R> *dotw = nullptr, bool in32_ready = false, bool store = true)
{
    if (ko.precond == SFDX_PRECOND_MG) {
        const Dims &d = D.g.d;
        // store false (with dotw): out keeps the V-cycle's edge part (or, with the
        // complex64 V-cycle, nothing) and only the dot of z + (i/wmu) G psi is formed
        const bool st_out = store || !dotw;
        if (sizeof(R) == 8 && ko.mg32) {
            // FP64 Krylov, complex64 V-cycle; the recombination z + (i/wmu) G psi is done in FP64
            zf *in32 = pool_vec<zf>(ctx, tag + "/mg_in", d.NL), *z32 = pool_vec<zf>(ctx, tag + "/mg_z", d.NL);
            zf *p32 = pool_vec<zf>(ctx, tag + "/mg_p", d.NP);
            if (!in32_ready) launch_convert<double, float>(d.NL, (const zd *)in, in32, ctx->st);
            mg_vcycle<float>(ctx, *ko.mg32, ko.mg_level, (float)wmu, in32, z32, p32);
            launch_hyb_combine(*(const DevGrid<double> *)&D.g, zd(0.0, 1.0 / (double)wmu), p32, z32, (zd *)out, ctx->st,
                               (const zd *)dotw, ctx->rb.partials, st_out);
        } else {
            MGHier<R> &H = *(MGHier<R> *)(sizeof(R) == 8 ? (void *)ko.mg64 : (void *)ko.mg32);
            cx<R> *pn = pool_vec<cx<R>>(ctx, tag + "/mg_pn", d.NP);
            mg_vcycle<R>(ctx, H, ko.mg_level, wmu, in, out, pn);
            launch_grad_add<R>(D.g, cx<R>((R)0, (R)1 / wmu), pn, out, ctx->st, dotw, ctx->rb.partials, st_out);
        }
        return dotw != nullptr;
    }
    if (sizeof(R) == 8 && ko.lo) {
        LevelDev<float> &F = *ko.lo;
        const Dims &d = F.g.d;
        zf *in32 = pool_vec<zf>(ctx, tag + "/mp_in", d.NL), *z32 = pool_vec<zf>(ctx, tag + "/mp_z", d.NL);
        zf *tn = pool_vec<zf>(ctx, tag + "/mp_t", d.NP), *ph = pool_vec<zf>(ctx, tag + "/mp_phi", d.NP);
        if (ko.precond == SFDX_PRECOND_HYBRID && ko.nodal == 1) {
            // edge SSOR in complex64; the nodal part (one Jacobi pass over P, where
            // the air's large potentials live) in complex128
            LevelDev<double> &D64 = *(LevelDev<double> *)&D;
            ensure_P(ctx, D64, tag);
            if (!D64.Ps32) {
                D64.Ps32 = pool_vec<float>(ctx, tag + "/Ps32", (size_t)10 * d.NP);
                launch_psym32(d.NP, D64.P, D64.Ps32, ctx->st);
            }
            zd *tn64 = pool_vec<zd>(ctx, tag + "/mp_t64", d.NP), *ph64 = pool_vec<zd>(ctx, tag + "/mp_phi64", d.NP);
            launch_hybrid_edge32(D64.g, F.g, (double)wmu, D64.Ps32, (const zd *)in, (zd *)out, in32, z32, tn64, ph64,
                                 (float)ko.omega, ctx->st, (const zd *)dotw, ctx->rb.partials, !in32_ready,
                                 store || !dotw);
            return dotw != nullptr;
        }
        if (ko.precond == SFDX_PRECOND_HYBRID) {
            ensure_P(ctx, F, tag);
            launch_hybrid_mixed(*(DevGrid<double> *)&D.g, F.g, (double)wmu, F.P, (const zd *)in, (zd *)out, in32, z32,
                                tn, ph, (float)ko.omega, ctx->st, ko.nodal, (const zd *)dotw, ctx->rb.partials);
            return dotw != nullptr;
        }
        launch_convert<double, float>(d.NL, (const zd *)in, in32, ctx->st);
        launch_ssor<float>(F.g, (float)wmu, in32, z32, (float)ko.omega, ctx->st);
        launch_convert<float, double>(d.NL, z32, (zd *)out, ctx->st);
        return false;
    }
    if (ko.precond == SFDX_PRECOND_HYBRID) {
        ensure_P(ctx, D, tag);
        const std::string p = tag + (sizeof(R) == 8 ? "/hy64" : "/hy32");
        cx<R> *tn = pool_vec<cx<R>>(ctx, p + "t", D.g.d.NP), *ph = pool_vec<cx<R>>(ctx, p + "phi", D.g.d.NP);
        launch_hybrid<R>(D.g, wmu, D.P, in, out, tn, ph, (R)ko.omega, ctx->st, ko.nodal, dotw, ctx->rb.partials,
                         store || !dotw);
        return dotw != nullptr;
    }
    launch_ssor<R>(D.g, wmu, in, out, (R)ko.omega, ctx->st);
    return false;
}

// Solve A_s x = f on the interior edges (boundary entries of x held fixed; f may
// be null = 0) to ||r/ve|| <= eps * bnorm.  COCR (Sogabe & Zhang 2007) with the
// symmetric multicolour SSOR preconditioner; unconjugated inner products.
// After a divergence correction either the recurrences restart (dc_restart) or
// r, z = M^-1 r and A z are recomputed and the search directions are kept.
template <typename R>
SolveInfo cocr(sfdx_ctx *ctx, LevelDev<R> &D, const std::string &tag, R wmu, const cx<R> *f, cx<R> *x, double bnorm,
               double eps, int maxit, const KOpts &ko_in)
{
    KOpts ko = ko_in;  // the complex64 preconditioner may be dropped below
    const Dims &d = D.g.d;
    const R omega = (R)ko.omega;
    const std::string p = tag + (sizeof(R) == 8 ? "/k64" : "/k32");
    cx<R> *r = pool_vec<cx<R>>(ctx, p + "r", d.NL), *z = pool_vec<cx<R>>(ctx, p + "z", d.NL);
    cx<R> *Az = pool_vec<cx<R>>(ctx, p + "Az", d.NL), *pv = pool_vec<cx<R>>(ctx, p + "p", d.NL);
    cx<R> *Ap = pool_vec<cx<R>>(ctx, p + "Ap", d.NL), *q = pool_vec<cx<R>>(ctx, p + "q", d.NL);
    // With the complex64-edge hybrid preconditioner, the update of Ap also
    // writes the complex64 copy the preconditioner reads (apply_prec's "/mp_in"
    // buffer), so apply_prec(Ap) skips its conversion pass.  Re-evaluated at
    // every use: the complex64 preconditioner may be dropped during the solve.
    auto ap32_of = [&]() -> zf * {
        if (sizeof(R) != 8) return nullptr;
        if (ko.lo && ko.precond == SFDX_PRECOND_HYBRID && ko.nodal == 1) return pool_vec<zf>(ctx, tag + "/mp_in", d.NL);
        if (ko.mg32 && ko.precond == SFDX_PRECOND_MG) return pool_vec<zf>(ctx, tag + "/mg_in", d.NL);  // the V-cycle's input
        return nullptr;
    };
    // The preconditioner pieces the update forms q = M Ap from (apply_prec's
    // buffers of the same tags): complex64 SSOR output and complex128 nodal
    // correction of the complex64-edge hybrid, nodal correction of the hybrid
    zf *const mp_z = sizeof(R) == 8 && ko.lo ? pool_vec<zf>(ctx, tag + "/mp_z", d.NL) : nullptr;
    zd *const mp_phi = sizeof(R) == 8 && ko.lo ? pool_vec<zd>(ctx, tag + "/mp_phi64", d.NP) : nullptr;
    cx<R> *const hy_phi = pool_vec<cx<R>>(ctx, tag + (sizeof(R) == 8 ? "/hy64phi" : "/hy32phi"), d.NP);
    // the V-cycle's pieces (apply_prec's buffers of the same tags): its complex64
    // edge and nodal parts, or its nodal part in precision R
    const bool mg = ko.precond == SFDX_PRECOND_MG;
    zf *const mg_z32 = sizeof(R) == 8 && mg && ko.mg32 ? pool_vec<zf>(ctx, tag + "/mg_z", d.NL) : nullptr;
    zf *const mg_p32 = sizeof(R) == 8 && mg && ko.mg32 ? pool_vec<zf>(ctx, tag + "/mg_p", d.NP) : nullptr;
    cx<R> *const mg_pn = mg ? pool_vec<cx<R>>(ctx, tag + "/mg_pn", d.NP) : nullptr;
    SliceLaunch L = edge_launch(d);
    const int64_t nbe = L.nblocks();
    const double sc = 2.0 * sizeof(R), vecb = (double)d.NL * sc;
    // preconditioner bytes: the edge SSOR model (two sweeps of stencil traffic);
    // the multigrid V-cycle is not counted
    const double b_spmv = 2 * vecb + (double)d.NC * 6 * sizeof(R), b_ssor = mg ? 0.0 : 2 * (b_spmv + vecb);
    SolveInfo info;
    if (ko.dc_at_start) {
        info.pcg_iters += div_correct<R>(ctx, D, tag, x, ko.corr_tol);
        info.dc_calls++;
    }
    // 0: fresh start, 1: soft restart after DC, 2: iterate
    int state = 0;
    bool fresh = false;
    double rel0 = 0;
    // Lagged convergence test: the scalars of iteration i are copied to pinned
    // memory asynchronously and examined only after iteration i+1 has been
    // queued, so the host never drains the stream between iterations.  A stop
    // decision taken on lagged values is re-checked on the newest values.
    double cur[SL_COUNT];
    bool pending = false;
    int pend_slot = 0;
    StreamGraph graph;
    // Safeguards: restart from the true residual on a breakdown (non-finite
    // scalars; the update of x is skipped on the device), on stagnation (no new
    // best residual for 200 iterations; with the complex64 preconditioner for
    // stall_window's count, 20-400 by the convergence rate: the recurrence
    // residual has reached its attainable accuracy or the recurrences drift;
    // with the one-level complex64 preconditioner the residual wanders for a
    // few hundred iterations between new bests at high frequencies, and a
    // 200-iteration window restarted it up to 12 times per solve on the
    // 128x128x96 model above 1 Hz, while a merge-ratio V-cycle's recurrences
    // froze at 6.5e-8 after ~100 iterations of a solve that takes ~70) and, with the
    // complex64 preconditioner (restart_drop > 0), on divergence (residual 100x
    // above the best since the last restart).  After 3 such restarts in a row that did not
    // halve the best residual, the complex64 preconditioner is replaced by the
    // complex128 one (a fixed linear operator); without it the solve stops with
    // SFDX_WARN_NOCONV and the last iterate.
    double best = 1e300, best_at_restart = 1e300;
    int since_best = 0, idle_restarts = 0, it_restart = 0, it_best = 0;
    const bool guard = ko.restart_drop > 0;
    while (true) {
        if (state != 2) {
            fresh = true;
            pending = false;
            launch_spmv<R>(D.g, wmu, x, r, (R)-1, 0, nullptr, ctx->rb, ctx->st);
            if (f) launch_axpby<R>(d.NL, cx<R>(1), f, cx<R>(1), r, ctx->st);
            apply_prec<R>(ctx, D, tag, wmu, r, z, ko);
            const int64_t nbs = launch_spmv<R>(D.g, wmu, z, Az, (R)1, 1, nullptr, ctx->rb, ctx->st);
            // The direction p is kept one step behind: the update kernel (mode 6)
            // forms p = z + beta p itself, so here only Ap is advanced.
            if (state == 0) {
                reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_COCR_RHO, ctx->st);  // also beta = 0
                launch_copy<R>(d.NL, z, pv, ctx->st);
                launch_cocr_ap<R>(d.NL, ctx->rb, Ap, Az, ap32_of(), true, ctx->st);
            } else {
                reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_COCR_BETA, ctx->st);
                launch_cocr_ap<R>(d.NL, ctx->rb, Ap, Az, ap32_of(), false, ctx->st);
            }
            launch_vec<R>(5, D.g, ctx->rb, r, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
            reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
            read_slots(ctx);
            std::memcpy(cur, ctx->hslots, sizeof(cur));
            info.bytes += 2 * b_spmv + b_ssor + 6 * vecb;
            state = 2;
        }
        info.relres = std::sqrt(cur[SL_RN]) / bnorm;
        const bool bad = !(info.relres == info.relres) || std::isinf(info.relres);
        const bool fresh_now = fresh;  // this residual was recomputed from x
        if (fresh) { rel0 = info.relres; fresh = false; best = info.relres; since_best = 0; it_restart = it_best = info.iters; }
        if (info.relres < best) { best = info.relres; since_best = 0; it_best = info.iters; }
        const bool diverging = guard && info.relres > 100 * best;
        const bool stalled = since_best > (guard ? stall_window(rel0, best, it_restart, it_best) : 200);
        const bool restart = (ko.restart_drop > 0 && info.relres < ko.restart_drop * rel0) || diverging || stalled;
        if (bad || info.relres < eps || info.iters >= maxit || cur[SL_BREAK] != 0 || restart) {
            if (pending) {  // decide on the newest iteration instead
                std::memcpy(cur, wait_slots(ctx, pend_slot), sizeof(cur));
                pending = false;
                continue;
            }
            // The recurrence residual drifts from b - A x (in every precision; with a
            // complex64 preconditioner by much more): convergence is accepted, and
            // the final residual reported, only on a residual recomputed from x.
            if (!bad && info.relres < eps) {
                if (!fresh_now) {
                    state = 0;
                    continue;
                }
                info.status = 0;
                break;
            }
            if (info.iters >= maxit) {
                if (!fresh_now && !bad) {
                    state = 0;
                    continue;
                }
                info.status = SFDX_WARN_NOCONV;
                break;
            }
            const bool planned = !bad && cur[SL_BREAK] == 0 && !stalled && !diverging;
            if (ko.verbose > 1)
                fprintf(stderr, "  cocr restart at it %d relres %.3e best %.3e |b| %.3e (%s)\n", info.iters,
                        info.relres, best, bnorm,
                        planned ? "planned" : stalled ? "stalled" : diverging ? "diverging" : "breakdown");
            if (!planned) {
                // no progress since the last restart: count it
                idle_restarts = (best < 0.5 * best_at_restart) ? 0 : idle_restarts + 1;
                best_at_restart = std::min(best_at_restart, best);
                if (idle_restarts >= 3) {
                    if (ko.lo || ko.mg32) {
                        ko.lo = nullptr;
                        ko.mg32 = nullptr;
                        ko.restart_drop = 0;
                        info.fallback = 1;
                        idle_restarts = 0;
                        graph.reset();  // the captured iteration used the complex64 preconditioner
                    } else if (ko.precond == SFDX_PRECOND_MG) {
                        // the V-cycle can diverge when the coarse grids no longer
                        // resolve the skin depth (measured at 25 and 63 Hz on the
                        // 128x128x96 model): continue with the hybrid preconditioner
                        ko.precond = SFDX_PRECOND_HYBRID;
                        info.fallback = 1;
                        idle_restarts = 0;
                        graph.reset();
                    } else {
                        info.status = SFDX_WARN_NOCONV;
                        info.stalled = 1;
                        break;
                    }
                }
            } else {
                // the preconditioner is applied in complex64: its rounding is not a
                // fixed linear operator, so the short recurrences drift; restart
                // from the true FP64 residual after each restart_drop reduction
                best_at_restart = std::min(best_at_restart, best);
            }
            state = 0;
            info.restarts++;
            continue;
        }
        if (ko.verbose > 1 && info.iters % 50 == 0) fprintf(stderr, "  cocr it %d relres %.3e\n", info.iters, info.relres);
        auto part1 = [&]() {
            int64_t nbd = nbe;
            // q = M Ap, with (Ap, q) fused into the preconditioner's last kernel when possible
            zf *ap32 = ap32_of();
            if (sizeof(R) == 8 && ko.precond == SFDX_PRECOND_MG && ko.mg32) {
                // complex64 V-cycle: only the dot (Ap, q) here; the update forms q
                // again from the V-cycle's complex64 edge and nodal parts
                apply_prec<R>(ctx, D, tag, wmu, Ap, q, ko, Ap, ap32 != nullptr, false);
                reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_COCR_ALPHA, ctx->st);
                launch_cocr_upd_hyb(*(DevGrid<double> *)&D.g, ctx->rb, (zd *)x, (zd *)pv, (const zd *)Ap, (zd *)r,
                                    (zd *)z, mg_z32, mg_p32, (double)wmu, ctx->st);
                reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
                return;
            }
            if (ko.precond == SFDX_PRECOND_MG && !ko.mg32) {
                // V-cycle in precision R: q holds its edge part, the update adds s G psi
                apply_prec<R>(ctx, D, tag, wmu, Ap, q, ko, Ap, false, false);
                reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_COCR_ALPHA, ctx->st);
                launch_cocr_upd_grad<R>(D.g, ctx->rb, x, pv, Ap, r, z, q, mg_pn, wmu, ctx->st);
                reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
                return;
            }
            if (ap32) {
                // complex64-edge hybrid: q is not stored; the update forms it again
                // from the preconditioner's complex64 SSOR output and nodal correction
                apply_prec<R>(ctx, D, tag, wmu, Ap, q, ko, Ap, true, false);
                reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_COCR_ALPHA, ctx->st);
                launch_cocr_upd_hyb(*(DevGrid<double> *)&D.g, ctx->rb, (zd *)x, (zd *)pv, (const zd *)Ap, (zd *)r,
                                    (zd *)z, mp_z, mp_phi, (double)wmu, ctx->st);
                reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
                return;
            }
            if (ko.precond == SFDX_PRECOND_HYBRID && !ko.lo && !ko.mg32) {
                // hybrid preconditioner: q holds the edge SSOR output only; the
                // update adds s G phi itself (q = M Ap is not stored)
                apply_prec<R>(ctx, D, tag, wmu, Ap, q, ko, Ap, false, false);
                reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_COCR_ALPHA, ctx->st);
                launch_cocr_upd_grad<R>(D.g, ctx->rb, x, pv, Ap, r, z, q, hy_phi, wmu, ctx->st);
                reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
                return;
            }
            if (!apply_prec<R>(ctx, D, tag, wmu, Ap, q, ko, Ap, ap32 != nullptr))
                launch_dot<R>(0, d.NL, Ap, q, ctx->rb, ctx->st, nbd);
            reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_COCR_ALPHA, ctx->st);
            launch_vec<R>(6, D.g, ctx->rb, x, pv, Ap, r, q, z, ctx->st);
            reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
        };
        auto part2 = [&]() {
            const int64_t nbs = launch_spmv<R>(D.g, wmu, z, Az, (R)1, 1, nullptr, ctx->rb, ctx->st);
            reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_COCR_BETA, ctx->st);
            launch_cocr_ap<R>(d.NL, ctx->rb, Ap, Az, ap32_of(), false, ctx->st);
        };
        since_best++;
        if (ko.use_graph && ko.corr_intv <= 0 && pending) {
            // steady state: the whole iteration is one captured graph, replayed
            // (every buffer and the nodal operator exist after the first iteration)
            if (!graph.exec) graph.capture(ctx->st, [&]() { part1(); part2(); });
            graph.launch(ctx->st);
            info.iters++;
            info.bytes += b_ssor + 2 * vecb + 9 * vecb + b_spmv + 6 * vecb;
            const int slot = info.iters & 1;
            post_slots(ctx, slot);
            std::memcpy(cur, wait_slots(ctx, pend_slot), sizeof(cur));
            pend_slot = slot;
            continue;
        }
        part1();
        info.iters++;
        info.bytes += b_ssor + 2 * vecb + 9 * vecb;
        if (ko.corr_intv > 0 && info.iters % ko.corr_intv == 0) {
            info.pcg_iters += div_correct<R>(ctx, D, tag, x, ko.corr_tol);
            info.dc_calls++;
            // COCR's coupled recurrences assume r = b - A x along one Krylov
            // sequence; after the correction only a full restart is consistent
            // (keeping p and Ap diverges with the hybrid preconditioner)
            state = (ko.dc_restart || ko.precond != SFDX_PRECOND_SSOR) ? 0 : 1;
            continue;
        }
        part2();
        info.bytes += b_spmv + 6 * vecb;
        const int slot = info.iters & 1;
        post_slots(ctx, slot);
        if (pending) std::memcpy(cur, wait_slots(ctx, pend_slot), sizeof(cur));
        pending = true;
        pend_slot = slot;
    }
    return info;
}

// COCR with complex64 directions in split form (SFDX_PREC_MIXED_SPLIT, with
// the multigrid V-cycle).  The solution x and the residual r stay complex128;
// the directions z and p (each an edge part and node values, v = v_e + (i/wmu)
// G v_psi, which is the form the complex64 V-cycle returns), Ap and Az are
// complex64.  Az = A (z_e + (i/wmu) G z_psi) is formed with the curl part
// acting on z_e alone (K G = 0), so the large gradient components of
// resistive cells never meet the curl operator in complex64, and the fused
// dot (z, Az) takes the assembled u = z_e + (i/wmu) G z_psi.  Restarts from
// the true FP64 residual after every restart_drop reduction, on stagnation and
// on divergence; convergence is accepted only on a recomputed residual.
// After 3 restarts in a row that did not halve the best residual it returns
// with fallback set and SFDX_WARN_NOCONV, leaving x as the iterate.
SolveInfo cocr_split(sfdx_ctx *ctx, LevelDev<double> &D, LevelDev<float> &F, MGHier<float> &H, int mg_level,
                     const std::string &tag, double wmu, zd *x, double bnorm, double eps, int maxit, const KOpts &ko)
{
    const Dims &d = D.g.d;
    const std::string p = tag + "/ks";
    zd *r = pool_vec<zd>(ctx, p + "r", d.NL);
    zf *r32 = pool_vec<zf>(ctx, p + "r32", d.NL);
    zf *ze = pool_vec<zf>(ctx, p + "ze", d.NL), *pe = pool_vec<zf>(ctx, p + "pe", d.NL);
    zf *qe = pool_vec<zf>(ctx, p + "qe", d.NL), *Ap = pool_vec<zf>(ctx, p + "Ap", d.NL);
    zf *Az = pool_vec<zf>(ctx, p + "Az", d.NL);
    zf *zpsi = pool_vec<zf>(ctx, p + "zp", d.NP), *ppsi = pool_vec<zf>(ctx, p + "pp", d.NP);
    zf *qpsi = pool_vec<zf>(ctx, p + "qp", d.NP);
    const float w32 = (float)wmu;
    const zf s32(0.0f, (float)(1.0 / wmu));
    const int64_t nbe = edge_launch(d).nblocks();
    const double v16 = (double)d.NL * 16, v8 = (double)d.NL * 8, n8 = (double)d.NP * 8;
    const double b_prod = 3 * v8 + n8 + (double)d.NC * 24;           // split product: z_e, z_psi, sigma in, Az out
    const double b_iter = b_prod + (2 * v8 + n8)                       // the dot (q, Ap)
                          + (4 * v16 + 6 * v8 + 6 * n8)                // split update (x, r; p_e, z_e, q_e, Ap; nodes)
                          + 3 * v8;                                    // Ap = Az + beta Ap
    SolveInfo info;
    int state = 0;
    bool fresh = false;
    double rel0 = 0, cur[SL_COUNT];
    bool pending = false;
    int pend_slot = 0;
    StreamGraph graph;
    double best = 1e300, best_at_restart = 1e300;
    int since_best = 0, idle_restarts = 0, it_restart = 0, it_best = 0;
    while (true) {
        if (state != 2) {
            fresh = true;
            pending = false;
            launch_spmv<double>(D.g, wmu, x, r, -1.0, 0, nullptr, ctx->rb, ctx->st);
            launch_convert<double, float>(d.NL, r, r32, ctx->st);
            mg_vcycle<float>(ctx, H, mg_level, w32, r32, ze, zpsi);
            const int64_t nbs = launch_split_product(F.g, w32, ze, zpsi, Az, ctx->rb, ctx->st);
            reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_COCR_RHO, ctx->st);  // also beta = 0
            launch_copy<float>(d.NL, ze, pe, ctx->st);
            launch_copy<float>(d.NP, zpsi, ppsi, ctx->st);
            launch_cocr_ap<float>(d.NL, ctx->rb, Ap, Az, nullptr, true, ctx->st);
            launch_vec<double>(5, D.g, ctx->rb, r, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
            reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
            read_slots(ctx);
            std::memcpy(cur, ctx->hslots, sizeof(cur));
            info.bytes += (2 * v16 + (double)d.NC * 48) + 2 * v8 + b_prod + 2 * (v8 + n8) + 3 * v8 + v16;
            state = 2;
        }
        info.relres = std::sqrt(cur[SL_RN]) / bnorm;
        const bool bad = !(info.relres == info.relres) || std::isinf(info.relres);
        const bool fresh_now = fresh;
        if (fresh) { rel0 = info.relres; fresh = false; best = info.relres; since_best = 0; it_restart = it_best = info.iters; }
        if (info.relres < best) { best = info.relres; since_best = 0; it_best = info.iters; }
        const bool diverging = info.relres > 100 * best;
        const bool stalled = since_best > stall_window(rel0, best, it_restart, it_best);
        const bool restart = (ko.restart_drop > 0 && info.relres < ko.restart_drop * rel0) || diverging || stalled;
        if (bad || info.relres < eps || info.iters >= maxit || cur[SL_BREAK] != 0 || restart) {
            if (pending) {  // decide on the newest iteration instead
                std::memcpy(cur, wait_slots(ctx, pend_slot), sizeof(cur));
                pending = false;
                continue;
            }
            if (!bad && (info.relres < eps || info.iters >= maxit)) {
                if (!fresh_now) {  // accept or stop only on a recomputed residual
                    state = 0;
                    continue;
                }
                info.status = info.relres < eps ? 0 : SFDX_WARN_NOCONV;
                break;
            }
            const bool planned = !bad && cur[SL_BREAK] == 0 && !stalled && !diverging;
            if (!planned) {
                idle_restarts = (best < 0.5 * best_at_restart) ? 0 : idle_restarts + 1;
                best_at_restart = std::min(best_at_restart, best);
                if (idle_restarts >= 3) {
                    info.fallback = 1;
                    info.status = SFDX_WARN_NOCONV;
                    if (!fresh_now) {  // leave with the residual of the returned iterate
                        launch_spmv<double>(D.g, wmu, x, r, -1.0, 0, nullptr, ctx->rb, ctx->st);
                        launch_vec<double>(5, D.g, ctx->rb, r, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
                        reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
                        read_slots(ctx);
                        info.relres = std::sqrt(ctx->hslots[SL_RN]) / bnorm;
                    }
                    break;
                }
            } else {
                best_at_restart = std::min(best_at_restart, best);
            }
            state = 0;
            info.restarts++;
            continue;
        }
        if (ko.verbose > 1 && info.iters % 50 == 0) fprintf(stderr, "  cocr32 it %d relres %.3e\n", info.iters, info.relres);
        auto part1 = [&]() {
            mg_vcycle<float>(ctx, H, mg_level, w32, Ap, qe, qpsi);
            launch_grad_add<float>(F.g, s32, qpsi, qe, ctx->st, Ap, ctx->rb.partials, false);  // (q, Ap) only
            reduce_final(ctx->rb, nbe, 2, SL_TMP, OP_COCR_ALPHA, ctx->st);
            launch_cocr_split_upd(D.g, ctx->rb, x, r, Ap, ze, qe, pe, ppsi, zpsi, qpsi, wmu, ctx->st);
            reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
        };
        auto part2 = [&]() {
            const int64_t nbs = launch_split_product(F.g, w32, ze, zpsi, Az, ctx->rb, ctx->st);
            reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_COCR_BETA, ctx->st);
            launch_cocr_ap<float>(d.NL, ctx->rb, Ap, Az, nullptr, false, ctx->st);
        };
        since_best++;
        if (ko.use_graph && pending) {
            // steady state: the whole iteration is one captured graph, replayed
            if (!graph.exec) graph.capture(ctx->st, [&]() { part1(); part2(); });
            graph.launch(ctx->st);
        } else {
            part1();
            part2();
        }
        info.iters++;
        info.bytes += b_iter;
        const int slot = info.iters & 1;
        post_slots(ctx, slot);
        if (pending) std::memcpy(cur, wait_slots(ctx, pend_slot), sizeof(cur));
        pending = true;
        pend_slot = slot;
    }
    return info;
}

// Right-preconditioned BiCGStab (Bicg_stab_p, src/solvers.f90:522-719) on A_s.
// Divergence correction every corr_intv iterations followed by r = f - Ax,
// rhat = r; with dc_restart = 0 rho, alpha, omega, p, v carry over as in the
// reference (src/solvers.f90:630-646).
template <typename R>
SolveInfo bicgstab(sfdx_ctx *ctx, LevelDev<R> &D, const std::string &tag, R wmu, const cx<R> *f, cx<R> *x,
                   double bnorm, double eps, int maxit, const KOpts &ko_in)
{
    KOpts ko = ko_in;  // the complex64 preconditioner may be dropped below
    const Dims &d = D.g.d;
    const R omega = (R)ko.omega;
    const std::string p = tag + (sizeof(R) == 8 ? "/b64" : "/b32");
    cx<R> *r = pool_vec<cx<R>>(ctx, p + "r", d.NL), *rh = pool_vec<cx<R>>(ctx, p + "rh", d.NL);
    cx<R> *pv = pool_vec<cx<R>>(ctx, p + "p", d.NL), *y = pool_vec<cx<R>>(ctx, p + "y", d.NL);
    cx<R> *v = pool_vec<cx<R>>(ctx, p + "v", d.NL), *s = pool_vec<cx<R>>(ctx, p + "s", d.NL);
    cx<R> *z = pool_vec<cx<R>>(ctx, p + "z", d.NL), *t = pool_vec<cx<R>>(ctx, p + "t", d.NL);
    SliceLaunch L = edge_launch(d);
    const int64_t nbe = L.nblocks();
    const double vecb = (double)d.NL * 2.0 * sizeof(R);
    const double b_spmv = 2 * vecb + (double)d.NC * 6 * sizeof(R), b_ssor = 2 * (b_spmv + vecb);
    SolveInfo info;
    if (ko.dc_at_start) {
        info.pcg_iters += div_correct<R>(ctx, D, tag, x, ko.corr_tol);
        info.dc_calls++;
    }
    int state = 0;  // 0 fresh, 1 residual refresh keeping p, v and the scalars, 2 iterate
    // With a complex64 preconditioner the recursive residual drifts and the
    // iteration can stall: as in cocr, restart from the true residual after
    // each restart_drop reduction, on stagnation (200 iterations without a new
    // best) or divergence (100x above the best), accept convergence only on a
    // recomputed residual, and after 3 unproductive restarts continue with the
    // complex128 preconditioner.
    bool fresh = false;
    double rel0 = 0, best = 1e300, best_at_restart = 1e300;
    int since_best = 0, idle_restarts = 0;
    while (true) {
        if (state != 2) {
            launch_spmv<R>(D.g, wmu, x, r, (R)-1, 0, nullptr, ctx->rb, ctx->st);
            if (f) launch_axpby<R>(d.NL, cx<R>(1), f, cx<R>(1), r, ctx->st);
            launch_copy<R>(d.NL, r, rh, ctx->st);
            if (state == 0) {
                SFDX_CHECK(hipMemsetAsync(pv, 0, d.NL * sizeof(cx<R>), ctx->st));
                SFDX_CHECK(hipMemsetAsync(v, 0, d.NL * sizeof(cx<R>), ctx->st));
                reduce_final(ctx->rb, 0, 0, SL_TMP, OP_BICG_INIT, ctx->st);
            }
            launch_vec<R>(5, D.g, ctx->rb, r, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
            reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
            read_slots(ctx);
            state = 2;
            fresh = true;
        }
        info.relres = std::sqrt(ctx->hslots[SL_RN]) / bnorm;
        if (!(info.relres == info.relres) || std::isinf(info.relres)) { info.status = SFDX_ERR_NAN; break; }
        if (ko.verbose > 1 && info.iters % 50 == 0) fprintf(stderr, "  bicgstab it %d relres %.3e\n", info.iters, info.relres);
        const bool lo32 = ko.lo || ko.mg32;
        const bool fresh_now = fresh;
        if (fresh) { rel0 = info.relres; best = info.relres; since_best = 0; fresh = false; }
        if (info.relres < best) { best = info.relres; since_best = 0; }
        // accept convergence (and report the final residual) only on the true
        // residual b - A x, recomputed from the iterate
        if (info.relres < eps) {
            if (!fresh_now) { state = 0; continue; }
            info.status = 0;
            break;
        }
        if (info.iters >= maxit) {
            if (!fresh_now) { state = 0; continue; }
            info.status = SFDX_WARN_NOCONV;
            break;
        }
        if (lo32) {
            const bool stalled = since_best > 200, diverging = info.relres > 100 * best;
            const bool planned = ko.restart_drop > 0 && info.relres < ko.restart_drop * rel0;
            if (stalled || diverging || planned) {
                if (!planned) {
                    idle_restarts = (best < 0.5 * best_at_restart) ? 0 : idle_restarts + 1;
                    if (idle_restarts >= 3) {
                        ko.lo = nullptr;
                        ko.mg32 = nullptr;
                        ko.restart_drop = 0;
                        info.fallback = 1;
                        idle_restarts = 0;
                    }
                }
                best_at_restart = std::min(best_at_restart, best);
                state = 0;
                info.restarts++;
                continue;
            }
        }
        since_best++;
        if (ctx->hslots[SL_BREAK] != 0) { info.status = SFDX_ERR_BREAKDOWN; break; }
        if (ko.corr_intv > 0 && info.iters > 0 && info.iters % ko.corr_intv == 0 &&
            info.dc_calls < info.iters / ko.corr_intv + (ko.dc_at_start ? 1 : 0)) {
            info.pcg_iters += div_correct<R>(ctx, D, tag, x, ko.corr_tol);
            info.dc_calls++;
            state = ko.dc_restart ? 0 : 1;
            continue;
        }
        int64_t nbd;
        launch_dot<R>(1, d.NL, rh, r, ctx->rb, ctx->st, nbd);
        reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_BICG_RHO, ctx->st);
        launch_vec<R>(4, D.g, ctx->rb, pv, r, v, nullptr, nullptr, nullptr, ctx->st);
        apply_prec<R>(ctx, D, tag, wmu, pv, y, ko);
        int64_t nbs = launch_spmv<R>(D.g, wmu, y, v, (R)1, 2, rh, ctx->rb, ctx->st);
        reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_BICG_ALPHA, ctx->st);
        launch_vec<R>(1, D.g, ctx->rb, s, r, v, nullptr, nullptr, nullptr, ctx->st);
        apply_prec<R>(ctx, D, tag, wmu, s, z, ko);
        nbs = launch_spmv<R>(D.g, wmu, z, t, (R)1, 3, s, ctx->rb, ctx->st);
        reduce_final(ctx->rb, nbs, 3, SL_TMP, OP_BICG_OMEGA, ctx->st);
        launch_vec<R>(2, D.g, ctx->rb, x, y, z, r, s, t, ctx->st);
        reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
        read_slots(ctx);
        info.iters++;
        info.bytes += 2 * b_spmv + 2 * b_ssor + 16 * vecb;
    }
    return info;
}

// Right-preconditioned GPBiCG (pgpbicg, src/solvers.f90:137-315; Zhang 1997) on
// A_s with conjugated inner products as the reference's dot_product.  All
// scalars stay on the device; the host reads the residual once per iteration.
// A divergence correction every corr_intv iterations is followed by a full
// restart (the reference resets r and r0 only, src/solvers.f90:223-235).
template <typename R>
SolveInfo gpbicg(sfdx_ctx *ctx, LevelDev<R> &D, const std::string &tag, R wmu, const cx<R> *f, cx<R> *x,
                 double bnorm, double eps, int maxit, const KOpts &ko_in)
{
    KOpts ko = ko_in;  // the complex64 preconditioner may be dropped below
    const Dims &d = D.g.d;
    const std::string pf = tag + (sizeof(R) == 8 ? "/g64" : "/g32");
    GPVecs<R> v;
    cx<R> **all[] = {&v.p, &v.rp, &v.u, &v.w, &v.at, &v.ap, &v.y, &v.t, &v.r, &v.q, &v.tp, &v.tmp, &v.z, &v.r0};
    for (int q = 0; q < 14; q++) *all[q] = pool_vec<cx<R>>(ctx, pf + std::to_string(q), d.NL);
    v.x = x;
    SliceLaunch L = edge_launch(d);
    const int64_t nbe = L.nblocks();
    const double vecb = (double)d.NL * 2.0 * sizeof(R);
    const double b_spmv = 2 * vecb + (double)d.NC * 6 * sizeof(R), b_ssor = 2 * (b_spmv + vecb);
    SolveInfo info;
    if (ko.dc_at_start) {
        info.pcg_iters += div_correct<R>(ctx, D, tag, x, ko.corr_tol);
        info.dc_calls++;
    }
    bool start = true;
    int step = 0;  // iterations since the last (re)start
    // complex64 preconditioner: the same restart policy as bicgstab and cocr
    bool fresh = false;
    double rel0 = 0, best = 1e300, best_at_restart = 1e300;
    int since_best = 0, idle_restarts = 0;
    while (true) {
        if (start) {
            launch_spmv<R>(D.g, wmu, x, v.r, (R)-1, 0, nullptr, ctx->rb, ctx->st);
            if (f) launch_axpby<R>(d.NL, cx<R>(1), f, cx<R>(1), v.r, ctx->st);
            launch_copy<R>(d.NL, v.r, v.r0, ctx->st);
            cx<R> *zero[] = {v.p, v.u, v.w, v.at, v.ap, v.t, v.tp, v.z};
            for (cx<R> *zv : zero) SFDX_CHECK(hipMemsetAsync(zv, 0, d.NL * sizeof(cx<R>), ctx->st));
            int64_t nbd;
            launch_dot<R>(1, d.NL, v.r0, v.r, ctx->rb, ctx->st, nbd);
            reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_GP_INIT, ctx->st);
            launch_vec<R>(5, D.g, ctx->rb, v.r, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
            reduce_final(ctx->rb, nbe, 1, SL_TMP, OP_RN, ctx->st);
            read_slots(ctx);
            start = false;
            step = 0;
            info.relres = std::sqrt(ctx->hslots[SL_RN]) / bnorm;
            if (info.relres >= eps) apply_prec<R>(ctx, D, tag, wmu, v.r, v.rp, ko);
            fresh = true;
        } else {
            info.relres = std::sqrt(ctx->hslots[SL_RN]) / bnorm;
        }
        if (!(info.relres == info.relres) || std::isinf(info.relres)) { info.status = SFDX_ERR_NAN; break; }
        if (ko.verbose > 1 && info.iters % 50 == 0) fprintf(stderr, "  gpbicg it %d relres %.3e\n", info.iters, info.relres);
        const bool lo32 = ko.lo || ko.mg32;
        const bool fresh_now = fresh;
        if (fresh) { rel0 = info.relres; best = info.relres; since_best = 0; fresh = false; }
        if (info.relres < best) { best = info.relres; since_best = 0; }
        // accept convergence (and report the final residual) only on the true
        // residual b - A x, recomputed from the iterate
        if (info.relres < eps) {
            if (!fresh_now) { start = true; continue; }
            info.status = 0;
            break;
        }
        if (info.iters >= maxit) {
            if (!fresh_now) { start = true; continue; }
            info.status = SFDX_WARN_NOCONV;
            break;
        }
        if (lo32) {
            const bool stalled = since_best > 200, diverging = info.relres > 100 * best;
            const bool planned = ko.restart_drop > 0 && info.relres < ko.restart_drop * rel0;
            if (stalled || diverging || planned) {
                if (!planned) {
                    idle_restarts = (best < 0.5 * best_at_restart) ? 0 : idle_restarts + 1;
                    if (idle_restarts >= 3) {
                        ko.lo = nullptr;
                        ko.mg32 = nullptr;
                        ko.restart_drop = 0;
                        info.fallback = 1;
                        idle_restarts = 0;
                    }
                }
                best_at_restart = std::min(best_at_restart, best);
                start = true;
                info.restarts++;
                continue;
            }
        }
        since_best++;
        if (ctx->hslots[SL_BREAK] != 0) { info.status = SFDX_ERR_BREAKDOWN; break; }
        if (ko.corr_intv > 0 && info.iters > 0 && step > 0 && info.iters % ko.corr_intv == 0) {
            info.pcg_iters += div_correct<R>(ctx, D, tag, x, ko.corr_tol);
            info.dc_calls++;
            start = true;
            continue;
        }
        launch_gpbicg<R>(0, D.g, ctx->rb, v, ctx->st);                         // w, p
        const int64_t nbs = launch_spmv<R>(D.g, wmu, v.p, v.ap, (R)1, 2, v.r0, ctx->rb, ctx->st);  // ap, (r0,ap)
        reduce_final(ctx->rb, nbs, 2, SL_TMP, OP_BICG_ALPHA, ctx->st);  // alpha = rho/(r0,ap)
        launch_gpbicg<R>(1, D.g, ctx->rb, v, ctx->st);                         // y, t
        apply_prec<R>(ctx, D, tag, wmu, v.ap, v.tmp, ko);                      // tmp = M ap
        launch_gpbicg<R>(2, D.g, ctx->rb, v, ctx->st);                         // q, tp
        launch_spmv<R>(D.g, wmu, v.tp, v.at, (R)1, 0, nullptr, ctx->rb, ctx->st);
        launch_gpbicg<R>(3, D.g, ctx->rb, v, ctx->st);                         // 5 inner products
        reduce_final(ctx->rb, nbe, 8, SL_TMP, step == 0 ? OP_GP_OMEGA1 : OP_GP_OMEGA, ctx->st);
        launch_gpbicg<R>(4, D.g, ctx->rb, v, ctx->st);                         // u, z, x, r
        reduce_final(ctx->rb, nbe, 3, SL_TMP, OP_GP_BETA, ctx->st);
        read_slots(ctx);
        info.iters++;
        step++;
        info.bytes += 2 * b_spmv + 2 * b_ssor + 30 * vecb;
        const double rel = std::sqrt(ctx->hslots[SL_RN]) / bnorm;
        if (rel >= eps && ctx->hslots[SL_BREAK] == 0) apply_prec<R>(ctx, D, tag, wmu, v.r, v.rp, ko);
    }
    return info;
}

template <typename R>
SolveInfo krylov(sfdx_ctx *ctx, int solver, LevelDev<R> &D, const std::string &tag, R wmu, const cx<R> *f, cx<R> *x,
                 double bnorm, double eps, int maxit, const KOpts &ko)
{
    if (solver == SFDX_SOLVER_COCR) {
        SolveInfo I = cocr<R>(ctx, D, tag, wmu, f, x, bnorm, eps, maxit, ko);
        if (I.status == SFDX_WARN_NOCONV && I.stalled && I.iters < maxit) {
            // COCR's short recurrences can stagnate (measured: relres 1.35e-6 at
            // 25 Hz on the 128x128x96 anisotropic model, restarts do not help):
            // continue from the current iterate with GPBiCG and the complex128
            // preconditioner for the remaining iteration budget
            KOpts k2 = ko;
            k2.lo = nullptr;
            k2.mg32 = nullptr;
            k2.restart_drop = 0;
            k2.dc_at_start = 0;
            if (k2.precond == SFDX_PRECOND_MG) k2.precond = SFDX_PRECOND_HYBRID;
            SolveInfo J = gpbicg<R>(ctx, D, tag, wmu, f, x, bnorm, eps, maxit - I.iters, k2);
            J.iters += I.iters;
            J.restarts += I.restarts;
            J.dc_calls += I.dc_calls;
            J.pcg_iters += I.pcg_iters;
            J.bytes += I.bytes;
            J.fallback = std::max(J.fallback, I.fallback);
            J.switched = 1;
            return J;
        }
        return I;
    }
    if (solver == SFDX_SOLVER_GPBICG) return gpbicg<R>(ctx, D, tag, wmu, f, x, bnorm, eps, maxit, ko);
    return bicgstab<R>(ctx, D, tag, wmu, f, x, bnorm, eps, maxit, ko);
}

KOpts kopts_of(const sfdx_opts &o)
{
    KOpts k;
    k.corr_intv = o.corr_intv;
    k.corr_tol = o.corr_tol;
    k.omega = o.ssor_omega;
    k.verbose = o.verbose;
    k.dc_restart = o.dc_restart;
    k.dc_at_start = o.dc_at_start;
    k.precond = o.precond;
    k.use_graph = o.use_graph;
    k.nodal = o.nodal_smoother;
    return k;
}

// Right-hand-side norm of a level: ||b_int|| (unscaled, = ||A_ib g / ve||) and
// ||g||, for the chosen residual convention.
double rhs_norm(sfdx_ctx *ctx, LevelDev<double> &D, const std::string &tag, const zd *edc, int pol, int norm_mode,
                double wmu, int full)
{
    const Dims &d = D.g.d;
    zd *g = pool_vec<zd>(ctx, tag + "/bn_g", d.NL), *r = pool_vec<zd>(ctx, tag + "/bn_r", d.NL);
    launch_dirichlet<double>(d, edc, pol, g, 1, ctx->st, full);
    launch_spmv<double>(D.g, wmu, g, r, -1.0, 0, nullptr, ctx->rb, ctx->st);
    SliceLaunch L = edge_launch(d);
    launch_vec<double>(5, D.g, ctx->rb, r, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
    reduce_final(ctx->rb, L.nblocks(), 1, SL_TMP, OP_RN, ctx->st);
    read_slots(ctx);
    double bi2 = ctx->hslots[SL_RN];
    if (norm_mode == SFDX_NORM_INTERIOR) return std::sqrt(bi2);
    int64_t nbd;
    launch_dot<double>(1, d.NL, g, g, ctx->rb, ctx->st, nbd);
    reduce_final(ctx->rb, nbd, 2, SL_TMP, OP_NONE, ctx->st);
    read_slots(ctx);
    return std::sqrt(bi2 + ctx->hslots[SL_TMP]);
}

struct LevelSolveResult {
    SolveInfo info;
    int refine = 0;
};

// Solve one level for one polarisation; x (FP64) holds the initial guess on the
// interior and receives the Dirichlet data first.
LevelSolveResult solve_level(sfdx_ctx *ctx, const sfdx_opts &o, LevelDev<double> &D64, LevelDev<float> *D32,
                             const std::string &tag, double freq, const zd *edc, int pol, zd *x, double eps, int maxit,
                             MGHier<double> *h64, MGHier<float> *h32, int level)
{
    const Dims &d = D64.g.d;
    const double wmu = 2 * 3.141592653589793 * freq * kMu0;
    launch_dirichlet<double>(d, edc, pol, x, 0, ctx->st, o.full_background);
    const double bnorm = rhs_norm(ctx, D64, tag, edc, pol, o.norm, wmu, o.full_background);
    KOpts ko = kopts_of(o);
    ko.mg64 = h64;
    ko.mg32 = h32;
    ko.mg_level = level;
    LevelSolveResult res;
    if (o.precision == SFDX_PREC_FP64 || !D32) {
        res.info = krylov<double>(ctx, o.solver, D64, tag, wmu, nullptr, x, bnorm, eps, maxit, ko);
        return res;
    }
    if (o.precision == SFDX_PREC_MIXED_SPLIT && h32 && ko.precond == SFDX_PRECOND_MG &&
        o.solver == SFDX_SOLVER_COCR) {
        KOpts ks = ko;
        ks.restart_drop = o.restart_drop;
        res.info = cocr_split(ctx, D64, *D32, *h32, level, tag, wmu, x, bnorm, eps, maxit, ks);
        res.refine = res.info.restarts;
        if (res.info.status != 0 && res.info.fallback && res.info.iters < maxit) {
            // continue from the iterate with complex128 directions (SFDX_PREC_MIXED)
            KOpts km = ko;
            km.lo = D32;
            km.restart_drop = o.restart_drop;
            const SolveInfo s2 =
                krylov<double>(ctx, o.solver, D64, tag, wmu, nullptr, x, bnorm, eps, maxit - res.info.iters, km);
            res.info.iters += s2.iters;
            res.info.restarts += s2.restarts;
            res.info.bytes += s2.bytes;
            res.info.status = s2.status;
            res.info.relres = s2.relres;
            res.info.switched += s2.switched;
            res.refine = res.info.restarts;
        }
        return res;
    }
    if (o.precision == SFDX_PREC_MIXED || o.precision == SFDX_PREC_MIXED_SPLIT) {
        KOpts km = ko;
        km.lo = D32;
        km.restart_drop = o.restart_drop;
        res.info = krylov<double>(ctx, o.solver, D64, tag, wmu, nullptr, x, bnorm, eps, maxit, km);
        res.refine = res.info.restarts;
        return res;
    }
    // Mixed precision: FP64 residual, divergence correction of the iterate and
    // update; complex64 inner solves of A_s d = r whose own corrections keep
    // G^T M d = 0 (valid because the FP64 iterate is corrected first).
    zd *r64 = pool_vec<zd>(ctx, tag + "/ir_r", d.NL);
    zf *r32 = pool_vec<zf>(ctx, tag + "/ir_r32", d.NL), *d32 = pool_vec<zf>(ctx, tag + "/ir_d32", d.NL);
    SliceLaunch L = edge_launch(d);
    int total_iters = 0, dc_calls = 0, pcg_iters = 0;
    double bytes = 0;
    KOpts kin = ko;
    kin.dc_at_start = 0;
    for (int step = 0;; step++) {
        if (o.corr_intv > 0 && (step > 0 || o.dc_at_start)) {
            pcg_iters += div_correct<double>(ctx, D64, tag, x, o.corr_tol);
            dc_calls++;
        }
        launch_spmv<double>(D64.g, wmu, x, r64, -1.0, 0, nullptr, ctx->rb, ctx->st);
        launch_vec<double>(5, D64.g, ctx->rb, r64, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->st);
        reduce_final(ctx->rb, L.nblocks(), 1, SL_TMP, OP_RN, ctx->st);
        read_slots(ctx);
        const double rel = std::sqrt(ctx->hslots[SL_RN]) / bnorm;
        bytes += (double)d.NL * 16 * 3 + (double)d.NC * 48;
        res.info.relres = rel;
        if (!(rel == rel)) { res.info.status = SFDX_ERR_NAN; break; }
        if (rel < eps) { res.info.status = 0; break; }
        if (step >= o.max_refine || total_iters >= maxit) { res.info.status = SFDX_WARN_NOCONV; break; }
        launch_convert<double, float>(d.NL, r64, r32, ctx->st);
        SFDX_CHECK(hipMemsetAsync(d32, 0, d.NL * sizeof(zf), ctx->st));
        const double rn = std::sqrt(ctx->hslots[SL_RN]);
        const double inner = std::max(o.inner_tol, 0.5 * eps * bnorm / rn);
        const int cap = std::min(maxit - total_iters, o.inner_maxit > 0 ? o.inner_maxit : maxit);
        SolveInfo si = krylov<float>(ctx, o.solver, *D32, tag, (float)wmu, r32, d32, rn, inner, cap, kin);
        total_iters += si.iters;
        dc_calls += si.dc_calls;
        pcg_iters += si.pcg_iters;
        bytes += si.bytes;
        if (si.status > 0 && si.status != SFDX_ERR_BREAKDOWN) { res.info.status = si.status; break; }
        launch_add_lo(d.NL, x, d32, ctx->st);
        res.refine = step + 1;
        if (o.verbose) fprintf(stderr, "  refine %d: outer %.3e inner its %d (%.2e)\n", step, rel, si.iters, si.relres);
    }
    res.info.iters = total_iters;
    res.info.dc_calls = dc_calls;
    res.info.pcg_iters = pcg_iters;
    res.info.bytes = bytes;
    return res;
}

// prolongation tables of one fine level on the device
ProlongDev upload_prolong(sfdx_ctx *ctx, const HostLevel &fine, int m, const std::string &tag)
{
    ProlongHost H;
    prolong_tables(fine, m, H);
    ProlongDev T;
    for (int q = 0; q < 3; q++) {
        const std::string p = tag + "/pt" + std::to_string(m) + std::to_string(q);
        T.nb[q] = pool_vec<int>(ctx, p + "nb", H.nb[q].size());
        T.eb[q] = pool_vec<int>(ctx, p + "eb", H.eb[q].size());
        T.nw[q] = pool_vec<double>(ctx, p + "nw", H.nw[q].size());
        T.ew[q] = pool_vec<double>(ctx, p + "ew", H.ew[q].size());
        SFDX_CHECK(hipMemcpyAsync(T.nb[q], H.nb[q].data(), H.nb[q].size() * sizeof(int), hipMemcpyHostToDevice, ctx->st));
        SFDX_CHECK(hipMemcpyAsync(T.eb[q], H.eb[q].data(), H.eb[q].size() * sizeof(int), hipMemcpyHostToDevice, ctx->st));
        SFDX_CHECK(hipMemcpyAsync(T.nw[q], H.nw[q].data(), H.nw[q].size() * 8, hipMemcpyHostToDevice, ctx->st));
        SFDX_CHECK(hipMemcpyAsync(T.ew[q], H.ew[q].data(), H.ew[q].size() * 8, hipMemcpyHostToDevice, ctx->st));
    }
    SFDX_CHECK(hipStreamSynchronize(ctx->st));
    return T;
}

// extint_v on the device: out(level i) = I4[(1+1/4) u1 - 1/4 I2 u0]
void extint_device(sfdx_ctx *ctx, const HostLevel &fine, const HostLevel &mid, const zd *u0, const zd *u1, zd *out,
                   const std::string &tag)
{
    const Dims df = make_dims(fine.nx, fine.ny, fine.nz), dm = make_dims(mid.nx, mid.ny, mid.nz);
    const Dims dcs = make_dims(mid.nx / 2, mid.ny / 2, mid.nz / 2);
    ProlongDev T2 = upload_prolong(ctx, mid, 2, tag + "/mid");
    ProlongDev T4 = upload_prolong(ctx, fine, 4, tag + "/fine");
    zd *w = pool_vec<zd>(ctx, tag + "/extw", dm.NL);
    launch_prolong<double>(dm, dcs, 2, T2, u0, w, ctx->st);
    launch_extrap<double>(dm.NL, 1.25, u1, 0.25, w, ctx->st);
    launch_prolong<double>(df, dm, 4, T4, w, out, ctx->st);
}

bool mesh_ok(const HostLevel &F, int L)
{
    const int f = 1 << L;
    return F.nx % f == 0 && F.ny % f == 0 && F.nz % f == 0 && F.nair % (1 << (L - 1)) == 0 && F.nx / f >= 1;
}

// Smallest cell width of a grid.
double min_width(const HostLevel &L)
{
    double h = 1e300;
    for (const std::vector<double> *w : {&L.a, &L.b, &L.c}) h = std::min(h, *std::min_element(w->begin(), w->end()));
    return h;
}

// V-cycle grids of the merge-ratio coarsening over one level (mg_merge_ratio
// > 0): grid L is the level itself (g[0]); each coarser grid merges two
// consecutive cells of an axis only where both are at most mg_merge_ratio
// times the grid's smallest width, and never across the air-earth interface.
// Coarsening ends at a grid of at most kMergeMinCells cells, when no axis
// merges or an axis would keep fewer than 2 cells, or at mg_levels grids (0:
// kMaxMergeGrids).  On stretched meshes the long cells of the padding stay
// unmerged until the short ones have grown to their size, so every coarse
// grid represents the error that point smoothing leaves along their long
// sides.  A short cell that cannot merge (next to the interface, between
// long cells) ends the coarsening at a larger coarsest grid; that grid gets
// one smoothing pair per kMergeMinCells cells (at least mg_coarse_pairs).
// Letting the reference width grow past such a cell instead (doubling it per
// grid) merged long cells against the anisotropy: on a 192x192x144 model with
// 54 km padding cells the fine level then took 462/492 iterations, against
// 127/118 with the coarsening stopped at 27x27x43 and 16 pairs there.
constexpr int kMaxMergeGrids = 12;
constexpr int64_t kMergeMinCells = 2048;
struct MergeHier {
    std::vector<HostLevel> g;  // g[n]: n coarsenings below the level
    std::vector<Starts> st;    // st[n]: starts of g[n + 1] over g[n]
    std::vector<LevelDev<double>> X64 = std::vector<LevelDev<double>>(kMaxMergeGrids + 1);  // n >= 1
    std::vector<LevelDev<float>> X32 = std::vector<LevelDev<float>>(kMaxMergeGrids + 1);
    MGHier<double> H64;
    MGHier<float> H32;
    int L = 0;

    // D32: the level's complex64 copy when its V-cycle runs in complex64
    void build(sfdx_ctx *ctx, const HostLevel &top, LevelDev<double> &D64, LevelDev<float> *D32, const sfdx_opts *o,
               const std::string &tag)
    {
        g.assign(1, top);
        const int cap = o->mg_levels > 0 ? std::min(o->mg_levels, kMaxMergeGrids) : kMaxMergeGrids;
        while ((int)g.size() < cap && (int64_t)g.back().ncell() > kMergeMinCells) {
            const HostLevel &G = g.back();
            const double thr = o->mg_merge_ratio * min_width(G);
            Starts s = {merge_starts(G.a, thr, -1), merge_starts(G.b, thr, -1), merge_starts(G.c, thr, G.nair)};
            if (s[0].size() == G.a.size() + 1 && s[1].size() == G.b.size() + 1 && s[2].size() == G.c.size() + 1) break;
            if (s[0].size() < 3 || s[1].size() < 3 || s[2].size() < 3) break;
            const int n = (int)g.size();
            const std::string t = tag + "_" + std::to_string(n);
            double *sd = nullptr;  // the grid's sigma stays on the device
            HostLevel C = coarsen_by_starts(ctx, G, n == 1 ? D64 : X64[n - 1], s, "crs" + t, &sd, false);
            g.push_back(std::move(C));
            st.push_back(s);
            X64[n] = level_from_device<double>(ctx, g.back(), sd, t);
            if (D32) X32[n] = level_from_device<float>(ctx, g.back(), sd, t, o->sigma_floor32);
        }
        L = (int)g.size();
        H64.name = H32.name = "fm" + tag + "_";
        H64.lev.assign(L + 1, nullptr);
        H32.lev.assign(L + 1, nullptr);
        H64.tr.assign(L + 1, MGTables{});
        const int pairs = o->mg_coarse_pairs > 0 ? o->mg_coarse_pairs : 4;
        const int64_t nc = (int64_t)g.back().ncell();
        H64.coarse_pairs = H32.coarse_pairs = std::max<int>(pairs, (int)((nc + kMergeMinCells - 1) / kMergeMinCells));
        H64.nodal = H32.nodal = o->nodal_smoother;
        for (int n = 0; n < L; n++) {
            H64.lev[L - n] = n == 0 ? &D64 : &X64[n];
            if (D32) H32.lev[L - n] = n == 0 ? D32 : &X32[n];
            if (n + 1 < L) H64.tr[L - n] = mg_tables(ctx, g[n], g[n + 1], st[n], "fmt" + tag + "_" + std::to_string(n));
            if (o->verbose)
                fprintf(stderr, "%s merge-ratio grid %d: %dx%dx%d\n", tag.c_str(), L - n, g[n].nx, g[n].ny, g[n].nz);
        }
        H32.tr = H64.tr;
    }
};

// Grids of the V-cycle hierarchy over a fine grid (the fine grid included):
// halve while every dimension divides and keeps at least 2 cells, capped by
// mg_levels when that is set.
int mg_grid_count(const HostLevel &F, const sfdx_opts *o)
{
    int m = 1;
    while (m < SFDX_MAX_LEVELS && F.nx % (1 << m) == 0 && F.ny % (1 << m) == 0 && F.nz % (1 << m) == 0 &&
           std::min(std::min(F.nx, F.ny), F.nz) >> m >= 2)
        m++;
    return o->mg_levels > 0 ? std::min(o->mg_levels, m) : m;
}

// SFDX_PRECOND_AUTO: the preconditioner of each frequency's solves.
// With the merge-ratio V-cycle (mg_merge_ratio > 0) the V-cycle on each
// Krylov level unless the skin depth in the most conductive earth cell is
// below half the level's smallest width: that level's grid under-resolves the
// field, and there the complex64 V-cycle converged slowly or not at all (a
// 32x32x24 block model at 1 kHz to 1e-10) while the hybrid preconditioner
// needs fewer and cheaper iterations on the coarser levels (the 128x128x96
// block model at 1 kHz, levels 1-3: 164-984 iterations against 757-1253 with
// the V-cycle; 16 sounding frequencies, whole solves 11.0 s against 13.2 s
// with the V-cycle on every level).  With the 2:1 V-cycle (mg_merge_ratio = 0) the V-cycle
// only where
//  (a) its first coarse grid resolves that skin depth, delta(sigma_max, f) >=
//      the smallest cell width of that grid (twice the fine grid's); below it
//      the rediscretised coarse operators no longer describe the fine grid's
//      diffusion and the V-cycle stops converging (on the 128x128x96 block
//      model from ~150 Hz up);
//  (b) the field is not quasi-static over the model: delta at the median earth
//      conductivity is at most the larger horizontal extent of the mesh; in
//      that regime (the 128x128x96 block model at 1-2.5 mHz) the complex64
//      V-cycle was measured to stall COCR's recurrences between restarts;
//  (c) no cell is more than kAutoMaxAspect times longer than its shortest
//      side: the 2:1 grids then cannot represent the error that point
//      smoothing leaves along the long side of a stretched cell; on a 96^3
//      model with 354:1 cells it took 1.8x the iterations of the hybrid
//      preconditioner, on the 128x128x96 model (133:1) 2.9x fewer;
// and the one-level hybrid preconditioner elsewhere.  sigma_max and the median
// are taken over the diagonal tensor components of the cells below the air
// layers.
constexpr double kAutoMaxAspect = 200.0;


struct AutoPrecond {
    double sig_max = 0, sig_med = 0, extent = 0, h_min = 0, h_coarse = 0, aspect = 0;
    bool usable = false;  // the fine grid coarsens at least once
    bool merge = false;   // merge-ratio grids
    AutoPrecond(const HostLevel &F, int grids, bool merge_grids)
    {
        usable = grids >= 2;
        merge = merge_grids;
        // largest cell aspect ratio: a cell takes one width per axis, so it is
        // the largest width of one axis over the smallest of another
        const std::vector<double> *w[3] = {&F.a, &F.b, &F.c};
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++)
                if (p != q)
                    aspect = std::max(aspect, *std::max_element(w[p]->begin(), w[p]->end()) /
                                                  *std::min_element(w[q]->begin(), w[q]->end()));
        const int64_t nc = F.ncell(), nce = (int64_t)F.nx * F.ny * (F.nz - F.nair);
        std::vector<double> mean(std::max<int64_t>(nce, 1), 0.0);
        for (int64_t q = 0; q < nce; q++) {
            const int64_t ie = (int64_t)F.nx * F.ny * F.nair + q;
            const double sx = F.S()[ie], sy = F.S()[nc + ie], sz = F.S()[2 * nc + ie];
            sig_max = std::max(sig_max, std::max(sx, std::max(sy, sz)));
            mean[q] = (sx + sy + sz) / 3;
        }
        std::nth_element(mean.begin(), mean.begin() + mean.size() / 2, mean.end());
        sig_med = mean[mean.size() / 2];
        double wa = 0, wb = 0;
        for (double v : F.a) wa += v;
        for (double v : F.b) wb += v;
        extent = std::max(wa, wb);
        // smallest width of the fine grid and of its first 2:1 coarse grid
        h_min = h_coarse = 1e300;
        for (const std::vector<double> *v : {&F.a, &F.b, &F.c}) {
            for (size_t s = 0; s < v->size(); s++) h_min = std::min(h_min, (*v)[s]);
            for (size_t s = 0; s + 2 <= v->size(); s += 2) h_coarse = std::min(h_coarse, (*v)[s] + (*v)[s + 1]);
        }
    }
    static double skin_depth(double sigma, double f) { return std::sqrt(2.0 / (2 * M_PI * f * 4e-7 * M_PI * sigma)); }
    // merge-ratio grids: the V-cycle on a level of smallest width h
    bool use_mg_on(double f, double h) const { return usable && sig_max > 0 && skin_depth(sig_max, f) >= 0.5 * h; }
    bool use_mg(double f) const
    {
        if (!usable || !(sig_max > 0)) return false;
        if (merge) return use_mg_on(f, h_min);
        return aspect <= kAutoMaxAspect && sig_med > 0 && skin_depth(sig_max, f) >= h_coarse &&
               skin_depth(sig_med, f) <= extent;
    }
};

HostLevel host_level(const sfdx_grid *g, const double *sigma)
{
    HostLevel L;
    L.nx = g->nx; L.ny = g->ny; L.nz = g->nz; L.nair = g->nair;
    L.a.assign(g->a, g->a + g->nx);
    L.b.assign(g->b, g->b + g->ny);
    L.c.assign(g->c, g->c + g->nz);
    L.ext = sigma;  // read in place for the duration of the call
    return L;
}

// ex (optional): fields 2*nl*nf; surf (optional): surface fields of each
// frequency evaluated on the device, nf*2*5*nx*ny (see post_process_surface)
int fwd_solve_impl(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq,
                   const sfdx_opts *o, double *ex, sfdx_stats *st, zc *surf = nullptr)
{
    Timer total;
    sfdx_stats S;
    std::memset(&S, 0, sizeof(S));
    const HostLevel fine = host_level(g, sigma);
    double probe = 0;  // NaN if any entry is NaN or infinite (a branch-free loop)
    for (size_t q = 0; q < 6 * fine.ncell(); q++) probe += fine.S()[q] * 0.0;
    if (probe != 0) { g_err = "non-finite sigma"; return SFDX_ERR_NAN; }
    if (o->precond < SFDX_PRECOND_SSOR || o->precond > SFDX_PRECOND_AUTO) {
        g_err = "precond must be 0..3";
        return SFDX_ERR_ARG;
    }
    // the V-cycle per frequency (SFDX_PRECOND_MG: all; SFDX_PRECOND_AUTO: by
    // the skin-depth rule); its grids are built when any frequency uses it
    const bool auto_pc = o->precond == SFDX_PRECOND_AUTO;
    std::vector<char> mg_freq(nf, o->precond == SFDX_PRECOND_MG);
    std::unique_ptr<AutoPrecond> rule_p;
    if (auto_pc) rule_p.reset(new AutoPrecond(fine, mg_grid_count(fine, o), o->mg_merge_ratio > 0));
    if (auto_pc) {
        const AutoPrecond &rule = *rule_p;
        for (int fi = 0; fi < nf; fi++) {
            mg_freq[fi] = rule.use_mg(freq[fi]);
            if (o->verbose)
                fprintf(stderr, "f=%g: %s (skin depth %.4g m at sigma_max %.4g S/m, smallest width %.4g m, first coarse "
                                "width %.4g m; %.4g m at median sigma %.4g S/m, extent %.4g m; cell aspect %.4g)\n",
                        freq[fi], mg_freq[fi] ? "multigrid V-cycle" : "hybrid", AutoPrecond::skin_depth(rule.sig_max, freq[fi]),
                        rule.sig_max, rule.h_min, rule.h_coarse, AutoPrecond::skin_depth(rule.sig_med, freq[fi]),
                        rule.sig_med, rule.extent, rule.aspect);
        }
    }
    const bool use_mg = std::find(mg_freq.begin(), mg_freq.end(), 1) != mg_freq.end();
    const int Lv = o->use_excmg ? o->level : 1;
    if (o->use_excmg && (Lv < 3 || Lv > SFDX_MAX_LEVELS || !mesh_ok(fine, Lv))) {
        g_err = "grid dimensions must be divisible by 2^level and nair by 2^(level-1)";
        return SFDX_ERR_MESH;
    }
    if (o->solver != SFDX_SOLVER_COCR && o->solver != SFDX_SOLVER_BICGSTAB && o->solver != SFDX_SOLVER_GPBICG) {
        g_err = "solver not supported on the device path";
        return SFDX_ERR_UNSUPPORTED;
    }
    if (o->spmv_tile < -1 || o->spmv_tile > 4) {
        g_err = "spmv_tile must be -1 (autotune) or 0..4";
        return SFDX_ERR_ARG;
    }
    // level hierarchy: lev[i] for i = 1..Lv (Lv = fine); the fine level is
    // uploaded first and the coarser ones are averaged from it on the device
    std::vector<HostLevel> lev(Lv + 1);
    lev[Lv] = fine;
    std::vector<LevelDev<double>> D64(Lv + 1);
    std::vector<LevelDev<float>> D32(Lv + 1);
    if (o->verbose) fprintf(stderr, "set-up: host copy and checks %.3fs\n", total.s());
    D64[Lv] = upload_level<double>(ctx, fine, "L" + std::to_string(Lv));
    if (o->verbose) fprintf(stderr, "set-up: fine level uploaded %.3fs\n", total.s());
    std::vector<double *> levsig(Lv + 1, nullptr);  // device sigma of the coarser levels
    for (int i = Lv - 1; i >= 1; i--)
        lev[i] = coarsen_device(ctx, fine, D64[Lv], Lv - i, "crs" + std::to_string(i), &levsig[i]);
    // levels solved by the Krylov method on the device: with coarse_solver = 1
    // also the two coarsest EXCMG levels (instead of the host direct solves)
    const bool coarse_dev = o->use_excmg && o->coarse_solver == 1;
    const int first_dev = o->use_excmg ? (coarse_dev ? 1 : 3) : Lv;
    const int first_up = use_mg ? 1 : first_dev;  // levels resident on the device
    // The complex64 preconditioner serves the fine level only.  The coarser
    // Krylov levels (8x and 64x fewer unknowns) take the complex128 one: with
    // complex64 the 64x64x48 level of the 128x128x96 model needed 1.7x the
    // iterations.  A complex64 V-cycle from the fine level still visits every
    // grid, so with multigrid all levels get their complex64 data.
    const bool lo_fine = o->precision != SFDX_PREC_FP64;
    const bool flex = use_mg && o->mg_merge_ratio > 0;
    if (flex && o->mg_merge_ratio < 1) {
        g_err = "mg_merge_ratio must be 0 or >= 1";
        return SFDX_ERR_ARG;
    }
    for (int i = first_up; i <= Lv; i++) {
        if (i < Lv) D64[i] = level_from_device<double>(ctx, lev[i], levsig[i], "L" + std::to_string(i));
        if (lo_fine && (i == Lv || (use_mg && !flex)))
            D32[i] = level_from_device<float>(ctx, lev[i], D64[i].sig64, "L" + std::to_string(i), o->sigma_floor32);
    }
    // Multigrid hierarchy mg[1..Lmg] (mg[Lmg] = fine) of 2:1 grids
    // (mg_merge_ratio = 0); the EXCMG level i is mg[i + off].  Grids coarser
    // than the EXCMG levels are uploaded here; with fewer grids than EXCMG
    // levels (off < 0) the V-cycle stops above the coarsest EXCMG level, and
    // EXCMG levels below its coarsest grid are solved with the hybrid
    // preconditioner.
    int Lmg = Lv;
    if (use_mg && !flex) Lmg = std::max(mg_grid_count(fine, o), 2);
    const int off = Lmg - Lv;
    std::vector<HostLevel> mgh(Lmg + 1);
    std::vector<LevelDev<double>> X64(std::max(off, 0) + 1);
    std::vector<LevelDev<float>> X32(std::max(off, 0) + 1);
    MGHier<double> H64;
    MGHier<float> H32;
    if (use_mg && !flex) {
        H64.lev.assign(Lmg + 1, nullptr);
        H32.lev.assign(Lmg + 1, nullptr);
        H64.tr.assign(Lmg + 1, MGTables{});
        H64.coarse_pairs = H32.coarse_pairs = o->mg_coarse_pairs > 0 ? o->mg_coarse_pairs : 4;
        H64.nodal = H32.nodal = o->nodal_smoother;
        for (int j = 1; j <= Lmg; j++) {
            if (j > off) {
                mgh[j] = lev[j - off];
                H64.lev[j] = &D64[j - off];
                if (o->precision != SFDX_PREC_FP64) H32.lev[j] = &D32[j - off];
            } else {
                double *sd = nullptr;
                mgh[j] = coarsen_device(ctx, fine, D64[Lv], Lmg - j, "crsm" + std::to_string(j), &sd);
                X64[j] = level_from_device<double>(ctx, mgh[j], sd, "M" + std::to_string(j));
                H64.lev[j] = &X64[j];
                if (o->precision != SFDX_PREC_FP64) {
                    X32[j] = level_from_device<float>(ctx, mgh[j], sd, "M" + std::to_string(j), o->sigma_floor32);
                    H32.lev[j] = &X32[j];
                }
            }
        }
        for (int j = 2; j <= Lmg; j++) {
            const Starts st = {uniform_starts(mgh[j].nx, 1), uniform_starts(mgh[j].ny, 1), uniform_starts(mgh[j].nz, 1)};
            H64.tr[j] = mg_tables(ctx, mgh[j], mgh[j - 1], st, "mgt" + std::to_string(j));
        }
        H32.tr = H64.tr;
    }
    // With mg_merge_ratio > 0 every Krylov level runs the V-cycle over grids of
    // its own (MergeHier): coarsenings that merge cells only where they are short.
    if (o->verbose) fprintf(stderr, "set-up: EXCMG levels %.3fs\n", total.s());
    std::vector<MergeHier> MH(flex ? Lv + 1 : 0);
    for (int i = flex ? first_dev : Lv + 1; i <= Lv; i++) {
        const bool lo_i = lo_fine && i == Lv;
        // a level that merges no cells gets a one-grid V-cycle (the smoother)
        MH[i].build(ctx, lev[i], D64[i], lo_i ? &D32[i] : nullptr, o, "F" + std::to_string(i));
        if (o->verbose) fprintf(stderr, "set-up: level %d grids %.3fs\n", i, total.s());
    }
    ensure_partials(ctx, edge_launch(make_dims(fine.nx, fine.ny, fine.nz)).nblocks() + 64);
    for (int i = first_up; i <= Lv; i++) {
        choose_tile(ctx, D64[i].g, "L" + std::to_string(i) + "/d", o->spmv_tile, o->verbose);
        if (lo_fine && i == Lv) choose_tile(ctx, D32[i].g, "L" + std::to_string(i) + "/f", o->spmv_tile, o->verbose);
    }
    // edge SSOR of the hybrid preconditioner and of the multigrid smoothers:
    // lumped mass only unless asked for the cross-orientation terms (the plain
    // SSOR preconditioner, the reference's, keeps the full operator)
    const bool xmass = o->precond == SFDX_PRECOND_SSOR || o->ssor_cross_mass != 0;
    for (int i = first_up; i <= Lv; i++) {
        D64[i].g.ssor_xmass = xmass;
        D32[i].g.ssor_xmass = xmass;
    }
    for (int j = 1; j <= off; j++) {
        X64[j].g.ssor_xmass = xmass;
        X32[j].g.ssor_xmass = xmass;
    }
    for (MergeHier &M : MH)
        for (int n = 1; n < M.L; n++) {
            M.X64[n].g.ssor_xmass = xmass;
            M.X32[n].g.ssor_xmass = xmass;
        }
    S.spmv_tile = D64[Lv].g.spmv_variant;
    const int64_t NLf = make_dims(fine.nx, fine.ny, fine.nz).NL;
    // surface node plane per column for the device post-process (topography
    // carved into sigma), or the flat datum
    int *planes_dev = nullptr;
    std::vector<int> planes;
    if (surf && o->surface_sigma > 0) {
        surface_planes(fine, o->surface_sigma, planes);
        for (int v : planes)
            if (v >= fine.nz) {
                g_err = "a cell column holds no cell above surface_sigma (no ground surface)";
                return SFDX_ERR_ARG;
            }
        planes_dev = pool_vec<int>(ctx, "surf_planes", planes.size());
        SFDX_CHECK(hipMemcpyAsync(planes_dev, planes.data(), planes.size() * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->st));
        SFDX_CHECK(hipStreamSynchronize(ctx->st));
    }
    int worst = 0;
    if (o->verbose) fprintf(stderr, "set-up (levels, grids, tables) %.3fs\n", total.s());
    for (int fi = 0; fi < nf; fi++) {
        const double ff = freq[fi];
        // device solution history: u[pol][level]
        std::vector<std::array<zd *, 2>> u(Lv + 1, {nullptr, nullptr});
        auto ubuf = [&](int i, int pol) {
            const Dims d = make_dims(lev[i].nx, lev[i].ny, lev[i].nz);
            return pool_vec<zd>(ctx, "u" + std::to_string(i) + "_" + std::to_string(pol), d.NL);
        };
        if (o->use_excmg && !coarse_dev) {
            // levels 1 and 2: direct solves (u1u2, src/fdfe_main.f90:1228-1256)
            for (int i = 1; i <= 2; i++) {
                const std::vector<zc> edc = background_field(lev[i], ff, o->full_background != 0);
                CoarseSystem CS;
                coarse_system(lev[i], ff, edc, CS);
                std::vector<zc> sol;
                if (!direct_solve(CS, sol, g_err)) return SFDX_ERR_DIRECT;
                const Dims d = make_dims(lev[i].nx, lev[i].ny, lev[i].nz);
                // both host copies stay alive until the stream has drained
                std::vector<zc> full[2];
                for (int pol = 0; pol < 2; pol++) {
                    full[pol].assign(d.NL, 0.0);
                    for (int64_t e = 0; e < d.NL; e++) {
                        int ii, jj, kk;
                        int oo = edge_decode(d, e, ii, jj, kk);
                        if (CS.imap[e] >= 0) full[pol][e] = sol[(size_t)pol * CS.n + CS.imap[e]];
                        else dirichlet_value(d, edc.data(), pol, oo, ii, jj, kk, full[pol][e], o->full_background != 0);
                    }
                    u[i][pol] = ubuf(i, pol);
                    SFDX_CHECK(hipMemcpyAsync(u[i][pol], full[pol].data(), d.NL * sizeof(zd), hipMemcpyHostToDevice,
                                              ctx->st));
                }
                SFDX_CHECK(hipStreamSynchronize(ctx->st));
            }
        }
        double tol = o->eps * std::pow(10.0, first_dev - 1 - Lv);  // (:1312, :1323)
        for (int i = first_dev; i <= Lv; i++) {
            tol *= 10;                                   // level tolerance eps*10^(i-L) (:1331)
            if (!o->use_excmg) tol = o->eps;
            // The two coarsest levels replace direct solves, which are exact; the
            // Krylov solves stop at 1e-12 at the most, below which their FP64
            // recurrences stall (a 5-level solve to eps 1e-10 asks 1e-14 of
            // level 1), and the EXCMG guess gains nothing from more digits.
            const double tol_i = (coarse_dev && i <= 2) ? std::max(tol, kCoarseTolFloor) : tol;
            const std::vector<zc> edc_h = background_field(lev[i], ff, o->full_background != 0);
            const Dims d = make_dims(lev[i].nx, lev[i].ny, lev[i].nz);
            zd *edc = pool_vec<zd>(ctx, "edc" + std::to_string(i), edc_h.size());
            SFDX_CHECK(hipMemcpyAsync(edc, edc_h.data(), edc_h.size() * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
            const std::string tag = "L" + std::to_string(i);
            // initial guesses of both polarisations on the context's stream
            zd *xs[2];
            for (int pol = 0; pol < 2; pol++) {
                xs[pol] = ubuf(i, pol);
                if (o->use_excmg && i >= 3)
                    extint_device(ctx, lev[i], lev[i - 1], u[i - 2][pol], u[i - 1][pol], xs[pol], tag);
                else SFDX_CHECK(hipMemsetAsync(xs[pol], 0, d.NL * sizeof(zd), ctx->st));
            }
            const int maxit = o->use_excmg ? o->maxit_excmg : o->maxit_plain;
            // levels below the multigrid hierarchy's coarsest grid: hybrid
            // with merge-ratio grids SFDX_PRECOND_AUTO decides per level: a coarser
            // EXCMG level whose own cells under-resolve the skin depth takes
            // the hybrid preconditioner
            const bool mg_here = mg_freq[fi] && (flex || i + off >= 1) &&
                                 !(auto_pc && flex && i < Lv && !rule_p->use_mg_on(ff, min_width(lev[i])));
            sfdx_opts oi = *o;
            if ((use_mg || auto_pc) && !mg_here) oi.precond = SFDX_PRECOND_HYBRID;
            if (auto_pc && mg_here) oi.precond = SFDX_PRECOND_MG;
            LevelDev<float> *d32 = (lo_fine && i == Lv) ? &D32[i] : nullptr;
            MGHier<double> *h64 = mg_here ? (flex ? &MH[i].H64 : &H64) : nullptr;
            MGHier<float> *h32 = (mg_here && lo_fine && i == Lv) ? (flex ? &MH[i].H32 : &H32) : nullptr;
            const int mgl = flex ? MH[i].L : i + off;
            const bool two = o->pol_streams == 2;
            if (two) prebuild_level(ctx, D64[i], d32, h64, h32, tag, lo_fine && i == Lv);
            SFDX_CHECK(hipStreamSynchronize(ctx->st));
            auto solve_pol = [&](sfdx_ctx *c, int pol) {
                LevelSolveResult R = solve_level(c, oi, D64[i], d32, tag, ff, edc, pol, xs[pol], tol_i, maxit, h64, h32,
                                                 mgl);
                SFDX_CHECK(hipStreamSynchronize(c->st));
                return R;
            };
            LevelSolveResult Rs[2];
            double secs[2] = {0, 0}, level_secs = 0;
            if (two) {
                // the second polarisation on the helper context's stream, in a
                // host thread of its own; both only read the shared level data
                sfdx_ctx *w = pol_helper(ctx);
                std::exception_ptr err;
                std::string err_text;
                Timer tl;
                std::thread th([&]() {
                    try {
                        SFDX_CHECK(hipSetDevice(ctx->device));
                        Timer t1;
                        Rs[1] = solve_pol(w, 1);
                        secs[1] = t1.s();
                    } catch (...) {
                        err = std::current_exception();
                        err_text = g_err;
                    }
                });
                try {
                    Rs[0] = solve_pol(ctx, 0);
                    secs[0] = tl.s();
                } catch (...) {
                    th.join();
                    throw;
                }
                th.join();
                if (err) {
                    if (!err_text.empty()) g_err = err_text;
                    std::rethrow_exception(err);
                }
                level_secs = tl.s();
            } else {
                for (int pol = 0; pol < 2; pol++) {
                    Timer tl;
                    Rs[pol] = solve_pol(ctx, pol);
                    secs[pol] = tl.s();
                    level_secs += secs[pol];
                }
            }
            if (i == Lv) {
                S.fine_seconds += level_secs;
                S.mg_freqs += mg_here;
            }
            if (o->verbose) fprintf(stderr, "level %d done at %.3fs\n", i, total.s());
            for (int pol = 0; pol < 2; pol++) {
                const LevelSolveResult &R = Rs[pol];
                if (i == Lv) {
                    S.fine_bytes += R.info.bytes;
                    S.fine_iters_total += R.info.iters;
                    S.refine_steps[pol] = R.refine;
                }
                S.fallback_fp64 += R.info.fallback;
                S.solver_switches += R.info.switched;
                S.iters[i - 1][pol] = R.info.iters;
                S.relres[i - 1][pol] = R.info.relres;
                S.dc_calls[i - 1][pol] = R.info.dc_calls;
                S.pcg_iters[i - 1][pol] = R.info.pcg_iters;
                if (o->verbose)
                    fprintf(stderr, "f=%g level %d (%dx%dx%d) pol %d: %d its relres %.3e dc %d pcg %d %.3fs\n", ff, i,
                            d.nx, d.ny, d.nz, pol, R.info.iters, R.info.relres, R.info.dc_calls, R.info.pcg_iters,
                            secs[pol]);
                if (R.info.status > 0) {
                    g_err = "Krylov solve failed on level " + std::to_string(i);
                    return R.info.status;
                }
                if (R.info.status < 0) worst = R.info.status;
                u[i][pol] = xs[pol];
            }
        }
        if (ex)
            for (int pol = 0; pol < 2; pol++)
                SFDX_CHECK(hipMemcpyAsync(ex + ((size_t)fi * 2 * NLf + (size_t)pol * NLf) * 2, u[Lv][pol],
                                          NLf * sizeof(zd), hipMemcpyDeviceToHost, ctx->st));
        if (surf) {
            // only the surface fields leave the device (10 nx*ny values, not 2 nl)
            const Dims df = make_dims(fine.nx, fine.ny, fine.nz);
            const size_t ns = (size_t)fine.nx * fine.ny;
            zd *sd = pool_vec<zd>(ctx, "surf", 10 * ns);
            const double om = 2 * 3.141592653589793 * ff * kMu0;
            for (int pol = 0; pol < 2; pol++)
                launch_surface_fields(df, fine.nair, planes_dev, D64[Lv].a64, D64[Lv].b64, D64[Lv].c64, om,
                                      u[Lv][pol], sd + (size_t)pol * 5 * ns, ctx->st);
            SFDX_CHECK(hipMemcpyAsync(surf + (size_t)fi * 10 * ns, sd, 10 * ns * sizeof(zd), hipMemcpyDeviceToHost,
                                      ctx->st));
        }
        SFDX_CHECK(hipStreamSynchronize(ctx->st));
    }
    S.total_seconds = total.s();
    S.status = worst;
    if (st) *st = S;
    return worst;
}

// Misfit gradient with respect to log cell-tensor scale factors (sfdx_adjoint).
int adjoint_impl(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq, int nrec,
                 const double *recv, int rd, const double *obs, const sfdx_opts *o, double *misfit, double *grad,
                 double *ex_out, sfdx_stats *st)
{
    const HostLevel fine = host_level(g, sigma);
    const Dims d = make_dims(fine.nx, fine.ny, fine.nz);
    PostGrid pg{g->nx, g->ny, g->nz, g->nair, g->a, g->b, g->c};
    std::vector<int> planes;
    if (o->surface_sigma > 0) {
        surface_planes(fine, o->surface_sigma, planes);
        for (int v : planes)
            if (v >= fine.nz) {
                g_err = "a cell column holds no cell above surface_sigma (no ground surface)";
                return SFDX_ERR_ARG;
            }
        pg.surf = planes.data();
    }
    LevelDev<double> D64 = upload_level<double>(ctx, fine, "ADJ");
    LevelDev<float> D32;
    const bool adj_mixed = o->precision == SFDX_PREC_MIXED || o->precision == SFDX_PREC_MIXED_SPLIT;
    if (adj_mixed) D32 = level_from_device<float>(ctx, fine, D64.sig64, "ADJ", o->sigma_floor32);
    ensure_partials(ctx, edge_launch(d).nblocks() + 64);
    choose_tile(ctx, D64.g, "ADJ/d", o->spmv_tile, o->verbose);
    if (adj_mixed) choose_tile(ctx, D32.g, "ADJ/f", o->spmv_tile, o->verbose);
    const bool xmass = o->precond == SFDX_PRECOND_SSOR || o->ssor_cross_mass != 0;
    D64.g.ssor_xmass = D32.g.ssor_xmass = xmass;
    double *gdev = pool_vec<double>(ctx, "adj/g", d.NC);
    SFDX_CHECK(hipMemsetAsync(gdev, 0, d.NC * sizeof(double), ctx->st));
    // Adjoint solves: no divergence correction (the adjoint source is not
    // divergence-free).  With the multigrid options and merge-ratio grids they
    // run the V-cycle over grids of their own (SFDX_PRECOND_AUTO: at the
    // frequencies its rule picks), else the hybrid preconditioner.
    KOpts ko = kopts_of(*o);
    ko.corr_intv = 0;
    ko.dc_at_start = 0;
    const bool adj_mg = (ko.precond == SFDX_PRECOND_MG || ko.precond == SFDX_PRECOND_AUTO) && o->mg_merge_ratio > 0;
    const bool adj_auto = ko.precond == SFDX_PRECOND_AUTO;
    if (ko.precond == SFDX_PRECOND_MG || ko.precond == SFDX_PRECOND_AUTO) ko.precond = SFDX_PRECOND_HYBRID;
    if (adj_mixed) {
        ko.lo = &D32;
        ko.restart_drop = o->restart_drop;
    }
    MergeHier MA;
    if (adj_mg) {
        MA.build(ctx, fine, D64, adj_mixed ? &D32 : nullptr, o, "ADJ");
        for (int n = 1; n < MA.L; n++) MA.X64[n].g.ssor_xmass = MA.X32[n].g.ssor_xmass = xmass;
    }
    const AutoPrecond rule(fine, mg_grid_count(fine, o), true);
    sfdx_stats S;
    std::memset(&S, 0, sizeof(S));
    int worst = SFDX_OK;
    double phi = 0;
    SliceLaunch L = edge_launch(d);
    // Every frequency is an independent item (forward solve, misfit and its
    // field derivative, the two adjoint solves), run on up to kAdjWorkers
    // worker contexts at once, higher frequencies first (they take more
    // iterations); a single frequency runs on the context itself.  Each item
    // keeps its fields and adjoint solutions, and the gradient is summed
    // afterwards on the context's stream in frequency and polarisation order:
    // the same sums as solving one frequency after the other.
    constexpr int kAdjWorkers = 6;
    const int nw = std::min(nf, kAdjWorkers);
    std::vector<sfdx_ctx *> wk(nw, ctx);
    if (nw > 1) {
        while ((int)ctx->workers.size() < nw) {
            sfdx_ctx *w = nullptr;
            const int rc = sfdx_create(ctx->device, &w);
            if (rc != SFDX_OK) return rc;
            ctx->workers.push_back(w);
        }
        for (int w = 0; w < nw; w++) {
            wk[w] = ctx->workers[w];
            ensure_partials(wk[w], L.nblocks() + 64);
        }
    }
    std::vector<zd *> edf(nf), lamf(nf), wdw(nw);
    for (int f = 0; f < nf; f++) {
        edf[f] = pool_vec<zd>(ctx, "adj/e" + std::to_string(f), 2 * d.NL);
        lamf[f] = pool_vec<zd>(ctx, "adj/l" + std::to_string(f), 2 * d.NL);
    }
    for (int w = 0; w < nw; w++) wdw[w] = pool_vec<zd>(ctx, "adj/w" + std::to_string(w), 2 * d.NL);
    // the level data that the solves build on first use, built once here
    prebuild_level(ctx, D64, adj_mixed ? &D32 : nullptr, adj_mg ? &MA.H64 : nullptr,
                   adj_mg && adj_mixed ? &MA.H32 : nullptr, "ADJ", adj_mixed);
    struct ItemOut {
        int rc = 0, bad = 0;
        double phf = 0, bnorm[2] = {0, 0};
        sfdx_stats sf;
        SolveInfo si[2];
        std::string err;
        std::exception_ptr ex;
    };
    std::vector<ItemOut> out(nf);
    auto run_item = [&](sfdx_ctx *c, zd *wd, int f) {
        ItemOut &it = out[f];
        std::memset(&it.sf, 0, sizeof(it.sf));
        std::vector<zc> exf((size_t)2 * d.NL), wf((size_t)2 * d.NL);
        sfdx_opts oj = *o;
        if (nw > 1) oj.pol_streams = 1;  // the items already run concurrently
        it.rc = fwd_solve_impl(c, g, sigma, 1, freq + f, &oj, (double *)exf.data(), &it.sf);
        if (it.rc > 0) {
            it.err = g_err;
            return;
        }
        if (ex_out) std::memcpy(ex_out + (size_t)f * 2 * d.NL * 2, exf.data(), (size_t)2 * d.NL * sizeof(zc));
        it.bad = misfit_fields(pg, 1, freq + f, nrec, recv, rd, obs + (size_t)f * nrec * 4, exf.data(), it.phf,
                               wf.data());
        if (it.bad) return;
        zd *ed = edf[f], *lam = lamf[f];
        SFDX_CHECK(hipMemcpyAsync(ed, exf.data(), (size_t)2 * d.NL * sizeof(zd), hipMemcpyHostToDevice, c->st));
        SFDX_CHECK(hipMemcpyAsync(wd, wf.data(), (size_t)2 * d.NL * sizeof(zd), hipMemcpyHostToDevice, c->st));
        const double wmu = 2 * 3.141592653589793 * freq[f] * kMu0;
        for (int pol = 0; pol < 2; pol++) {
            zd *w = wd + (size_t)pol * d.NL;
            launch_zero_boundary(d, w, c->st);  // the Dirichlet values are data, not unknowns
            launch_vec<double>(5, D64.g, c->rb, w, nullptr, nullptr, nullptr, nullptr, nullptr, c->st);
            reduce_final(c->rb, L.nblocks(), 1, SL_TMP, OP_RN, c->st);
            read_slots(c);
            it.bnorm[pol] = std::sqrt(c->hslots[SL_RN]);
            SFDX_CHECK(hipMemsetAsync(lam + (size_t)pol * d.NL, 0, d.NL * sizeof(zd), c->st));
        }
        KOpts kf = ko;
        if (adj_mg && (!adj_auto || rule.use_mg(freq[f]))) {
            kf.precond = SFDX_PRECOND_MG;
            kf.mg64 = &MA.H64;
            kf.mg32 = adj_mixed ? &MA.H32 : nullptr;
            kf.mg_level = MA.L;
        }
        SFDX_CHECK(hipStreamSynchronize(c->st));  // the host copies above end here
        auto solve_pol = [&](sfdx_ctx *cc, int pol) {
            SolveInfo r;
            if (!(it.bnorm[pol] > 0)) return r;
            r = krylov<double>(cc, o->solver, D64, "ADJ", wmu, wd + (size_t)pol * d.NL, lam + (size_t)pol * d.NL,
                               it.bnorm[pol], o->eps, o->maxit_plain, kf);
            SFDX_CHECK(hipStreamSynchronize(cc->st));
            return r;
        };
        const bool two = nw == 1 && o->pol_streams == 2 && it.bnorm[0] > 0 && it.bnorm[1] > 0;
        if (two) {
            // the two polarisations' adjoint solves at once, the second on the
            // helper context's stream in a host thread of its own
            sfdx_ctx *hw = pol_helper(c);
            std::exception_ptr err;
            std::string err_text;
            std::thread th([&]() {
                try {
                    SFDX_CHECK(hipSetDevice(ctx->device));
                    it.si[1] = solve_pol(hw, 1);
                } catch (...) {
                    err = std::current_exception();
                    err_text = g_err;
                }
            });
            try {
                it.si[0] = solve_pol(c, 0);
            } catch (...) {
                th.join();
                throw;
            }
            th.join();
            if (err) {
                if (!err_text.empty()) g_err = err_text;
                std::rethrow_exception(err);
            }
        } else {
            for (int pol = 0; pol < 2; pol++) it.si[pol] = solve_pol(c, pol);
        }
    };
    if (nw == 1) {
        for (int f = 0; f < nf; f++) run_item(ctx, wdw[0], f);
    } else {
        std::vector<int> order(nf);
        for (int f = 0; f < nf; f++) order[f] = f;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return freq[x] > freq[y]; });
        std::atomic<int> next{0};
        auto worker = [&](int w) {
            for (int k = next++; k < nf; k = next++) {
                ItemOut &it = out[order[k]];
                try {
                    SFDX_CHECK(hipSetDevice(ctx->device));
                    run_item(wk[w], wdw[w], order[k]);
                } catch (...) {
                    it.ex = std::current_exception();
                    it.err = g_err;
                }
            }
        };
        std::vector<std::thread> th;
        for (int w = 0; w < nw; w++) th.emplace_back(worker, w);
        for (auto &t : th) t.join();
    }
    for (int f = 0; f < nf; f++) {
        ItemOut &it = out[f];
        if (it.ex) {
            if (!it.err.empty()) g_err = it.err;
            std::rethrow_exception(it.ex);
        }
        if (it.rc > 0) {
            g_err = it.err;
            return it.rc;
        }
        if (it.rc < 0) worst = it.rc;
        const sfdx_stats &sf = it.sf;
        S.total_seconds += sf.total_seconds;
        S.fine_seconds += sf.fine_seconds;
        S.fine_iters_total += sf.fine_iters_total;
        S.fallback_fp64 += sf.fallback_fp64;
        S.solver_switches += sf.solver_switches;
        S.mg_freqs += sf.mg_freqs;
        S.spmv_tile = sf.spmv_tile;
        if (it.bad) {
            g_err = "station " + std::to_string(it.bad) + " lies outside the interpolation area";
            return SFDX_ERR_ARG;
        }
        phi += it.phf;
        const double wmu = 2 * 3.141592653589793 * freq[f] * kMu0;
        for (int pol = 0; pol < 2; pol++) {
            if (!(it.bnorm[pol] > 0)) continue;
            S.fallback_fp64 += it.si[pol].fallback;
            S.solver_switches += it.si[pol].switched;
            if (it.si[pol].status > 0) {
                g_err = "adjoint solve failed (frequency " + std::to_string(f) + ")";
                return it.si[pol].status;
            }
            if (it.si[pol].status < 0) worst = it.si[pol].status;
            launch_cell_grad(D64.g, wmu, edf[f] + (size_t)pol * d.NL, lamf[f] + (size_t)pol * d.NL, gdev, ctx->st);
        }
    }
    SFDX_CHECK(hipMemcpyAsync(grad, gdev, d.NC * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    SFDX_CHECK(hipStreamSynchronize(ctx->st));
    *misfit = phi;
    S.status = worst;
    if (st) *st = S;
    return worst;
}

// Device address of `scratch` on `device` (the current device), looked up once.
double *noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        SFDX_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = (double *)p;
    }
    return g_scratch_dev[device];
}

template <typename F>
int guarded(F &&fn)
{
    try {
        return fn();
    } catch (const HipError &e) {
        g_err = e.what;
        return SFDX_ERR_HIP;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SFDX_ERR_ARG;
    }
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

const char *sfdx_last_error(void) { return g_err.c_str(); }

void sfdx_default_opts(sfdx_opts *o)
{
    std::memset(o, 0, sizeof(*o));
    o->level = 3;
    o->use_excmg = 1;
    o->solver = SFDX_SOLVER_BICGSTAB;
    o->eps = 1e-8;
    o->corr_intv = 200;
    o->corr_tol = 1e-3;
    o->ssor_omega = 1.0;
    o->precision = SFDX_PREC_FP64;
    o->norm = SFDX_NORM_REFERENCE;
    o->maxit_plain = 5000;
    o->maxit_excmg = 1000;
    o->inner_tol = 1e-5;
    o->max_refine = 20;
    o->verbose = 0;
    o->dc_restart = 0;
    o->dc_at_start = 0;
    o->inner_maxit = 400;
    o->precond = SFDX_PRECOND_AUTO;
    o->mg_levels = 0;
    o->mg_coarse_pairs = 4;
    o->restart_drop = 1e-3;
    o->use_graph = 1;
    o->nodal_smoother = 1;
    o->coarse_solver = 1;
    o->full_background = 0;
    o->spmv_tile = 0;
    o->surface_sigma = 0;
    o->ssor_cross_mass = 0;
    o->pol_streams = 2;
    o->mg_merge_ratio = 3;
}

int sfdx_create(int device, sfdx_ctx **out)
{
    if (!out) return SFDX_ERR_ARG;
    return guarded([&]() {
        auto *c = new sfdx_ctx();
        c->device = device;
        SFDX_CHECK(hipSetDevice(device));
        SFDX_CHECK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        SFDX_CHECK(hipMalloc(&c->rb.slots, SL_COUNT * sizeof(double)));
        SFDX_CHECK(hipMemset(c->rb.slots, 0, SL_COUNT * sizeof(double)));
        SFDX_CHECK(hipHostMalloc(&c->hslots, SL_COUNT * sizeof(double), hipHostMallocDefault));
        SFDX_CHECK(hipHostMalloc(&c->hlag, 2 * SL_COUNT * sizeof(double), hipHostMallocDefault));
        for (int q = 0; q < 2; q++) SFDX_CHECK(hipEventCreateWithFlags(&c->lag_ev[q], hipEventDisableTiming));
        noalloc_scratch_init(device);
        *out = c;
        return SFDX_OK;
    });
}

int sfdx_destroy(sfdx_ctx *c)
{
    if (!c) return SFDX_OK;
    for (sfdx_ctx *w : c->workers) sfdx_destroy(w);
    c->workers.clear();
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    c->pool.release();
    if (c->rb.partials) (void)hipFree(c->rb.partials);
    if (c->rb.slots) (void)hipFree(c->rb.slots);
    if (c->hslots) (void)hipHostFree(c->hslots);
    if (c->hlag) (void)hipHostFree(c->hlag);
    for (int q = 0; q < 2; q++)
        if (c->lag_ev[q]) (void)hipEventDestroy(c->lag_ev[q]);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return SFDX_OK;
}

int sfdx_fwd_solve(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq,
                   const sfdx_opts *o, double *ex, sfdx_stats *st)
{
    if (!ctx || !g || !sigma || !freq || !ex || nf <= 0) return SFDX_ERR_ARG;
    sfdx_opts def;
    if (!o) { sfdx_default_opts(&def); o = &def; }
    if (nf > 1 && o->pol_streams == 2) {
        // several frequencies: independent items on concurrent streams, as a
        // batch of this one model (results identical to the frequency loop)
        sfdx_job J;
        std::memset(&J, 0, sizeof(J));
        J.grid = *g;
        J.sigma = sigma;
        J.nf = nf;
        J.freq = freq;
        J.ex = ex;
        return sfdx_solve_batch(ctx, 1, &J, o, 0, st);
    }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        return fwd_solve_impl(ctx, g, sigma, nf, freq, o, ex, st);
    });
}

int sfdx_solve_batch(sfdx_ctx *ctx, int njobs, const sfdx_job *jobs, const sfdx_opts *o, int nstreams,
                     sfdx_stats *st)
{
    if (!ctx || njobs <= 0 || !jobs) return SFDX_ERR_ARG;
    for (int j = 0; j < njobs; j++)
        if (!jobs[j].sigma || !jobs[j].freq || !jobs[j].ex || jobs[j].nf <= 0 || !jobs[j].grid.a || !jobs[j].grid.b ||
            !jobs[j].grid.c)
            return SFDX_ERR_ARG;
    sfdx_opts def;
    if (!o) { sfdx_default_opts(&def); o = &def; }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        struct Item {
            int j, f;
            double cost, freq;
        };
        std::vector<Item> items;
        for (int j = 0; j < njobs; j++) {
            const sfdx_grid &g = jobs[j].grid;
            const double nl = (double)make_dims(g.nx, g.ny, g.nz).NL;
            for (int f = 0; f < jobs[j].nf; f++) items.push_back({j, f, nl, jobs[j].freq[f]});
        }
        // longest processing time first: larger grids first, and on grids of one
        // size higher frequencies first (the skin depth shrinks towards the cell
        // widths and the iteration counts grow: on a 128x128x96 model 65
        // iterations at 0.1 Hz, about 180 at 100 Hz and over 1000 at 1 kHz)
        std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
            return x.cost != y.cost ? x.cost > y.cost : x.freq > y.freq;
        });
        // default 6: on the C4 batch (16 frequencies of a 128x128x96 model and 8
        // models of 48^3 to 96^3) 6 streams took 8.67-8.78 s, 5 or 7 about 9.0 s,
        // 8 9.0-9.2 s, 4 10.2 s (the device has 4 hardware queues per process)
        const int ns = std::max(1, std::min({nstreams > 0 ? nstreams : 6, 16, (int)items.size()}));
        while ((int)ctx->workers.size() < ns) {
            sfdx_ctx *w = nullptr;
            const int rc = sfdx_create(ctx->device, &w);
            if (rc != SFDX_OK) return rc;
            ctx->workers.push_back(w);
        }
        std::vector<int> rcs(items.size(), 0);
        std::vector<std::string> errs(items.size());
        std::vector<sfdx_stats> sts(items.size());
        std::atomic<int> next{0};
        auto worker = [&](int w) {
            (void)hipSetDevice(ctx->device);
            for (int k = next++; k < (int)items.size(); k = next++) {
                const Item &it = items[k];
                const sfdx_job &J = jobs[it.j];
                sfdx_opts oj = *o;
                if (J.level > 0) oj.level = J.level;
                oj.pol_streams = 1;  // the items already run concurrently
                const int64_t nl = make_dims(J.grid.nx, J.grid.ny, J.grid.nz).NL;
                double *exf = J.ex + (size_t)it.f * 2 * nl * 2;
                std::memset(&sts[k], 0, sizeof(sfdx_stats));
                rcs[k] = guarded([&]() { return fwd_solve_impl(ctx->workers[w], &J.grid, J.sigma, 1, J.freq + it.f, &oj, exf, &sts[k]); });
                if (rcs[k] > 0) errs[k] = g_err;
            }
        };
        std::vector<std::thread> th;
        for (int w = 0; w < ns; w++) th.emplace_back(worker, w);
        for (auto &t : th) t.join();
        int worst = SFDX_OK;
        for (size_t k = 0; k < items.size(); k++) {
            if (rcs[k] > 0) {
                g_err = "job " + std::to_string(items[k].j) + " frequency " + std::to_string(items[k].f) + ": " + errs[k];
                return rcs[k];
            }
            if (rcs[k] < 0) worst = rcs[k];
        }
        if (st) {
            for (int j = 0; j < njobs; j++) std::memset(&st[j], 0, sizeof(sfdx_stats));
            for (size_t k = 0; k < items.size(); k++) {
                sfdx_stats &S = st[items[k].j];
                const sfdx_stats &T = sts[k];
                S.total_seconds += T.total_seconds;
                S.fine_seconds += T.fine_seconds;
                S.fine_bytes += T.fine_bytes;
                S.fine_iters_total += T.fine_iters_total;
                S.fallback_fp64 += T.fallback_fp64;
                S.solver_switches += T.solver_switches;
                S.mg_freqs += T.mg_freqs;
                S.spmv_tile = T.spmv_tile;
                if (T.status < 0) S.status = T.status;
                if (items[k].f == jobs[items[k].j].nf - 1) {
                    std::memcpy(S.iters, T.iters, sizeof(S.iters));
                    std::memcpy(S.relres, T.relres, sizeof(S.relres));
                    std::memcpy(S.dc_calls, T.dc_calls, sizeof(S.dc_calls));
                    std::memcpy(S.pcg_iters, T.pcg_iters, sizeof(S.pcg_iters));
                    std::memcpy(S.refine_steps, T.refine_steps, sizeof(S.refine_steps));
                }
            }
        }
        return worst;
    });
}

int sfdx_fwd_response(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq,
                      const sfdx_opts *o, int nrec, const double *recv, int out_type, int rd, double *results,
                      double *tipper, sfdx_stats *st)
{
    if (!ctx || !g || !sigma || !freq || !recv || !results || nf <= 0 || nrec <= 0) return SFDX_ERR_ARG;
    sfdx_opts def;
    if (!o) { sfdx_default_opts(&def); o = &def; }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        std::vector<zc> surf((size_t)nf * 10 * g->nx * g->ny);
        const int rc = fwd_solve_impl(ctx, g, sigma, nf, freq, o, nullptr, st, surf.data());
        if (rc > 0) return rc;
        const PostGrid pg{g->nx, g->ny, g->nz, g->nair, g->a, g->b, g->c};
        const int bad = post_process_surface(pg, nf, freq, nrec, recv, surf.data(), out_type, rd, results, tipper);
        if (bad) {
            g_err = "station " + std::to_string(bad) + " lies outside the interpolation area";
            return (int)SFDX_ERR_ARG;
        }
        return rc;
    });
}

int sfdx_misfit(const sfdx_grid *g, int nf, const double *freq, int nrec, const double *recv, int rd,
                const double *obs, const double *ex, double *misfit, double *w)
{
    if (!g || !freq || !recv || !obs || !ex || !misfit || nf <= 0 || nrec <= 0) return SFDX_ERR_ARG;
    return guarded([&]() {
        const PostGrid pg{g->nx, g->ny, g->nz, g->nair, g->a, g->b, g->c};
        const int bad = misfit_fields(pg, nf, freq, nrec, recv, rd, obs, (const zc *)ex, *misfit, (zc *)w);
        if (bad) {
            g_err = "station " + std::to_string(bad) + " lies outside the interpolation area";
            return (int)SFDX_ERR_ARG;
        }
        return (int)SFDX_OK;
    });
}

int sfdx_adjoint(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq, int nrec,
                 const double *recv, int rd, const double *obs, const sfdx_opts *o, double *misfit, double *grad,
                 double *ex, sfdx_stats *st)
{
    if (!ctx || !g || !sigma || !freq || !recv || !obs || !misfit || !grad || nf <= 0 || nrec <= 0) return SFDX_ERR_ARG;
    sfdx_opts def;
    if (!o) { sfdx_default_opts(&def); o = &def; }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        return adjoint_impl(ctx, g, sigma, nf, freq, nrec, recv, rd, obs, o, misfit, grad, ex, st);
    });
}

int sfdx_post_process(const sfdx_grid *g, int nf, const double *freq, int nrec, const double *recv, const double *ex,
                      int out_type, int rd, double *results, double *tipper)
{
    return sfdx_post_process_topo(g, nullptr, nf, freq, nrec, recv, ex, out_type, rd, results, tipper);
}

int sfdx_post_process_topo(const sfdx_grid *g, const int *surf, int nf, const double *freq, int nrec,
                           const double *recv, const double *ex, int out_type, int rd, double *results, double *tipper)
{
    if (!g || !freq || !recv || !ex || !results) return SFDX_ERR_ARG;
    if (surf)
        for (size_t q = 0; q < (size_t)g->nx * g->ny; q++)
            if (surf[q] < 0 || surf[q] >= g->nz) {
                g_err = "surface plane of column " + std::to_string(q) + " outside 0..nz-1";
                return SFDX_ERR_ARG;
            }
    PostGrid pg{g->nx, g->ny, g->nz, g->nair, g->a, g->b, g->c, surf};
    int s = post_process(pg, nf, freq, nrec, recv, (const zc *)ex, out_type, rd, results, tipper);
    if (s) { g_err = "station " + std::to_string(s) + " lies outside the interpolation area"; return SFDX_ERR_ARG; }
    return SFDX_OK;
}

int sfdx_surface_planes(const sfdx_grid *g, const double *sigma, double air_max, int *surf)
{
    if (!g || !g->a || !g->b || !g->c || !sigma || !surf || g->nx <= 0 || g->ny <= 0 || g->nz <= 0)
        return SFDX_ERR_ARG;
    return guarded([&]() {
        const HostLevel L = host_level(g, sigma);
        std::vector<int> planes;
        surface_planes(L, air_max, planes);
        std::copy(planes.begin(), planes.end(), surf);
        for (int v : planes)
            if (v >= g->nz) {
                g_err = "a cell column holds no cell above air_max (no ground surface)";
                return (int)SFDX_ERR_ARG;
            }
        return (int)SFDX_OK;
    });
}

int sfdx_mesh_coarsen(const sfdx_grid *g, const double *sigma, int k, double *a_c, double *b_c, double *c_c,
                      double *sigma_c)
{
    if (!g || !g->a || !g->b || !g->c || !sigma || !a_c || !b_c || !c_c || !sigma_c || k < 0 || k > 16)
        return SFDX_ERR_ARG;
    const int f = 1 << k;
    if (g->nx % f || g->ny % f || g->nz % f) {
        g_err = "grid dimensions must be divisible by 2^k";
        return SFDX_ERR_MESH;
    }
    return guarded([&]() {
        const HostLevel C = coarsen(host_level(g, sigma), k);
        std::copy(C.a.begin(), C.a.end(), a_c);
        std::copy(C.b.begin(), C.b.end(), b_c);
        std::copy(C.c.begin(), C.c.end(), c_c);
        std::copy(C.sig.begin(), C.sig.end(), sigma_c);
        return (int)SFDX_OK;
    });
}

int sfdx_mesh_coarsen_device(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int k, double *a_c,
                             double *b_c, double *c_c, double *sigma_c)
{
    if (!ctx || !g || !g->a || !g->b || !g->c || !sigma || !a_c || !b_c || !c_c || !sigma_c || k < 0 || k > 16)
        return SFDX_ERR_ARG;
    const int f = 1 << k;
    if (g->nx % f || g->ny % f || g->nz % f) {
        g_err = "grid dimensions must be divisible by 2^k";
        return SFDX_ERR_MESH;
    }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        const HostLevel F = host_level(g, sigma);
        const LevelDev<double> D = upload_level<double>(ctx, F, "mc");
        const HostLevel C = coarsen_device(ctx, F, D, k, "mc");
        std::copy(C.a.begin(), C.a.end(), a_c);
        std::copy(C.b.begin(), C.b.end(), b_c);
        std::copy(C.c.begin(), C.c.end(), c_c);
        std::copy(C.sig.begin(), C.sig.end(), sigma_c);
        return (int)SFDX_OK;
    });
}

int sfdx_direct_solve(const sfdx_grid *g, const double *sigma, double freq, double *u)
{
    if (!g || !sigma || !u) return SFDX_ERR_ARG;
    return guarded([&]() {
        const HostLevel L = host_level(g, sigma);
        const std::vector<zc> edc = background_field(L, freq);
        CoarseSystem CS;
        coarse_system(L, freq, edc, CS);
        std::vector<zc> sol;
        if (!direct_solve(CS, sol, g_err)) return SFDX_ERR_DIRECT;
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        zc *out = (zc *)u;
        for (int pol = 0; pol < 2; pol++)
            for (int64_t e = 0; e < d.NL; e++) {
                int ii, jj, kk;
                int oo = edge_decode(d, e, ii, jj, kk);
                zc v(0);
                if (CS.imap[e] >= 0) v = sol[(size_t)pol * CS.n + CS.imap[e]];
                else dirichlet_value(d, edc.data(), pol, oo, ii, jj, kk, v);
                out[(size_t)pol * d.NL + e] = v;
            }
        return SFDX_OK;
    });
}

int sfdx_mt1d_anl(int nz, const double *conv, const double *c, double freq, double *ex, double *hx, double *imp)
{
    if (!conv || !c || !ex) return SFDX_ERR_ARG;
    mt1d_anl(nz, conv, c, freq, (zc *)ex, (zc *)hx, (zc *)imp);
    return SFDX_OK;
}

int sfdx_apply_operator(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, int precision,
                        const double *x, double *y)
{
    if (!ctx || !g || !sigma || !x || !y) return SFDX_ERR_ARG;
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel L = host_level(g, sigma);
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        ensure_partials(ctx, edge_launch(d).nblocks() + 64);
        const double wmu = 2 * 3.141592653589793 * freq * kMu0;
        if (precision == SFDX_PREC_FP64) {
            LevelDev<double> D = upload_level<double>(ctx, L, "op");
            zd *xd = pool_vec<zd>(ctx, "op/x", d.NL), *yd = pool_vec<zd>(ctx, "op/y", d.NL);
            SFDX_CHECK(hipMemcpyAsync(xd, x, d.NL * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
            launch_spmv<double>(D.g, wmu, xd, yd, 1.0, 0, nullptr, ctx->rb, ctx->st);
            SFDX_CHECK(hipMemcpyAsync(y, yd, d.NL * sizeof(zd), hipMemcpyDeviceToHost, ctx->st));
        } else {
            LevelDev<float> D = upload_level<float>(ctx, L, "op");
            std::vector<zf> xf(d.NL), yf(d.NL);
            for (int64_t e = 0; e < d.NL; e++) xf[e] = zf((float)x[2 * e], (float)x[2 * e + 1]);
            zf *xd = pool_vec<zf>(ctx, "op/xf", d.NL), *yd = pool_vec<zf>(ctx, "op/yf", d.NL);
            SFDX_CHECK(hipMemcpyAsync(xd, xf.data(), d.NL * sizeof(zf), hipMemcpyHostToDevice, ctx->st));
            launch_spmv<float>(D.g, (float)wmu, xd, yd, 1.0f, 0, nullptr, ctx->rb, ctx->st);
            SFDX_CHECK(hipMemcpyAsync(yf.data(), yd, d.NL * sizeof(zf), hipMemcpyDeviceToHost, ctx->st));
            SFDX_CHECK(hipStreamSynchronize(ctx->st));
            for (int64_t e = 0; e < d.NL; e++) { y[2 * e] = yf[e].re; y[2 * e + 1] = yf[e].im; }
        }
        SFDX_CHECK(hipStreamSynchronize(ctx->st));
        return SFDX_OK;
    });
}

int sfdx_split_residual(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, int precision,
                        const double *r, const double *z, const double *psi, double *t)
{
    if (!ctx || !g || !sigma || !r || !z || !psi || !t) return SFDX_ERR_ARG;
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel L = host_level(g, sigma);
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        const double wmu = 2 * 3.141592653589793 * freq * kMu0;
        auto run = [&](auto tag) {
            using R = decltype(tag);
            using C = cx<R>;
            LevelDev<R> D = upload_level<R>(ctx, L, "sr");
            std::vector<C> hr(d.NL), hz(d.NL), hp(d.NP), ht(d.NL);
            for (int64_t e = 0; e < d.NL; e++) {
                hr[e] = C((R)r[2 * e], (R)r[2 * e + 1]);
                hz[e] = C((R)z[2 * e], (R)z[2 * e + 1]);
            }
            for (int64_t n = 0; n < d.NP; n++) hp[n] = C((R)psi[2 * n], (R)psi[2 * n + 1]);
            const std::string tg = sizeof(R) == 8 ? "sr/d" : "sr/f";
            C *rd = pool_vec<C>(ctx, tg + "r", d.NL), *zd_ = pool_vec<C>(ctx, tg + "z", d.NL);
            C *pd = pool_vec<C>(ctx, tg + "p", d.NP), *td = pool_vec<C>(ctx, tg + "t", d.NL);
            SFDX_CHECK(hipMemcpyAsync(rd, hr.data(), d.NL * sizeof(C), hipMemcpyHostToDevice, ctx->st));
            SFDX_CHECK(hipMemcpyAsync(zd_, hz.data(), d.NL * sizeof(C), hipMemcpyHostToDevice, ctx->st));
            SFDX_CHECK(hipMemcpyAsync(pd, hp.data(), d.NP * sizeof(C), hipMemcpyHostToDevice, ctx->st));
            launch_resid_split<R>(D.g, (R)wmu, zd_, pd, rd, td, ctx->st);
            SFDX_CHECK(hipMemcpyAsync(ht.data(), td, d.NL * sizeof(C), hipMemcpyDeviceToHost, ctx->st));
            SFDX_CHECK(hipStreamSynchronize(ctx->st));
            for (int64_t e = 0; e < d.NL; e++) { t[2 * e] = ht[e].re; t[2 * e + 1] = ht[e].im; }
        };
        if (precision == SFDX_PREC_FP64) run(0.0);
        else run(0.0f);
        return SFDX_OK;
    });
}

int sfdx_apply_operator_noalloc(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, const double *x,
                                double *y, void *stream)
{
    if (!ctx || !g || !g->a || !g->b || !g->c || !sigma || !x || !y || g->nx <= 0 || g->ny <= 0 || g->nz <= 0)
        return SFDX_ERR_ARG;
    if (g->nx > kNoallocMaxDim || g->ny > kNoallocMaxDim || g->nz > kNoallocMaxDim) {
        g_err = "sfdx_apply_operator_noalloc: nx, ny and nz must not exceed " + std::to_string(kNoallocMaxDim);
        return SFDX_ERR_ARG;
    }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        double *sc = g_scratch_dev[ctx->device];
        if (!sc) throw std::runtime_error("scratch address not initialised on this device");
        const hipStream_t st = stream ? (hipStream_t)stream : ctx->st;
        DevGrid<double> D;
        D.d = make_dims(g->nx, g->ny, g->nz);
        D.a = const_cast<double *>(g->a);
        D.b = const_cast<double *>(g->b);
        D.c = const_cast<double *>(g->c);
        D.ra = sc;
        D.rb = sc + kNoallocMaxDim;
        D.rc = sc + 2 * kNoallocMaxDim;
        D.sig = const_cast<double *>(sigma);
        launch_reciprocal(g->nx, g->a, D.ra, st);
        launch_reciprocal(g->ny, g->b, D.rb, st);
        launch_reciprocal(g->nz, g->c, D.rc, st);
        const double wmu = 2 * 3.141592653589793 * freq * kMu0;
        launch_spmv<double>(D, wmu, (const zd *)x, (zd *)y, 1.0, 0, nullptr, ctx->rb, st);
        SFDX_CHECK(hipGetLastError());
        return (int)SFDX_OK;
    });
}

int sfdx_ssor_apply(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, double omega, int precision,
                    const double *x, double *y)
{
    if (!ctx || !g || !sigma || !x || !y) return SFDX_ERR_ARG;
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel L = host_level(g, sigma);
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        const double wmu = 2 * 3.141592653589793 * freq * kMu0;
        if (precision == SFDX_PREC_FP64) {
            LevelDev<double> D = upload_level<double>(ctx, L, "op");
            zd *xd = pool_vec<zd>(ctx, "op/x", d.NL), *yd = pool_vec<zd>(ctx, "op/y", d.NL);
            SFDX_CHECK(hipMemcpyAsync(xd, x, d.NL * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
            launch_ssor<double>(D.g, wmu, xd, yd, omega, ctx->st);
            SFDX_CHECK(hipMemcpyAsync(y, yd, d.NL * sizeof(zd), hipMemcpyDeviceToHost, ctx->st));
            SFDX_CHECK(hipStreamSynchronize(ctx->st));
        } else {
            LevelDev<float> D = upload_level<float>(ctx, L, "op");
            std::vector<zf> xf(d.NL), yf(d.NL);
            for (int64_t e = 0; e < d.NL; e++) xf[e] = zf((float)x[2 * e], (float)x[2 * e + 1]);
            zf *xd = pool_vec<zf>(ctx, "op/xf", d.NL), *yd = pool_vec<zf>(ctx, "op/yf", d.NL);
            SFDX_CHECK(hipMemcpyAsync(xd, xf.data(), d.NL * sizeof(zf), hipMemcpyHostToDevice, ctx->st));
            launch_ssor<float>(D.g, (float)wmu, xd, yd, (float)omega, ctx->st);
            SFDX_CHECK(hipMemcpyAsync(yf.data(), yd, d.NL * sizeof(zf), hipMemcpyDeviceToHost, ctx->st));
            SFDX_CHECK(hipStreamSynchronize(ctx->st));
            for (int64_t e = 0; e < d.NL; e++) { y[2 * e] = yf[e].re; y[2 * e + 1] = yf[e].im; }
        }
        return SFDX_OK;
    });
}

int sfdx_hybrid_apply(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, double omega, int nodal,
                      int precision, const double *x, double *y)
{
    if (!ctx || !g || !sigma || !x || !y || (nodal != 0 && nodal != 1)) return SFDX_ERR_ARG;
    if (precision == SFDX_PREC_MIXED && nodal != 1) return SFDX_ERR_ARG;
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel L = host_level(g, sigma);
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        const double wmu = 2 * 3.141592653589793 * freq * kMu0;
        LevelDev<double> D = upload_level<double>(ctx, L, "op");
        D.g.ssor_xmass = false;  // the solvers' default edge SSOR (lumped mass)
        ensure_P(ctx, D, "op");
        zd *xd = pool_vec<zd>(ctx, "op/x", d.NL), *yd = pool_vec<zd>(ctx, "op/y", d.NL);
        zd *tn = pool_vec<zd>(ctx, "op/tn", d.NP), *ph = pool_vec<zd>(ctx, "op/ph", d.NP);
        SFDX_CHECK(hipMemcpyAsync(xd, x, d.NL * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
        if (precision == SFDX_PREC_FP64) {
            launch_hybrid<double>(D.g, wmu, D.P, xd, yd, tn, ph, omega, ctx->st, nodal);
        } else {
            // the mixed path's preconditioner: edge SSOR in complex64, nodal part
            // in complex128 on the FP32 symmetric nodal operator
            LevelDev<float> F = upload_level<float>(ctx, L, "op");
            F.g.ssor_xmass = false;
            float *ps32 = pool_vec<float>(ctx, "op/Ps32", (size_t)10 * d.NP);
            launch_psym32(d.NP, D.P, ps32, ctx->st);
            zf *in32 = pool_vec<zf>(ctx, "op/in32", d.NL), *z32 = pool_vec<zf>(ctx, "op/z32", d.NL);
            launch_hybrid_edge32(D.g, F.g, wmu, ps32, xd, yd, in32, z32, tn, ph, (float)omega, ctx->st);
        }
        SFDX_CHECK(hipMemcpyAsync(y, yd, d.NL * sizeof(zd), hipMemcpyDeviceToHost, ctx->st));
        SFDX_CHECK(hipStreamSynchronize(ctx->st));
        return SFDX_OK;
    });
}

int sfdx_div_correct(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double *x, double corr_tol, int *pcg_iters)
{
    if (!ctx || !g || !sigma || !x) return SFDX_ERR_ARG;
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel L = host_level(g, sigma);
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        ensure_partials(ctx, edge_launch(d).nblocks() + 64);
        LevelDev<double> D = upload_level<double>(ctx, L, "dcop");
        zd *xd = pool_vec<zd>(ctx, "dcop/x", d.NL);
        SFDX_CHECK(hipMemcpyAsync(xd, x, d.NL * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
        int it = div_correct<double>(ctx, D, "dcop", xd, corr_tol);
        SFDX_CHECK(hipMemcpyAsync(x, xd, d.NL * sizeof(zd), hipMemcpyDeviceToHost, ctx->st));
        SFDX_CHECK(hipStreamSynchronize(ctx->st));
        if (pcg_iters) *pcg_iters = it;
        return SFDX_OK;
    });
}

int sfdx_extint_v(sfdx_ctx *ctx, const sfdx_grid *g, const double *u0, const double *u1, double *out)
{
    if (!ctx || !g || !u0 || !u1 || !out) return SFDX_ERR_ARG;
    if (g->nx % 8 || g->ny % 8 || g->nz % 8 || g->nair % 2) { g_err = "grid not divisible by 8"; return SFDX_ERR_MESH; }
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel F;
        F.nx = g->nx; F.ny = g->ny; F.nz = g->nz; F.nair = g->nair;
        F.a.assign(g->a, g->a + g->nx); F.b.assign(g->b, g->b + g->ny); F.c.assign(g->c, g->c + g->nz);
        HostLevel M;
        M.nx = F.nx / 2; M.ny = F.ny / 2; M.nz = F.nz / 2; M.nair = F.nair / 2;
        for (int i = 0; i < M.nx; i++) M.a.push_back(F.a[2 * i] + F.a[2 * i + 1]);
        for (int i = 0; i < M.ny; i++) M.b.push_back(F.b[2 * i] + F.b[2 * i + 1]);
        for (int i = 0; i < M.nz; i++) M.c.push_back(F.c[2 * i] + F.c[2 * i + 1]);
        const Dims df = make_dims(F.nx, F.ny, F.nz), dm = make_dims(M.nx, M.ny, M.nz), d0 = make_dims(M.nx / 2, M.ny / 2, M.nz / 2);
        zd *u0d = pool_vec<zd>(ctx, "ext/u0", d0.NL), *u1d = pool_vec<zd>(ctx, "ext/u1", dm.NL), *od = pool_vec<zd>(ctx, "ext/o", df.NL);
        SFDX_CHECK(hipMemcpyAsync(u0d, u0, d0.NL * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
        SFDX_CHECK(hipMemcpyAsync(u1d, u1, dm.NL * sizeof(zd), hipMemcpyHostToDevice, ctx->st));
        extint_device(ctx, F, M, u0d, u1d, od, "ext");
        SFDX_CHECK(hipMemcpyAsync(out, od, df.NL * sizeof(zd), hipMemcpyDeviceToHost, ctx->st));
        SFDX_CHECK(hipStreamSynchronize(ctx->st));
        return SFDX_OK;
    });
}

int sfdx_bench_kernels(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, int precision, int reps,
                       double *ms)
{
    if (!ctx || !g || !sigma || !ms || reps <= 0) return SFDX_ERR_ARG;
    return guarded([&]() {
        SFDX_CHECK(hipSetDevice(ctx->device));
        HostLevel L = host_level(g, sigma);
        const Dims d = make_dims(L.nx, L.ny, L.nz);
        ensure_partials(ctx, edge_launch(d).nblocks() + 64);
        hipEvent_t e0, e1;
        SFDX_CHECK(hipEventCreate(&e0));
        SFDX_CHECK(hipEventCreate(&e1));
        auto run = [&](auto tagR) {
            using R = decltype(tagR);
            LevelDev<R> D = upload_level<R>(ctx, L, "bench");
            const R wmu = (R)(2 * 3.141592653589793 * freq * kMu0);
            std::vector<cx<R>> hx(d.NL);
            for (int64_t e = 0; e < d.NL; e++) hx[e] = cx<R>((R)std::sin(0.001 * e), (R)std::cos(0.002 * e));
            // 12 edge vectors in rotation (0.9 GB in complex128): every timed call
            // reads and writes vectors that the previous calls did not touch, so
            // the 256 MB last-level cache cannot serve them as it would a
            // repeated call on the same data
            constexpr int NV = 12;
            cx<R> *v[NV];
            for (int q = 0; q < NV; q++) {
                v[q] = pool_vec<cx<R>>(ctx, "bench/v" + std::to_string(q) + (sizeof(R) == 8 ? "d" : "f"), d.NL);
                SFDX_CHECK(hipMemcpyAsync(v[q], hx.data(), d.NL * sizeof(cx<R>), hipMemcpyHostToDevice, ctx->st));
                launch_dirichlet<R>(d, (const cx<R> *)nullptr, 2, v[q], 0, ctx->st);
            }
            auto V = [&](int q) { return v[((q % NV) + NV) % NV]; };
            auto timed = [&](auto &&fn) {
                for (int q = 0; q < 3; q++) fn(q);  // warm-up
                SFDX_CHECK(hipEventRecord(e0, ctx->st));
                for (int q = 0; q < reps; q++) fn(q);
                SFDX_CHECK(hipEventRecord(e1, ctx->st));
                SFDX_CHECK(hipEventSynchronize(e1));
                float t;
                SFDX_CHECK(hipEventElapsedTime(&t, e0, e1));
                return (double)t / reps;
            };
            // SpMV
            ms[0] = timed([&](int q) { launch_spmv<R>(D.g, wmu, V(2 * q), V(2 * q + 1), (R)1, 0, nullptr, ctx->rb, ctx->st); });
            // SSOR (12 colour sweeps)
            ms[1] = timed([&](int q) { launch_ssor<R>(D.g, wmu, V(2 * q), V(2 * q + 1), (R)1, ctx->st); });
            // COCR vector update (reads 6 vectors, writes 3)
            ms[2] = timed([&](int q) {
                launch_vec<R>(0, D.g, ctx->rb, V(6 * q), V(6 * q + 1), V(6 * q + 2), V(6 * q + 3), V(6 * q + 4),
                              V(6 * q + 5), ctx->st);
            });
            // plain copy of one edge vector (2 * NL * sizeof(cx) bytes): calibrates
            // hardware byte counters against a kernel of known traffic
            ms[4] = timed([&](int q) { launch_copy<R>(d.NL, V(2 * q), V(2 * q + 1), ctx->st); });
            // the same bytes with the bandwidth-tuned streaming copy (4 or 8
            // loads in flight per thread, non-temporal stores; the faster one):
            // the attainable streaming rate
            const size_t cb = (size_t)d.NL * sizeof(cx<R>) / 16 * 16;
            ms[3] = std::min(timed([&](int q) { launch_stream_copy(cb, V(2 * q), V(2 * q + 1), 4, ctx->st); }),
                             timed([&](int q) { launch_stream_copy(cb, V(2 * q), V(2 * q + 1), 8, ctx->st); }));
            // hybrid preconditioner (SSOR + nodal auxiliary correction)
            ensure_P(ctx, D, "bench");
            cx<R> *tn = pool_vec<cx<R>>(ctx, std::string("bench/tn") + (sizeof(R) == 8 ? "d" : "f"), d.NP);
            cx<R> *ph = pool_vec<cx<R>>(ctx, std::string("bench/ph") + (sizeof(R) == 8 ? "d" : "f"), d.NP);
            // (with the nodal Jacobi polynomial, the solver's default)
            ms[5] = timed([&](int q) { launch_hybrid<R>(D.g, wmu, D.P, V(2 * q), V(2 * q + 1), tn, ph, (R)1, ctx->st, 1); });
            // the one-row-per-thread gather form of the SpMV, the default kernel
            // (variant 0), the k-marching tile variants 1-3 (default chunk) and
            // the 32x8 k-marching tile over 4-plane chunks
            ms[6] = timed([&](int q) { launch_spmv_gather<R>(D.g, wmu, V(2 * q), V(2 * q + 1), (R)1, 0, nullptr, ctx->rb, ctx->st); });
            for (int u = 0; u < 4; u++)
                ms[7 + u] = timed([&](int q) {
                    launch_spmv_march<R>(D.g, wmu, V(2 * q), V(2 * q + 1), (R)1, 0, nullptr, ctx->rb, ctx->st, 0, u);
                });
            ms[11] = timed([&](int q) {
                launch_spmv_march<R>(D.g, wmu, V(2 * q), V(2 * q + 1), (R)1, 0, nullptr, ctx->rb, ctx->st, 4, 4);
            });
            // the k-marching 32x8 tile with prefetch (in complex128 the default is the cell-plane march)
            ms[12] = timed([&](int q) {
                launch_spmv_march<R>(D.g, wmu, V(2 * q), V(2 * q + 1), (R)1, 0, nullptr, ctx->rb, ctx->st, 0, 4);
            });
            // the V-cycle's split residual t = r - A (z + (i/wmu) G psi) (psi: node values)
            {
                std::vector<cx<R>> hp(d.NP);
                for (int64_t n = 0; n < d.NP; n++) hp[n] = cx<R>((R)std::cos(0.003 * n), (R)std::sin(0.001 * n));
                cx<R> *psi = pool_vec<cx<R>>(ctx, std::string("bench/psi") + (sizeof(R) == 8 ? "d" : "f"), d.NP);
                SFDX_CHECK(hipMemcpyAsync(psi, hp.data(), d.NP * sizeof(cx<R>), hipMemcpyHostToDevice, ctx->st));
                ms[13] = timed([&](int q) { launch_resid_split<R>(D.g, wmu, V(3 * q), psi, V(3 * q + 1), V(3 * q + 2), ctx->st); });
                SFDX_CHECK(hipStreamSynchronize(ctx->st));
            }
        };
        if (precision == SFDX_PREC_FP64) run(0.0); else run(0.0f);
        SFDX_CHECK(hipEventDestroy(e0));
        SFDX_CHECK(hipEventDestroy(e1));
        return SFDX_OK;
    });
}

}  // extern "C"
