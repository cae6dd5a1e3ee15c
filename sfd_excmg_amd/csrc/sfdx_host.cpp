// sfdx_host.cpp -- host-side FP64 parts of the forward solve: the analytic 1-D
// anisotropic background, sigma coarsening, EXCMG weight tables, the
// coarse-level direct solve (MKL PARDISO) and the impedance post-process.
#include "sfdx_host.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace sfdx {

static const double kPi = 3.141592653589793238462643;
const double kMu0 = 4.0 * 3.141592653589793238462643e-7;  // src/global_para.f90:14

// ------------------------------------------------------------------ 1-D field
// Analytic 1-D anisotropic layered field, restating mt1d_anl
// (sfd_excmg/src/model_selec.f90:266-401, Pek & Santos 2002).  Layer i carries
// the effective horizontal tensor A = sigma_h - sigma_hz sigma_zh / sigma_zz; its
// two principal wavenumbers propagate the 4-vector (Ex, Ey, Hx, Hy) upward from
// the basement (the last layer continued as a half-space) by 4x4 layer
// matrices.  The top field is normalised to E = (1,0) for xy and (0,1) for yx.
typedef std::complex<double> zc;
typedef std::array<std::array<zc, 4>, 4> M4;

static M4 mul4(const M4 &A, const M4 &B)
{
    M4 C{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            zc s = 0;
            for (int k = 0; k < 4; k++) s += A[i][k] * B[k][j];
            C[i][j] = s;
        }
    return C;
}

// The reference's propagator (below) multiplies layer matrices whose cosh and
// sinh grow like exp(k h); the fields it returns are ratios of those products.
// A column whose summed decay sqrt(w mu a / 2) h (a: the smaller horizontal
// conductivity of each layer, after the sigma_z reduction) passes ~700
// overflows and returns NaN, which models with tens of km of deep layers do
// at a few hundred Hz.  Below a summed decay of kColumnDecay the fields are
// zero to double precision relative to the top (e^-40 ~ 4e-18), so the column
// is cut there: the layer that crosses it is thinned to reach it exactly and
// continued as the half-space below, the nodes under the cut get zero fields
// and imp refers to the cut bottom.  Columns that do not reach it (every
// column of the benchmark workloads) are solved unchanged.
static const double kColumnDecay = 40.0;

static void mt1d_column(int nz, const double *conv, const double *c, double freq, zc *ex, zc *hx, zc *imp,
                        zc *exc);

void mt1d_anl(int nz, const double *conv, const double *c, double freq, zc *ex, zc *hx, zc *imp, zc *exc)
{
    const double w = 2 * kPi * freq;
    auto C = [&](int layer, int comp) { return conv[(size_t)comp * nz + layer]; };
    double decay = 0;
    int ne = nz;
    double cut_h = 0;
    for (int L = 0; L < nz && ne == nz; L++) {
        const double axx = C(L, 0) - C(L, 4) * C(L, 4) / C(L, 2);
        const double ayy = C(L, 1) - C(L, 5) * C(L, 5) / C(L, 2);
        const double axy = C(L, 3) - C(L, 4) * C(L, 5) / C(L, 2);
        const double amin = (axx + ayy) / 2 - std::sqrt((axx - ayy) * (axx - ayy) / 4 + axy * axy);
        const double kr = std::sqrt(w * kMu0 * std::max(amin, 0.0) / 2);
        if (decay + kr * c[L] > kColumnDecay) {
            ne = L + 1;
            cut_h = (kColumnDecay - decay) / kr;
        }
        decay += kr * c[L];
    }
    if (ne == nz) {
        mt1d_column(nz, conv, c, freq, ex, hx, imp, exc);
        return;
    }
    std::vector<double> cv((size_t)6 * ne), cc(c, c + ne);
    for (int q = 0; q < 6; q++)
        for (int L = 0; L < ne; L++) cv[(size_t)q * ne + L] = C(L, q);
    cc[ne - 1] = cut_h;
    std::vector<zc> e(2 * ne + 2), h(2 * ne + 2), x(2 * ne + 2);
    mt1d_column(ne, cv.data(), cc.data(), freq, e.data(), h.data(), imp, x.data());
    for (int mode = 0; mode < 2; mode++)
        for (int n = 0; n <= nz; n++) {
            const size_t o = (size_t)mode * (nz + 1) + n, t = (size_t)mode * (ne + 1) + n;
            ex[o] = n <= ne ? e[t] : zc(0);
            if (hx) hx[o] = n <= ne ? h[t] : zc(0);
            if (exc) exc[o] = n <= ne ? x[t] : zc(0);
        }
}

static void mt1d_column(int nz, const double *conv, const double *c, double freq, zc *ex, zc *hx, zc *imp, zc *exc)
{
    const double w = 2 * kPi * freq;
    const zc iwm(0.0, w * kMu0);
    const zc root = std::sqrt(-iwm);
    std::vector<M4> prod(nz + 1);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) prod[nz][i][j] = (i == j) ? 1.0 : 0.0;
    M4 basement{};
    auto C = [&](int layer, int comp) { return conv[(size_t)comp * nz + layer]; };
    for (int L = nz - 1; L >= 0; L--) {
        const double axx = C(L, 0) - C(L, 4) * C(L, 4) / C(L, 2);
        const double ayy = C(L, 1) - C(L, 5) * C(L, 5) / C(L, 2);
        const double axy = C(L, 3) - C(L, 4) * C(L, 5) / C(L, 2);
        const double sum = axx + ayy, dif = axx - ayy;
        const double disc = std::copysign(std::sqrt(dif * dif + 4 * axy * axy), dif >= 0 ? 1.0 : -1.0);
        const double a1 = (sum + disc) / 2, a2 = (sum - disc) / 2;
        const zc kp = root * std::sqrt(zc(a1, 0)), km = root * std::sqrt(zc(a2, 0));
        const zc xp = -kp / iwm, xm = -km / iwm;
        double qp = 0, qm = 0;
        if (axy != 0) {
            qp = 2 * axy / (dif + disc);
            qm = (dif - disc) / axy / 2;
        }
        const double qq = qp * qm, dq = 1 - qq;
        if (L == nz - 1) {
            basement[0][1] = 1.0; basement[1][1] = qp; basement[2][1] = -xp * qp; basement[3][1] = xp;
            basement[0][3] = qm;  basement[1][3] = 1.0; basement[2][3] = -xm; basement[3][3] = xm * qm;
        }
        const zc ep = std::exp(kp * c[L]), em = std::exp(km * c[L]);
        const zc sp = (ep - 1.0 / ep) / 2.0, cp = (ep + 1.0 / ep) / 2.0;
        const zc sm = (em - 1.0 / em) / 2.0, cm = (em + 1.0 / em) / 2.0;
        M4 s = {{{cp - qq * cm, -qm * (cp - cm), qm * (sp / xp - sm / xm), sp / xp - qq * sm / xm},
                 {qp * (cp - cm), -qq * cp + cm, qq * sp / xp - sm / xm, qp * (sp / xp - sm / xm)},
                 {-qp * (xp * sp - xm * sm), qq * xp * sp - xm * sm, -qq * cp + cm, -qp * (cp - cm)},
                 {xp * sp - qq * xm * sm, -qm * (xp * sp - xm * sm), qm * (cp - cm), cp - qq * cm}}};
        for (auto &row : s)
            for (auto &v : row) v /= dq;
        prod[L] = mul4(s, prod[L + 1]);
    }
    const M4 top = mul4(prod[0], basement);
    zc bottom[8];
    for (int mode = 0; mode < 2; mode++) {
        const zc e0 = mode == 0 ? zc(1, 0) : zc(0, 1);
        const zc g11 = top[0][1], g21 = top[1][1], g12 = top[0][3], g22 = top[1][3];
        const zc det = g11 * g22 - g12 * g21;
        const zc cb[4] = {0.0, (g22 * e0.real() - g12 * e0.imag()) / det, 0.0, (-g21 * e0.real() + g11 * e0.imag()) / det};
        zc mc[4];
        for (int r = 0; r < 4; r++) {
            mc[r] = 0;
            for (int q = 0; q < 4; q++) mc[r] += basement[r][q] * cb[q];
        }
        for (int n = 0; n <= nz; n++) {
            zc f[4];
            for (int r = 0; r < 4; r++) {
                f[r] = 0;
                for (int q = 0; q < 4; q++) f[r] += prod[n][r][q] * mc[q];
            }
            if (mode == 0) { ex[n] = f[0]; if (hx) hx[n] = f[3]; if (exc) exc[n] = f[1]; }
            else { ex[nz + 1 + n] = f[1]; if (hx) hx[nz + 1 + n] = f[2]; if (exc) exc[nz + 1 + n] = f[0]; }
            if (n == nz)
                for (int r = 0; r < 4; r++) bottom[4 * mode + r] = f[r];
        }
    }
    if (imp) {
        const zc t = bottom[2] * bottom[7] - bottom[3] * bottom[6];
        imp[0] = (bottom[0] * bottom[7] - bottom[4] * bottom[3]) / t;
        imp[1] = (bottom[4] * bottom[2] - bottom[0] * bottom[6]) / t;
        imp[2] = (bottom[1] * bottom[7] - bottom[5] * bottom[3]) / t;
        imp[3] = (bottom[5] * bottom[2] - bottom[1] * bottom[6]) / t;
    }
}

// Background column of a model: cell (1,1) of every layer, diagonal only
// (sfd_excmg/src/fdfe_main.f90:938-942): 2*nz+2 values.  With full, the whole
// tensor of that column, and the cross-polar components appended (4*nz+4
// values, the layout dirichlet_value reads with full).
std::vector<zc> background_field(const HostLevel &L, double freq, bool full)
{
    const int nz = L.nz;
    std::vector<double> conv((size_t)6 * nz, 0.0);
    const int ncomp = full ? 6 : 3;
    for (int k = 0; k < nz; k++)
        for (int q = 0; q < ncomp; q++)
            conv[(size_t)q * nz + k] = L.S()[(size_t)q * L.ncell() + (size_t)k * L.nx * L.ny];
    std::vector<zc> e(full ? 4 * nz + 4 : 2 * nz + 2);
    mt1d_anl(nz, conv.data(), L.c.data(), freq, e.data(), nullptr, nullptr, full ? e.data() + 2 * nz + 2 : nullptr);
    return e;
}

// ------------------------------------------------------------------ coarsening
// Volume-weighted average over 2^k blocks (sfd_excmg/src/fdfe_main.f90:18-84,
// with the fine-cell index taken inside the block).
HostLevel coarsen(const HostLevel &F, int k)
{
    const int f = 1 << k;
    HostLevel C;
    C.nx = F.nx / f; C.ny = F.ny / f; C.nz = F.nz / f; C.nair = F.nair / f;
    C.a.assign(C.nx, 0.0); C.b.assign(C.ny, 0.0); C.c.assign(C.nz, 0.0);
    for (int i = 0; i < C.nx; i++) for (int t = 0; t < f; t++) C.a[i] += F.a[i * f + t];
    for (int i = 0; i < C.ny; i++) for (int t = 0; t < f; t++) C.b[i] += F.b[i * f + t];
    for (int i = 0; i < C.nz; i++) for (int t = 0; t < f; t++) C.c[i] += F.c[i * f + t];
    const size_t ncf = F.ncell(), ncc = C.ncell();
    C.sig.assign(6 * ncc, 0.0);
    for (int K = 0; K < C.nz; K++)
        for (int J = 0; J < C.ny; J++)
            for (int Ii = 0; Ii < C.nx; Ii++) {
                double acc[6] = {0, 0, 0, 0, 0, 0};
                for (int r = 0; r < f; r++)
                    for (int s = 0; s < f; s++)
                        for (int t = 0; t < f; t++) {
                            const int fi = Ii * f + t, fj = J * f + s, fk = K * f + r;
                            const size_t ie = ((size_t)fk * F.ny + fj) * F.nx + fi;
                            const double v = F.a[fi] * F.b[fj] * F.c[fk];
                            for (int q = 0; q < 6; q++) acc[q] += F.S()[q * ncf + ie] * v;
                        }
                const double V = C.a[Ii] * C.b[J] * C.c[K];
                const size_t ic = ((size_t)K * C.ny + J) * C.nx + Ii;
                for (int q = 0; q < 6; q++) C.sig[q * ncc + ic] = acc[q] / V;
            }
    return C;
}

// ------------------------------------------------------------------ topography
// The reference's topography branch of post_process (src/fdfe_main.f90:604-614,
// 656-666) takes, per column, the first cell k1 whose centre lies below the
// ground.  With cells above the ground carved to air that is the first
// non-air cell, and its top node plane is the ground surface.
void surface_planes(const HostLevel &L, double air_max, std::vector<int> &surf)
{
    const size_t nxy = (size_t)L.nx * L.ny;
    surf.assign(nxy, L.nz);
    for (size_t s = 0; s < nxy; s++)
        for (int k = 0; k < L.nz; k++)
            if (L.S()[(size_t)k * nxy + s] > air_max) {
                surf[s] = k;
                break;
            }
}

// ------------------------------------------------------------------ EXCMG tables
// Lagrange basis of degree n on points xs (src/excmg.f90:754-772).
static double lagrange(int n, const double *xs, int id, double x)
{
    double num = 1, den = 1;
    for (int q = 0; q <= n; q++)
        if (q != id) { num *= x - xs[q]; den *= xs[id] - xs[q]; }
    return num / den;
}

// Node coordinates the reference interpolation uses (src/excmg.f90:367-378):
// x and y centred, z measured from the top of the air.
static std::vector<double> coords(const std::vector<double> &w, double origin)
{
    std::vector<double> x(w.size() + 1);
    x[0] = origin;
    for (size_t i = 0; i < w.size(); i++) x[i + 1] = x[i] + w[i];
    return x;
}

// One dimension of the block prolongation with merge m on the fine widths w:
// blocks of 2m fine cells = m coarse cells.  Transverse (node) weights use the
// m+1 coarse nodes of the block; along-edge weights use the m points placed, as
// cal_coef does (src/excmg.f90:683-685), at the midpoints of fine edges
// 0,2,..,2m-2 of the block.
void prolong_tables_1d(const std::vector<double> &w, double origin, int m, std::vector<int> &nbase,
                       std::vector<double> &nw, std::vector<int> &ebase, std::vector<double> &ew)
{
    const int n = (int)w.size(), B = 2 * m, nblk = n / B;
    const std::vector<double> x = coords(w, origin);
    nbase.assign(n + 1, 0); nw.assign((size_t)(n + 1) * 5, 0.0);
    ebase.assign(n, 0); ew.assign((size_t)n * 4, 0.0);
    for (int q = 0; q <= n; q++) {
        const int blk = std::min(q / B, nblk - 1);
        double cn[5];
        for (int t = 0; t <= m; t++) cn[t] = x[blk * B + 2 * t];
        nbase[q] = blk * m;
        for (int t = 0; t <= m; t++) nw[(size_t)q * 5 + t] = lagrange(m, cn, t, x[q]);
    }
    for (int q = 0; q < n; q++) {
        const int blk = q / B;
        double ce[4];
        for (int t = 0; t < m; t++) ce[t] = (x[blk * B + 2 * t] + x[blk * B + 2 * t + 1]) / 2;
        ebase[q] = blk * m;
        const double mid = (x[q] + x[q + 1]) / 2;
        for (int t = 0; t < m; t++) ew[(size_t)q * 4 + t] = lagrange(m - 1, ce, t, mid);
    }
}

void prolong_tables(const HostLevel &fine, int m, ProlongHost &T)
{
    double sa = 0, sb = 0, sc = 0;
    for (double v : fine.a) sa += v;
    for (double v : fine.b) sb += v;
    for (int i = 0; i < fine.nair; i++) sc += fine.c[i];
    prolong_tables_1d(fine.a, -sa / 2, m, T.nb[0], T.nw[0], T.eb[0], T.ew[0]);
    prolong_tables_1d(fine.b, -sb / 2, m, T.nb[1], T.nw[1], T.eb[1], T.ew[1]);
    prolong_tables_1d(fine.c, -sc, m, T.nb[2], T.nw[2], T.eb[2], T.ew[2]);
}

// ------------------------------------------------------------------ coarse level
// CSR (0-based, upper triangle, sorted) of the interior block of A_s and the
// right-hand sides -A_s,ib g for both polarisations.
void coarse_system(const HostLevel &L, double freq, const std::vector<zc> &edc, CoarseSystem &S)
{
    const Dims d = make_dims(L.nx, L.ny, L.nz);
    const double wmu = 2 * kPi * freq * kMu0;
    const bool full = edc.size() == (size_t)(4 * L.nz + 4);  // full-tensor boundary data
    S.imap.assign(d.NL, -1);
    S.inner.clear();
    for (int64_t e = 0; e < d.NL; e++) {
        int i, j, k;
        int o = edge_decode(d, e, i, j, k);
        if (edge_interior(d, o, i, j, k)) { S.imap[e] = (int)S.inner.size(); S.inner.push_back(e); }
    }
    const int n = (int)S.inner.size();
    S.n = n;
    S.ia.assign(n + 1, 0); S.ja.clear(); S.val.clear();
    S.rhs.assign((size_t)2 * n, 0.0);
    std::vector<std::pair<int, zc>> row;
    for (int r = 0; r < n; r++) {
        const int64_t e = S.inner[r];
        int i, j, k;
        int o = edge_decode(d, e, i, j, k);
        Row<double> rw;
        row_coeffs(d, L.a.data(), L.b.data(), L.c.data(), L.S(), o, i, j, k, rw);
        row.clear();
        row.push_back({r, zc(rw.kgd, -wmu * rw.kmd)});
        for (int s = 0; s < 12; s++) {
            const zc cf(rw.kg[s], -wmu * rw.km[s]);
            const int64_t nb = rw.nb[s];
            if (S.imap[nb] >= 0) {
                if (S.imap[nb] >= r) row.push_back({S.imap[nb], cf});
            } else {
                int ii, jj, kk;
                int oo = edge_decode(d, nb, ii, jj, kk);
                for (int pol = 0; pol < 2; pol++) {
                    zc g(0);
                    dirichlet_value(d, edc.data(), pol, oo, ii, jj, kk, g, full);
                    S.rhs[(size_t)pol * n + r] -= cf * g;
                }
            }
        }
        std::sort(row.begin(), row.end(), [](const std::pair<int, zc> &x, const std::pair<int, zc> &y) { return x.first < y.first; });
        for (auto &pr : row) { S.ja.push_back(pr.first); S.val.push_back(pr.second); }
        S.ia[r + 1] = (int)S.ja.size();
    }
}

// MKL PARDISO bound at run time (libmkl_rt); mtype 6 = complex symmetric.
typedef void (*pardiso_fn)(void *, const int *, const int *, const int *, const int *, const int *, const void *,
                           const int *, const int *, int *, const int *, int *, const int *, void *, void *, int *);
typedef int (*mkl_int_fn)(int);
typedef void (*mkl_void_fn)(int);

// PARDISO's parallel factorisation sums its updates in an order that depends on
// the number of threads (measured: the 78,936-unknown coarse level of the
// 128x128x96 model moves by 2e-7 relative between 1, 3 and 8 threads), and
// MKL's dynamic threading may pick fewer threads than requested under load.
// The coarse solutions seed the EXCMG initial guess, so they are pinned: the
// dynamic adjustment is switched off and every solve runs on a fixed number
// of threads set thread-locally (SFDX_PARDISO_THREADS, default min(16, cores)),
// which makes repeated solves on one machine bit-identical whatever the
// caller's MKL_NUM_THREADS or load.
struct MklFns {
    pardiso_fn pardiso = nullptr;
    mkl_int_fn set_num_threads_local = nullptr;
    int threads = 1;
};

static const MklFns *load_mkl(std::string &err)
{
    static std::mutex mu;  // batched solves call this from several threads
    std::lock_guard<std::mutex> lock(mu);
    static MklFns fns;
    static std::string cached_err;
    static bool tried = false;
    if (tried) {
        err = cached_err;
        return fns.pardiso ? &fns : nullptr;
    }
    tried = true;
    std::vector<std::string> names;
    if (const char *env = std::getenv("SFDX_MKL_RT")) names.push_back(env);
    names.push_back("libmkl_rt.so.1");
    names.push_back("libmkl_rt.so");
    names.push_back("/opt/conda/lib/libmkl_rt.so.1");
    for (auto &nm : names) {
        void *h = dlopen(nm.c_str(), RTLD_NOW | RTLD_GLOBAL);
        if (!h) continue;
        fns.pardiso = (pardiso_fn)dlsym(h, "pardiso");
        if (!fns.pardiso) continue;
        // the mixed-case names are MKL's C entry points (the lower- and upper-case
        // ones take their arguments by reference)
        fns.set_num_threads_local = (mkl_int_fn)dlsym(h, "MKL_Set_Num_Threads_Local");
        if (mkl_void_fn set_dynamic = (mkl_void_fn)dlsym(h, "MKL_Set_Dynamic")) set_dynamic(0);
        int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
        if (const char *env = std::getenv("SFDX_PARDISO_THREADS")) nt = std::max(1, std::atoi(env));
        fns.threads = nt;
        return &fns;
    }
    cached_err = "MKL PARDISO (libmkl_rt) not found; set SFDX_MKL_RT";
    err = cached_err;
    return nullptr;
}

int direct_solve_threads()
{
    std::string err;
    const MklFns *m = load_mkl(err);
    return m ? m->threads : 0;
}

bool direct_solve(const CoarseSystem &S, std::vector<zc> &sol, std::string &err)
{
    const MklFns *mkl = load_mkl(err);
    if (!mkl) return false;
    pardiso_fn pardiso = mkl->pardiso;
    // fixed thread count for this thread's MKL calls (restored on return)
    const int prev_threads = mkl->set_num_threads_local ? mkl->set_num_threads_local(mkl->threads) : 0;
    struct Restore {
        const MklFns *m;
        int prev;
        ~Restore()
        {
            if (m->set_num_threads_local) m->set_num_threads_local(prev);
        }
    } restore{mkl, prev_threads};
    void *pt[64];
    std::memset(pt, 0, sizeof(pt));
    int iparm[64];
    std::memset(iparm, 0, sizeof(iparm));
    iparm[0] = 1;   // explicit parameters
    iparm[1] = 2;   // nested dissection (METIS)
    iparm[7] = 2;   // iterative refinement steps
    iparm[9] = 8;   // pivot perturbation 1e-8
    iparm[34] = 1;  // 0-based ia/ja
    const int maxfct = 1, mnum = 1, mtype = 6, nrhs = 2, msglvl = 0;
    int phase = 12, error = 0, idum = 0;
    const int n = S.n;
    pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, S.val.data(), S.ia.data(), S.ja.data(), &idum, &nrhs, iparm,
            &msglvl, nullptr, nullptr, &error);
    bool ok = error == 0;
    if (ok) {
        sol.assign((size_t)2 * n, 0.0);
        phase = 33;
        std::vector<zc> b = S.rhs;
        pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, S.val.data(), S.ia.data(), S.ja.data(), &idum, &nrhs,
                iparm, &msglvl, b.data(), sol.data(), &error);
        ok = error == 0;
    }
    if (!ok) err = "PARDISO error " + std::to_string(error);
    phase = -1;
    int e2 = 0;
    pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, nullptr, S.ia.data(), S.ja.data(), &idum, &nrhs, iparm, &msglvl,
            nullptr, nullptr, &e2);
    return ok;
}

// ------------------------------------------------------------------ post-process
// Surface E and H at cell centres (src/fdfe_main.f90:597-683) and biquadratic
// interpolation to stations (arho_interp2, src/fdfe_main.f90:718-790), then
// Z = E H^-1 from the two polarisations (src/fdfe_main.f90:687-705).
// The surface of column (i,j) is node plane m = g.plane(i,j): nair on the flat
// datum; with topography the top of the column's first earth cell.  The
// one-sided z difference spans the cell below the surface, c[m] (the
// reference's c(nair+1) on the flat datum; it keeps c(nair+1) under
// topography too, and takes the plane below the first earth cell's top,
// k1 instead of k1-1, which on a flat surface would be one plane under the
// datum it uses without topography).
struct Surface {
    int nx, ny;
    std::vector<zc> f[5][2];  // ex, ey, hx, hy, hz for each polarisation, index i + nx*j
};

static void surface_fields(const PostGrid &g, double freq, const zc *x, int pol, Surface &S)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t n1 = (int64_t)nx * (ny + 1) * (nz + 1), n2 = (int64_t)ny * (nz + 1) * (nx + 1);
    const zc kap(0.0, 2 * kPi * freq * kMu0);
    for (int q = 0; q < 5; q++) S.f[q][pol].assign((size_t)nx * ny, 0.0);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int m = g.plane(i, j);
            const int64_t idx = (int64_t)m * nx * (ny + 1) + (int64_t)j * nx + i;
            const int64_t idx2 = n1 + (int64_t)m * ny * (nx + 1) + (int64_t)j * (nx + 1) + i;
            const int64_t idx3 = n1 + n2 + (int64_t)m * (nx + 1) * (ny + 1) + (int64_t)j * (nx + 1) + i;
            const zc ex = (x[idx] + x[idx + nx]) / 2.0;
            const zc ey = (x[idx2] + x[idx2 + 1]) / 2.0;
            const double cz = g.c[m];
            const zc exz = ((x[idx + (int64_t)nx * (ny + 1)] + x[idx + (int64_t)nx * (ny + 1) + nx]) / 2.0 - ex) / cz;
            const zc ezx = ((x[idx3 + 1] + x[idx3 + nx + 2]) / 2.0 - (x[idx3 + nx + 1] + x[idx3]) / 2.0) / g.a[i];
            const zc eyz = ((x[idx2 + (int64_t)ny * (nx + 1)] + x[idx2 + (int64_t)ny * (nx + 1) + 1]) / 2.0 - ey) / cz;
            const zc ezy = (-(x[idx3] + x[idx3 + 1]) / 2.0 + (x[idx3 + nx + 1] + x[idx3 + nx + 2]) / 2.0) / g.b[j];
            const zc exy = (x[idx + nx] - x[idx]) / g.b[j];
            const zc eyx = (x[idx2 + 1] - x[idx2]) / g.a[i];
            const size_t s = (size_t)i + (size_t)nx * j;
            S.f[0][pol][s] = ex;
            S.f[1][pol][s] = ey;
            S.f[2][pol][s] = (ezy - eyz) / kap;
            S.f[3][pol][s] = (exz - ezx) / kap;
            S.f[4][pol][s] = (exy - eyx) / kap;
        }
}

static void quad3(double x, const double *x0, double *w)
{
    w[0] = (x - x0[1]) * (x - x0[2]) / (x0[0] - x0[1]) / (x0[0] - x0[2]);
    w[1] = (x - x0[0]) * (x - x0[2]) / (x0[1] - x0[0]) / (x0[1] - x0[2]);
    w[2] = (x - x0[0]) * (x - x0[1]) / (x0[2] - x0[0]) / (x0[2] - x0[1]);
}

static int station_interp(const PostGrid &g, const Surface &S, int pol, int rd, const double *recv, int nrec,
                          std::vector<zc> &out)
{
    const int nx = g.nx, ny = g.ny;
    std::vector<double> ax(nx), by(ny);
    double sa = 0, sb = 0;
    for (int i = 0; i < nx; i++) sa += g.a[i];
    for (int i = 0; i < ny; i++) sb += g.b[i];
    double acc = 0;
    for (int i = 0; i < nx; i++) { ax[i] = -sa / 2 + acc + g.a[i] / 2; acc += g.a[i]; }
    acc = 0;
    for (int i = 0; i < ny; i++) { by[i] = -sb / 2 + acc + g.b[i] / 2; acc += g.b[i]; }
    out.assign((size_t)nrec * 5, 0.0);
    for (int r = 0; r < nrec; r++) {
        const double xr = recv[(size_t)r * 3], yr = recv[(size_t)r * 3 + 1];
        int ir = -1, jr = -1;
        for (int q = 0; q + 1 < nx; q++)
            if (xr >= ax[q] && xr <= ax[q + 1]) { ir = q; break; }
        for (int q = 0; q + 1 < ny; q++)
            if (yr >= by[q] && yr <= by[q + 1]) { jr = q; break; }
        if (ir < 1 || jr < 1 || ir + 1 >= nx || jr + 1 >= ny) return r + 1;  // station outside usable area
        double w1[3], w2[3];
        quad3(xr, &ax[ir - 1], w1);
        quad3(yr, &by[jr - 1], w2);
        for (int mm = 0; mm < 3; mm++)
            for (int nn = 0; nn < 3; nn++) {
                const double wt = w1[mm] * w2[nn];
                int I = ir - 1 + mm, J = jr - 1 + nn;  // rd = 1
                if (rd == 2) { I = jr - 1 + mm; J = ir - 1 + nn; }
                const int Ih = ir - 1 + mm, Jh = jr - 1 + nn;
                if (I >= nx || J >= ny) return r + 1;
                for (int q = 0; q < 4; q++) out[(size_t)r * 5 + q] += wt * S.f[q][pol][(size_t)I + (size_t)nx * J];
                out[(size_t)r * 5 + 4] += wt * S.f[4][pol][(size_t)Ih + (size_t)nx * Jh];
            }
    }
    return 0;
}

// Station responses of one frequency from its surface fields (both polarisations).
static int respond(const PostGrid &g, double freq, const Surface &S, int nrec, const double *recv, int out_type, int rd,
                   double *results, double *tipper)
{
    std::vector<zc> s1, s2;
    int st = station_interp(g, S, 0, rd, recv, nrec, s1);
    if (!st) st = station_interp(g, S, 1, rd, recv, nrec, s2);
    if (st) return st;
    const double om = 2 * kPi * freq * kMu0;
    for (int r = 0; r < nrec; r++) {
        const zc *A1 = &s1[(size_t)r * 5], *A2 = &s2[(size_t)r * 5];
        const zc t = A1[2] * A2[3] - A2[2] * A1[3];
        const zc zxy = (A2[0] * A1[2] - A1[0] * A2[2]) / t;
        const zc zyx = (A1[1] * A2[3] - A2[1] * A1[3]) / t;
        double *o = results + (size_t)r * 4;
        if (out_type == 2) {
            o[0] = std::abs(zxy * zxy * zc(0, 1)) / om;
            o[1] = std::abs(zyx * zyx * zc(0, 1)) / om;
            o[2] = -std::atan(zxy.imag() / zxy.real()) * 180.0 / kPi;
            o[3] = -std::atan(zyx.imag() / zyx.real()) * 180.0 / kPi;
        } else {
            o[0] = zxy.real(); o[1] = zxy.imag(); o[2] = zyx.real(); o[3] = zyx.imag();
        }
        if (tipper) {
            const zc tx = (A2[3] * A1[4] - A1[3] * A2[4]) / t;
            const zc ty = (A1[2] * A2[4] - A2[2] * A1[4]) / t;
            double *tp = tipper + (size_t)r * 4;
            tp[0] = tx.real(); tp[1] = tx.imag(); tp[2] = ty.real(); tp[3] = ty.imag();
        }
    }
    return 0;
}

int post_process(const PostGrid &g, int nf, const double *freq, int nrec, const double *recv, const zc *ex,
                 int out_type, int rd, double *results, double *tipper)
{
    const Dims d = make_dims(g.nx, g.ny, g.nz);
    for (int k = 0; k < nf; k++) {
        const zc *x1 = ex + (size_t)k * 2 * d.NL;
        const zc *x2 = x1 + d.NL;
        Surface S;
        surface_fields(g, freq[k], x1, 0, S);
        surface_fields(g, freq[k], x2, 1, S);
        const int st = respond(g, freq[k], S, nrec, recv, out_type, rd, results + (size_t)k * nrec * 4,
                               tipper ? tipper + (size_t)k * nrec * 4 : nullptr);
        if (st) return st;
    }
    return 0;
}

int post_process_surface(const PostGrid &g, int nf, const double *freq, int nrec, const double *recv, const zc *surf,
                         int out_type, int rd, double *results, double *tipper)
{
    const size_t ns = (size_t)g.nx * g.ny;
    for (int k = 0; k < nf; k++) {
        Surface S;
        for (int pol = 0; pol < 2; pol++)
            for (int q = 0; q < 5; q++) {
                const zc *src = surf + (((size_t)k * 2 + pol) * 5 + q) * ns;
                S.f[q][pol].assign(src, src + ns);
            }
        const int st = respond(g, freq[k], S, nrec, recv, out_type, rd, results + (size_t)k * nrec * 4,
                               tipper ? tipper + (size_t)k * nrec * 4 : nullptr);
        if (st) return st;
    }
    return 0;
}

// ------------------------------------------------------------------ misfit and its field gradient
// Transposes of surface_fields / station_interp: coefficients of dPhi = Re(sum
// g dfield) are scattered back onto the edge values with the same (real or
// complex) weights, unconjugated.
static void station_interp_T(const PostGrid &g, int rd, const double *recv, int nrec, const std::vector<zc> &gs,
                             std::vector<zc> (&gsurf)[5])
{
    const int nx = g.nx, ny = g.ny;
    std::vector<double> ax(nx), by(ny);
    double sa = 0, sb = 0;
    for (int i = 0; i < nx; i++) sa += g.a[i];
    for (int i = 0; i < ny; i++) sb += g.b[i];
    double acc = 0;
    for (int i = 0; i < nx; i++) { ax[i] = -sa / 2 + acc + g.a[i] / 2; acc += g.a[i]; }
    acc = 0;
    for (int i = 0; i < ny; i++) { by[i] = -sb / 2 + acc + g.b[i] / 2; acc += g.b[i]; }
    for (int q = 0; q < 5; q++) gsurf[q].assign((size_t)nx * ny, 0.0);
    for (int r = 0; r < nrec; r++) {
        const double xr = recv[(size_t)r * 3], yr = recv[(size_t)r * 3 + 1];
        int ir = -1, jr = -1;
        for (int q = 0; q + 1 < nx; q++)
            if (xr >= ax[q] && xr <= ax[q + 1]) { ir = q; break; }
        for (int q = 0; q + 1 < ny; q++)
            if (yr >= by[q] && yr <= by[q + 1]) { jr = q; break; }
        double w1[3], w2[3];
        quad3(xr, &ax[ir - 1], w1);
        quad3(yr, &by[jr - 1], w2);
        for (int mm = 0; mm < 3; mm++)
            for (int nn = 0; nn < 3; nn++) {
                const double wt = w1[mm] * w2[nn];
                int I = ir - 1 + mm, J = jr - 1 + nn;
                if (rd == 2) { I = jr - 1 + mm; J = ir - 1 + nn; }
                const int Ih = ir - 1 + mm, Jh = jr - 1 + nn;
                for (int q = 0; q < 4; q++) gsurf[q][(size_t)I + (size_t)nx * J] += wt * gs[(size_t)r * 5 + q];
                gsurf[4][(size_t)Ih + (size_t)nx * Jh] += wt * gs[(size_t)r * 5 + 4];
            }
    }
}

static void surface_fields_T(const PostGrid &g, double freq, const std::vector<zc> (&gs)[5], zc *w)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t n1 = (int64_t)nx * (ny + 1) * (nz + 1), n2 = (int64_t)ny * (nz + 1) * (nx + 1);
    const zc kap(0.0, 2 * kPi * freq * kMu0);
    const int64_t P = (int64_t)nx * (ny + 1), Q = (int64_t)ny * (nx + 1);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const size_t s = (size_t)i + (size_t)nx * j;
            const int m = g.plane(i, j);
            const int64_t idx = (int64_t)m * nx * (ny + 1) + (int64_t)j * nx + i;
            const int64_t idx2 = n1 + (int64_t)m * ny * (nx + 1) + (int64_t)j * (nx + 1) + i;
            const int64_t idx3 = n1 + n2 + (int64_t)m * (nx + 1) * (ny + 1) + (int64_t)j * (nx + 1) + i;
            const double cz = g.c[m], ai = g.a[i], bj = g.b[j];
            // coefficients of the intermediate quantities
            const zc gHx = gs[2][s] / kap, gHy = gs[3][s] / kap, gHz = gs[4][s] / kap;
            const zc g_ezy = gHx, g_eyz = -gHx, g_exz = gHy, g_ezx = -gHy, g_exy = gHz, g_eyx = -gHz;
            const zc g_ex = gs[0][s] - g_exz / cz;   // ex enters ex and exz
            const zc g_ey = gs[1][s] - g_eyz / cz;   // ey enters ey and eyz
            // ex = (x[idx] + x[idx+nx]) / 2
            w[idx] += g_ex / 2.0;
            w[idx + nx] += g_ex / 2.0;
            // ey = (x[idx2] + x[idx2+1]) / 2
            w[idx2] += g_ey / 2.0;
            w[idx2 + 1] += g_ey / 2.0;
            // exz = ((x[idx+P] + x[idx+P+nx]) / 2 - ex) / cz
            w[idx + P] += g_exz / (2.0 * cz);
            w[idx + P + nx] += g_exz / (2.0 * cz);
            // ezx = ((x[idx3+1] + x[idx3+nx+2]) / 2 - (x[idx3+nx+1] + x[idx3]) / 2) / a_i
            w[idx3 + 1] += g_ezx / (2.0 * ai);
            w[idx3 + nx + 2] += g_ezx / (2.0 * ai);
            w[idx3 + nx + 1] -= g_ezx / (2.0 * ai);
            w[idx3] -= g_ezx / (2.0 * ai);
            // eyz = ((x[idx2+Q] + x[idx2+Q+1]) / 2 - ey) / cz
            w[idx2 + Q] += g_eyz / (2.0 * cz);
            w[idx2 + Q + 1] += g_eyz / (2.0 * cz);
            // ezy = (-(x[idx3] + x[idx3+1]) / 2 + (x[idx3+nx+1] + x[idx3+nx+2]) / 2) / b_j
            w[idx3] -= g_ezy / (2.0 * bj);
            w[idx3 + 1] -= g_ezy / (2.0 * bj);
            w[idx3 + nx + 1] += g_ezy / (2.0 * bj);
            w[idx3 + nx + 2] += g_ezy / (2.0 * bj);
            // exy = (x[idx+nx] - x[idx]) / b_j ; eyx = (x[idx2+1] - x[idx2]) / a_i
            w[idx + nx] += g_exy / bj;
            w[idx] -= g_exy / bj;
            w[idx2 + 1] += g_eyx / ai;
            w[idx2] -= g_eyx / ai;
        }
}

int misfit_fields(const PostGrid &g, int nf, const double *freq, int nrec, const double *recv, int rd,
                  const double *obs, const zc *ex, double &misfit, zc *w)
{
    const Dims d = make_dims(g.nx, g.ny, g.nz);
    const double ln10 = std::log(10.0);
    misfit = 0;
    if (w) std::fill(w, w + (size_t)nf * 2 * d.NL, zc(0.0));
    for (int k = 0; k < nf; k++) {
        const zc *x1 = ex + (size_t)k * 2 * d.NL;
        const zc *x2 = x1 + d.NL;
        Surface S;
        surface_fields(g, freq[k], x1, 0, S);
        surface_fields(g, freq[k], x2, 1, S);
        std::vector<zc> s1, s2;
        int st = station_interp(g, S, 0, rd, recv, nrec, s1);
        if (!st) st = station_interp(g, S, 1, rd, recv, nrec, s2);
        if (st) return st;
        const double om = 2 * kPi * freq[k] * kMu0;
        std::vector<zc> g1((size_t)nrec * 5, 0.0), g2((size_t)nrec * 5, 0.0);
        for (int r = 0; r < nrec; r++) {
            const zc *A1 = &s1[(size_t)r * 5], *A2 = &s2[(size_t)r * 5];
            const zc Ex1 = A1[0], Ey1 = A1[1], Hx1 = A1[2], Hy1 = A1[3];
            const zc Ex2 = A2[0], Ey2 = A2[1], Hx2 = A2[2], Hy2 = A2[3];
            const zc t = Hx1 * Hy2 - Hx2 * Hy1;
            const zc N = Ex2 * Hx1 - Ex1 * Hx2, M = Ey1 * Hy2 - Ey2 * Hy1;
            const zc zxy = N / t, zyx = M / t;
            const double *ob = obs + ((size_t)k * nrec + r) * 4;
            zc cz[2];
            const zc Z[2] = {zxy, zyx};
            for (int q = 0; q < 2; q++) {
                const double rho = std::norm(Z[q]) / om;
                const double phi = -std::atan(Z[q].imag() / Z[q].real());
                const double rr = std::log10(rho) - std::log10(ob[q]);
                const double rp = phi - ob[2 + q] * kPi / 180.0;
                misfit += 0.5 * (rr * rr + rp * rp);
                // d log10 rho = (2/ln10) Re(dZ/Z), d phi = -Im(dZ/Z)
                cz[q] = zc(2.0 * rr / ln10, rp) / Z[q];
            }
            if (!w) continue;
            const zc t2 = t * t;
            // d zxy / d(Ex1, Ex2, Hx1, Hx2, Hy1, Hy2); d zyx / d(Ey1, Ey2, Hx1, Hx2, Hy1, Hy2)
            zc *G1 = &g1[(size_t)r * 5], *G2 = &g2[(size_t)r * 5];
            G1[0] += cz[0] * (-Hx2 / t);
            G2[0] += cz[0] * (Hx1 / t);
            G1[2] += cz[0] * (Ex2 / t - N * Hy2 / t2);
            G2[2] += cz[0] * (-Ex1 / t + N * Hy1 / t2);
            G1[3] += cz[0] * (N * Hx2 / t2);
            G2[3] += cz[0] * (-N * Hx1 / t2);
            G1[1] += cz[1] * (Hy2 / t);
            G2[1] += cz[1] * (-Hy1 / t);
            G1[2] += cz[1] * (-M * Hy2 / t2);
            G2[2] += cz[1] * (M * Hy1 / t2);
            G1[3] += cz[1] * (-Ey2 / t + M * Hx2 / t2);
            G2[3] += cz[1] * (Ey1 / t - M * Hx1 / t2);
        }
        if (!w) continue;
        std::vector<zc> gsurf[5];
        station_interp_T(g, rd, recv, nrec, g1, gsurf);
        surface_fields_T(g, freq[k], gsurf, w + (size_t)k * 2 * d.NL);
        station_interp_T(g, rd, recv, nrec, g2, gsurf);
        surface_fields_T(g, freq[k], gsurf, w + (size_t)k * 2 * d.NL + d.NL);
    }
    return 0;
}

}  // namespace sfdx
