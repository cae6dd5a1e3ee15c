// sfdx_internal.h -- types shared by the kernels and the host orchestration.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "sfdx_stencil.h"

namespace sfdx {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R re, im;
    __host__ __device__ cx() : re(0), im(0) {}
    __host__ __device__ cx(R r) : re(r), im(0) {}
    __host__ __device__ cx(R r, R i) : re(r), im(i) {}
    template <typename Q>
    __host__ __device__ explicit cx(const cx<Q> &o) : re((R)o.re), im((R)o.im) {}
};
template <typename R> __host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return cx<R>(a.re + b.re, a.im + b.im); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return cx<R>(a.re - b.re, a.im - b.im); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator-(cx<R> a) { return cx<R>(-a.re, -a.im); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) { return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return cx<R>(s * a.re, s * a.im); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, R s) { return cx<R>(s * a.re, s * a.im); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator/(cx<R> a, cx<R> b)
{
    R den = b.re * b.re + b.im * b.im;
    return cx<R>((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
}
template <typename R> __host__ __device__ __forceinline__ cx<R> conj(cx<R> a) { return cx<R>(a.re, -a.im); }
template <typename R> __host__ __device__ __forceinline__ R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
// a + b*c
template <typename R> __host__ __device__ __forceinline__ cx<R> fma_c(cx<R> a, cx<R> b, cx<R> c)
{
    return cx<R>(a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re);
}

typedef cx<double> zd;
typedef cx<float> zf;

#define SFDX_CHECK(x)                                                                \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) throw sfdx::HipError(e_, #x, __FILE__, __LINE__);      \
    } while (0)

struct HipError {
    hipError_t code;
    std::string what;
    HipError(hipError_t c, const char *expr, const char *file, int line);
};

// Geometry of one grid level on the device (precision R for the kernels).
template <typename R>
struct DevGrid {
    Dims d;
    R *a = nullptr, *b = nullptr, *c = nullptr;     // widths (device)
    R *ra = nullptr, *rb = nullptr, *rc = nullptr;  // reciprocal widths (device)
    R *sig = nullptr;                               // 6*NC (device)
    // complex128 levels: sigma as three 16-byte pieces per cell, (xx,yy),
    // (zz,xy), (xz,yz), each a plane of NC pieces (the cell-plane march reads it)
    const double *sigp = nullptr;
    int spmv_variant = 0;                           // SpMV kernel variant (launch_spmv_march)
    bool ssor_xmass = true;                         // launch_ssor: with the cross-orientation mass terms
};

// Reduction scratch: partial sums per block, final slots.
struct RedBuf {
    double *partials = nullptr;  // [max_blocks * 8]
    int64_t max_blocks = 0;
    double *slots = nullptr;     // [64] device scalars
    double *host_slots = nullptr;
};

// Launch geometry for edge kernels: one y-block per (orientation, k) slice.
struct SliceLaunch {
    dim3 grid, block;
    int nslices;
    int64_t nblocks() const { return (int64_t)grid.x * grid.y; }
};

SliceLaunch edge_launch(const Dims &d);
SliceLaunch node_launch(const Dims &d);
SliceLaunch colour_launch(const Dims &d, int colour);

}  // namespace sfdx
