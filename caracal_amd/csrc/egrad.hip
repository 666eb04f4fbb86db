// Drop-in analytic surfaces: the bead-batched plugin contract of the
// reference (egrad_h3, src/egrad_h3.f:29-77; egrad_ch4h, src/egrad_ch4h.f:74-129;
// egrad_oh3, src/egrad_oh3.f:33-190; egrad_brh2, src/egrad_brh2.f:33-79;
// egrad_o3, src/egrad_o3.f:33-56).
// One lane per bead.  A block stages its 256 beads through LDS so the global
// reads and writes are coalesced although each bead's coordinates are one
// contiguous 3*natoms-double record.
#include <mutex>

#include "common.hpp"
#include "pes_brh2.hpp"
#include "pes_ch4h.hpp"
#include "pes_h3.hpp"
#include "pes_o3.hpp"
#include "pes_oh3.hpp"

namespace cam {
namespace {

constexpr int kBlock = 256;

// Surfaces evaluated one lane per bead.
template <typename T>
struct LaneH3 {
    static constexpr int NC = 9;
    __device__ static void egrad(const double *x, double &v, double *g) { h3::egrad_bead<T>(x, v, g); }
};
struct LaneOH3 {
    static constexpr int NC = 12;
    __device__ static void egrad(const double *x, double &v, double *g) { oh3::egrad_bead(x, v, g); }
};
struct LaneBrH2 {
    static constexpr int NC = 9;
    __device__ static void egrad(const double *x, double &v, double *g) { brh2::egrad_bead(x, v, g); }
};
struct LaneO3 {
    static constexpr int NC = 9;
    __device__ static void egrad(const double *x, double &v, double *g) { o3::egrad_bead(x, v, g); }
};

constexpr int surface_natoms(int surf)
{
    return surf == CAM_SURFACE_H3 || surf == CAM_SURFACE_BRH2 || surf == CAM_SURFACE_O3 ? 3
           : surf == CAM_SURFACE_OH3                                                  ? 4
                                                                                      : 6;
}

template <typename P>
__global__ __launch_bounds__(kBlock) void egrad_lane_kernel(const double *__restrict__ q, int nbeads,
                                                            double *__restrict__ v, double *__restrict__ g)
{
    constexpr int NC = P::NC;
    __shared__ double tile[kBlock * NC];
    const long b0 = (long)blockIdx.x * kBlock;
    const int nb = (int)min((long)kBlock, (long)nbeads - b0);
    for (int i = threadIdx.x; i < nb * NC; i += kBlock) tile[i] = q[b0 * NC + i];
    __syncthreads();
    double x[NC], gr[NC], vb = 0.0;
    const bool act = (int)threadIdx.x < nb;
    if (act) {
#pragma unroll
        for (int c = 0; c < NC; ++c) x[c] = tile[threadIdx.x * NC + c];
        P::egrad(x, vb, gr);
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int c = 0; c < NC; ++c) tile[threadIdx.x * NC + c] = gr[c];
        v[b0 + threadIdx.x] = vb;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb * NC; i += kBlock) g[b0 * NC + i] = tile[i];
}

// CH4H: four lanes per bead (see ch4h::egrad_quad); a block holds 64 beads.
template <typename T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void egrad_ch4h_kernel(const double *__restrict__ q, int nbeads,
                                                            double *__restrict__ v, double *__restrict__ g)
{
    constexpr int NC = 18, BPB = kBlock / 4;
    __shared__ double tile[BPB * NC];
    const long b0 = (long)blockIdx.x * BPB;
    const int nb = (int)min((long)BPB, (long)nbeads - b0);
    for (int i = threadIdx.x; i < nb * NC; i += kBlock) tile[i] = q[b0 * NC + i];
    __syncthreads();
    const int bead = threadIdx.x >> 2, r = threadIdx.x & 3;
    const bool act = bead < nb;
    const int bl = act ? bead : 0;                       // idle quads evaluate bead 0 of the block
    const int ah = ch4h::hidx(r);
    const int ax = r == 0 ? ch4h::kC : ch4h::kB;
    double own[6], gown[6], vb;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        own[d] = tile[bl * NC + 3 * ah + d];
        own[3 + d] = r < 2 ? tile[bl * NC + 3 * ax + d] : 0.0;
    }
    ch4h::egrad_quad<T>(own, vb, gown);
    __syncthreads();
    if (act) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            tile[bead * NC + 3 * ah + d] = gown[d];
            if (r < 2) tile[bead * NC + 3 * ax + d] = gown[3 + d];
        }
        if (r == 0) v[b0 + bead] = vb;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb * NC; i += kBlock) g[b0 * NC + i] = tile[i];
}

template <int SURF>
int launch_egrad(const double *q, int nbeads, double *v, double *g, int precision, hipStream_t s)
{
    if (nbeads == 0) return CAM_OK;
    if constexpr (SURF == CAM_SURFACE_H3) {
        const int grid = (nbeads + kBlock - 1) / kBlock;
        if (precision == CAM_PREC_MIXED)
            hipLaunchKernelGGL(egrad_lane_kernel<LaneH3<float>>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
        else
            hipLaunchKernelGGL(egrad_lane_kernel<LaneH3<double>>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
    } else if constexpr (SURF == CAM_SURFACE_OH3) {
        const int grid = (nbeads + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(egrad_lane_kernel<LaneOH3>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
    } else if constexpr (SURF == CAM_SURFACE_BRH2) {
        const int grid = (nbeads + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(egrad_lane_kernel<LaneBrH2>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
    } else if constexpr (SURF == CAM_SURFACE_O3) {
        const int grid = (nbeads + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(egrad_lane_kernel<LaneO3>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
    } else {
        const int grid = (nbeads + kBlock / 4 - 1) / (kBlock / 4);
        if (precision == CAM_PREC_MIXED)
            hipLaunchKernelGGL(egrad_ch4h_kernel<float>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
        else
            hipLaunchKernelGGL(egrad_ch4h_kernel<double>, dim3(grid), dim3(kBlock), 0, s, q, nbeads, v, g);
    }
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

template <int SURF>
int egrad_dev(const double *q, int natoms, int nbeads, double *v, double *g, int precision, void *stream)
{
    constexpr int NA = surface_natoms(SURF);
    CAM_CHECK_ARG(natoms == NA, "natoms must be %d for this surface (got %d)", NA, natoms);
    CAM_CHECK_ARG(nbeads >= 0, "nbeads must be >= 0 (got %d)", nbeads);
    CAM_CHECK_ARG(nbeads == 0 || (q && v && g), "null device pointer");
    CAM_CHECK_ARG(precision == CAM_PREC_FP64 || precision == CAM_PREC_MIXED, "bad precision %d", precision);
    CAM_CHECK_ARG(SURF == CAM_SURFACE_H3 || SURF == CAM_SURFACE_CH4H || precision == CAM_PREC_FP64,
                  "the OH3, BrH2 and O3 surfaces are evaluated in FP64 only");
    clear_error();
    return launch_egrad<SURF>(q, nbeads, v, g, precision, (hipStream_t)stream);
}

// Host-pointer form: stage through device buffers owned by this call.
template <int SURF>
int egrad_host(const double *q, int natoms, int nbeads, double *v, double *g)
{
    constexpr int NA = surface_natoms(SURF);
    CAM_CHECK_ARG(natoms == NA, "natoms must be %d for this surface (got %d)", NA, natoms);
    CAM_CHECK_ARG(nbeads >= 0, "nbeads must be >= 0 (got %d)", nbeads);
    CAM_CHECK_ARG(nbeads == 0 || (q && v && g), "null host pointer");
    clear_error();
    if (nbeads == 0) return CAM_OK;
    const size_t nq = (size_t)nbeads * 3 * NA;
    double *dq = nullptr, *dv = nullptr, *dg = nullptr;
    int rc = CAM_OK;
    auto cleanup = [&]() {
        if (dq) (void)hipFree(dq);
        if (dv) (void)hipFree(dv);
        if (dg) (void)hipFree(dg);
    };
    auto step = [&]() -> int {
        CAM_HIP(hipMalloc(&dq, nq * sizeof(double)));
        CAM_HIP(hipMalloc(&dv, (size_t)nbeads * sizeof(double)));
        CAM_HIP(hipMalloc(&dg, nq * sizeof(double)));
        CAM_HIP(hipMemcpy(dq, q, nq * sizeof(double), hipMemcpyHostToDevice));
        int r = launch_egrad<SURF>(dq, nbeads, dv, dg, CAM_PREC_FP64, nullptr);
        if (r != CAM_OK) return r;
        CAM_HIP(hipMemcpy(v, dv, (size_t)nbeads * sizeof(double), hipMemcpyDeviceToHost));
        CAM_HIP(hipMemcpy(g, dg, nq * sizeof(double), hipMemcpyDeviceToHost));
        return CAM_OK;
    };
    rc = step();
    cleanup();
    return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free host-pointer form.  The staging buffers of egrad_host live
// in one file-scope device array per GPU instead of per-call hipMalloc'd
// memory: q, then v, then dV/dq of up to kNoAllocBeads beads of the largest
// surface (6 atoms).  In host code `scratch` names a host-side stand-in, so
// its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is used.  Calls on one device
// are serialised by a mutex (the array is shared by every call there); the
// array keeps the last call's contents, which no call reads before writing.
constexpr int kNoAllocBeads = 16384;
constexpr size_t kNoAllocDoubles = (size_t)kNoAllocBeads * (2 * 18 + 1);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex[kMaxDevices];
double *g_noalloc_ptr[kMaxDevices] = {};

template <int SURF>
int egrad_host_noalloc(const double *q, int natoms, int nbeads, double *v, double *g)
{
    constexpr int NA = surface_natoms(SURF);
    CAM_CHECK_ARG(natoms == NA, "natoms must be %d for this surface (got %d)", NA, natoms);
    CAM_CHECK_ARG(nbeads >= 0 && nbeads <= kNoAllocBeads,
                  "nbeads must be in 0..%d for the allocation-free entry point (got %d)", kNoAllocBeads, nbeads);
    CAM_CHECK_ARG(nbeads == 0 || (q && v && g), "null host pointer");
    clear_error();
    if (nbeads == 0) return CAM_OK;
    int dev = 0;
    CAM_HIP(hipGetDevice(&dev));
    CAM_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_noalloc_mutex[dev]);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        CAM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[dev] = static_cast<double *>(ptr);
    }
    const size_t nq = (size_t)nbeads * 3 * NA;
    double *dq = g_noalloc_ptr[dev];
    double *dv = dq + (size_t)kNoAllocBeads * 18;
    double *dg = dv + kNoAllocBeads;
    CAM_HIP(hipMemcpy(dq, q, nq * sizeof(double), hipMemcpyHostToDevice));
    const int r = launch_egrad<SURF>(dq, nbeads, dv, dg, CAM_PREC_FP64, nullptr);
    if (r != CAM_OK) return r;
    CAM_HIP(hipMemcpy(v, dv, (size_t)nbeads * sizeof(double), hipMemcpyDeviceToHost));
    CAM_HIP(hipMemcpy(g, dg, nq * sizeof(double), hipMemcpyDeviceToHost));
    return CAM_OK;
}

}  // namespace
}  // namespace cam

extern "C" {
int cam_egrad_noalloc(int surface, const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    if (surface == CAM_SURFACE_H3) return cam::egrad_host_noalloc<CAM_SURFACE_H3>(q, natoms, nbeads, v, dvdq);
    if (surface == CAM_SURFACE_CH4H) return cam::egrad_host_noalloc<CAM_SURFACE_CH4H>(q, natoms, nbeads, v, dvdq);
    if (surface == CAM_SURFACE_OH3) return cam::egrad_host_noalloc<CAM_SURFACE_OH3>(q, natoms, nbeads, v, dvdq);
    if (surface == CAM_SURFACE_BRH2) return cam::egrad_host_noalloc<CAM_SURFACE_BRH2>(q, natoms, nbeads, v, dvdq);
    if (surface == CAM_SURFACE_O3) return cam::egrad_host_noalloc<CAM_SURFACE_O3>(q, natoms, nbeads, v, dvdq);
    CAM_CHECK_ARG(false, "unknown surface %d", surface);
}
int cam_egrad_h3(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    return cam::egrad_host<CAM_SURFACE_H3>(q, natoms, nbeads, v, dvdq);
}
int cam_egrad_ch4h(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    return cam::egrad_host<CAM_SURFACE_CH4H>(q, natoms, nbeads, v, dvdq);
}
int cam_egrad_oh3(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    return cam::egrad_host<CAM_SURFACE_OH3>(q, natoms, nbeads, v, dvdq);
}
int cam_egrad_oh3_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq, int precision, void *stream)
{
    return cam::egrad_dev<CAM_SURFACE_OH3>(q, natoms, nbeads, v, dvdq, precision, stream);
}
int cam_egrad_brh2(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    return cam::egrad_host<CAM_SURFACE_BRH2>(q, natoms, nbeads, v, dvdq);
}
int cam_egrad_brh2_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq, int precision, void *stream)
{
    return cam::egrad_dev<CAM_SURFACE_BRH2>(q, natoms, nbeads, v, dvdq, precision, stream);
}
int cam_egrad_o3(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    return cam::egrad_host<CAM_SURFACE_O3>(q, natoms, nbeads, v, dvdq);
}
int cam_egrad_o3_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq, int precision, void *stream)
{
    return cam::egrad_dev<CAM_SURFACE_O3>(q, natoms, nbeads, v, dvdq, precision, stream);
}
int cam_egrad_h3_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq, int precision, void *stream)
{
    return cam::egrad_dev<CAM_SURFACE_H3>(q, natoms, nbeads, v, dvdq, precision, stream);
}
int cam_egrad_ch4h_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq, int precision, void *stream)
{
    return cam::egrad_dev<CAM_SURFACE_CH4H>(q, natoms, nbeads, v, dvdq, precision, stream);
}
}
