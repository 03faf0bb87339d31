// tof.hip -- time-of-flight Siddon forward / backward projection and fused OSEM data step for gfx950.
//
// Semantics are the reference CPU projector's with param.TOF (projectorType123Implementation4,
// source/cpp/projector_functions.h:559-657, :1641-1691, :1741-1742, :1945-1991, :2030-2035): a
// measurement array of a call holds tof_bins blocks of n values (bin b of measurement i at i + b * n).
// Per visited voxel the bins' weights w_b (a 4-point trapezoid of a Gaussian over the voxel's stretch,
// tof_weight) are normalised by their sum floored at epps; the forward projection adds
// (len * x[v]) * w_b / sum to bin b, the backprojection adds sum_b (len * weight) * (w_b / sum) * y_b
// to rhs[v] (and the sum of those factors to the sensitivity).
//
// One lane per LOR, one wave per workgroup. A lane's per-bin accumulators and weights live in LDS,
// laid out [bin][lane] (conflict-free), so any bin count up to kTofMax runs without register arrays
// indexed at run time (which would go to scratch memory).
//
// Several rays per LOR (:1441-1446, :2036-2056): the forward projection sums each ray's bins times that
// ray's weight (1 / (L nRays), its own attenuation factor, norm, ...) into a fourth per-lane LDS array,
// in ray order; the backprojection traces every ray with its own weight.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "geom.hpp"
#include "siddon_device.hpp"

namespace omg {

constexpr int kTofBlock = 64;

// Per-lane cache of the Gaussians at the last trapezoid point of the previous voxel: on a Siddon walk
// that point, D - copysign(len, DD), is bitwise the next voxel's D, so its values are the next
// voxel's first terms (a quarter of the exponentials saved; keyed by the exact argument, so the
// results do not change).
struct TofCache {
    float* x;   // the cached argument (one float per lane), NaN when empty
    float* p;   // p[b * kTofBlock]: Gaussian of bin b at *x
};

// TOFLoop (projector_functions.h:575-586): the bins' weights for a voxel of length `elem` at distance
// D into wl[b * kTofBlock], dX = elem / 3.f; returns their sum floored at epps. Each weight is
// TOFWeight's p(D) + 2 p(D - dX) + 2 p(D - 2 dX) + p(D - elem) (signs by DD), times dX.
__device__ __forceinline__ float tof_weights(const SidArgs& a, float elem, float D, float DD, float* wl,
                                             TofCache c) {
    const float dX = elem / 3.f;
    const float dXs = copysignf(dX, DD);
    const float x3 = D - copysignf(elem, DD);
    const bool reuse = D == *c.x;
    float sum = 0.f;
    for (uint32_t b = 0; b < a.tof_bins; b++) {
        const float mu = a.tof_c[b];
        float out = reuse ? c.p[b * kTofBlock] : tof_pdf(a, D, mu);
        out += tof_pdf(a, D - dXs * 1.f, mu) * 2.f;
        out += tof_pdf(a, D - dXs * 2.f, mu) * 2.f;
        const float p3 = tof_pdf(a, x3, mu);
        out += p3;
        c.p[b * kTofBlock] = p3;
        const float v = out * dX;
        wl[b * kTofBlock] = v;
        sum += v;
    }
    *c.x = x3;
    return sum < a.tof_epps ? a.tof_epps : sum;
}

// The backprojection's weights (rhs, projector_functions.h:632-657) take dX = elem / 3.0 in double
// rounded to float (TRAPZ_BINS is a double literal there); recomputed only where that, or the length
// the weight sum used, differs from what wl holds.
__device__ __forceinline__ void tof_bp_weights(const SidArgs& a, float len, float esum, float D, float DD, float* wl) {
    const float dXb = (float)((double)len / 3.0);
    if (esum == len && dXb == len / 3.f) return;
    for (uint32_t b = 0; b < a.tof_bins; b++) wl[b * kTofBlock] = tof_weight(a, len, D, DD, a.tof_c[b], dXb) * dXb;
}

// FP walk: acc[b] += (len * src[v]) * w_b / sum over the LOR's voxels (denominator, :609-630).
__device__ __forceinline__ bool tof_gather(const SidArgs& a, int64_t lo, const float* __restrict__ src, float* acc,
                                           float* wl, TofCache c, float& w, int ray = 0) {
    for (uint32_t b = 0; b < a.tof_bins; b++) acc[b * kTofBlock] = 0.f;
    *c.x = __builtin_nanf("");
    return siddon_tof_trace(a, lo, w, [&](uint32_t ind, float len, float D, float DD, float esum) {
        const float apu = len * src[ind];
        const float sum = tof_weights(a, esum, D, DD, wl, c);
        if (esum != len) tof_weights(a, len, D, DD, wl, c);
        for (uint32_t b = 0; b < a.tof_bins; b++) acc[b * kTofBlock] = acc[b * kTofBlock] + apu * wl[b * kTofBlock] / sum;
    }, ray);
}

// BP walk: rhs[v] += sum_b (len w) (w_b / sum) yv[b]; sens[v] += sum_b (len w) (w_b / sum).
__device__ __forceinline__ void tof_scatter(const SidArgs& a, int64_t lo, const float* yv, float* wl, TofCache c,
                                            float* __restrict__ rhs, float* __restrict__ sens, int ray = 0) {
    float w = 0.f;
    *c.x = __builtin_nanf("");
    siddon_tof_trace(a, lo, w, [&](uint32_t ind, float len, float D, float DD, float esum) {
        const float sum = tof_weights(a, esum, D, DD, wl, c);
        tof_bp_weights(a, len, esum, D, DD, wl);
        const float lw = len * w;
        float yax = 0.f, val = 0.f;
        for (uint32_t b = 0; b < a.tof_bins; b++) {
            const float ap = lw * (wl[b * kTofBlock] / sum);
            val += ap;
            yax += ap * yv[b * kTofBlock];
        }
        unsafeAtomicAdd(rhs + ind, yax);
        if (sens) unsafeAtomicAdd(sens + ind, val);
    }, ray);
}

// Several rays: tot[b] = sum over the rays (in order) of the ray's bin b times its weight; false when
// every ray misses the image.
__device__ __forceinline__ bool tof_gather_rays(const SidArgs& a, int64_t lo, const float* __restrict__ src, float* acc,
                                                float* wl, TofCache c, float* tot) {
    for (uint32_t b = 0; b < a.tof_bins; b++) tot[b * kTofBlock] = 0.f;
    bool any = false;
    for (uint32_t ray = 0; ray < n_rays(a); ray++) {
        float w = 0.f;
        if (!tof_gather(a, lo, src, acc, wl, c, w, (int)ray)) continue;
        any = true;
        for (uint32_t b = 0; b < a.tof_bins; b++) tot[b * kTofBlock] = tot[b * kTofBlock] + acc[b * kTofBlock] * w;
    }
    return any;
}

__device__ __forceinline__ TofCache tof_cache(const SidArgs& a, float* lds) {
    return TofCache{lds + 3 * a.tof_bins * kTofBlock + threadIdx.x, lds + 2 * a.tof_bins * kTofBlock + threadIdx.x};
}

// the rays' bin sums, after the cache's argument slot
__device__ __forceinline__ float* tof_tot(const SidArgs& a, float* lds) {
    return lds + (3 * a.tof_bins + 1) * kTofBlock + threadIdx.x;
}

// y[lo + b n] = w * acc[b]
__global__ __launch_bounds__(kTofBlock) void k_tof_fp(SidArgs a, const float* __restrict__ img, float* __restrict__ y) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kTofBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    float* acc = lds + threadIdx.x;
    float* wl = lds + a.tof_bins * kTofBlock + threadIdx.x;
    const TofCache c = tof_cache(a, lds);
    if (n_rays(a) > 1) {
        float* tot = tof_tot(a, lds);
        tof_gather_rays(a, lo, img, acc, wl, c, tot);
        for (uint32_t b = 0; b < a.tof_bins; b++) y[lo + (int64_t)b * a.nmeas] = tot[b * kTofBlock];
        return;
    }
    float w = 0.f;
    const bool hit = tof_gather(a, lo, img, acc, wl, c, w);
    for (uint32_t b = 0; b < a.tof_bins; b++) y[lo + (int64_t)b * a.nmeas] = hit ? acc[b * kTofBlock] * w : 0.f;
}

__global__ __launch_bounds__(kTofBlock) void k_tof_bp(SidArgs a, const float* __restrict__ yin, float* __restrict__ rhs,
                                                      float* __restrict__ sens) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kTofBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    float* yv = lds + threadIdx.x;
    float* wl = lds + a.tof_bins * kTofBlock + threadIdx.x;
    const TofCache c = tof_cache(a, lds);
    bool any = false;
    for (uint32_t b = 0; b < a.tof_bins; b++) {
        const float v = yin[lo + (int64_t)b * a.nmeas];
        yv[b * kTofBlock] = v;
        any |= v != 0.f;
    }
    if (!any && sens == nullptr) return;   // adds of +0 leave rhs unchanged
    for (uint32_t ray = 0; ray < n_rays(a); ray++) tof_scatter(a, lo, yv, wl, c, rhs, sens, (int)ray);
}

// Fused data step: yhat_b = w acc_b; r_b = meas_b / ((yhat_b + epps) + add); NaN/Inf -> epps; BP of r.
__global__ __launch_bounds__(kTofBlock) void k_tof_osem(SidArgs a, const float* __restrict__ img,
                                                        const float* __restrict__ meas, float* __restrict__ rhs,
                                                        float* __restrict__ sens, float epps) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kTofBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    float* acc = lds + threadIdx.x;
    float* wl = lds + a.tof_bins * kTofBlock + threadIdx.x;
    const TofCache c = tof_cache(a, lds);
    if (n_rays(a) > 1) {   // every ray missing the image: nothing to backproject, as with one ray
        float* tot = tof_tot(a, lds);
        if (!tof_gather_rays(a, lo, img, acc, wl, c, tot)) return;
        for (uint32_t b = 0; b < a.tof_bins; b++) {
            float r = meas[lo + (int64_t)b * a.nmeas] / ratio_den(a, lo, tot[b * kTofBlock], epps);
            if (isnan(r) || isinf(r)) r = epps;
            tot[b * kTofBlock] = r;
        }
        for (uint32_t ray = 0; ray < n_rays(a); ray++) tof_scatter(a, lo, tot, wl, c, rhs, sens, (int)ray);
        return;
    }
    float w = 0.f;
    if (!tof_gather(a, lo, img, acc, wl, c, w)) return;
    for (uint32_t b = 0; b < a.tof_bins; b++) {
        const float yhat = acc[b * kTofBlock] * w;
        float r = meas[lo + (int64_t)b * a.nmeas] / ratio_den(a, lo, yhat, epps);
        if (isnan(r) || isinf(r)) r = epps;
        acc[b * kTofBlock] = r;
    }
    tof_scatter(a, lo, acc, wl, c, rhs, sens);
}

// LDS per workgroup: [acc or yv: bins x lanes][wl: bins x lanes][cache p: bins x lanes][cache x: lanes]
// and, with several rays per LOR, [tot: bins x lanes]: at most (4 * 64 + 1) * 64 * 4 B = 65,792 B.
static size_t tof_lds(const SidArgs& a) {
    return (3 * (size_t)a.tof_bins + 1 + (a.rays_xy * a.rays_z > 1 ? (size_t)a.tof_bins : 0)) * kTofBlock * sizeof(float);
}

// Lifts the three kernels' dynamic LDS limit above the default 64 KiB once (several rays, > 63 bins).
static int tof_lds_attr() {
    static bool done = false;
    if (done) return OMG_OK;
    const int bytes = (int)((4 * (size_t)kTofMax + 1) * kTofBlock * sizeof(float));
    OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_tof_fp),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_tof_bp),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_tof_osem),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done = true;
    return OMG_OK;
}

int tof_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s) {
    const SidArgs a = make_sid_args(g, first, n);
    if (int rc = tof_lds_attr()) return rc;
    hipLaunchKernelGGL(k_tof_fp, dim3(grid_for(n, kTofBlock)), dim3(kTofBlock), tof_lds(a), s, a, img, y);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int tof_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens,
                 hipStream_t s) {
    const SidArgs a = make_sid_args(g, first, n);
    if (int rc = tof_lds_attr()) return rc;
    hipLaunchKernelGGL(k_tof_bp, dim3(grid_for(n, kTofBlock)), dim3(kTofBlock), tof_lds(a), s, a, yin, rhs, sens);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int tof_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
             float* sens, float epps, hipStream_t s, const float* add) {
    SidArgs a = make_sid_args(g, first, n);
    a.add = add;
    if (int rc = tof_lds_attr()) return rc;
    hipLaunchKernelGGL(k_tof_osem, dim3(grid_for(n, kTofBlock)), dim3(kTofBlock), tof_lds(a), s, a, img, meas, rhs,
                       sens, epps);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

}  // namespace omg
