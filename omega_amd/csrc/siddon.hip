// siddon.hip -- improved Siddon forward/backprojection kernels and their C-ABI entry points.
//
// Launch shape: one lane per measurement, 256 lanes (4 wave64) per workgroup.
// Measurements are in OMEGA data order (radial bin fastest, then angle, then
// plane), so the 64 lanes of a wave trace 64 neighbouring parallel chords of one
// plane and their gathers/scatters fall on shared cache lines.
//
// Backprojection scatters with hardware float atomics (global_atomic_add_f32,
// no CAS loop: built with -munsafe-fp-atomics) exactly where the reference CPU
// uses `omp atomic` (source/cpp/projector_functions.h:652-656).
#include "geom.hpp"

namespace omg {

constexpr int kBlock = 256;

// ---- forward projection -------------------------------------------------------------------------
// The gather of step i is issued before the accumulation of step i-1 (a two-deep software
// pipeline), while the sum itself keeps the traversal order of the reference.
__global__ __launch_bounds__(kBlock) void k_siddon_fp(SidArgs a, const float* __restrict__ img, float* __restrict__ y) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    float ax = 0.f, w = 0.f;
    float pv = 0.f, pl = 0.f;  // pending value and length
    bool pend = false;
    const bool hit = siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
        const float v = img[ind];
        if (pend) ax += pl * pv;
        pv = v; pl = len; pend = true;
    });
    if (pend) ax += pl * pv;
    y[lo] = hit ? (a.mode == MODE_CT_FLAT ? ax : ax * w) : 0.f;
}

// ---- backprojection -----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_siddon_bp(SidArgs a, const float* __restrict__ yin, float* __restrict__ rhs,
                                                      float* __restrict__ sens) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    const float yi = yin[lo];
    if (yi == 0.f && sens == nullptr) return;  // adds of +0 leave rhs unchanged
    float w = 0.f;
    if (sens) {
        siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
            const float lw = len * w;
            unsafeAtomicAdd(rhs + ind, yi * lw);
            unsafeAtomicAdd(sens + ind, lw);
        });
    } else {
        siddon_trace(a, lo, w, [&](uint32_t ind, float len) { unsafeAtomicAdd(rhs + ind, yi * (len * w)); });
    }
}

// ---- fused OSEM data step: FP, ratio, BP for one LOR ---------------------------------------------
__global__ __launch_bounds__(kBlock) void k_siddon_osem(SidArgs a, const float* __restrict__ img,
                                                        const float* __restrict__ meas, float* __restrict__ rhs,
                                                        float* __restrict__ sens, float epps) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    float ax = 0.f, w = 0.f, pv = 0.f, pl = 0.f;
    bool pend = false;
    const bool hit = siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
        const float v = img[ind];
        if (pend) ax += pl * pv;
        pv = v; pl = len; pend = true;
    });
    if (!hit) return;
    if (pend) ax += pl * pv;
    const float yhat = ax * w;
    float r = meas[lo] / (yhat + epps);
    if (isnan(r) || isinf(r)) r = epps;
    if (r == 0.f && sens == nullptr) return;
    if (sens) {
        siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
            const float lw = len * w;
            unsafeAtomicAdd(rhs + ind, r * lw);
            unsafeAtomicAdd(sens + ind, lw);
        });
    } else {
        siddon_trace(a, lo, w, [&](uint32_t ind, float len) { unsafeAtomicAdd(rhs + ind, r * (len * w)); });
    }
}

// ---- traversal dump -----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_siddon_trace(SidArgs a, int32_t* __restrict__ counts,
                                                         const int64_t* __restrict__ offsets, uint32_t* __restrict__ idx,
                                                         float* __restrict__ len) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    float w = 0.f;
    int32_t c = 0;
    if (offsets) {
        int64_t o = offsets[lo];
        siddon_trace(a, lo, w, [&](uint32_t ind, float l) { idx[o] = ind; len[o] = l; o++; c++; });
    } else {
        siddon_trace(a, lo, w, [&](uint32_t, float) { c++; });
    }
    counts[lo] = c;
}

}  // namespace omg

using namespace omg;

static int siddon_common_checks(const omg_geom* g, int64_t first, int64_t n) {
    OMG_REQUIRE(g != nullptr, "null geometry handle");
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON || g->d.projector_type == OMG_PROJ_ORTH ||
                    g->d.projector_type == OMG_PROJ_BDD,
                "unknown projector type %d", g->d.projector_type);
    return check_range(g, first, n);
}

// Forward declarations of the orthogonal / BDD dispatchers (other translation units).
int orth_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s);
int orth_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens, hipStream_t s);
int bdd_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s);
int bdd_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens, hipStream_t s);

extern "C" int omg_forward_project(const omg_geom* g, int64_t first, int64_t n_meas, const float* img, float* y,
                                   void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(img && y, "null image or output pointer");
    if (n_meas == 0) return OMG_OK;
    hipStream_t s = as_stream(stream);
    if (g->d.projector_type == OMG_PROJ_ORTH) return orth_forward(g, first, n_meas, img, y, s);
    if (g->d.projector_type == OMG_PROJ_BDD) return bdd_forward(g, first, n_meas, img, y, s);
    SidArgs a = make_sid_args(g, first, n_meas);
    hipLaunchKernelGGL(k_siddon_fp, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, s, a, img, y);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_backward_project(const omg_geom* g, int64_t first, int64_t n_meas, const float* y, float* rhs,
                                    float* sens, void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(y && rhs, "null input or output pointer");
    if (n_meas == 0) return OMG_OK;
    hipStream_t s = as_stream(stream);
    if (g->d.projector_type == OMG_PROJ_ORTH) return orth_backward(g, first, n_meas, y, rhs, sens, s);
    if (g->d.projector_type == OMG_PROJ_BDD) return bdd_backward(g, first, n_meas, y, rhs, sens, s);
    SidArgs a = make_sid_args(g, first, n_meas);
    hipLaunchKernelGGL(k_siddon_bp, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, s, a, y, rhs, sens);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_osem_data_step(const omg_geom* g, int64_t first, int64_t n_meas, const float* img,
                                  const float* meas, float* rhs, float* sens, float epps, void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(img && meas && rhs, "null pointer argument");
    OMG_REQUIRE(g->d.mode != OMG_MODE_CT_FLAT, "the fused OSEM data step is for emission (PET) data");
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON, "the fused OSEM data step is implemented for projector_type 1");
    if (n_meas == 0) return OMG_OK;
    SidArgs a = make_sid_args(g, first, n_meas);
    hipLaunchKernelGGL(k_siddon_osem, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, as_stream(stream), a, img, meas,
                       rhs, sens, epps);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_siddon_trace(const omg_geom* g, int64_t first, int64_t n_meas, int32_t* counts,
                                const int64_t* offsets, uint32_t* idx, float* len, void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON, "trace lists are produced by projector_type 1");
    OMG_REQUIRE(counts != nullptr, "null counts");
    OMG_REQUIRE(offsets == nullptr || (idx && len), "offsets given without idx/len outputs");
    if (n_meas == 0) return OMG_OK;
    SidArgs a = make_sid_args(g, first, n_meas);
    hipLaunchKernelGGL(k_siddon_trace, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, as_stream(stream), a, counts,
                       offsets, idx, len);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}
