// ctosem.hip -- OSEM / OSL-OSEM for transmission (CT) data.
//
// The reference's subset loop for CT data with OSEM / OSL-OSEM (reconstructionAF.h:322-327,1224-1260;
// computeForwardProject.h:76-90; subiterStep.h:111-121,219-231), without randoms:
//   iteration 0: S_k = A_k^T y_k (the backprojection of the measured counts), S_k < epps -> epps;
//   yhat = A_k x; yhat < 0 -> 0; r = exp(-yhat);  rhs = A_k^T r;
//   OSEM:     x = x / S_k * (flat * rhs)
//   OSL-OSEM: x = x / (S_k + dU) * (flat * rhs), dU = beta * grad RDP(x; gamma, epps), zero boundary;
//   positivity x < epps -> epps (reconstructionAF.h:1461-1465).
// Any projector of the geometry (Siddon, branchless distance-driven) through the library's own FP / BP.
#include "geom.hpp"

namespace omg {

// r = exp(-max(yhat, 0)) in place (computeForwardProject.h:80-88), evaluated in double and rounded once to
// float: the float result is then the correctly rounded exponential (but for double-rounding ties), the same
// on any device or host, which the distance-driven projector needs -- its summed-area-table forward
// projection turns one-ulp differences of the image into much larger differences of the projections.
__global__ __launch_bounds__(256) void k_ct_ratio(float* __restrict__ v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float a = v[i];
    if (a < 0.f) a = 0.f;
    v[i] = (float)exp(-(double)a);
}

// x = x / S * (flat * rhs), or x / (S + dU) * (flat * rhs); positivity.
__global__ __launch_bounds__(256) void k_ct_update(uint32_t N, float* __restrict__ x, const float* __restrict__ S,
                                                   const float* __restrict__ rhs, const float* __restrict__ dU,
                                                   float flat, float epps, int pos) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const float num = flat * rhs[n];
    float v = dU ? x[n] / (S[n] + dU[n]) * num : x[n] / S[n] * num;
    if (pos && v < epps) v = epps;
    x[n] = v;
}

__global__ __launch_bounds__(256) void k_ct_clamp(float* __restrict__ s, uint32_t N, float epps) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n < N && s[n] < epps) s[n] = epps;
}

__global__ __launch_bounds__(256) void k_ct_zero(float* __restrict__ p, uint32_t N) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n < N) p[n] = 0.f;
}

}  // namespace omg

using namespace omg;

extern "C" int omg_ct_osem_subset_iteration(const omg_geom* g, int64_t first, int64_t n_meas, float* x,
                                            const float* meas, float* S, int32_t compute_sens, float flat, float epps,
                                            float beta, float gamma, int32_t positivity, void* stream) {
    OMG_REQUIRE(g != nullptr, "null geometry handle");
    OMG_REQUIRE(x && meas && S, "null pointer argument");
    OMG_REQUIRE(g->d.mode == OMG_MODE_CT_FLAT, "the transmission OSEM iteration is for CT data");
    OMG_REQUIRE(flat > 0.f, "flat (the blank-scan value) must be positive, got %g", (double)flat);
    int rc = check_range(g, first, n_meas);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    omg_geom* gm = const_cast<omg_geom*>(g);
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    rc = work_reserve(gm, N);
    if (rc) return rc;
    rc = q_reserve(gm, (size_t)(n_meas > 0 ? n_meas : 1));
    if (rc) return rc;
    float* dU = work_ptr(g);
    float* rhs = work_ptr(g) + N;
    float* yhat = q_ptr(g);
    const dim3 gN(grid_for(N, 256));
    if (compute_sens) {
        hipLaunchKernelGGL(k_ct_zero, gN, dim3(256), 0, s, S, N);
        OMG_CHECK_LAUNCH();
        if (n_meas > 0) {
            rc = omg_backward_project(g, first, n_meas, meas, S, nullptr, stream);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_ct_clamp, gN, dim3(256), 0, s, S, N, epps);
        OMG_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ct_zero, gN, dim3(256), 0, s, rhs, N);
    OMG_CHECK_LAUNCH();
    if (n_meas > 0) {
        rc = omg_forward_project(g, first, n_meas, x, yhat, stream);
        if (rc) return rc;
        hipLaunchKernelGGL(k_ct_ratio, dim3(grid_for(n_meas, 256)), dim3(256), 0, s, yhat, n_meas);
        OMG_CHECK_LAUNCH();
        rc = omg_backward_project(g, first, n_meas, yhat, rhs, nullptr, stream);
        if (rc) return rc;
    }
    const float* dUp = nullptr;
    if (beta != 0.f) {
        hipLaunchKernelGGL(k_ct_zero, gN, dim3(256), 0, s, dU, N);
        OMG_CHECK_LAUNCH();
        rc = omg_rdp_gradient(x, dU, g->d.Nx, g->d.Ny, g->d.Nz, gamma, epps, beta, 0, stream);
        if (rc) return rc;
        dUp = dU;
    }
    hipLaunchKernelGGL(k_ct_update, gN, dim3(256), 0, s, N, x, S, rhs, dUp, flat, epps, (int)(positivity != 0));
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}
