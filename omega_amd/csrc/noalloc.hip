// noalloc.hip -- the OSEM / OSL-OSEM subset iteration with its work buffers in a fixed device array.
//
// omg_osem_subset_iteration_ex keeps its work buffers (the OSL gradient, the rhs image, the slice map and
// the planar kernels' per-workgroup partial slices) in the geometry handle and grows them with hipMalloc
// on first use. Callers that must not allocate (code under HIP graph capture, serving loops) call
// omg_osem_subset_iteration_noalloc instead: the same kernels, in the same order, with every work buffer
// carved out of `scratch`, a file-scope device array of kScratchFloats floats that each GPU holds until
// the process ends. Every call on one GPU shares that GPU's array: calls that overlap on one GPU race.
// Nothing in it has to start at zero: each call clears or fully writes what it reads.
#include "geom.hpp"

#include <mutex>

namespace omg {

constexpr size_t kScratchFloats = (size_t)16 << 20;   // 64 MiB
constexpr int kMaxDevices = 64;

int omg_osem_subset_iteration_impl(const omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas,
                                   const float* additive, float* S, int32_t compute_sens, float epps, float beta,
                                   float gamma, int32_t positivity, void* stream);

}  // namespace omg

static __device__ __align__(256) float scratch[omg::kScratchFloats];

namespace {

std::mutex g_scratch_mu;
float* g_scratch_dev[omg::kMaxDevices] = {};   // device address of `scratch` per device, once looked up

// Device address of the current device's copy of `scratch` (looked up on the first call on that device).
int scratch_address(int dev, float** out) {
    using namespace omg;
    OMG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d is outside the %d devices this library tracks", dev,
                kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return OMG_ERR_HIP;
        }
        g_scratch_dev[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_dev[dev];
    return OMG_OK;
}

}  // namespace

using namespace omg;

extern "C" int omg_osem_subset_iteration_noalloc(const omg_geom* g, int64_t first, int64_t n_meas, float* x,
                                                 const float* meas, const float* additive, float* S,
                                                 int32_t compute_sens, float epps, float beta, float gamma,
                                                 int32_t positivity, void* stream) {
    OMG_REQUIRE(g != nullptr, "null geometry handle");
    int dev = -1;
    OMG_CHECK_HIP(hipGetDevice(&dev));
    OMG_REQUIRE(dev == g->device, "the current device (%d) is not the geometry's device (%d)", dev, g->device);
    float* base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc) return rc;
    const FixedRegion region{base, kScratchFloats};
    FixedRegionGuard guard(&region);
    return omg_osem_subset_iteration_impl(g, first, n_meas, x, meas, additive, S, compute_sens, epps, beta, gamma,
                                          positivity, stream);
}
