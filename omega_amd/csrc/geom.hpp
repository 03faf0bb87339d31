// geom.hpp -- the opaque geometry handle behind omg_geom*.
#pragma once
#include "common.hpp"
#include "siddon_device.hpp"

#include <vector>

struct omg_geom {
    omg_geom_desc d;          // caller's description (host table pointers cleared after upload)
    float* x_dev = nullptr;   // coordinate tables on the device
    float* z_dev = nullptr;
    float* norm_dev = nullptr;
    float* atten_dev = nullptr;      // per-measurement attenuation factors (norm's layout), or null
    int64_t atten_ray_stride = 0;    // factors of an image with several rays: one block of them per ray
    float* scat_dev = nullptr;       // per-measurement multiplicative correction (norm's layout), or null
    float* vol_dev = nullptr;        // projector 3: the volume-of-intersection table, or null
    float* tof_c = nullptr;          // TOF bin centres (tof_bins floats), or null
    uint32_t* index_dev = nullptr;   // measurement-level subsets: sinogram bin per measurement (subset order)
    float bmaxx = 0.f, bmaxy = 0.f, bmaxz = 0.f;
    int device = 0;
    // Planar fast path: every plane's LORs have zs == zd and the sinogram is in PET mode.
    bool planar = false;
    // Route every call through the one-lane-per-LOR kernels (set from omg_geom_desc.reserved[0]).
    bool force_generic = false;
    // Every LOR of the (planar) sinogram has both endpoints outside the image with margin, so every
    // crossing parameter of its walk lies in [0, 1] (the planar kernels' short length expression applies).
    bool ring = false;
    // Batched frames: y-fastest accumulator stack of direct omg_frames_osem_step calls (zero between calls).
    float* frames_t = nullptr;
    size_t frames_t_floats = 0;
    // Orthogonal projector: direct-plane sinogram whose tubes stay in their slice (orth.hip planar kernels).
    bool orth_planar = false;
    // Volume projector: direct-plane PET sinogram (orth.hip k_orth_vol_planar for the fused step).
    bool vol_planar = false;
    // Projectors 2 and 3 on a direct-plane PET sinogram: the tube lists' LDS-slice kernel applies.
    bool tube_planar = false;
    // Projectors 2 and 3: the tube lists of the measurement ranges used so far, kept on the device
    // (orth.hip; omg_geom_desc.reserved[4] sets the budget).
    struct TubeCache* tubes = nullptr;
    // Workgroups per plane for the planar kernels (omg_geom_desc.reserved[1]; 0 = automatic).
    int32_t planar_groups = 0;
    // Library-owned scratch (floats) for partial images; grown on demand, freed with the handle.
    float* scratch = nullptr;
    size_t scratch_floats = 0;
    // Library-owned work buffers of the fused OSEM iteration: [dU: N][rhs: N][slice map: Nz int32].
    float* work = nullptr;
    size_t work_floats = 0;
    // Host copy of the z table (plane -> slice maps are checked on the host).
    std::vector<float> z_host;
    // Branchless distance-driven projector: per-projection back-projection geometry (16 floats each).
    float* bdd_geom5 = nullptr;
    // Voxel-driven planar backprojection (vbp.hip): usable when every angle column is a parallel-beam set
    // of evenly spaced LORs with endpoints outside the image; per-angle parameters and q/w work buffers.
    bool vbp_ok = false;
    struct VbpAngle* vbp_angles = nullptr;
    float* vbp_q = nullptr;
    size_t vbp_q_floats = 0;
};

namespace omg {

// Kernel argument block for Siddon-family launches covering measurements [first, first + n).
// Projectors 2 and 3 share the tube sweep (orth.hip); they differ in the voxel and LOR weights.
static inline bool is_tube(const omg_geom* g) {
    return g->d.projector_type == OMG_PROJ_ORTH || g->d.projector_type == OMG_PROJ_VOLUME;
}

static inline SidArgs make_sid_args(const omg_geom* g, int64_t first, int64_t n) {
    SidArgs a;
    a.Nx = g->d.Nx; a.Ny = g->d.Ny; a.Nz = g->d.Nz; a.Nyx = g->d.Nx * g->d.Ny;
    a.dx = g->d.dx; a.dy = g->d.dy; a.dz = g->d.dz;
    a.bx = g->d.bx; a.by = g->d.by; a.bz = g->d.bz;
    a.bmaxx = g->bmaxx; a.bmaxy = g->bmaxy; a.bmaxz = g->bmaxz;
    a.size_x = g->d.size_x; a.size_y = g->d.size_y;
    a.mode = g->d.mode; a.ct_pitch = g->d.ct_pitch;
    a.plane_offset = first;
    a.nmeas = n;
    a.dPitchZ = g->d.dPitchZ;
    a.global_factor = g->d.global_factor;
    a.x = g->x_dev; a.z = g->z_dev;
    a.pitch = g->d.Nx;
    a.index = g->index_dev;
    a.add = nullptr;
    const int64_t per = (g->d.mode == MODE_LIST6 || g->index_dev) ? 1 : (int64_t)g->d.size_x * g->d.size_y;
    a.norm = g->norm_dev ? g->norm_dev + first * per : nullptr;
    a.atten = g->atten_dev ? g->atten_dev + first * per : nullptr;
    a.scat = g->scat_dev ? g->scat_dev + first * per : nullptr;
    a.atten_lor = g->d.atten_mode == 2;
    a.atten_ray_stride = g->atten_ray_stride;
    a.tof_bins = g->d.tof_bins > 1 ? g->d.tof_bins : 1u;
    a.tof_sigma = g->d.tof_sigma;
    a.tof_pf = g->d.tof_sigma > 0.f ? (float)0.3989422804014327 / g->d.tof_sigma : 0.f;
    a.tof_epps = g->d.tof_epps;
    a.tof_c = g->tof_c;
    a.rays_xy = g->d.rays_xy > 1 ? g->d.rays_xy : 1u;
    a.rays_z = g->d.rays_z > 1 ? g->d.rays_z : 1u;
    a.cr_xy = g->d.cr_xy;
    a.cr_z = g->d.cr_z;
    return a;
}

// Time-of-flight projections (tof.hip): measurement arrays hold tof_bins blocks of n values.
int tof_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s);
int tof_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens,
                 hipStream_t s);
int tof_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
             float* sens, float epps, hipStream_t s, const float* add);

// Writes f[i] = exp(-sum len * mu) over the Siddon walk of measurements [0, n) (siddon.hip).
int atten_factors(const omg_geom* g, const float* mu, float* f, int64_t n, hipStream_t s);

// Interpolation projector (projector type 4, interp.hip): forward and backward projection, the fused
// emission data step FP -> ratio -> BP, and attenuation factors by its own sampling walk.
int interp_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s);
int interp_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens,
                    hipStream_t s);
int interp_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
                float* sens, float epps, hipStream_t s, const float* add);
int interp_atten_factors(const omg_geom* g, const float* mu, float* f, int64_t n, hipStream_t s);

// Validates a [first, first + n) measurement range against the tables; returns OMG_OK or OMG_ERR_ARG.
int check_range(const omg_geom* g, int64_t first, int64_t n);

// Work buffers of one call come from the handle (grown with hipMalloc on demand) unless the calling
// thread holds a fixed region (the no-allocation entry point, noalloc.hip): then they are carved out of
// it as [work: 2N + Nz floats, rounded up to 64][scratch / planar q buffers: the rest], and a call that
// needs more than the region holds fails with OMG_ERR_ARG instead of allocating.
struct FixedRegion {
    float* base;      // device address of the region
    size_t floats;    // its size
};
// Installs `r` for the calling thread until the guard ends (nested use is not supported).
struct FixedRegionGuard {
    explicit FixedRegionGuard(const FixedRegion* r);
    ~FixedRegionGuard();
};
// True while the calling thread holds a fixed region (no device allocation may happen).
bool fixed_region_active();

// Ensures scratch_ptr(g) holds at least `floats` floats (hipMalloc on growth; synchronises the device then).
int scratch_reserve(omg_geom* g, size_t floats);
float* scratch_ptr(const omg_geom* g);

// Ensures work_ptr(g) holds the fused-iteration buffers for an image of N voxels.
int work_reserve(omg_geom* g, uint32_t N);
float* work_ptr(const omg_geom* g);

// Ensures q_ptr(g) holds `floats` floats of per-measurement work memory (the voxel-driven planar path's q and w
// buffers, the forward projections of the transmission OSEM iteration and of projector 4's fused data step).
int q_reserve(omg_geom* g, size_t floats);
float* q_ptr(const omg_geom* g);

// ---- direct-plane (planar) launches, shared by the Siddon and orthogonal projectors (siddon.hip) ----
// Workgroups of a planar launch over [first, first + n): n_planes planes, `chunk` LORs per workgroup,
// `blocks` = n_planes * groups workgroups; block b owns plane b % n_planes, LORs from (b / n_planes) * chunk.
struct PlanarShape { uint32_t n_planes, chunk, blocks; };
PlanarShape planar_shape(const omg_geom* g, int64_t n);
constexpr int kPlanarTPB = 640;   // lanes per planar workgroup (10 wave64)
// Adds the partial slices in g->scratch into rhs (and sens when non-null).
int planar_reduce_partials(const omg_geom* g, int64_t first, int64_t n, float* rhs, float* sens, hipStream_t s);
// EM / OSL update of x from the partial slices of a fused planar launch (needs work_reserve).
int planar_update(const omg_geom* g, int64_t first, int64_t n_meas, float* x, float* S, const float* dU,
                  int32_t compute_sens, float epps, int32_t positivity, hipStream_t s);
// Kernel timing hooks (omg_timing_enable / omg_timing_collect): HIP events around one launch.
void timing_begin(hipStream_t s, hipEvent_t* stop);
void timing_end(hipStream_t s, hipEvent_t stop);
// True when no two planes of [first, first + n_planes) lie in the same image slice.
bool unique_slices(const omg_geom* g, int64_t first, uint32_t n_planes);

}  // namespace omg
