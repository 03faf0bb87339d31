// frames.hip -- batched dynamic-study OSEM: every frame of a study advances through one subset step in
// one launch.
//
// The reference reconstructs the frames (timesteps) of a dynamic study one after another inside each
// subset: FP -> ratio -> BP per frame, then the update of all frames (reconstructionAF.h:1089-1453,
// per-frame data :429-436). The frames share the geometry, so the Siddon traversal of a LOR -- the
// expensive part -- is the same for all of them. Here a lane traces its LOR once per tile of FT frames
// and carries FT sums:
//   * stacks are frame-major: x[frame][voxel], rhs[frame][voxel], meas[frame][lor] (each frame in the
//     subset order of a single-frame run), so every frame keeps the cache footprint of a single-frame
//     reconstruction and a tile of FT frames touches FT such images (an interleaved [voxel][frame] layout
//     was measured 2x slower: each load then pulls a cache line of other frames' data);
//   * the sensitivity image of a subset is frame-independent (omg_frames_sensitivity) and computed once;
//   * a persistent grid walks (LOR block, frame tile) work items, so one launch covers all frames;
//   * the update kernel clears rhs behind itself, so a whole iteration (16 subset steps) replays from a
//     captured hipGraph with no memset nodes.
// Per frame, each FP sum, ratio and BP term is computed exactly as the single-frame fused kernel
// computes it (k_siddon_osem in siddon.hip), so frames agree with single-frame runs up to the order of
// the float atomics. Subsets are those of the geometry: plane subsets (types 8-11) or the measurement
// index table of types 1-7, which siddon_setup reads per LOR as the single-frame kernels do.
#include "geom.hpp"
#include "rdp_device.hpp"

#include <algorithm>
#include <vector>

namespace omg {

// traversal steps traced before their loads are consumed: 8 like the single-frame gather for up to two
// frames per lane, 4 above (register budget)
template <int FT>
constexpr int frame_batch() { return FT <= 2 ? 8 : 4; }

template <int FT>
__device__ __forceinline__ void load_frames(const float* __restrict__ p, size_t stride, float (&v)[FT]) {
#pragma unroll
    for (int f = 0; f < FT; f++) v[f] = p[(size_t)f * stride];
}

// Sum of len * x[f][voxel] along LOR lo for the FT frames at x, x + N, ..., in traversal order.
template <int FT>
__device__ __forceinline__ bool frames_gather(const SidArgs& g, int64_t lo, const float* __restrict__ x, uint32_t N,
                                              float (&ax)[FT], float& weight) {
#pragma unroll
    for (int f = 0; f < FT; f++) ax[f] = 0.f;
    Walk w;
    if (!siddon_setup<false>(g, lo, w, weight)) return false;
    if (w.perp) {
        perp_walk(w, [&](uint32_t ind, float len) {
            float v[FT];
            load_frames<FT>(x + ind, N, v);
#pragma unroll
            for (int f = 0; f < FT; f++) ax[f] += len * v[f];
        });
        return true;
    }
    constexpr int kFB = frame_batch<FT>();
    while (w.left > 0u) {
        uint32_t ind[kFB];
        float len[kFB], v[kFB][FT];
#pragma unroll
        for (int k = 0; k < kFB; k++) walk_step<false>(g, w, ind[k], len[k]);
#pragma unroll
        for (int k = 0; k < kFB; k++) load_frames<FT>(x + (len[k] > 0.f ? ind[k] : 0u), N, v[k]);
#pragma unroll
        for (int k = 0; k < kFB; k++)
#pragma unroll
            for (int f = 0; f < FT; f++) ax[f] = (len[k] > 0.f) ? ax[f] + len[k] * v[k][f] : ax[f];
    }
    return true;
}

// One subset step of all frames: FP, ratio, BP into rhs[frame][voxel]. Work item = (frame tile, LOR block).
// `x`, `rhs` point at frame 0 of their stacks (frame stride N); `meas` at frame 0's first LOR of the subset
// (frame stride meas_stride).
// rhs_t (or null): y-fastest stack [frame][z][x][y] that x-major LORs accumulate into (siddon_trace_split),
// folded into rhs by k_frames_update_t.
template <int FT, int RPB>
__global__ __launch_bounds__(RPB) void k_frames_osem(SidArgs a, uint32_t N, const float* __restrict__ x,
                                                     const float* __restrict__ meas, int64_t meas_stride,
                                                     float* __restrict__ rhs, float* __restrict__ rhs_t, float epps,
                                                     uint32_t lor_blocks, uint32_t n_items) {
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t tile = item / lor_blocks;
        const uint32_t lb = item - tile * lor_blocks;
        const int64_t lo = (int64_t)lb * RPB + threadIdx.x;
        if (lo >= a.nmeas) continue;
        const uint32_t f0 = tile * FT;
        float ax[FT], w = 0.f;
        if (!frames_gather<FT>(a, lo, x + (size_t)f0 * N, N, ax, w)) continue;
        float r[FT];
        bool any = false;
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const float yhat = ax[f] * w;
            float q = meas[(size_t)(f0 + f) * meas_stride + lo] / (yhat + epps);
            if (isnan(q) || isinf(q)) q = epps;
            r[f] = q;
            any |= (q != 0.f);
        }
        if (!any) continue;
        // frames whose ratio is 0 (no counts in this LOR) add nothing: skip their atomics, as the
        // single-frame kernel skips the whole LOR
        if (rhs_t) {
            siddon_trace_split(a, lo, w, [&](uint32_t ind, float len, bool t) {
                const float lw = len * w;
                float* p = (t ? rhs_t : rhs) + (size_t)f0 * N + ind;
#pragma unroll
                for (int f = 0; f < FT; f++)
                    if (r[f] != 0.f) unsafeAtomicAdd(p + (size_t)f * N, r[f] * lw);
            });
        } else {
            siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
                const float lw = len * w;
                float* p = rhs + (size_t)f0 * N + ind;
#pragma unroll
                for (int f = 0; f < FT; f++)
                    if (r[f] != 0.f) unsafeAtomicAdd(p + (size_t)f * N, r[f] * lw);
            });
        }
    }
}

// yhat[f][lo] = A_k x_f for every frame of a tile (0 for LORs that miss the image); frame stride y_stride.
template <int FT>
__global__ __launch_bounds__(256) void k_frames_fp(SidArgs a, uint32_t N, const float* __restrict__ x,
                                                   float* __restrict__ yhat, int64_t y_stride) {
    const int64_t lo = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (lo >= a.nmeas) return;
    const uint32_t f0 = blockIdx.y * FT;
    float ax[FT], w = 0.f;
    const bool hit = frames_gather<FT>(a, lo, x + (size_t)f0 * N, N, ax, w);
#pragma unroll
    for (int f = 0; f < FT; f++) yhat[(size_t)(f0 + f) * y_stride + lo] = hit ? ax[f] * w : 0.f;
}

// out[f] += sum_i y log(yhat) - yhat over [F][n] stacks, terms with y == 0 contribute -yhat (FP64 sums).
__global__ __launch_bounds__(256) void k_frames_ll(const float* __restrict__ y, const float* __restrict__ yhat,
                                                   int64_t n, uint32_t F, double* __restrict__ out) {
    const uint32_t f = blockIdx.y;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double yi = y[(size_t)f * n + i], hi = yhat[(size_t)f * n + i];
        s += (yi != 0.0 ? yi * log(hi) : 0.0) - hi;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out + f, s);
}

// S += A_k^T 1 (frame-independent sensitivity of one subset).
__global__ __launch_bounds__(256) void k_frames_sens(SidArgs a, float* __restrict__ S) {
    const int64_t lo = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (lo >= a.nmeas) return;
    float w = 0.f;
    siddon_trace(a, lo, w, [&](uint32_t ind, float len) { unsafeAtomicAdd(S + ind, len * w); });
}

__global__ __launch_bounds__(256) void k_frames_clamp(float* __restrict__ S, uint32_t N, float epps) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n < N && S[n] < epps) S[n] = epps;
}

__global__ __launch_bounds__(256) void k_frames_fill(float* __restrict__ p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// dU[f][n] = beta * grad RDP(x_f)[n], zero outside the volume: the single-frame gradient (omg_rdp_gradient,
// boundary 0; RDP.h:4-78) on every frame of the stack, lanes summed in the same order.
__global__ __launch_bounds__(256) void k_frames_rdp(const float* __restrict__ u, float* __restrict__ dU, uint32_t Nx,
                                                    uint32_t Ny, uint32_t Nz, uint32_t F, float gamma, float epps,
                                                    float beta) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t Nxy = Nx * Ny;
    const size_t N = (size_t)Nxy * Nz;
    if (i >= N * F) return;
    const uint32_t n = (uint32_t)(i % N);
    const uint32_t z = n / Nxy, rem = n - z * Nxy, y = rem / Nx, x = rem - y * Nx;
    const float uj = u[i];
    const float xp = (x + 1 < Nx) ? u[i + 1] : 0.f;
    const float xm = (x > 0) ? u[i - 1] : 0.f;
    const float yp = (y + 1 < Ny) ? u[i + Nx] : 0.f;
    const float ym = (y > 0) ? u[i - Nx] : 0.f;
    const float zp = (z + 1 < Nz) ? u[i + Nxy] : 0.f;
    const float zm = (z > 0) ? u[i - Nxy] : 0.f;
    float l0 = rdp_term(uj, xp, gamma, epps);
    l0 = l0 + rdp_term(uj, yp, gamma, epps);
    l0 = l0 + rdp_term(uj, zp, gamma, epps);
    float l1 = rdp_term(uj, xm, gamma, epps);
    l1 = l1 + rdp_term(uj, ym, gamma, epps);
    l1 = l1 + rdp_term(uj, zm, gamma, epps);
    if (isnan(l0)) l0 = 0.f;
    if (isnan(l1)) l1 = 0.f;
    dU[i] = 0.f + beta * (l0 + l1);
}

// OSEM: x[f][n] = x / S[n] * (rhs + epps) (subiterStep.h:111-121, EM algorithms.h:12-15);
// OSL (dU != null): x = x / (S[n] + dU) * rhs (subiterStep.h:219-231, OSL algorithms.h:24-27);
// positivity x < epps -> epps; rhs = 0 behind.
__global__ __launch_bounds__(256) void k_frames_update(uint32_t N, uint32_t F, float* __restrict__ x,
                                                       const float* __restrict__ S, float* __restrict__ rhs,
                                                       const float* __restrict__ dU, float epps, int pos) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)N * F) return;
    const float s = S[i % N];
    float xn = dU ? x[i] / (s + dU[i]) * rhs[i] : x[i] / s * (rhs[i] + epps);
    if (pos && xn < epps) xn = epps;
    x[i] = xn;
    rhs[i] = 0.f;
}

// k_frames_update with the y-fastest accumulator rhs_t folded in. Per 32 x 32 (x, y) tile of one
// (frame, slice) plane, rhs_t's tile is read along y, transposed through LDS and cleared behind; then
// x, rhs (cleared behind) and dU are read along x. grid = (ceil(Nx / 32), ceil(Ny / 32), Nz * F),
// block = (32, 8).
__global__ __launch_bounds__(256) void k_frames_update_t(uint32_t Nx, uint32_t Ny, uint32_t Nz, float* __restrict__ x,
                                                         const float* __restrict__ S, float* __restrict__ rhs,
                                                         float* __restrict__ rhs_t, const float* __restrict__ dU,
                                                         float epps, int pos) {
    __shared__ float tile[32][33];
    const uint32_t x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const size_t Nyx = (size_t)Nx * Ny;
    const size_t off = (size_t)blockIdx.z * Nyx;              // (frame, slice) plane
    const size_t soff = (size_t)(blockIdx.z % Nz) * Nyx;      // the slice in S
    for (uint32_t k = threadIdx.y; k < 32; k += 8) {
        const uint32_t xx = x0 + k, yy = y0 + threadIdx.x;
        float v = 0.f;
        if (xx < Nx && yy < Ny) {
            float* p = rhs_t + off + (size_t)xx * Ny + yy;
            v = *p;
            *p = 0.f;
        }
        tile[k][threadIdx.x] = v;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.y; k < 32; k += 8) {
        const uint32_t yy = y0 + k, xx = x0 + threadIdx.x;
        if (xx >= Nx || yy >= Ny) continue;
        const size_t i = off + (size_t)yy * Nx + xx;
        const float r = rhs[i] + tile[threadIdx.x][k];
        const float sv = S[soff + (size_t)yy * Nx + xx];
        float xn = dU ? x[i] / (sv + dU[i]) * r : x[i] / sv * (r + epps);
        if (pos && xn < epps) xn = epps;
        x[i] = xn;
        rhs[i] = 0.f;
    }
}

struct FramesStep {
    SidArgs a;
    uint32_t F, N;
    float* x;
    const float* meas;
    int64_t meas_stride;
    const float* S;
    float* rhs;
    float* rhs_t;       // y-fastest accumulator stack [F][N] (zero between steps), or null
    float* dU;          // OSL-RDP work stack [F][N], or null for OSEM
    float beta, gamma;
    uint32_t Nx, Ny, Nz;
    float epps;
    int pos;
    int rpb, ft, grid;
};

// Workgroups of one resident wave of k_frames_osem<FT, RPB> on the current device.
template <int FT, int RPB>
static int resident_blocks() {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_frames_osem<FT, RPB>, RPB, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return cus * per_cu;
}

template <int RPB>
static int resident_blocks_ft(int ft) {
    switch (ft) {
        case 1: return resident_blocks<1, RPB>();
        case 2: return resident_blocks<2, RPB>();
        case 4: return resident_blocks<4, RPB>();
        default: return resident_blocks<8, RPB>();
    }
}

static int resident_blocks_for(int rpb, int ft) {
    switch (rpb) {
        case 64: return resident_blocks_ft<64>(ft);
        case 128: return resident_blocks_ft<128>(ft);
        default: return resident_blocks_ft<256>(ft);
    }
}

template <int FT, int RPB>
static hipError_t launch_data(const FramesStep& st, hipStream_t s) {
    const uint32_t lor_blocks = (uint32_t)((st.a.nmeas + RPB - 1) / RPB);
    const uint32_t n_items = lor_blocks * (st.F / FT);
    uint32_t grid = (uint32_t)st.grid;
    if (grid > n_items) grid = n_items;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_frames_osem<FT, RPB>), dim3(grid), dim3(RPB), 0, s, st.a, st.N, st.x, st.meas,
                       st.meas_stride, st.rhs, st.rhs_t, st.epps, lor_blocks, n_items);
    return hipGetLastError();
}

template <int RPB>
static hipError_t launch_data_ft(const FramesStep& st, hipStream_t s) {
    switch (st.ft) {
        case 1: return launch_data<1, RPB>(st, s);
        case 2: return launch_data<2, RPB>(st, s);
        case 4: return launch_data<4, RPB>(st, s);
        default: return launch_data<8, RPB>(st, s);
    }
}

static int launch_step(const FramesStep& st, hipStream_t s) {
    const size_t NF = (size_t)st.N * st.F;
    if (st.dU) {
        hipLaunchKernelGGL(k_frames_rdp, dim3((unsigned)((NF + 255) / 256)), dim3(256), 0, s, st.x, st.dU, st.Nx, st.Ny,
                           st.Nz, st.F, st.gamma, st.epps, st.beta);
        OMG_CHECK_LAUNCH();
    }
    hipError_t e;
    switch (st.rpb) {
        case 64: e = launch_data_ft<64>(st, s); break;
        case 128: e = launch_data_ft<128>(st, s); break;
        default: e = launch_data_ft<256>(st, s); break;
    }
    if (e != hipSuccess) {
        set_error("frames OSEM launch failed: %s", hipGetErrorString(e));
        return OMG_ERR_HIP;
    }
    if (st.rhs_t) {
        const dim3 tg((st.Nx + 31) / 32, (st.Ny + 31) / 32, st.Nz * st.F);
        hipLaunchKernelGGL(k_frames_update_t, tg, dim3(32, 8), 0, s, st.Nx, st.Ny, st.Nz, st.x, st.S, st.rhs, st.rhs_t,
                           st.dU, st.epps, st.pos);
    } else {
        hipLaunchKernelGGL(k_frames_update, dim3((unsigned)((NF + 255) / 256)), dim3(256), 0, s, st.N, st.F, st.x,
                           st.S, st.rhs, st.dU, st.epps, st.pos);
    }
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

// The frames kernels accumulate x-major LORs into a y-fastest stack on PET sinograms unless
// omg_geom_desc.reserved[3] != 0 (as the single-frame one-lane-per-LOR kernels do).
static bool frames_split(const omg_geom* g) { return g->d.mode == OMG_MODE_PET_SINO && g->d.reserved[3] == 0; }

static int make_step(const omg_geom* g, int64_t first, int64_t n_meas, int32_t n_frames, float* x, const float* meas,
                     int64_t meas_stride, const float* S, float* rhs, float* rhs_t, float* dU, float beta, float gamma,
                     float epps, int32_t positivity, int32_t rays_per_block, int32_t frames_per_tile,
                     int32_t grid_blocks, FramesStep& st) {
    OMG_REQUIRE(g != nullptr, "null geometry");
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON, "batched frames run the improved Siddon projector");
    OMG_REQUIRE(g->d.mode != OMG_MODE_CT_FLAT, "batched frames are emission (PET) data");
    OMG_REQUIRE(g->d.tof_bins <= 1, "batched frames do not take time-of-flight data");
    OMG_REQUIRE(std::max(g->d.rays_xy, 1u) * std::max(g->d.rays_z, 1u) == 1, "batched frames trace one ray per LOR");
    OMG_REQUIRE(x && meas && S && rhs, "null pointer argument");
    OMG_REQUIRE(n_frames >= 1, "n_frames must be >= 1");
    OMG_REQUIRE(rays_per_block == 64 || rays_per_block == 128 || rays_per_block == 256,
                "rays_per_block must be 64, 128 or 256");
    OMG_REQUIRE(frames_per_tile == 1 || frames_per_tile == 2 || frames_per_tile == 4 || frames_per_tile == 8,
                "frames_per_tile must be 1, 2, 4 or 8");
    OMG_REQUIRE(n_frames % frames_per_tile == 0, "n_frames (%d) must be a multiple of frames_per_tile (%d)", n_frames,
                frames_per_tile);
    OMG_REQUIRE(grid_blocks >= 0, "grid_blocks must be >= 0 (0 = one wave of resident workgroups)");
    OMG_REQUIRE(beta == 0.f || dU != nullptr, "OSL-RDP (beta != 0) needs the dU work stack");
    OMG_REQUIRE(n_frames == 1 || meas_stride >= n_meas, "meas_stride must cover the subset's measurements");
    const int rc = check_range(g, first, n_meas);
    if (rc) return rc;
    st.a = make_sid_args(g, first, n_meas);
    st.F = (uint32_t)n_frames;
    st.N = g->d.Nx * g->d.Ny * g->d.Nz;
    st.x = x; st.meas = meas; st.meas_stride = meas_stride; st.S = S; st.rhs = rhs; st.rhs_t = rhs_t;
    st.dU = beta != 0.f ? dU : nullptr;
    st.beta = beta; st.gamma = gamma;
    st.Nx = g->d.Nx; st.Ny = g->d.Ny; st.Nz = g->d.Nz;
    st.epps = epps; st.pos = positivity != 0;
    st.rpb = rays_per_block; st.ft = frames_per_tile;
    st.grid = grid_blocks > 0 ? grid_blocks : resident_blocks_for(rays_per_block, frames_per_tile);
    return OMG_OK;
}

}  // namespace omg

using namespace omg;

struct omg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float* rhs_t = nullptr;   // the graph's own y-fastest accumulator stack (frames_split), zeroed by the graph
};

// Ensures g->frames_t holds n floats of y-fastest accumulator, zero (allocated and cleared on growth).
static int frames_t_reserve(omg_geom* g, size_t n, hipStream_t s) {
    if (g->frames_t_floats >= n) return OMG_OK;
    OMG_CHECK_HIP(hipDeviceSynchronize());
    OMG_CHECK_HIP(hipFree(g->frames_t));
    g->frames_t = nullptr;
    g->frames_t_floats = 0;
    OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g->frames_t), n * sizeof(float)));
    hipLaunchKernelGGL(k_frames_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g->frames_t, n, 0.f);
    OMG_CHECK_LAUNCH();
    g->frames_t_floats = n;
    return OMG_OK;
}

extern "C" int omg_frames_sensitivity(const omg_geom* g, int64_t first, int64_t n_meas, float* S, float epps,
                                      void* stream) {
    OMG_REQUIRE(g != nullptr && S != nullptr, "null argument");
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON && g->d.mode != OMG_MODE_CT_FLAT,
                "batched frames run the improved Siddon projector on emission data");
    OMG_REQUIRE(g->d.tof_bins <= 1, "batched frames do not take time-of-flight data");
    OMG_REQUIRE(std::max(g->d.rays_xy, 1u) * std::max(g->d.rays_z, 1u) == 1, "batched frames trace one ray per LOR");
    const int rc = check_range(g, first, n_meas);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    hipLaunchKernelGGL(k_frames_fill, dim3(grid_for(N, 256)), dim3(256), 0, s, S, (size_t)N, 0.f);
    OMG_CHECK_LAUNCH();
    if (n_meas > 0) {
        const SidArgs a = make_sid_args(g, first, n_meas);
        hipLaunchKernelGGL(k_frames_sens, dim3(grid_for(n_meas, 256)), dim3(256), 0, s, a, S);
        OMG_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_frames_clamp, dim3(grid_for(N, 256)), dim3(256), 0, s, S, N, epps);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_frames_forward(const omg_geom* g, int64_t first, int64_t n_meas, int32_t n_frames, const float* x,
                                  float* yhat, int64_t yhat_stride, void* stream) {
    OMG_REQUIRE(g != nullptr && x && yhat && n_frames >= 1, "bad arguments");
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON && g->d.mode != OMG_MODE_CT_FLAT,
                "batched frames run the improved Siddon projector on emission data");
    OMG_REQUIRE(g->d.tof_bins <= 1, "batched frames do not take time-of-flight data");
    OMG_REQUIRE(std::max(g->d.rays_xy, 1u) * std::max(g->d.rays_z, 1u) == 1, "batched frames trace one ray per LOR");
    OMG_REQUIRE(n_frames == 1 || yhat_stride >= n_meas, "yhat_stride must cover the subset's measurements");
    const int rc = check_range(g, first, n_meas);
    if (rc) return rc;
    if (n_meas == 0) return OMG_OK;
    const SidArgs a = make_sid_args(g, first, n_meas);
    hipStream_t s = as_stream(stream);
    const unsigned gx = grid_for(n_meas, 256);
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    if (n_frames % 4 == 0)
        hipLaunchKernelGGL(k_frames_fp<4>, dim3(gx, (unsigned)(n_frames / 4)), dim3(256), 0, s, a, N, x, yhat,
                           yhat_stride);
    else
        hipLaunchKernelGGL(k_frames_fp<1>, dim3(gx, (unsigned)n_frames), dim3(256), 0, s, a, N, x, yhat, yhat_stride);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_frames_loglik(const float* y, const float* yhat, int64_t n, int32_t n_frames, double* out,
                                 void* stream) {
    OMG_REQUIRE(out && n >= 0 && n_frames >= 1 && (n == 0 || (y && yhat)), "bad arguments");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_frames_fill, dim3(1), dim3(256), 0, s, reinterpret_cast<float*>(out), (size_t)(2 * n_frames),
                       0.f);
    OMG_CHECK_LAUNCH();
    if (n == 0) return OMG_OK;
    const unsigned gx = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_frames_ll, dim3(gx, (unsigned)n_frames), dim3(256), 0, s, y, yhat, n, (uint32_t)n_frames, out);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_frames_osem_step(const omg_geom* g, int64_t first, int64_t n_meas, int32_t n_frames, float* x,
                                    const float* meas, int64_t meas_stride, const float* S, float* rhs, float* dU,
                                    float beta, float gamma, float epps, int32_t positivity, int32_t rays_per_block,
                                    int32_t frames_per_tile, int32_t grid_blocks, void* stream) {
    FramesStep st;
    float* rhs_t = nullptr;
    if (g && n_frames >= 1 && frames_split(g)) {
        const size_t NF = (size_t)g->d.Nx * g->d.Ny * g->d.Nz * (size_t)n_frames;
        const int rc = frames_t_reserve(const_cast<omg_geom*>(g), NF, as_stream(stream));
        if (rc) return rc;
        rhs_t = g->frames_t;
    }
    int rc = make_step(g, first, n_meas, n_frames, x, meas, meas_stride, S, rhs, rhs_t, dU, beta, gamma, epps,
                       positivity, rays_per_block, frames_per_tile, grid_blocks, st);
    if (rc) return rc;
    return launch_step(st, as_stream(stream));
}

extern "C" int omg_frames_graph_create(const omg_geom* g, int32_t n_subsets, const int64_t* first,
                                       const int64_t* n_meas, const int64_t* meas_offset, int32_t n_frames, float* x,
                                       const float* meas, int64_t meas_stride, const float* S, float* rhs, float* dU,
                                       float beta,
                                       float gamma, float epps, int32_t positivity, int32_t rays_per_block,
                                       int32_t frames_per_tile, int32_t grid_blocks, omg_graph** out) {
    OMG_REQUIRE(out != nullptr, "null output handle");
    *out = nullptr;
    OMG_REQUIRE(n_subsets >= 1 && first && n_meas && meas_offset, "subset tables are required");
    OMG_REQUIRE(g != nullptr, "null geometry");
    OMG_REQUIRE(g->d.tof_bins <= 1, "batched frames do not take time-of-flight data");
    OMG_REQUIRE(std::max(g->d.rays_xy, 1u) * std::max(g->d.rays_z, 1u) == 1, "batched frames trace one ray per LOR");
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    OMG_REQUIRE(n_frames >= 1, "n_frames must be >= 1");
    float* rhs_t = nullptr;
    const size_t NF = (size_t)N * (size_t)n_frames;
    if (frames_split(g)) OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&rhs_t), NF * sizeof(float)));
    std::vector<FramesStep> steps((size_t)n_subsets);
    for (int32_t k = 0; k < n_subsets; k++) {
        int rc = meas_offset[k] >= 0 ? OMG_OK : OMG_ERR_ARG;
        if (rc) set_error("negative measurement offset");
        if (!rc)
            rc = make_step(g, first[k], n_meas[k], n_frames, x, meas + meas_offset[k], meas_stride, S + (size_t)k * N,
                           rhs, rhs_t, dU, beta, gamma, epps, positivity, rays_per_block, frames_per_tile,
                           grid_blocks, steps[(size_t)k]);
        if (rc) {
            hipFree(rhs_t);
            return rc;
        }
    }
    hipStream_t cs;
    OMG_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    omg_graph* h = new omg_graph;
    h->rhs_t = rhs_t;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = OMG_OK;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_frames_fill, dim3((unsigned)((NF + 255) / 256)), dim3(256), 0, cs, rhs, NF, 0.f);
        if (rhs_t)
            hipLaunchKernelGGL(k_frames_fill, dim3((unsigned)((NF + 255) / 256)), dim3(256), 0, cs, rhs_t, NF, 0.f);
        for (const FramesStep& st : steps) {
            rc = launch_step(st, cs);
            if (rc) break;
        }
        const hipError_t e2 = hipStreamEndCapture(cs, &h->graph);
        if (!rc && e2 != hipSuccess) e = e2;
    }
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0);
    hipStreamDestroy(cs);
    if (rc || e != hipSuccess) {
        if (!rc) set_error("hipGraph capture of the frames iteration failed: %s", hipGetErrorString(e));
        if (h->exec) hipGraphExecDestroy(h->exec);
        if (h->graph) hipGraphDestroy(h->graph);
        hipFree(h->rhs_t);
        delete h;
        return rc ? rc : OMG_ERR_HIP;
    }
    *out = h;
    return OMG_OK;
}

extern "C" int omg_graph_launch(omg_graph* h, void* stream) {
    OMG_REQUIRE(h != nullptr && h->exec != nullptr, "null graph");
    OMG_CHECK_HIP(hipGraphLaunch(h->exec, as_stream(stream)));
    return OMG_OK;
}

extern "C" int omg_graph_destroy(omg_graph* h) {
    if (!h) return OMG_OK;
    if (h->exec) hipGraphExecDestroy(h->exec);
    if (h->graph) hipGraphDestroy(h->graph);
    hipFree(h->rhs_t);
    delete h;
    return OMG_OK;
}
