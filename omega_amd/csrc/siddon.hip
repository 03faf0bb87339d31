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

#include <utility>
#include <vector>

namespace omg {

constexpr int kBlock = 256;

// Kernel timing around the projector launches, the counterpart of the reference's verbose >= 3
// HIP event timers (source/cpp/ProjectorClass.h:3594-3599). Off unless enabled.
struct TimingState {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t used = 0;
    hipError_t err = hipSuccess;   // first event failure since the last collect
};
static TimingState g_timing;

void timing_begin(hipStream_t s, hipEvent_t* stop) {
    *stop = nullptr;
    if (!g_timing.on) return;
    if (g_timing.used == g_timing.ev.size()) {
        hipEvent_t a, b;
        hipError_t e = hipEventCreate(&a);
        if (e == hipSuccess) e = hipEventCreate(&b);
        if (e != hipSuccess) {
            g_timing.err = e;
            return;
        }
        g_timing.ev.push_back({a, b});
    }
    auto& pr = g_timing.ev[g_timing.used];
    const hipError_t e = hipEventRecord(pr.first, s);
    if (e != hipSuccess) {
        g_timing.err = e;
        return;
    }
    g_timing.used++;
    *stop = pr.second;
}

void timing_end(hipStream_t s, hipEvent_t stop) {
    if (stop) hipEventRecord(stop, s);
}


__global__ void k_fill_f32(float* __restrict__ p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- forward projection -------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_siddon_fp(SidArgs a, const float* __restrict__ img, float* __restrict__ y) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    if (n_rays(a) > 1) {   // the rays' weighted sums, added in ray order (projector_functions.h:2036-2056)
        float total = 0.f;
        for (uint32_t ray = 0; ray < n_rays(a); ray++) {
            float ax, w = 0.f;
            if (siddon_gather<false>(a, lo, img, ax, w, (int)ray)) total += ax * w;
        }
        y[lo] = total;
        return;
    }
    float ax, w = 0.f;
    const bool hit = siddon_gather<false>(a, lo, img, ax, w);
    y[lo] = hit ? (a.mode == MODE_CT_FLAT ? ax : ax * w) : 0.f;
}

// ---- attenuation factors ---------------------------------------------------------------------------
// f[lo] = exp(jelppi), jelppi = sum of len * -mu[voxel] over the LOR's walk in traversal order, with
// the walk's (clipped) lengths: the reference's forward-projection sum (compute_attenuation,
// projector_functions.h:597-601, :1938-1939, :2026-2027). An axis-parallel LOR integrates mu over its
// whole image row with the voxel size as length (perpendicular_elements, :1193-1205). A LOR that misses
// the image gets 1. `a` carries no weights (its atten and norm are null). With several rays per LOR,
// blockIdx.y is the ray and its factors form block ray of f (each ray's own walk, :1480, :1826, :2027).
__global__ __launch_bounds__(kBlock) void k_atten_factor(SidArgs a, const float* __restrict__ mu,
                                                         float* __restrict__ f) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    const int ray = (int)blockIdx.y;
    Walk w;
    float weight, jel = 0.f;
    if (siddon_setup<false>(a, lo, w, weight, false, ray)) {
        if (w.perp) {
            for (uint32_t i = 0; i < w.p_nrun; i++) jel += w.p_d * -mu[w.p_row + i * w.p_rstride];
        } else {
            while (w.left > 0u) {
                uint32_t ind;
                float len;
                walk_step<false>(a, w, ind, len);
                if (len > 0.f) jel += len * -mu[ind];
            }
        }
    }
    f[(int64_t)ray * a.nmeas + lo] = expf(jel);
}

// f holds n_rays blocks of n factors.
int atten_factors(const omg_geom* g, const float* mu, float* f, int64_t n, hipStream_t s) {
    SidArgs a = make_sid_args(g, 0, n);
    a.norm = nullptr;
    a.atten = nullptr;
    a.scat = nullptr;
    hipLaunchKernelGGL(k_atten_factor, dim3(grid_for(n, kBlock), a.rays_xy * a.rays_z), dim3(kBlock), 0, s, a, mu, f);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

// ---- backprojection -----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_siddon_bp(SidArgs a, const float* __restrict__ yin, float* __restrict__ rhs,
                                                      float* __restrict__ sens) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    const float yi = yin[lo];
    if (yi == 0.f && sens == nullptr) return;  // adds of +0 leave rhs unchanged
    float w = 0.f;
    if (n_rays(a) > 1) {   // each ray with its own weight
        for (uint32_t ray = 0; ray < n_rays(a); ray++)
            siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
                const float lw = len * w;
                unsafeAtomicAdd(rhs + ind, yi * lw);
                if (sens) unsafeAtomicAdd(sens + ind, lw);
            }, (int)ray);
        return;
    }
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

// The same with x-major LORs accumulated into y-fastest buffers rhs_t / sens_t (siddon_trace_split),
// added into rhs / sens by k_add_transposed afterwards.
template <bool SENS>
__global__ __launch_bounds__(kBlock) void k_siddon_bp_split(SidArgs a, const float* __restrict__ yin,
                                                            float* __restrict__ rhs, float* __restrict__ sens,
                                                            float* __restrict__ rhs_t, float* __restrict__ sens_t) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    const float yi = yin[lo];
    if (!SENS && yi == 0.f) return;
    float w = 0.f;
    const auto visit = [&](uint32_t ind, float len, bool t) {
        const float lw = len * w;
        unsafeAtomicAdd((t ? rhs_t : rhs) + ind, yi * lw);
        if (SENS) unsafeAtomicAdd((t ? sens_t : sens) + ind, lw);
    };
    if (n_rays(a) > 1) {
        for (uint32_t ray = 0; ray < n_rays(a); ray++) siddon_trace_split(a, lo, w, visit, (int)ray);
        return;
    }
    siddon_trace_split(a, lo, w, visit);
}

// dst[x + y*Nx + z*Nx*Ny] += src_t[y + x*Ny + z*Nx*Ny], through 32 x 32 LDS tiles (both sides coalesced).
// grid = (ceil(Nx / 32), ceil(Ny / 32), Nz), block = (32, 8).
__global__ __launch_bounds__(256) void k_add_transposed(const float* __restrict__ src_t, float* __restrict__ dst,
                                                        uint32_t Nx, uint32_t Ny) {
    __shared__ float tile[32][33];
    const uint32_t x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const size_t zoff = (size_t)blockIdx.z * Nx * Ny;
    for (uint32_t k = threadIdx.y; k < 32; k += 8) {   // read: y fastest
        const uint32_t x = x0 + k, y = y0 + threadIdx.x;
        tile[k][threadIdx.x] = (x < Nx && y < Ny) ? src_t[zoff + (size_t)x * Ny + y] : 0.f;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.y; k < 32; k += 8) {   // write: x fastest
        const uint32_t y = y0 + k, x = x0 + threadIdx.x;
        if (x < Nx && y < Ny) dst[zoff + (size_t)y * Nx + x] += tile[threadIdx.x][k];
    }
}

// ---- fused OSEM data step: FP, ratio, BP for one LOR ---------------------------------------------
// With rhs_t (/ sens_t) non-null, x-major LORs accumulate into y-fastest buffers, which the caller adds
// back (k_add_transposed).
__global__ __launch_bounds__(kBlock) void k_siddon_osem(SidArgs a, const float* __restrict__ img,
                                                        const float* __restrict__ meas, float* __restrict__ rhs,
                                                        float* __restrict__ sens, float* __restrict__ rhs_t,
                                                        float* __restrict__ sens_t, float epps) {
    const int64_t lo = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lo >= a.nmeas) return;
    const uint32_t nr = n_rays(a);
    float ax, w = 0.f, yhat;
    if (nr > 1) {   // all rays missing the image: nothing to backproject, as with one ray
        yhat = 0.f;
        bool any = false;
        for (uint32_t ray = 0; ray < nr; ray++) {
            if (siddon_gather<false>(a, lo, img, ax, w, (int)ray)) {
                yhat += ax * w;
                any = true;
            }
        }
        if (!any) return;
    } else {
        if (!siddon_gather<false>(a, lo, img, ax, w)) return;
        yhat = ax * w;
    }
    float r = meas[lo] / ratio_den(a, lo, yhat, epps);
    if (isnan(r) || isinf(r)) r = epps;
    if (r == 0.f && sens == nullptr) return;
    if (nr > 1) {
        for (uint32_t ray = 0; ray < nr; ray++) {
            if (rhs_t) {
                siddon_trace_split(a, lo, w, [&](uint32_t ind, float len, bool t) {
                    const float lw = len * w;
                    unsafeAtomicAdd((t ? rhs_t : rhs) + ind, r * lw);
                    if (sens) unsafeAtomicAdd((t ? sens_t : sens) + ind, lw);
                }, (int)ray);
            } else {
                siddon_trace(a, lo, w, [&](uint32_t ind, float len) {
                    const float lw = len * w;
                    unsafeAtomicAdd(rhs + ind, r * lw);
                    if (sens) unsafeAtomicAdd(sens + ind, lw);
                }, (int)ray);
            }
        }
        return;
    }
    if (rhs_t) {
        siddon_trace_split(a, lo, w, [&](uint32_t ind, float len, bool t) {
            const float lw = len * w;
            unsafeAtomicAdd((t ? rhs_t : rhs) + ind, r * lw);
            if (sens) unsafeAtomicAdd((t ? sens_t : sens) + ind, lw);
        });
    } else if (sens) {
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

// ---- planar LDS kernels ---------------------------------------------------------------------------
// For sinograms whose planes are direct (zs == zd), every LOR of a plane lies in one image slice.
// A workgroup owns (plane, contiguous LOR chunk of that plane):
//   FP  stage the slice (Nx*Ny floats) in LDS and gather from it;
//   BP  accumulate the workgroup's share of rhs (and of the sensitivity image) for the slice in
//       LDS, then store it as a partial slice; k_planar_reduce / k_planar_update sum the partial
//       slices of a plane in workgroup order.
// The LDS accumulator is 64-bit fixed point: on gfx950 a float LDS atomic (ds_add_f32) retires at
// ~0.33 lanes per clock per CU even without bank conflicts, a 64-bit integer one (ds_add_u64) at
// ~5, so the float version is bound by the atomic rate. Each term t = r * (len * w) (the float
// product of the reference's backprojection) is added as round(t * 2^F) with F chosen per
// workgroup so that no sum can leave +-2^62: with B = (LORs of the workgroup) * max |r w| *
// (voxel diagonal) * (1 + 2^-10) >= any voxel's sum of |t|, F = 62 - e where B < 2^e. A term is
// represented to 2^-F absolute (exactly when |t| 2^F >= 2^23), integer sums are exact and do not
// depend on the order in which the atomics retire, and the one float rounding left is that of the
// final sum. Inputs must be finite (an Inf or NaN term has no fixed-point value).
// Block b handles plane b % n_planes: consecutive blocks go to different XCDs, and with 8 planes
// per launch all blocks of one plane share one XCD's L2.

template <int TPB, bool FP, bool BP, bool SENS, bool RING>
__global__ __launch_bounds__(TPB) void k_siddon_planar(SidArgs a, uint32_t n_planes, uint32_t chunk,
                                                       const float* __restrict__ img, const float* __restrict__ yin,
                                                       float* __restrict__ yout, float* __restrict__ rhs,
                                                       float* __restrict__ sens, float epps) {
    extern __shared__ float lds[];
    const uint32_t plane = blockIdx.x % n_planes, grp = blockIdx.x / n_planes;
    const uint32_t per = a.size_x * a.size_y;
    const uint32_t l0 = grp * chunk;
    if (l0 >= per) return;
    const uint32_t l1 = min(per, l0 + chunk);
    const uint32_t Nx = a.Nx, Nyx = a.Nyx, P = a.pitch;   // LDS rows are P = Nx + 1 entries apart
    const uint32_t S = a.Ny * P;                          // LDS entries per slice
    const int64_t base = (int64_t)plane * per;
    const float zs = a.z[(a.plane_offset + plane) * 2];
    const int32_t tk = (int32_t)(fabsf(zs - a.bz) / a.dz);
    if (tk < 0 || (uint32_t)tk >= a.Nz) {  // every LOR of this plane misses the image
        if (FP && !BP)
            for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) yout[base + l] = 0.f;
        return;
    }
    const size_t slice = (size_t)tk * Nyx;
    // LDS: [slice: S floats (FP) / S int64 (BP)][R: chunk floats][red: TPB / 64 floats]
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds);
    float* R = lds + (BP ? 2 : 1) * S;       // per-LOR ratio between the FP and BP phases
    float* red = R + chunk;
    float mr = 0.f, ms = 0.f;                // this lane's max |r w| and max w
    if (FP) {
        const float4* src = reinterpret_cast<const float4*>(img + slice);
        for (uint32_t i = threadIdx.x; i < Nyx / 4; i += TPB) {
            const float4 v = src[i];
            const uint32_t e = 4 * i, yy = e / Nx, xx = e - yy * Nx;   // Nx % 4 == 0: a float4 stays in one row
            float* d = lds + yy * P + xx;
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
        __syncthreads();
        for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
            const int64_t lo = base + l;
            float ax, w = 0.f;
            const bool hit = siddon_gather<true, RING>(a, lo, lds, ax, w);
            const float yv = hit ? ax * w : 0.f;
            if (!BP) {
                yout[lo] = yv;
            } else {
                float r = yin[lo] / ratio_den(a, lo, yv, epps);
                if (isnan(r) || isinf(r)) r = epps;
                R[l - l0] = hit ? r : 0.f;
                if (hit) {
                    mr = fmaxf(mr, fabsf(r * w));
                    ms = fmaxf(ms, w);
                }
            }
        }
    } else if (BP) {
        // bound pass: the LOR weights from the set-up alone
        for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
            const int64_t lo = base + l;
            Walk wk;
            float w = 0.f;
            if (siddon_setup<true>(a, lo, wk, w)) {
                mr = fmaxf(mr, fabsf(yin[lo] * w));
                ms = fmaxf(ms, w);
            }
        }
    }
    if (BP) {
        // |r (len w)| <= |r w| * voxel diagonal (float rounding inside the 2^-10 margin); a LOR adds at
        // most one term to a voxel, so a voxel's sum is bounded by (LORs of the workgroup) * that
        const float diag = sqrtf(a.dx * a.dx + a.dy * a.dy) * (1.f + 0x1p-10f);
        const float nl = (float)(l1 - l0);
        double up_r, down_r, up_s = 1.0, down_s = 1.0;
        const float br = block_max<TPB>(mr, red) * diag;   // (syncs: the FP slice is free after this)
        fixed_scale(br, br * nl, up_r, down_r);
        if (SENS) {
            __syncthreads();
            const float bs = block_max<TPB>(ms, red) * diag;
            fixed_scale(bs, bs * nl, up_s, down_s);
        }
        for (int pass = 0; pass < (SENS ? 2 : 1); pass++) {
            for (uint32_t i = threadIdx.x; i < S; i += TPB) acc[i] = 0ull;
            __syncthreads();
            for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
                const int64_t lo = base + l;
                const float r = FP ? R[l - l0] : yin[lo];
                if (pass == 0) {
                    if (r == 0.f) continue;
                    float w = 0.f;
                    siddon_scatter<true, RING>(a, lo, w, [&](uint32_t ind, float len) {
                        lds_add_fixed(acc + ind, r * (len * w), up_r);
                    });
                } else {
                    float w = 0.f;
                    siddon_scatter<true, RING>(a, lo, w, [&](uint32_t ind, float len) {
                        lds_add_fixed(acc + ind, len * w, up_s);
                    });
                }
            }
            __syncthreads();
            // this workgroup's partial slice -> scratch (plain coalesced stores)
            float* part = (pass == 0 ? rhs : sens) + (size_t)blockIdx.x * Nyx;
            const double down = pass == 0 ? down_r : down_s;
            for (uint32_t i = threadIdx.x; i < Nyx; i += TPB) {
                const uint32_t yy = i / Nx, j = i + yy;   // padded LDS index of voxel i
                part[i] = fixed_value(acc[j], down);
            }
            if (SENS && pass == 0) __syncthreads();
        }
    }
}

// Launch geometry of the planar kernels: one LOR per lane; a plane's LORs are split into
// workgroups of up to kPlanarTPB lanes (640 = 10 wave64; with the 132 KB int64 accumulator of a
// 128 x 128 slice one workgroup fills a CU).

// Sum of the per-workgroup partial slices of each plane, in workgroup order, added to the
// plane's image slice. Workgroup b = grp * n_planes + plane wrote partial slot b.
__global__ __launch_bounds__(256) void k_planar_reduce(SidArgs a, uint32_t n_planes, uint32_t groups,
                                                       const float* __restrict__ part, float* __restrict__ out) {
    const uint32_t Nyx = a.Nyx;
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    const uint32_t plane = blockIdx.y;
    if (v >= Nyx) return;
    const float zs = a.z[(a.plane_offset + plane) * 2];
    const int32_t tk = (int32_t)(fabsf(zs - a.bz) / a.dz);
    if (tk < 0 || (uint32_t)tk >= a.Nz) return;
    float s = 0.f;
    for (uint32_t g = 0; g < groups; g++) s += part[(size_t)(g * n_planes + plane) * Nyx + v];
    if (s != 0.f) unsafeAtomicAdd(out + (size_t)tk * Nyx + v, s);   // atomic: two planes may share a slice
}

PlanarShape planar_shape(const omg_geom* g, int64_t n) {
    const uint32_t per = g->d.size_x * g->d.size_y;
    PlanarShape s;
    s.n_planes = (uint32_t)(n / per);
    uint32_t groups = (per + kPlanarTPB - 1) / kPlanarTPB;
    if (g->planar_groups > 0) groups = (uint32_t)g->planar_groups;
    const uint32_t max_groups = (per + 63) / 64;
    if (groups > max_groups) groups = max_groups;
    if (groups < 1) groups = 1;
    s.chunk = (per + groups - 1) / groups;
    groups = (per + s.chunk - 1) / s.chunk;
    s.blocks = s.n_planes * groups;
    return s;
}

// Adds the partial slices in scratch_ptr(g) (layout of planar_shape(g, n): slot b = group * n_planes + plane;
// sensitivity partials after all rhs partials) of launch [first, first + n) into rhs (and sens).
int planar_reduce_partials(const omg_geom* g, int64_t first, int64_t n, float* rhs, float* sens, hipStream_t s) {
    const PlanarShape sh = planar_shape(g, n);
    const uint32_t Nyx = g->d.Nx * g->d.Ny;
    const uint32_t groups = sh.blocks / sh.n_planes;
    const SidArgs a = make_sid_args(g, first, n);
    const dim3 rg(grid_for(Nyx, 256), sh.n_planes);
    hipLaunchKernelGGL(k_planar_reduce, rg, dim3(256), 0, s, a, sh.n_planes, groups, scratch_ptr(g), rhs);
    OMG_CHECK_LAUNCH();
    if (sens) {
        hipLaunchKernelGGL(k_planar_reduce, rg, dim3(256), 0, s, a, sh.n_planes, groups,
                           scratch_ptr(g) + (size_t)sh.blocks * Nyx, sens);
        OMG_CHECK_LAUNCH();
    }
    return OMG_OK;
}

// The planar path applies to direct-plane PET sinograms with a square slice that fits twice in LDS.
bool planar_ok(const omg_geom* g) {
    return g->planar && g->d.Nx == g->d.Ny && g->d.Ny * (g->d.Nx + 1) <= 16512 && g->d.Nx % 4 == 0;
}

template <bool FP, bool BP, bool SENS>
static int launch_planar(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* yin, float* yout,
                         float* rhs, float* sens, float epps, hipStream_t s, const float* add = nullptr) {
    const PlanarShape sh = planar_shape(g, n);
    const uint32_t S = g->d.Ny * (g->d.Nx + 1);
    // [slice: S floats (FP only) or S int64 (BP)][R: chunk floats][block_max scratch]
    const size_t lds = BP ? (size_t)S * 8 + sizeof(float) * ((size_t)sh.chunk + kPlanarTPB / 64) : sizeof(float) * (size_t)S;
    const int ring = g->ring ? 1 : 0;
    auto kern = ring ? k_siddon_planar<kPlanarTPB, FP, BP, SENS, true> : k_siddon_planar<kPlanarTPB, FP, BP, SENS, false>;
    static size_t attr_bytes[2] = {0, 0};   // largest dynamic LDS size requested so far per instantiation
    if (lds > attr_bytes[ring]) {
        OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_bytes[ring] = lds;
    }
    SidArgs a = make_sid_args(g, first, n);
    a.pitch = g->d.Nx + 1;
    a.add = add;
    float* part = nullptr;
    float* part2 = nullptr;
    const uint32_t Nyx = g->d.Nx * g->d.Ny;
    if (BP) {
        const size_t need = (size_t)sh.blocks * Nyx * (SENS ? 2 : 1);
        int rc = scratch_reserve(const_cast<omg_geom*>(g), need);
        if (rc) return rc;
        part = scratch_ptr(g);
        part2 = SENS ? scratch_ptr(g) + (size_t)sh.blocks * Nyx : nullptr;
    }
    hipEvent_t stop;
    timing_begin(s, &stop);
    hipLaunchKernelGGL(kern, dim3(sh.blocks), dim3(kPlanarTPB), lds, s, a, sh.n_planes, sh.chunk, img, yin, yout,
                       BP ? part : rhs, BP ? part2 : sens, epps);
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    if (BP && !rhs) return OMG_OK;   // the caller reduces the partial slices itself
    if (BP) return planar_reduce_partials(g, first, n, rhs, SENS ? sens : nullptr, s);
    return OMG_OK;
}

// ---- OSEM / OSL-OSEM update that also reduces the planar partial slices -----------------------------
// For voxel n in slice z: if z is the slice of launch plane p (slice_plane[z] = p, else -1), rhs and,
// when sensitivity is computed, S are the workgroup-ordered sums of the plane's partial slices;
// otherwise rhs = 0 (and S = 0). S is clamped (S < epps -> epps, functions.hpp:2996) and stored.
// OSEM:     x = x / S * (rhs + epps)            (algorithms.h:12-15, subiterStep.h:111-121)
// OSL-OSEM: x = x / (S + dU) * rhs              (subiterStep.h:219-231), dU = beta * grad RDP(x)
// positivity: x < epps -> epps                  (reconstructionAF.h:1461-1465)
template <bool SENS, bool OSL>
__global__ __launch_bounds__(256) void k_planar_update(uint32_t Nyx, uint32_t N, uint32_t n_planes, uint32_t groups,
                                                       const int32_t* __restrict__ slice_plane,
                                                       const float* __restrict__ part, const float* __restrict__ part2,
                                                       float* __restrict__ x, float* __restrict__ S,
                                                       const float* __restrict__ dU, float epps, int pos) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const uint32_t z = n / Nyx, v = n - z * Nyx;
    const int32_t p = slice_plane[z];
    float rhs = 0.f, s;
    if (p >= 0) {
        for (uint32_t g = 0; g < groups; g++) rhs += part[(size_t)(g * n_planes + (uint32_t)p) * Nyx + v];
    }
    if (SENS) {
        s = 0.f;
        if (p >= 0)
            for (uint32_t g = 0; g < groups; g++) s += part2[(size_t)(g * n_planes + (uint32_t)p) * Nyx + v];
        if (s < epps) s = epps;
        S[n] = s;
    } else {
        s = S[n];
    }
    float xn;
    if (OSL) xn = x[n] / (s + dU[n]) * rhs;
    else xn = x[n] / s * (rhs + epps);
    if (pos && xn < epps) xn = epps;
    x[n] = xn;
}

// Sum over workgroups g = 0, 1, ... of plane p's partial slices at float4 index v, in that order;
// eight loads in flight per lane.
__device__ __forceinline__ float4 sum_partials(const float4* __restrict__ part, uint32_t groups, uint32_t n_planes,
                                               uint32_t p, uint32_t Nyx4, uint32_t v) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t g = 0;
    for (; g + 8 <= groups; g += 8) {
        float4 t[8];
#pragma unroll
        for (int k = 0; k < 8; k++) t[k] = part[(size_t)((g + k) * n_planes + p) * Nyx4 + v];
#pragma unroll
        for (int k = 0; k < 8; k++) { acc.x += t[k].x; acc.y += t[k].y; acc.z += t[k].z; acc.w += t[k].w; }
    }
    for (; g < groups; g++) {
        const float4 t = part[(size_t)(g * n_planes + p) * Nyx4 + v];
        acc.x += t.x; acc.y += t.y; acc.z += t.z; acc.w += t.w;
    }
    return acc;
}

// The same update four voxels per lane (float4 loads and stores; the caller checks 16-byte alignment
// and Nx % 4 == 0). grid = (ceil(Nyx / 4 / 256), Nz): the slice's launch plane is found per workgroup
// from the z table, so no slice -> plane map is built.
template <bool SENS, bool OSL>
__global__ __launch_bounds__(256) void k_planar_update4(SidArgs a, uint32_t n_planes, uint32_t groups,
                                                        const float4* __restrict__ part, const float4* __restrict__ part2,
                                                        float4* __restrict__ x, float4* __restrict__ S,
                                                        const float4* __restrict__ dU, float epps, int pos) {
    const uint32_t Nyx4 = a.Nyx / 4;
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    const uint32_t z = blockIdx.y;
    int32_t p = -1;
    for (uint32_t q = 0; q < n_planes; q++) {
        const float zs = a.z[(a.plane_offset + q) * 2];
        if ((int32_t)(fabsf(zs - a.bz) / a.dz) == (int32_t)z) { p = (int32_t)q; break; }
    }
    if (v >= Nyx4) return;
    const size_t n = (size_t)z * Nyx4 + v;
    float4 rhs = make_float4(0.f, 0.f, 0.f, 0.f), s;
    if (p >= 0) rhs = sum_partials(part, groups, n_planes, (uint32_t)p, Nyx4, v);
    if (SENS) {
        s = make_float4(0.f, 0.f, 0.f, 0.f);
        if (p >= 0) s = sum_partials(part2, groups, n_planes, (uint32_t)p, Nyx4, v);
        s.x = s.x < epps ? epps : s.x; s.y = s.y < epps ? epps : s.y;
        s.z = s.z < epps ? epps : s.z; s.w = s.w < epps ? epps : s.w;
        S[n] = s;
    } else {
        s = S[n];
    }
    const float4 xo = x[n];
    float4 xn;
    if (OSL) {
        const float4 d = dU[n];
        xn.x = xo.x / (s.x + d.x) * rhs.x; xn.y = xo.y / (s.y + d.y) * rhs.y;
        xn.z = xo.z / (s.z + d.z) * rhs.z; xn.w = xo.w / (s.w + d.w) * rhs.w;
    } else {
        xn.x = xo.x / s.x * (rhs.x + epps); xn.y = xo.y / s.y * (rhs.y + epps);
        xn.z = xo.z / s.z * (rhs.z + epps); xn.w = xo.w / s.w * (rhs.w + epps);
    }
    if (pos) {
        xn.x = xn.x < epps ? epps : xn.x; xn.y = xn.y < epps ? epps : xn.y;
        xn.z = xn.z < epps ? epps : xn.z; xn.w = xn.w < epps ? epps : xn.w;
    }
    x[n] = xn;
}

}  // namespace omg

using namespace omg;

extern "C" int omg_timing_enable(int32_t on) {
    g_timing.on = on != 0;
    return OMG_OK;
}

extern "C" int omg_timing_collect(double* total_ms, int64_t* launches) {
    OMG_REQUIRE(total_ms && launches, "null output pointer");
    if (g_timing.err != hipSuccess) {
        const hipError_t e = g_timing.err;
        g_timing.err = hipSuccess;
        g_timing.used = 0;
        set_error("recording timing events failed: %s", hipGetErrorString(e));
        return OMG_ERR_HIP;
    }
    double t = 0.0;
    for (size_t i = 0; i < g_timing.used; i++) {
        OMG_CHECK_HIP(hipEventSynchronize(g_timing.ev[i].second));
        float ms = 0.f;
        OMG_CHECK_HIP(hipEventElapsedTime(&ms, g_timing.ev[i].first, g_timing.ev[i].second));
        t += ms;
    }
    *total_ms = t;
    *launches = (int64_t)g_timing.used;
    g_timing.used = 0;
    return OMG_OK;
}

// One-lane-per-LOR launches on PET sinograms walk x-major LORs in y-fastest copies of the image and of
// the accumulators (siddon_setup's allow_trans); omg_geom_desc.reserved[3] != 0 turns this off.
static bool split_layout(const omg_geom* g) {
    return g->d.mode == OMG_MODE_PET_SINO && g->d.reserved[3] == 0;
}

static dim3 xy_tiles(const omg_geom* g) { return dim3((g->d.Nx + 31) / 32, (g->d.Ny + 31) / 32, g->d.Nz); }

// The forward projection gathers from the x-fastest image only: gathering x-major LORs from a y-fastest
// copy measured slower (9.6 ms against 7.4 ms for a C2 forward projection).
static int generic_fp(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s) {
    if (g->d.tof_bins > 1) return tof_forward(g, first, n, img, y, s);
    const SidArgs a = make_sid_args(g, first, n);
    hipEvent_t stop;
    timing_begin(s, &stop);
    hipLaunchKernelGGL(k_siddon_fp, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, a, img, y);
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

// FP -> ratio -> BP into rhs (and sens), both already holding what the caller wants added to.
static int generic_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
                        float* sens, float epps, hipStream_t s, const float* add) {
    if (g->d.tof_bins > 1) return tof_osem(g, first, n, img, meas, rhs, sens, epps, s, add);
    SidArgs a = make_sid_args(g, first, n);
    a.add = add;
    float *rhs_t = nullptr, *sens_t = nullptr;
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    if (split_layout(g)) {   // y-fastest accumulators only (see generic_fp)
        const int rc = scratch_reserve(const_cast<omg_geom*>(g), (size_t)N * (sens ? 2 : 1));
        if (rc) return rc;
        float* t = scratch_ptr(g);
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for((int64_t)N * (sens ? 2 : 1), 256)), dim3(256), 0, s, t,
                           N * (sens ? 2u : 1u), 0.f);
        OMG_CHECK_LAUNCH();
        rhs_t = t;
        sens_t = sens ? t + N : nullptr;
    }
    hipEvent_t stop;
    timing_begin(s, &stop);
    hipLaunchKernelGGL(k_siddon_osem, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, a, img, meas, rhs, sens,
                       rhs_t, sens_t, epps);
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    if (rhs_t) {
        hipLaunchKernelGGL(k_add_transposed, xy_tiles(g), dim3(32, 8), 0, s, rhs_t, rhs, g->d.Nx, g->d.Ny);
        OMG_CHECK_LAUNCH();
        if (sens) {
            hipLaunchKernelGGL(k_add_transposed, xy_tiles(g), dim3(32, 8), 0, s, sens_t, sens, g->d.Nx, g->d.Ny);
            OMG_CHECK_LAUNCH();
        }
    }
    return OMG_OK;
}

static int siddon_common_checks(const omg_geom* g, int64_t first, int64_t n) {
    OMG_REQUIRE(g != nullptr, "null geometry handle");
    OMG_REQUIRE(g->d.projector_type == OMG_PROJ_SIDDON || is_tube(g) || g->d.projector_type == OMG_PROJ_BDD ||
                    g->d.projector_type == OMG_PROJ_INTERP,
                "unknown projector type %d", g->d.projector_type);
    return check_range(g, first, n);
}

// Forward declarations of the orthogonal / BDD dispatchers (other translation units).
int orth_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s);
int orth_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens, hipStream_t s);
int orth_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
              float* sens, float epps, hipStream_t s, const float* add);
int orth_trace(const omg_geom* g, int64_t first, int64_t n, int32_t* counts, const int64_t* offsets, uint32_t* idx,
               float* w, hipStream_t s);
int orth_planar_fused(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas,
                      int32_t compute_sens, float epps, hipStream_t s, const float* add);
int vbp_subset_iteration(omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas, float* S,
                         int32_t compute_sens, float epps, const float* dU, int32_t positivity, const int32_t* map,
                         hipStream_t s, const float* add);
int bdd_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s);
int bdd_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens, hipStream_t s);

extern "C" int omg_forward_project(const omg_geom* g, int64_t first, int64_t n_meas, const float* img, float* y,
                                   void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(img && y, "null image or output pointer");
    if (n_meas == 0) return OMG_OK;
    hipStream_t s = as_stream(stream);
    if (is_tube(g)) return orth_forward(g, first, n_meas, img, y, s);
    if (g->d.projector_type == OMG_PROJ_BDD) return bdd_forward(g, first, n_meas, img, y, s);
    if (g->d.projector_type == OMG_PROJ_INTERP) return interp_forward(g, first, n_meas, img, y, s);
    if (planar_ok(g) && !g->force_generic) return launch_planar<true, false, false>(g, first, n_meas, img, nullptr, y, nullptr, nullptr, 0.f, s);
    return generic_fp(g, first, n_meas, img, y, s);
}

extern "C" int omg_backward_project(const omg_geom* g, int64_t first, int64_t n_meas, const float* y, float* rhs,
                                    float* sens, void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(y && rhs, "null input or output pointer");
    if (n_meas == 0) return OMG_OK;
    hipStream_t s = as_stream(stream);
    if (is_tube(g)) return orth_backward(g, first, n_meas, y, rhs, sens, s);
    if (g->d.projector_type == OMG_PROJ_BDD) return bdd_backward(g, first, n_meas, y, rhs, sens, s);
    if (g->d.projector_type == OMG_PROJ_INTERP) return interp_backward(g, first, n_meas, y, rhs, sens, s);
    if (planar_ok(g) && !g->force_generic) {
        if (sens) return launch_planar<false, true, true>(g, first, n_meas, nullptr, y, nullptr, rhs, sens, 0.f, s);
        return launch_planar<false, true, false>(g, first, n_meas, nullptr, y, nullptr, rhs, nullptr, 0.f, s);
    }
    if (g->d.tof_bins > 1) return tof_backward(g, first, n_meas, y, rhs, sens, s);
    SidArgs a = make_sid_args(g, first, n_meas);
    if (split_layout(g)) {
        // x-major LORs into y-fastest buffers (scratch), added back transposed
        const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
        rc = scratch_reserve(const_cast<omg_geom*>(g), (size_t)N * (sens ? 2 : 1));
        if (rc) return rc;
        float* rt = scratch_ptr(g);
        float* st = sens ? scratch_ptr(g) + N : nullptr;
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for((int64_t)N * (sens ? 2 : 1), 256)), dim3(256), 0, s, rt,
                           N * (sens ? 2u : 1u), 0.f);
        OMG_CHECK_LAUNCH();
        hipEvent_t stop;
        timing_begin(s, &stop);
        if (sens)
            hipLaunchKernelGGL(k_siddon_bp_split<true>, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, s, a, y, rhs,
                               sens, rt, st);
        else
            hipLaunchKernelGGL(k_siddon_bp_split<false>, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, s, a, y, rhs,
                               nullptr, rt, nullptr);
        timing_end(s, stop);
        OMG_CHECK_LAUNCH();
        const dim3 tg = xy_tiles(g);
        hipLaunchKernelGGL(k_add_transposed, tg, dim3(32, 8), 0, s, rt, rhs, g->d.Nx, g->d.Ny);
        OMG_CHECK_LAUNCH();
        if (sens) {
            hipLaunchKernelGGL(k_add_transposed, tg, dim3(32, 8), 0, s, st, sens, g->d.Nx, g->d.Ny);
            OMG_CHECK_LAUNCH();
        }
        return OMG_OK;
    }
    hipEvent_t stop;
    timing_begin(s, &stop);
    hipLaunchKernelGGL(k_siddon_bp, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, s, a, y, rhs, sens);
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_osem_data_step_ex(const omg_geom* g, int64_t first, int64_t n_meas, const float* img,
                                     const float* meas, const float* additive, float* rhs, float* sens, float epps,
                                     void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(img && meas && rhs, "null pointer argument");
    OMG_REQUIRE(g->d.mode != OMG_MODE_CT_FLAT, "the fused OSEM data step is for emission (PET) data");
    OMG_REQUIRE(g->d.projector_type != OMG_PROJ_BDD, "the fused OSEM data step is implemented for projector_type 1-4");
    if (n_meas == 0) return OMG_OK;
    hipStream_t s = as_stream(stream);
    if (is_tube(g)) return orth_osem(g, first, n_meas, img, meas, rhs, sens, epps, s, additive);
    if (g->d.projector_type == OMG_PROJ_INTERP) return interp_osem(g, first, n_meas, img, meas, rhs, sens, epps, s, additive);
    if (planar_ok(g) && !g->force_generic) {
        if (sens) return launch_planar<true, true, true>(g, first, n_meas, img, meas, nullptr, rhs, sens, epps, s, additive);
        return launch_planar<true, true, false>(g, first, n_meas, img, meas, nullptr, rhs, nullptr, epps, s, additive);
    }
    return generic_osem(g, first, n_meas, img, meas, rhs, sens, epps, s, additive);
}

extern "C" int omg_osem_data_step(const omg_geom* g, int64_t first, int64_t n_meas, const float* img,
                                  const float* meas, float* rhs, float* sens, float epps, void* stream) {
    return omg_osem_data_step_ex(g, first, n_meas, img, meas, nullptr, rhs, sens, epps, stream);
}

namespace omg {

// slice -> launch plane map for the planar update (-1: slice not in this launch)
__global__ void k_slice_plane(SidArgs a, uint32_t n_planes, int32_t* __restrict__ map) {
    const uint32_t zsl = blockIdx.x * 256 + threadIdx.x;
    if (zsl >= a.Nz) return;
    int32_t found = -1;
    for (uint32_t p = 0; p < n_planes; p++) {
        const float zs = a.z[(a.plane_offset + p) * 2];
        const int32_t tk = (int32_t)(fabsf(zs - a.bz) / a.dz);
        if (tk == (int32_t)zsl) { found = (int32_t)p; break; }
    }
    map[zsl] = found;
}

__global__ __launch_bounds__(256) void k_generic_update(uint32_t N, float* __restrict__ x, float* __restrict__ S,
                                                        float* __restrict__ rhs, const float* __restrict__ dU,
                                                        float epps, int clamp_s, int pos) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float s = S[n];
    if (clamp_s) {
        if (s < epps) s = epps;
        S[n] = s;
    }
    float xn = dU ? x[n] / (s + dU[n]) * rhs[n] : x[n] / s * (rhs[n] + epps);
    if (pos && xn < epps) xn = epps;
    x[n] = xn;
    rhs[n] = 0.f;
}

// True when the planes [first, first + n_planes) map to distinct image slices.
// EM / OSL update of image x from the partial slices of a planar fused launch [first, first + n):
// k_planar_update4 when x, S, dU and the partials are 16-byte aligned, else the slice map and
// k_planar_update. work_ptr(g) must hold the fused-iteration buffers (work_reserve).
int planar_update(const omg_geom* g, int64_t first, int64_t n_meas, float* x, float* S, const float* dU,
                  int32_t compute_sens, float epps, int32_t positivity, hipStream_t s) {
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    const PlanarShape sh = planar_shape(g, n_meas);
    const uint32_t Nyx = g->d.Nx * g->d.Ny;
    const float* part = scratch_ptr(g);
    const float* part2 = scratch_ptr(g) + (size_t)sh.blocks * Nyx;
    const uint32_t groups = sh.blocks / sh.n_planes;
    SidArgs a = make_sid_args(g, first, n_meas);
    auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
    if (al16(x) && al16(S) && (!dU || al16(dU)) && al16(part2)) {
        // Nx % 4 == 0 on the planar path, so Nyx % 4 == 0 and the partial slices stay 16-byte aligned
        const dim3 g4(grid_for(Nyx / 4, 256), g->d.Nz);
        const float4* p4 = reinterpret_cast<const float4*>(part);
        const float4* q4 = reinterpret_cast<const float4*>(part2);
        float4* x4 = reinterpret_cast<float4*>(x);
        float4* S4 = reinterpret_cast<float4*>(S);
        const float4* d4 = reinterpret_cast<const float4*>(dU);
        if (compute_sens && dU)
            hipLaunchKernelGGL((k_planar_update4<true, true>), g4, dim3(256), 0, s, a, sh.n_planes, groups, p4, q4, x4,
                               S4, d4, epps, (int)positivity);
        else if (compute_sens)
            hipLaunchKernelGGL((k_planar_update4<true, false>), g4, dim3(256), 0, s, a, sh.n_planes, groups, p4, q4,
                               x4, S4, d4, epps, (int)positivity);
        else if (dU)
            hipLaunchKernelGGL((k_planar_update4<false, true>), g4, dim3(256), 0, s, a, sh.n_planes, groups, p4, q4,
                               x4, S4, d4, epps, (int)positivity);
        else
            hipLaunchKernelGGL((k_planar_update4<false, false>), g4, dim3(256), 0, s, a, sh.n_planes, groups, p4, q4,
                               x4, S4, d4, epps, (int)positivity);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    int32_t* map = reinterpret_cast<int32_t*>(work_ptr(g) + 2 * (size_t)N);
    hipLaunchKernelGGL(k_slice_plane, dim3(grid_for(g->d.Nz, 256)), dim3(256), 0, s, a, sh.n_planes, map);
    OMG_CHECK_LAUNCH();
    dim3 grid(grid_for(N, 256));
    if (compute_sens && dU)
        hipLaunchKernelGGL((k_planar_update<true, true>), grid, dim3(256), 0, s, Nyx, N, sh.n_planes, groups, map, part,
                           part2, x, S, dU, epps, (int)positivity);
    else if (compute_sens)
        hipLaunchKernelGGL((k_planar_update<true, false>), grid, dim3(256), 0, s, Nyx, N, sh.n_planes, groups, map, part,
                           part2, x, S, dU, epps, (int)positivity);
    else if (dU)
        hipLaunchKernelGGL((k_planar_update<false, true>), grid, dim3(256), 0, s, Nyx, N, sh.n_planes, groups, map, part,
                           part2, x, S, dU, epps, (int)positivity);
    else
        hipLaunchKernelGGL((k_planar_update<false, false>), grid, dim3(256), 0, s, Nyx, N, sh.n_planes, groups, map, part,
                           part2, x, S, dU, epps, (int)positivity);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

bool unique_slices(const omg_geom* g, int64_t first, uint32_t n_planes) {
    std::vector<char> seen(g->d.Nz, 0);
    for (uint32_t p = 0; p < n_planes; p++) {
        const float zs = g->z_host[(size_t)(first + p) * 2];
        const int32_t tk = (int32_t)(fabsf(zs - g->d.bz) / g->d.dz);
        if (tk < 0 || (uint32_t)tk >= g->d.Nz) continue;
        if (seen[tk]) return false;
        seen[tk] = 1;
    }
    return true;
}

}  // namespace omg

extern "C" int omg_osem_subset_iteration(const omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas,
                                         float* S, int32_t compute_sens, float epps, float beta, float gamma,
                                         int32_t positivity, void* stream) {
    return omg_osem_subset_iteration_ex(g, first, n_meas, x, meas, nullptr, S, compute_sens, epps, beta, gamma,
                                        positivity, stream);
}

namespace omg {

// The subset iteration behind omg_osem_subset_iteration_ex and omg_osem_subset_iteration_noalloc (which
// holds a fixed work region while it calls this).
int omg_osem_subset_iteration_impl(const omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas,
                                   const float* additive, float* S, int32_t compute_sens, float epps, float beta,
                                   float gamma, int32_t positivity, void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(x && meas && S, "null pointer argument");
    OMG_REQUIRE(g->d.mode != OMG_MODE_CT_FLAT, "OSEM subset iterations here are for emission (PET) data");
    OMG_REQUIRE(g->d.projector_type != OMG_PROJ_BDD, "the fused OSEM iteration is implemented for projector_type 1-4");
    hipStream_t s = as_stream(stream);
    omg_geom* gm = const_cast<omg_geom*>(g);
    const uint32_t N = g->d.Nx * g->d.Ny * g->d.Nz;
    rc = work_reserve(gm, N);
    if (rc) return rc;
    float* dU = nullptr;
    if (beta != 0.f) {
        dU = work_ptr(g);
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(N, 256)), dim3(256), 0, s, dU, N, 0.f);
        OMG_CHECK_LAUNCH();
        rc = omg_rdp_gradient(x, dU, g->d.Nx, g->d.Ny, g->d.Nz, gamma, epps, beta, 0, stream);
        if (rc) return rc;
    }
    const uint32_t per = g->d.size_x * g->d.size_y;
    if (n_meas > 0 && g->d.projector_type == OMG_PROJ_SIDDON && planar_ok(g) && !g->force_generic &&
        unique_slices(g, first, (uint32_t)(n_meas / per))) {
        if (g->vbp_ok) {
            // scatter-free path: ray-driven FP -> q = r w, voxel-driven BP with the update fused (vbp.hip)
            int32_t* map = reinterpret_cast<int32_t*>(work_ptr(g) + 2 * (size_t)N);
            SidArgs a = make_sid_args(g, first, n_meas);
            hipLaunchKernelGGL(k_slice_plane, dim3(grid_for(g->d.Nz, 256)), dim3(256), 0, s, a,
                               (uint32_t)(n_meas / per), map);
            OMG_CHECK_LAUNCH();
            hipEvent_t stop;
            timing_begin(s, &stop);
            rc = vbp_subset_iteration(gm, first, n_meas, x, meas, S, compute_sens, epps, dU, positivity, map, s,
                                      additive);
            timing_end(s, stop);
            return rc;
        }
        if (compute_sens)
            rc = launch_planar<true, true, true>(g, first, n_meas, x, meas, nullptr, nullptr, nullptr, epps, s, additive);
        else
            rc = launch_planar<true, true, false>(g, first, n_meas, x, meas, nullptr, nullptr, nullptr, epps, s, additive);
        if (rc) return rc;
        return planar_update(g, first, n_meas, x, S, dU, compute_sens, epps, positivity, s);
    }
    if (n_meas > 0 && is_tube(g) && g->orth_planar && !g->force_generic &&
        unique_slices(g, first, (uint32_t)(n_meas / per))) {
        rc = orth_planar_fused(g, first, n_meas, x, meas, compute_sens, epps, s, additive);
        if (rc) return rc;
        return planar_update(g, first, n_meas, x, S, dU, compute_sens, epps, positivity, s);
    }
    // one-lane-per-LOR path: rhs (and S) accumulate with global atomics, then the update
    float* rhs = work_ptr(g) + N;
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(N, 256)), dim3(256), 0, s, rhs, N, 0.f);
    OMG_CHECK_LAUNCH();
    if (compute_sens) {
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(N, 256)), dim3(256), 0, s, S, N, 0.f);
        OMG_CHECK_LAUNCH();
    }
    if (n_meas > 0 && is_tube(g)) {
        rc = orth_osem(g, first, n_meas, x, meas, rhs, compute_sens ? S : nullptr, epps, s, additive);
        if (rc) return rc;
    } else if (n_meas > 0 && g->d.projector_type == OMG_PROJ_INTERP) {
        rc = interp_osem(g, first, n_meas, x, meas, rhs, compute_sens ? S : nullptr, epps, s, additive);
        if (rc) return rc;
    } else if (n_meas > 0) {
        rc = generic_osem(g, first, n_meas, x, meas, rhs, compute_sens ? S : nullptr, epps, s, additive);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_generic_update, dim3(grid_for(N, 256)), dim3(256), 0, s, N, x, S, rhs, dU, epps,
                       (int)(compute_sens != 0), (int)positivity);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

}  // namespace omg

extern "C" int omg_osem_subset_iteration_ex(const omg_geom* g, int64_t first, int64_t n_meas, float* x,
                                            const float* meas, const float* additive, float* S, int32_t compute_sens,
                                            float epps, float beta, float gamma, int32_t positivity, void* stream) {
    return omg_osem_subset_iteration_impl(g, first, n_meas, x, meas, additive, S, compute_sens, epps, beta, gamma,
                                          positivity, stream);
}

extern "C" int omg_siddon_trace(const omg_geom* g, int64_t first, int64_t n_meas, int32_t* counts,
                                const int64_t* offsets, uint32_t* idx, float* len, void* stream) {
    int rc = siddon_common_checks(g, first, n_meas);
    if (rc) return rc;
    OMG_REQUIRE(g->d.projector_type != OMG_PROJ_BDD, "trace lists are produced by projector_type 1 and 2");
    OMG_REQUIRE(std::max(g->d.rays_xy, 1u) * std::max(g->d.rays_z, 1u) == 1,
                "trace lists describe single-ray LORs (this geometry has several rays per LOR)");
    OMG_REQUIRE(counts != nullptr, "null counts");
    OMG_REQUIRE(offsets == nullptr || (idx && len), "offsets given without idx/len outputs");
    if (n_meas == 0) return OMG_OK;
    if (is_tube(g)) return orth_trace(g, first, n_meas, counts, offsets, idx, len, as_stream(stream));
    SidArgs a = make_sid_args(g, first, n_meas);
    hipLaunchKernelGGL(k_siddon_trace, dim3(grid_for(n_meas, kBlock)), dim3(kBlock), 0, as_stream(stream), a, counts,
                       offsets, idx, len);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}
