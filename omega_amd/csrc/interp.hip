// interp.hip -- the interpolation projector (projector type 4) for gfx950.
//
// Semantics of the reference's GPU kernels (source/opencl/projectorType4.cl), with the 3-D texture fetch
// (normalised coordinates, linear filtering, clamp to edge: samplerForw / samplerIm,
// general_opencl_functions.h:456-458) evaluated as exact FP32 trilinear interpolation -- nested lerps in x,
// then y, then z, u = p N - 0.5, corners clamped to the image -- instead of the texture unit's reduced-
// precision weights:
//   FP  (projectorType4Forward -DFP, :84-674): one lane per measurement; the ray's line clipped to the box
//       [b, b + N d]; samples every tStep = dL / L from the entry (dL <= 0: the Joseph variant, one sample per
//       voxel plane of the dominant axis, :434-461); the normalised position recomputed every 64 samples and
//       stepped by dp = v scale tStep between them (:531-541); ax += stepLen tex(p); PET y = ax temp with
//       temp = 1 / L * norm * scat * given-attenuation * global (* the image attenuation factor last,
//       :612-626), CT y = ax.
//   BP for emission data (projectorType4Forward -DBP): the same samples, each adding stepLen temp y to the
//       voxel containing it (truncated normalised coordinates; voxel 0 and indices past the image skipped,
//       :570-576), the image attenuation factor first in temp and, with it, the samples starting a tenth of a
//       step in (:485-510). Float atomics.
//   BP for CT (projectorType4Backward, :718-1172, no offset / FDK / helical): one lane per (x, y) column and
//       z group of 8 voxels (NVOXELS; the first voxel of a group from its centre, the next by v.z += dz and
//       lowerPart -= dz cP.z); every projection's panel sampled bilinearly where the ray from the source
//       through the voxel centre meets it; weight = L^3 / |v|^2 kerroin; out += yVar weight, sens += weight,
//       both over the samples with yVar != 0. No atomics.
// Attenuation images are turned into one factor per measurement at geometry creation, by the forward
// projection's own walk (exp of -sum stepLen mu(p), compute_attenuation, general_opencl_functions.h:1633).
// The library is built with -ffp-contract=off and IEEE division / square root, so the forward projection is
// bit-identical to the oracle's restatement; the oracle's exp is the host's, so attenuation factors agree to
// float rounding.
#include "geom.hpp"

namespace omg {

constexpr int kInterpBlock = 256;
constexpr uint32_t kInterpNV = 8;   // voxels per lane along z in the CT backprojection (NVOXELS)

struct IF3 { float x, y, z; };

struct InterpArgs {
    SidArgs a;
    IF3 b, bmax, scale;   // box corners, 1 / (N d)
    float dL;             // interpolation length (mm); <= 0: Joseph
    float kerroin;        // CT backprojection scale (dx dy dz) / (dPitchX dPitchY sourceToDetector)
    float dPitchX, dPitchY;
    int32_t atn_image;    // a.atten holds factors made from an attenuation image
};

__device__ __forceinline__ void tex_axis(float p, uint32_t N, uint32_t& i0, uint32_t& i1, float& a) {
    const float u = p * (float)N - 0.5f;
    const float f = floorf(u);
    a = u - f;
    i0 = (uint32_t)fminf(fmaxf(f, 0.f), (float)(N - 1));
    i1 = (uint32_t)fminf(fmaxf(f + 1.f, 0.f), (float)(N - 1));
}

// Trilinear sample of an Nx x Ny x Nz array (x fastest) at normalised coordinates, clamp to edge.
__device__ __forceinline__ float tex3(const float* __restrict__ img, uint32_t Nx, uint32_t Ny, uint32_t Nz, float px,
                                      float py, float pz) {
    uint32_t x0, x1, y0, y1, z0, z1;
    float a, b, c;
    tex_axis(px, Nx, x0, x1, a);
    tex_axis(py, Ny, y0, y1, b);
    tex_axis(pz, Nz, z0, z1, c);
    const size_t Nxy = (size_t)Nx * Ny;
    const float* k0 = img + (size_t)z0 * Nxy;
    const float* k1 = img + (size_t)z1 * Nxy;
    const size_t r0 = (size_t)y0 * Nx, r1 = (size_t)y1 * Nx;
    const float v000 = k0[r0 + x0], v100 = k0[r0 + x1], v010 = k0[r1 + x0], v110 = k0[r1 + x1];
    const float v001 = k1[r0 + x0], v101 = k1[r0 + x1], v011 = k1[r1 + x0], v111 = k1[r1 + x1];
    const float c00 = (1.f - a) * v000 + a * v100;
    const float c10 = (1.f - a) * v010 + a * v110;
    const float c01 = (1.f - a) * v001 + a * v101;
    const float c11 = (1.f - a) * v011 + a * v111;
    const float c0 = (1.f - b) * c00 + b * c10;
    const float c1 = (1.f - b) * c01 + b * c11;
    return (1.f - c) * c0 + c * c1;
}

// tex3 with its two z slices given (the planar forward projection's LDS copies): the same lerps.
__device__ __forceinline__ float tex_slices(const float* k0, const float* k1, uint32_t Nx, uint32_t Ny, float px,
                                            float py, float c) {
    uint32_t x0, x1, y0, y1;
    float a, b;
    tex_axis(px, Nx, x0, x1, a);
    tex_axis(py, Ny, y0, y1, b);
    const uint32_t r0 = y0 * Nx, r1 = y1 * Nx;
    const float v000 = k0[r0 + x0], v100 = k0[r0 + x1], v010 = k0[r1 + x0], v110 = k0[r1 + x1];
    const float v001 = k1[r0 + x0], v101 = k1[r0 + x1], v011 = k1[r1 + x0], v111 = k1[r1 + x1];
    const float c00 = (1.f - a) * v000 + a * v100;
    const float c10 = (1.f - a) * v010 + a * v110;
    const float c01 = (1.f - a) * v001 + a * v101;
    const float c11 = (1.f - a) * v011 + a * v111;
    const float c0 = (1.f - b) * c00 + b * c10;
    const float c1 = (1.f - b) * c01 + b * c11;
    return (1.f - c) * c0 + c * c1;
}

struct Samples { IF3 s, v; float t0, tStep, stepLen, L, tStart, tEnd; uint32_t n; };

__device__ __forceinline__ bool interp_samples(const InterpArgs& ia, int64_t lo, Samples& S, int ray = 0) {
    Ray r;
    fetch_ray(ia.a, lo, r, ray);
    S.s = IF3{r.xs, r.ys, r.zs};
    S.v = IF3{r.xd - r.xs, r.yd - r.ys, r.zd - r.zs};
    const IF3 tBack{(ia.b.x - S.s.x) / S.v.x, (ia.b.y - S.s.y) / S.v.y, (ia.b.z - S.s.z) / S.v.z};
    const IF3 tFront{(ia.bmax.x - S.s.x) / S.v.x, (ia.bmax.y - S.s.y) / S.v.y, (ia.bmax.z - S.s.z) / S.v.z};
    const float tStart = fmaxf(fmaxf(fminf(tFront.x, tBack.x), fminf(tFront.y, tBack.y)), fminf(tFront.z, tBack.z));
    const float tEnd = fminf(fminf(fmaxf(tFront.x, tBack.x), fmaxf(tFront.y, tBack.y)), fmaxf(tFront.z, tBack.z));
    if (!(tStart < tEnd)) return false;
    S.tStart = tStart;
    S.tEnd = tEnd;
    S.L = sqrtf(S.v.x * S.v.x + S.v.y * S.v.y + S.v.z * S.v.z);
    if (ia.dL <= 0.f) {
        const float crx = fabsf(S.v.x) * ia.scale.x * (float)ia.a.Nx;
        const float cry = fabsf(S.v.y) * ia.scale.y * (float)ia.a.Ny;
        const float crz = fabsf(S.v.z) * ia.scale.z * (float)ia.a.Nz;
        S.tStep = 1.f / fmaxf(fmaxf(crx, cry), crz);
        float sInd, vDim;
        if (crx >= cry && crx >= crz) {
            sInd = (S.s.x + S.v.x * tStart - ia.b.x) * ia.scale.x * (float)ia.a.Nx - 0.5f;
            vDim = S.v.x;
        } else if (cry >= crz) {
            sInd = (S.s.y + S.v.y * tStart - ia.b.y) * ia.scale.y * (float)ia.a.Ny - 0.5f;
            vDim = S.v.y;
        } else {
            sInd = (S.s.z + S.v.z * tStart - ia.b.z) * ia.scale.z * (float)ia.a.Nz - 0.5f;
            vDim = S.v.z;
        }
        S.t0 = vDim > 0.f ? tStart + (floorf(sInd) - sInd) * S.tStep : tStart + (sInd - floorf(sInd)) * S.tStep;
        S.stepLen = S.tStep * S.L;
    } else {
        S.tStep = ia.dL / S.L;
        S.stepLen = ia.dL;
        S.t0 = tStart;
    }
    S.n = S.t0 <= tEnd ? (uint32_t)((tEnd - S.t0) / S.tStep) + 1u : 0u;
    return true;
}

template <class V>
__device__ __forceinline__ void interp_walk(const InterpArgs& ia, const Samples& S, float t0, uint32_t n, V&& visit) {
    const IF3 dp{S.v.x * ia.scale.x * S.tStep, S.v.y * ia.scale.y * S.tStep, S.v.z * ia.scale.z * S.tStep};
    IF3 p{0.f, 0.f, 0.f};
    for (uint32_t ii = 0; ii < n; ii++) {
        if ((ii & 63u) == 0u) {
            const float t = t0 + (float)ii * S.tStep;
            p = IF3{S.v.x * t + S.s.x, S.v.y * t + S.s.y, S.v.z * t + S.s.z};
            p = IF3{(p.x - ia.b.x) * ia.scale.x, (p.y - ia.b.y) * ia.scale.y, (p.z - ia.b.z) * ia.scale.z};
        } else {
            p = IF3{p.x + dp.x, p.y + dp.y, p.z + dp.z};
        }
        visit(p);
    }
}

// temp of a PET measurement: 1 / (L nRays) (projectorType4.cl:612-617), the corrections, the image attenuation
// factor of this ray last (FP) or first (BP).
__device__ __forceinline__ float interp_weight(const InterpArgs& ia, int64_t lo, float L, bool fp, int ray = 0) {
    const SidArgs& g = ia.a;
    const uint32_t nr = n_rays(g);
    float temp = nr > 1 ? 1.f / (L * (float)nr) : 1.f / L;
    const bool atn = g.atten && ia.atn_image;
    const float fa = atn ? g.atten[(int64_t)ray * g.atten_ray_stride + lo] : 0.f;
    if (atn && !fp) temp *= fa;
    if (g.norm) temp *= g.norm[lo];
    if (g.scat) temp *= g.scat[lo];
    if (g.atten && !ia.atn_image) temp *= g.atten[lo];
    temp *= g.global_factor;
    if (atn && fp) temp *= fa;
    return temp;
}

__global__ __launch_bounds__(kInterpBlock) void k_interp_fp(InterpArgs ia, const float* __restrict__ img,
                                                            float* __restrict__ y) {
    const int64_t lo = (int64_t)blockIdx.x * kInterpBlock + threadIdx.x;
    if (lo >= ia.a.nmeas) return;
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz;
    if (ia.a.mode == MODE_CT_FLAT) {
        Samples S;
        if (!interp_samples(ia, lo, S)) { y[lo] = 0.f; return; }
        float ax = 0.f;
        interp_walk(ia, S, S.t0, S.n, [&](IF3 p) { ax += S.stepLen * tex3(img, Nx, Ny, Nz, p.x, p.y, p.z); });
        y[lo] = ax;
        return;
    }
    // PET: every ray's samples times its own weight, added in ray order (a ray missing the image adds nothing)
    float tot = 0.f;
    for (uint32_t ray = 0; ray < n_rays(ia.a); ray++) {
        Samples S;
        if (!interp_samples(ia, lo, S, (int)ray)) continue;
        float ax = 0.f;
        interp_walk(ia, S, S.t0, S.n, [&](IF3 p) { ax += S.stepLen * tex3(img, Nx, Ny, Nz, p.x, p.y, p.z); });
        tot += ax * interp_weight(ia, lo, S.L, true, (int)ray);
    }
    y[lo] = tot;
}

// exp(-sum stepLen mu(p)) over the forward projection's samples; 1 where the ray misses the image.
__global__ __launch_bounds__(kInterpBlock) void k_interp_atten(InterpArgs ia, const float* __restrict__ mu,
                                                               float* __restrict__ f) {
    const int64_t lo = (int64_t)blockIdx.x * kInterpBlock + threadIdx.x;
    if (lo >= ia.a.nmeas) return;
    const int ray = (int)blockIdx.y;   // several rays: ray r's factors at f[r * nmeas + lo]
    f += (int64_t)ray * ia.a.nmeas;
    Samples S;
    if (!interp_samples(ia, lo, S, ray)) { f[lo] = 1.f; return; }
    float jel = 0.f;
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz;
    interp_walk(ia, S, S.t0, S.n, [&](IF3 p) { jel += S.stepLen * -tex3(mu, Nx, Ny, Nz, p.x, p.y, p.z); });
    f[lo] = expf(jel);
}

// Emission-data backprojection: rhs += y stepLen temp, sens += stepLen temp into each sample's voxel.
__global__ __launch_bounds__(kInterpBlock) void k_interp_bp_rays(InterpArgs ia, const float* __restrict__ yin,
                                                                 float* __restrict__ rhs, float* __restrict__ sens) {
    const int64_t lo = (int64_t)blockIdx.x * kInterpBlock + threadIdx.x;
    if (lo >= ia.a.nmeas) return;
    const float yv = yin[lo];
    if (yv == 0.f && sens == nullptr) return;
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz;
    const uint64_t N = (uint64_t)Nx * Ny * Nz;
    for (uint32_t ray = 0; ray < n_rays(ia.a); ray++) {   // each ray with its own weight
        Samples S;
        if (!interp_samples(ia, lo, S, (int)ray)) continue;
        float t0 = S.t0;
        uint32_t n = S.n;
        if (ia.a.atten && ia.atn_image && ia.dL > 0.f) {
            t0 = S.tStart + S.tStep / 10.f;
            n = t0 <= S.tEnd ? (uint32_t)((S.tEnd - t0) / S.tStep) + 1u : 0u;
        }
        const float e = S.stepLen * interp_weight(ia, lo, S.L, false, (int)ray);
        interp_walk(ia, S, t0, n, [&](IF3 p) {
            const uint64_t ind = (uint64_t)(uint32_t)fmaxf(p.x * (float)Nx, 0.f) +
                                 (uint64_t)(uint32_t)fmaxf(p.y * (float)Ny, 0.f) * Nx +
                                 (uint64_t)(uint32_t)fmaxf(p.z * (float)Nz, 0.f) * Nx * Ny;
            if (ind == 0 || ind >= N) return;
            unsafeAtomicAdd(rhs + ind, yv * e);
            if (sens) unsafeAtomicAdd(sens + ind, e);
        });
    }
}

// ---- Time of flight on PET data (projectorType4.cl -DTOF) ----
// One lane per LOR, one wave per workgroup; a lane's bins (accumulators or measurements) and the current
// sample's bin weights live in LDS, laid out [bin][lane]. D starts at |v tStart| - L / 2 (TOFDis, :419-423),
// moved to the first Joseph sample by (t0 - tStart) L DDsign (:463-466); each sample takes the weights of
// its own stretch of length stepLen (TOFLoop with the weight-sum floor 1e-6, :566-568), D then moves by
// stepLen DDsign (:600-602).
constexpr int kInterpTofBlock = 64;

__device__ __forceinline__ float cl_sign4(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : (v == v ? v : 0.f)); }

struct InterpTof {
    const SidArgs& a;
    float* wl;
    float inv, pf;
    __device__ InterpTof(const SidArgs& args, float* w)
        : a(args), wl(w), inv(1.f / args.tof_sigma), pf(0.3989423f * (1.f / args.tof_sigma)) {}
    __device__ float pdf(float x, float mu) const {
        const float q = (x - mu) * inv;
        return pf * expf(-0.5f * q * q);
    }
    // TOFLoop / TOFWeight (general_opencl_functions.h:1074-1116, -DUSEMAD, TRAPZ_BINS 4) into wl; the sum
    // floored at 1e-6
    __device__ float loop(float sgn, float elem, float D) const {
        const float dX = elem / 3.f;
        const float dXs = dX * sgn;
        float s = 0.f;
        for (uint32_t b = 0; b < a.tof_bins; b++) {
            const float mu = a.tof_c[b];
            float out = pdf(D, mu);
            out += pdf(fmaf(-dXs, 1.f, D), mu) * 2.f;
            out += pdf(fmaf(-dXs, 2.f, D), mu) * 2.f;
            out += pdf(fmaf(-elem, sgn, D), mu);
            const float v = out * dX;
            wl[b * kInterpTofBlock] = v;
            s += v;
        }
        return s < 1e-6f ? 1e-6f : s;
    }
};

__device__ __forceinline__ float interp_tof_start(const InterpArgs& ia, const Samples& S, float& sgn) {
    const float ax = S.v.x * S.tStart, ay = S.v.y * S.tStart, az = S.v.z * S.tStart;
    float D = sqrtf(ax * ax + ay * ay + az * az) - S.L / 2.f;
    sgn = cl_sign4(D);
    if (ia.dL <= 0.f) D -= ((S.t0 - S.tStart) * S.L * sgn);
    return D;
}

// FP (denominator): ax_b += weight_b ((stepLen x(p)) / sum); y_b = ax_b temp.
__global__ __launch_bounds__(kInterpTofBlock) void k_interp_tof_fp(InterpArgs ia, const float* __restrict__ img,
                                                                   float* __restrict__ y) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kInterpTofBlock + threadIdx.x;
    if (lo >= ia.a.nmeas) return;
    const uint32_t nb = ia.a.tof_bins;
    float* tot = lds + threadIdx.x;                              // the LOR's bins, over its rays
    float* ax = lds + nb * kInterpTofBlock + threadIdx.x;        // one ray's bins
    const InterpTof t(ia.a, lds + 2 * nb * kInterpTofBlock + threadIdx.x);
    for (uint32_t b = 0; b < nb; b++) tot[b * kInterpTofBlock] = 0.f;
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz;
    for (uint32_t ray = 0; ray < n_rays(ia.a); ray++) {   // each ray's bins times its weight (:633-641)
        Samples S;
        if (!interp_samples(ia, lo, S, (int)ray)) continue;
        for (uint32_t b = 0; b < nb; b++) ax[b * kInterpTofBlock] = 0.f;
        float sgn;
        float D = interp_tof_start(ia, S, sgn);
        interp_walk(ia, S, S.t0, S.n, [&](IF3 p) {
            const float sum = t.loop(sgn, S.stepLen, D);
            const float dX = (S.stepLen * tex3(img, Nx, Ny, Nz, p.x, p.y, p.z)) / sum;
            for (uint32_t b = 0; b < nb; b++) ax[b * kInterpTofBlock] += t.wl[b * kInterpTofBlock] * dX;
            D -= S.stepLen * sgn;
        });
        const float temp = interp_weight(ia, lo, S.L, true, (int)ray);
        for (uint32_t b = 0; b < nb; b++) tot[b * kInterpTofBlock] += ax[b * kInterpTofBlock] * temp;
    }
    for (uint32_t b = 0; b < nb; b++) y[lo + (int64_t)b * ia.a.nmeas] = tot[b * kInterpTofBlock];
}

// BP (rhs): the sample's voxel gets sum_b ((stepLen temp) / sum) weight_b y_b (and the factors' sum in
// sens); a sample whose voxel is skipped (index 0 or past the image) leaves D where it was, as the kernel's
// `continue` there skips the update (:570-576).
__global__ __launch_bounds__(kInterpTofBlock) void k_interp_tof_bp(InterpArgs ia, const float* __restrict__ yin,
                                                                   float* __restrict__ rhs, float* __restrict__ sens) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kInterpTofBlock + threadIdx.x;
    if (lo >= ia.a.nmeas) return;
    const uint32_t nb = ia.a.tof_bins;
    float* yv = lds + threadIdx.x;
    bool any = false;
    for (uint32_t b = 0; b < nb; b++) {
        const float v = yin[lo + (int64_t)b * ia.a.nmeas];
        yv[b * kInterpTofBlock] = v;
        any |= v != 0.f;
    }
    if (!any && sens == nullptr) return;
    const InterpTof t(ia.a, lds + nb * kInterpTofBlock + threadIdx.x);
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz;
    const uint64_t N = (uint64_t)Nx * Ny * Nz;
    for (uint32_t ray = 0; ray < n_rays(ia.a); ray++) {
        Samples S;
        if (!interp_samples(ia, lo, S, (int)ray)) continue;
        float t0 = S.t0;
        uint32_t n = S.n;
        if (ia.a.atten && ia.atn_image && ia.dL > 0.f) {
            t0 = S.tStart + S.tStep / 10.f;
            n = t0 <= S.tEnd ? (uint32_t)((S.tEnd - t0) / S.tStep) + 1u : 0u;
        }
        float sgn;
        float D = interp_tof_start(ia, S, sgn);
        const float e = S.stepLen * interp_weight(ia, lo, S.L, false, (int)ray);
        interp_walk(ia, S, t0, n, [&](IF3 p) {
            const float sum = t.loop(sgn, S.stepLen, D);
            const uint64_t ind = (uint64_t)(uint32_t)fmaxf(p.x * (float)Nx, 0.f) +
                                 (uint64_t)(uint32_t)fmaxf(p.y * (float)Ny, 0.f) * Nx +
                                 (uint64_t)(uint32_t)fmaxf(p.z * (float)Nz, 0.f) * Nx * Ny;
            if (ind == 0 || ind >= N) return;
            const float dX = e / sum;
            float val = 0.f, yax = 0.f;
            for (uint32_t b = 0; b < nb; b++) {
                const float apu = dX * t.wl[b * kInterpTofBlock];
                val += apu;
                yax += apu * yv[b * kInterpTofBlock];
            }
            unsafeAtomicAdd(rhs + ind, yax);
            if (sens) unsafeAtomicAdd(sens + ind, val);
            D -= S.stepLen * sgn;
        });
    }
}

// Forward projection of direct-plane PET sinograms: every sample of a plane has the plane's z, so a
// workgroup (plane, chunk of its LORs) stages the two image slices its trilinear samples blend in LDS and
// gathers the corners there; values as k_interp_fp's (the same lerps on the same numbers).
constexpr int kInterpPlanarFpTPB = 512;

__global__ __launch_bounds__(kInterpPlanarFpTPB) void k_interp_fp_planar(InterpArgs ia, uint32_t n_planes,
                                                                        uint32_t chunk,
                                                                        const float* __restrict__ img,
                                                                        float* __restrict__ y) {
    extern __shared__ float slices[];   // [Nyx] slice z0, [Nyx] slice z1
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz, Nyx = Nx * Ny;
    const uint32_t per = ia.a.size_x * ia.a.size_y;
    const uint32_t plane = blockIdx.x % n_planes, grp = blockIdx.x / n_planes;
    // the walk's z for this plane: ((0 t + z) - bz) sz, constant along every LOR of the plane
    const float pz = ((0.f + ia.a.z[(ia.a.plane_offset + plane) * 2]) - ia.b.z) * ia.scale.z;
    uint32_t z0, z1;
    float c;
    tex_axis(pz, Nz, z0, z1, c);
    for (uint32_t v = threadIdx.x; v < Nyx; v += kInterpPlanarFpTPB) {
        slices[v] = img[(size_t)z0 * Nyx + v];
        slices[Nyx + v] = img[(size_t)z1 * Nyx + v];
    }
    __syncthreads();
    const uint32_t i = grp * chunk + threadIdx.x;
    if (i >= min(per, (grp + 1) * chunk)) return;
    const int64_t lo = (int64_t)plane * per + i;
    Samples S;
    if (!interp_samples(ia, lo, S)) { y[lo] = 0.f; return; }
    float ax = 0.f;
    interp_walk(ia, S, S.t0, S.n, [&](IF3 p) { ax += S.stepLen * tex_slices(slices, slices + Nyx, Nx, Ny, p.x, p.y, c); });
    y[lo] = 0.f + ax * interp_weight(ia, lo, S.L, true);
}

// Direct-plane PET sinograms (every LOR of a plane in one image slice: zs == zd, so every sample's z is the
// plane's): k_interp_bp_rays with the slice accumulated in LDS. A workgroup owns (plane, chunk of the plane's
// LORs), adds each sample's stepLen temp y (and stepLen temp) to its LDS slice, then adds the slice to rhs
// (and sens) with one global atomic per touched voxel, instead of one per sample. The slice is the one the
// walk's arithmetic gives for the plane's z, trunc(max(((0 t + z) - bz) sz Nz, 0)); a sample on the image's
// far x face of its last row has the linear index of the next slice's first voxel (x = Nx wraps), where the
// reference deposits it: such samples go to global memory directly.
constexpr int kInterpPlanarTPB = 512;

template <bool SENS>
__global__ __launch_bounds__(kInterpPlanarTPB) void k_interp_bp_planar(InterpArgs ia, uint32_t n_planes, uint32_t chunk,
                                                                       const float* __restrict__ yin,
                                                                       float* __restrict__ rhs,
                                                                       float* __restrict__ sens) {
    extern __shared__ float slice[];   // [Nyx] rhs, then [Nyx] sens
    const uint32_t Nx = ia.a.Nx, Ny = ia.a.Ny, Nz = ia.a.Nz, Nyx = Nx * Ny;
    const uint32_t per = ia.a.size_x * ia.a.size_y;
    const uint32_t plane = blockIdx.x % n_planes, grp = blockIdx.x / n_planes;
    const float pz = ((0.f + ia.a.z[(ia.a.plane_offset + plane) * 2]) - ia.b.z) * ia.scale.z;
    const uint64_t zb = (uint64_t)(uint32_t)fmaxf(pz * (float)Nz, 0.f);
    for (uint32_t v = threadIdx.x; v < (SENS ? 2 : 1) * Nyx; v += kInterpPlanarTPB) slice[v] = 0.f;
    __syncthreads();
    const uint64_t N = (uint64_t)Nyx * Nz;
    const uint32_t i_end = min(per, (grp + 1) * chunk);
    for (uint32_t i = grp * chunk + threadIdx.x; i < i_end; i += kInterpPlanarTPB) {
        const int64_t lo = (int64_t)plane * per + i;
        const float yv = yin[lo];
        if (yv == 0.f && !SENS) continue;
        Samples S;
        if (!interp_samples(ia, lo, S)) continue;
        float t0 = S.t0;
        uint32_t n = S.n;
        if (ia.a.atten && ia.atn_image && ia.dL > 0.f) {
            t0 = S.tStart + S.tStep / 10.f;
            n = t0 <= S.tEnd ? (uint32_t)((S.tEnd - t0) / S.tStep) + 1u : 0u;
        }
        const float e = S.stepLen * interp_weight(ia, lo, S.L, false);
        interp_walk(ia, S, t0, n, [&](IF3 p) {
            const uint64_t ind = (uint64_t)(uint32_t)fmaxf(p.x * (float)Nx, 0.f) +
                                 (uint64_t)(uint32_t)fmaxf(p.y * (float)Ny, 0.f) * Nx +
                                 (uint64_t)(uint32_t)fmaxf(p.z * (float)Nz, 0.f) * Nyx;
            if (ind == 0 || ind >= N) return;
            if (ind / Nyx == zb) {
                const uint32_t v = (uint32_t)(ind - zb * Nyx);
                atomicAdd(slice + v, yv * e);
                if (SENS) atomicAdd(slice + Nyx + v, e);
            } else {
                unsafeAtomicAdd(rhs + ind, yv * e);
                if (SENS) unsafeAtomicAdd(sens + ind, e);
            }
        });
    }
    __syncthreads();
    if (zb >= Nz) return;   // the plane lies outside the image: nothing was added to the slice
    for (uint32_t v = threadIdx.x; v < Nyx; v += kInterpPlanarTPB) {
        const float a = slice[v];
        if (a != 0.f) unsafeAtomicAdd(rhs + zb * Nyx + v, a);
        if (SENS) {
            const float b = slice[Nyx + v];
            if (b != 0.f) unsafeAtomicAdd(sens + zb * Nyx + v, b);
        }
    }
}

// CT backprojection, voxel-driven: one lane per (x, y, z group of kInterpNV voxels), projections
// [plane_offset, plane_offset + n_proj) of the panel stack yin (n_proj panels of size_x x size_y).
__global__ __launch_bounds__(kInterpBlock) void k_interp_bp_ct(InterpArgs ia, uint32_t n_proj,
                                                               const float* __restrict__ yin, float* __restrict__ out,
                                                               float* __restrict__ sens) {
    const SidArgs& g = ia.a;
    const uint32_t Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const uint32_t groups = (Nz + kInterpNV - 1) / kInterpNV;
    const uint64_t q = (uint64_t)blockIdx.x * kInterpBlock + threadIdx.x;
    if (q >= (uint64_t)Nx * Ny * groups) return;
    const uint32_t gi = (uint32_t)(q % Nx), gj = (uint32_t)((q / Nx) % Ny);
    const uint32_t gk = (uint32_t)(q / ((uint64_t)Nx * Ny)) * kInterpNV;
    float temp[kInterpNV], wsum[kInterpNV];
#pragma unroll
    for (uint32_t zz = 0; zz < kInterpNV; zz++) { temp[zz] = 0.f; wsum[zz] = 0.f; }
    const IF3 dV{(float)gi * g.dx + g.dx / 2.f + g.bx, (float)gj * g.dy + g.dy / 2.f + g.by,
                 (float)gk * g.dz + g.dz / 2.f + g.bz};
    const size_t NA = g.ct_pitch ? 6 : 2;
    const float kx = (float)g.size_x * ia.dPitchX, ky = (float)g.size_y * ia.dPitchY;
    const float ix = (float)g.size_x / 2.f, iy = (float)g.size_y / 2.f;
    for (uint32_t kk = 0; kk < n_proj; kk++) {
        const float* xp = g.x + (size_t)(g.plane_offset + kk) * 6;
        const float* up = g.z + (size_t)(g.plane_offset + kk) * NA;
        const IF3 s{xp[0], xp[1], xp[2]}, d1{xp[3], xp[4], xp[5]};
        IF3 apuX, apuY;
        if (g.ct_pitch) {
            apuX = IF3{up[0] * ix, up[1] * ix, up[2] * ix};
            apuY = IF3{up[3] * iy, up[4] * iy, up[5] * iy};
        } else {
            apuX = IF3{ix * up[0], ix * up[1], 0.f};
            apuY = IF3{0.f, 0.f, iy * ia.dPitchY};
        }
        const IF3 d2{apuX.x - apuY.x, apuX.y - apuY.y, apuX.z - apuY.z};
        const IF3 d3{d1.x - apuX.x - apuY.x, d1.y - apuX.y - apuY.y, d1.z - apuX.z - apuY.z};
        const float nx = sqrtf(apuX.x * apuX.x + apuX.y * apuX.y + apuX.z * apuX.z);
        const float ny = sqrtf(apuY.x * apuY.x + apuY.y * apuY.y + apuY.z * apuY.z);
        const IF3 normX{apuX.x / nx / kx, apuX.y / nx / kx, apuX.z / nx / kx};
        const IF3 normY{apuY.x / ny / ky, apuY.y / ny / ky, apuY.z / ny / ky};
        const IF3 e{d3.x - d1.x, d3.y - d1.y, d3.z - d1.z};
        const IF3 cP{d2.y * e.z - d2.z * e.y, d2.z * e.x - d2.x * e.z, d2.x * e.y - d2.y * e.x};
        const float pz = ((float)kk + 0.5f) / (float)n_proj;
        const float dApu = g.dz * cP.z;
        const IF3 sd{s.x - d1.x, s.y - d1.y, s.z - d1.z};
        const float upper = cP.x * sd.x + cP.y * sd.y + cP.z * sd.z;
        IF3 v{dV.x - s.x, dV.y - s.y, dV.z - s.z};
        float lower = -(v.x * cP.x + v.y * cP.y + v.z * cP.z);
        const float vApu = v.x * v.x + v.y * v.y;
#pragma unroll
        for (uint32_t zz = 0; zz < kInterpNV; zz++) {
            if (gk + zz < Nz) {
                const float t = upper / lower;
                IF3 p{s.x + v.x * t, s.y + v.y * t, s.z + v.z * t};
                const float l1 = vApu + v.z * v.z;
                const IF3 ps{p.x - s.x, p.y - s.y, p.z - s.z};
                const float L = sqrtf(ps.x * ps.x + ps.y * ps.y + ps.z * ps.z);
                const float weight = (L * L * L) / l1 * ia.kerroin;
                p = IF3{p.x - d3.x, p.y - d3.y, p.z - d3.z};
                const float px = p.x * normX.x + p.y * normX.y + p.z * normX.z;
                const float py = p.x * normY.x + p.y * normY.y + p.z * normY.z;
                float yv = 0.f;
                if (px <= 1.f && py <= 1.f && pz <= 1.f && px >= 0.f && py >= 0.f && pz >= 0.f)
                    yv = tex3(yin, g.size_x, g.size_y, n_proj, px, py, pz);
                if (yv != 0.f) {
                    temp[zz] += yv * weight;
                    wsum[zz] += weight;
                }
                v.z += g.dz;
                lower -= dApu;
            }
        }
    }
#pragma unroll
    for (uint32_t zz = 0; zz < kInterpNV; zz++) {
        if (gk + zz < Nz) {
            const size_t idx = (size_t)gi + (size_t)gj * Nx + (size_t)(gk + zz) * Nx * Ny;
            out[idx] += temp[zz];
            if (sens) sens[idx] += wsum[zz];
        }
    }
}

// r = meas / ((yhat + epps) + additive), NaN/Inf -> epps, in place on yhat.
// n values (TOF: bins x n_lor); the additive term is one value per LOR, added to every bin.
__global__ __launch_bounds__(256) void k_interp_ratio(const float* __restrict__ meas, float* __restrict__ q,
                                                      const float* __restrict__ add, int64_t n, int64_t n_lor,
                                                      float epps) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float d = q[i] + epps;
    if (add) d += add[i % n_lor];
    float r = meas[i] / d;
    if (isnan(r) || isinf(r)) r = epps;
    q[i] = r;
}

static InterpArgs make_interp_args(const omg_geom* g, int64_t first, int64_t n) {
    InterpArgs ia;
    ia.a = make_sid_args(g, first, n);
    const omg_geom_desc& d = g->d;
    // box far corner b + N d in double, rounded once; scale 1 / (d N) (init.py:705-735, coordinates.py:224-226)
    ia.b = IF3{d.bx, d.by, d.bz};
    ia.bmax = IF3{(float)((double)d.bx + (double)d.Nx * (double)d.dx), (float)((double)d.by + (double)d.Ny * (double)d.dy),
                  (float)((double)d.bz + (double)d.Nz * (double)d.dz)};
    ia.scale = IF3{(float)(1.0 / ((double)d.dx * d.Nx)), (float)(1.0 / ((double)d.dy * d.Ny)),
                   (float)(1.0 / ((double)d.dz * d.Nz))};
    ia.dL = d.interp_dl;
    ia.kerroin = d.interp_kerroin;
    ia.dPitchX = d.dPitchX;
    ia.dPitchY = d.dPitchY;
    ia.atn_image = d.atten_mode == 1;
    ia.a.atten_lor = d.atten_mode == 2;
    return ia;
}

// LDS of the TOF kernels: FP [bins: the LOR's sums][bins: one ray's sums][bins: weights], BP the first two
// The slice-in-LDS kernels: direct-plane sinograms without TOF or several rays (the geometry's planar flag),
// whole planes per call, two slices within one CU's LDS.
static bool interp_planar_ok(const omg_geom* g, int64_t n) {
    const int64_t per = (int64_t)g->d.size_x * g->d.size_y;
    return g->planar && !g->force_generic && per > 0 && n % per == 0 &&
           2 * (size_t)g->d.Nx * g->d.Ny * sizeof(float) <= 160 * 1024 - 256;
}

static int interp_planar_attr(size_t bytes) {   // above 64 KiB of LDS the kernels must opt in
    static size_t granted = 64 * 1024;
    if (bytes <= granted) return OMG_OK;
    OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_interp_bp_planar<true>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_interp_bp_planar<false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_interp_fp_planar),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    granted = bytes;
    return OMG_OK;
}

static size_t interp_tof_lds(const omg_geom* g) { return 3 * (size_t)g->d.tof_bins * kInterpTofBlock * sizeof(float); }

int interp_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s) {
    const InterpArgs ia = make_interp_args(g, first, n);
    hipEvent_t stop;
    timing_begin(s, &stop);
    if (g->d.tof_bins > 1) {
        hipLaunchKernelGGL(k_interp_tof_fp, dim3(grid_for(n, kInterpTofBlock)), dim3(kInterpTofBlock),
                           interp_tof_lds(g), s, ia, img, y);
    } else if (interp_planar_ok(g, n)) {   // one LOR per lane, a plane's LORs over several workgroups
        const uint32_t per = g->d.size_x * g->d.size_y, n_planes = (uint32_t)(n / per);
        const uint32_t groups = (per + kInterpPlanarFpTPB - 1) / kInterpPlanarFpTPB;
        const uint32_t chunk = (per + groups - 1) / groups;
        const size_t lds = 2 * (size_t)g->d.Nx * g->d.Ny * sizeof(float);
        if (int rc = interp_planar_attr(lds)) return rc;
        hipLaunchKernelGGL(k_interp_fp_planar, dim3(n_planes * groups), dim3(kInterpPlanarFpTPB), lds, s, ia,
                           n_planes, chunk, img, y);
    } else
        hipLaunchKernelGGL(k_interp_fp, dim3(grid_for(n, kInterpBlock)), dim3(kInterpBlock), 0, s, ia, img, y);
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int interp_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens,
                    hipStream_t s) {
    const InterpArgs ia = make_interp_args(g, first, n);
    hipEvent_t stop;
    timing_begin(s, &stop);
    if (g->d.mode == OMG_MODE_CT_FLAT) {
        const uint32_t per = g->d.size_x * g->d.size_y;
        const uint32_t nproj = (uint32_t)(n / per);
        const uint64_t lanes = (uint64_t)g->d.Nx * g->d.Ny * ((g->d.Nz + kInterpNV - 1) / kInterpNV);
        hipLaunchKernelGGL(k_interp_bp_ct, dim3((unsigned)((lanes + kInterpBlock - 1) / kInterpBlock)),
                           dim3(kInterpBlock), 0, s, ia, nproj, yin, rhs, sens);
    } else if (g->d.tof_bins > 1) {
        hipLaunchKernelGGL(k_interp_tof_bp, dim3(grid_for(n, kInterpTofBlock)), dim3(kInterpTofBlock),
                           interp_tof_lds(g), s, ia, yin, rhs, sens);
    } else if (interp_planar_ok(g, n)) {
        const uint32_t per = g->d.size_x * g->d.size_y, n_planes = (uint32_t)(n / per);
        const uint32_t groups = (per + kInterpPlanarTPB - 1) / kInterpPlanarTPB;   // one LOR per lane
        const uint32_t chunk = (per + groups - 1) / groups;
        const size_t lds = (sens ? 2 : 1) * (size_t)g->d.Nx * g->d.Ny * sizeof(float);
        if (int rc = interp_planar_attr(2 * (size_t)g->d.Nx * g->d.Ny * sizeof(float))) return rc;
        if (sens)
            hipLaunchKernelGGL(k_interp_bp_planar<true>, dim3(n_planes * groups), dim3(kInterpPlanarTPB), lds, s, ia,
                               n_planes, chunk, yin, rhs, sens);
        else
            hipLaunchKernelGGL(k_interp_bp_planar<false>, dim3(n_planes * groups), dim3(kInterpPlanarTPB), lds, s, ia,
                               n_planes, chunk, yin, rhs, sens);
    } else {
        hipLaunchKernelGGL(k_interp_bp_rays, dim3(grid_for(n, kInterpBlock)), dim3(kInterpBlock), 0, s, ia, yin, rhs,
                           sens);
    }
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

// FP -> ratio -> BP into rhs (and sens) for emission data, through a per-measurement work buffer.
int interp_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
                float* sens, float epps, hipStream_t s, const float* add) {
    const int64_t nq = n * (g->d.tof_bins > 1 ? (int64_t)g->d.tof_bins : 1);
    int rc = q_reserve(const_cast<omg_geom*>(g), (size_t)nq);
    if (rc) return rc;
    float* q = q_ptr(g);
    rc = interp_forward(g, first, n, img, q, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_interp_ratio, dim3(grid_for(nq, 256)), dim3(256), 0, s, meas, q, add, nq, n, epps);
    OMG_CHECK_LAUNCH();
    return interp_backward(g, first, n, q, rhs, sens, s);
}

// Attenuation image -> one factor per measurement by the interpolation walk (the geometry's own LORs).
int interp_atten_factors(const omg_geom* g, const float* mu, float* f, int64_t n, hipStream_t s) {
    const InterpArgs ia = make_interp_args(g, 0, n);
    hipLaunchKernelGGL(k_interp_atten, dim3(grid_for(n, kInterpBlock), ia.a.rays_xy * ia.a.rays_z), dim3(kInterpBlock),
                       0, s, ia, mu, f);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

}  // namespace omg
