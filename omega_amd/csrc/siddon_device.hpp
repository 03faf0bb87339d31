// siddon_device.hpp -- improved Siddon ray traversal for gfx950 (device side).
//
// One lane traces one LOR/detector pixel. The arithmetic is the reference CPU
// projector's, operation for operation (projectorType123Implementation4,
// source/cpp/projector_functions.h:1311-2063), so that the visited voxel list and
// the forward projection are bit-identical to OMEGA's CPU build in FP32. The
// library is compiled with -ffp-contract=off and IEEE-rounded div/sqrt for this.
//
// Differences in form (not in result) from the reference loop:
//   * the three-way "next plane crossing" choice (:1857-1922) is a single select
//     of (axis, t_next) followed by one length formula;
//   * the trace is a template over a visitor so FP, BP, fused OSEM and the
//     index-list dump share one traversal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace omg {

enum Mode : int32_t { MODE_PET_SINO = 0, MODE_CT_FLAT = 1, MODE_LIST6 = 2 };

struct SidArgs {
    uint32_t Nx, Ny, Nz, Nyx;
    float dx, dy, dz, bx, by, bz;
    float bmaxx, bmaxy, bmaxz;     // far FOV edges, computed as in projector_functions.h:1331-1333
    uint32_t size_x, size_y;       // detector rows (radial bins) / columns (angles)
    int32_t mode;
    int32_t ct_pitch;
    int64_t plane_offset;          // subset's first plane/projection (LOR for list mode)
    int64_t nmeas;                 // measurements in this call
    float dPitchZ;
    float global_factor;
    const float* __restrict__ x;   // coordinate tables (device)
    const float* __restrict__ z;
    const float* __restrict__ norm; // per-measurement normalisation, subset-local, or null
    const float* __restrict__ atten; // per-measurement attenuation factor, subset-local, or null
    const float* __restrict__ scat;  // per-measurement multiplicative (scatter) correction, subset-local, or null
    bool atten_lor;                  // atten holds given per-LOR factors (not factors computed from an image)
    int64_t atten_ray_stride;        // factors made from an image with several rays per LOR: ray r's factor of
                                     // measurement lo is atten[r * atten_ray_stride + lo]; else 0
    uint32_t tof_bins;               // time-of-flight bins (<= 1: no TOF)
    float tof_sigma, tof_pf, tof_epps;  // TOF standard deviation (mm), 0.3989422804014327f / sigma, weight-sum floor
    const float* __restrict__ tof_c; // TOF bin centres (mm), tof_bins floats
    uint32_t rays_xy, rays_z;        // rays per LOR, transaxial x axial (1 x 1: one ray)
    float cr_xy, cr_z;               // the pitches the rays' shifts are spread over (paramStruct dPitchXY, dPitchZ)
    uint32_t pitch;                // row pitch of slice-relative (planar) voxel indices: Nx, or Nx + 1 in LDS
    const uint32_t* __restrict__ index;  // PET sinogram: sinogram bin per measurement (subset types 1-5), or null;
                                         // plane_offset is then the call's first measurement in this table
    const float* __restrict__ add;  // additive term of the fused ratio (randoms / scatter), call-local, or null
};

__device__ __forceinline__ uint32_t n_rays(const SidArgs& a) { return a.rays_xy * a.rays_z; }

// Denominator of the fused EM ratio: (A x + epps) + additive (computeForwardProject.h:20-21, 66-69).
__device__ __forceinline__ float ratio_den(const SidArgs& a, int64_t lo, float yv, float epps) {
    float d = yv + epps;
    if (a.add) d += a.add[lo];
    return d;
}

struct Ray { float xs, ys, zs, xd, yd, zd; };

__device__ __forceinline__ uint32_t to_u32(float v) { return (uint32_t)(int64_t)v; }

// std::min / std::max semantics (first argument wins ties and unordered compares).
__device__ __forceinline__ float mn(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return (a < b) ? b : a; }

// Ray `ray` (axial index ray / rays_xy outer, transaxial index inner) of a multi-ray LOR: both
// endpoints move by (cr / (2n) - cr / 2) + (cr / n) * index, the transaxial shift on x and y alike
// (multirayCoordinateShiftZ / XY, projector_functions.h:439-461, applied Z first).
__device__ __forceinline__ void ray_shift(const SidArgs& a, int ray, Ray& r) {
    const int lorZ = ray / (int)a.rays_xy, lorXY = ray - lorZ * (int)a.rays_xy;
    if (a.rays_z > 1) {
        float interval = a.cr_z / (float)(a.rays_z * 2);
        r.zs += (interval - a.cr_z / 2.f);
        r.zd += (interval - a.cr_z / 2.f);
        interval *= 2.f;
        r.zs += interval * (float)lorZ;
        r.zd += interval * (float)lorZ;
    }
    if (a.rays_xy > 1) {
        float interval = a.cr_xy / (float)(a.rays_xy * 2);
        r.xs += (interval - a.cr_xy / 2.f);
        r.xd += (interval - a.cr_xy / 2.f);
        r.ys += (interval - a.cr_xy / 2.f);
        r.yd += (interval - a.cr_xy / 2.f);
        interval *= 2.f;
        r.xs += interval * (float)lorXY;
        r.xd += interval * (float)lorXY;
        r.ys += interval * (float)lorXY;
        r.yd += interval * (float)lorXY;
    }
}

__device__ __forceinline__ void fetch_ray_base(const SidArgs& a, int64_t lo, Ray& r);

__device__ __forceinline__ void fetch_ray(const SidArgs& a, int64_t lo, Ray& r, int ray = 0) {
    fetch_ray_base(a, lo, r);
    if (n_rays(a) > 1) ray_shift(a, ray, r);
}

__device__ __forceinline__ void fetch_ray_base(const SidArgs& a, int64_t lo, Ray& r) {
    if (a.mode == MODE_LIST6) {
        const float* p = a.x + (a.plane_offset + lo) * 6;
        r.xs = p[0]; r.ys = p[1]; r.zs = p[2]; r.xd = p[3]; r.yd = p[4]; r.zd = p[5];
        return;
    }
    const uint32_t per = a.size_x * a.size_y;
    if (a.index) {   // measurement-level subset: the bin comes from the index table (getIndex)
        const uint32_t b = a.index[a.plane_offset + lo];
        const uint32_t iz = b / per, rem = b - iz * per;
        const float4 q = *reinterpret_cast<const float4*>(a.x + (size_t)rem * 4);
        r.xs = q.x; r.ys = q.y; r.xd = q.z; r.yd = q.w;
        const float2 zz = *reinterpret_cast<const float2*>(a.z + (size_t)iz * 2);
        r.zs = zz.x; r.zd = zz.y;
        return;
    }
    const uint32_t l32 = (uint32_t)lo;      // a subset call holds < 2^32 measurements
    uint32_t iz = l32 / per;
    const uint32_t rem = l32 - iz * per;
    const uint32_t ix = rem % a.size_x;
    const uint32_t iy = rem / a.size_x;
    const int64_t izg = (int64_t)iz + a.plane_offset;
    if (a.mode == MODE_PET_SINO) {
        const float4 q = *reinterpret_cast<const float4*>(a.x + (size_t)(ix + iy * a.size_x) * 4);
        r.xs = q.x; r.ys = q.y; r.xd = q.z; r.yd = q.w;
        const float2 zz = *reinterpret_cast<const float2*>(a.z + izg * 2);
        r.zs = zz.x; r.zd = zz.y;
    } else {
        const float* p = a.x + izg * 6;
        r.xs = p[0]; r.ys = p[1]; r.zs = p[2]; r.xd = p[3]; r.yd = p[4]; r.zd = p[5];
        // (T)ix - (T)size_x / 2. + .5 is exact in float for the sizes used (projector_functions.h:480-481)
        const float i1 = (float)((double)ix - (double)a.size_x / 2. + .5);
        const float i2 = (float)((double)iy - (double)a.size_y / 2. + .5);
        if (a.ct_pitch) {
            const float* u = a.z + izg * 6;
            r.xd += u[0] * i1 + u[3] * i2;
            r.yd += u[1] * i1 + u[4] * i2;
            r.zd += u[2] * i1 + u[5] * i2;
        } else {
            const float2 u = *reinterpret_cast<const float2*>(a.z + izg * 2);
            r.xd += u.x * i1;
            r.yd += u.y * i1;
            r.zd += a.dPitchZ * i2;
        }
    }
}

__device__ __forceinline__ void axis_inc(float tc, float tmin_a, float tmax, float tmax_a, uint32_t& vmin, uint32_t& vmax,
                                         float& t0, int32_t& step, float diff, float b, float d, float s, uint32_t N) {
    if (tc == tmin_a) vmin = 1u;
    else { const float p = s + tc * diff; vmin = to_u32(ceilf((p - b) / d)); }
    if (tmax == tmax_a) vmax = N;
    else { const float p = s + tmax * diff; vmax = to_u32(floorf((p - b) / d)); }
    t0 += (float)vmin * d / diff;
    step = 1;
}

__device__ __forceinline__ void axis_dec(float tc, float tmin_a, float tmax, float tmax_a, uint32_t& vmin, uint32_t& vmax,
                                         float& t0, int32_t& step, float diff, float b, float d, float s, uint32_t N) {
    if (tc == tmin_a) vmax = N - 1u;
    else { const float p = s + tc * diff; vmax = to_u32(floorf((p - b) / d)); }
    if (tmax == tmax_a) vmin = 0u;
    else { const float p = s + tmax * diff; vmin = to_u32(ceilf((p - b) / d)); }
    t0 += (float)vmax * d / diff;
    step = -1;
}

__device__ __forceinline__ int32_t first_voxel(float pt, float diff, float d, float a) {
    return (int32_t)floorf((pt * diff - a) / d);
}

__device__ __forceinline__ bool preloop2(float b1, float b2, float diff1, float diff2, float max1, float max2, float d1, float d2,
                                         uint32_t N1, uint32_t N2, int32_t& i1, int32_t& i2, float& t1u, float& t2u,
                                         uint32_t& Np, float s2, float s1, float e2, float e1, float& tc, int32_t& u1,
                                         int32_t& u2, float& t10, float& t20) {
    const float a1 = b1 - s1, a2 = b2 - s2;
    t10 = a1 / diff1;
    t20 = a2 / diff2;
    const float t1b = (max1 - s1) / diff1, t2b = (max2 - s2) / diff2;
    const float t1min = mn(t10, t1b), t1max = mx(t10, t1b);
    const float t2min = mn(t20, t2b), t2max = mx(t20, t2b);
    tc = mx(t1min, t2min);
    const float tmax = mn(t1max, t2max);
    if (tc >= tmax) return true;
    uint32_t v1min = 0, v1max = 0, v2min = 0, v2max = 0;
    if (s1 < e1) axis_inc(tc, t1min, tmax, t1max, v1min, v1max, t10, u1, diff1, b1, d1, s1, N1);
    else axis_dec(tc, t1min, tmax, t1max, v1min, v1max, t10, u1, diff1, b1, d1, s1, N1);
    if (s2 < e2) axis_inc(tc, t2min, tmax, t2max, v2min, v2max, t20, u2, diff2, b2, d2, s2, N2);
    else axis_dec(tc, t2min, tmax, t2max, v2min, v2max, t20, u2, diff2, b2, d2, s2, N2);
    Np = v1max + 1u + v2max + 1u - v1min - v2min;
    const float pt = (mn(t10, t20) + tc) * 0.5f;   // x/2 and x*0.5 are the same IEEE result
    i1 = first_voxel(pt, diff1, d1, a1);
    i2 = first_voxel(pt, diff2, d2, a2);
    if (i1 < 0 || (uint32_t)i1 >= N1 || i2 < 0 || (uint32_t)i2 >= N2) return true;
    t1u = d1 / fabsf(diff1);
    t2u = d2 / fabsf(diff2);
    return false;
}

__device__ __forceinline__ bool preloop3(const SidArgs& g, const Ray& r, float xdf, float ydf, float zdf,
                                         int32_t& ti, int32_t& tj, int32_t& tk, float& tyu, float& txu, float& tzu,
                                         uint32_t& Np, float& tc, int32_t& iu, int32_t& ju, int32_t& ku,
                                         float& tx0, float& ty0, float& tz0) {
    const float ax = g.bx - r.xs, ay = g.by - r.ys, az = g.bz - r.zs;
    tx0 = ax / xdf; ty0 = ay / ydf; tz0 = az / zdf;
    const float txb = (g.bmaxx - r.xs) / xdf, tyb = (g.bmaxy - r.ys) / ydf, tzb = (g.bmaxz - r.zs) / zdf;
    const float txmin = mn(tx0, txb), txmax = mx(tx0, txb);
    const float tymin = mn(ty0, tyb), tymax = mx(ty0, tyb);
    const float tzmin = mn(tz0, tzb), tzmax = mx(tz0, tzb);
    tc = mx(mx(txmin, tzmin), tymin);
    const float tmax = mn(mn(txmax, tzmax), tymax);
    if (tc >= tmax) return true;
    uint32_t imin = 0, imax = 0, jmin = 0, jmax = 0, kmin = 0, kmax = 0;
    if (r.xs < r.xd) axis_inc(tc, txmin, tmax, txmax, imin, imax, tx0, iu, xdf, g.bx, g.dx, r.xs, g.Nx);
    else axis_dec(tc, txmin, tmax, txmax, imin, imax, tx0, iu, xdf, g.bx, g.dx, r.xs, g.Nx);
    if (r.ys < r.yd) axis_inc(tc, tymin, tmax, tymax, jmin, jmax, ty0, ju, ydf, g.by, g.dy, r.ys, g.Ny);
    else axis_dec(tc, tymin, tmax, tymax, jmin, jmax, ty0, ju, ydf, g.by, g.dy, r.ys, g.Ny);
    if (r.zs < r.zd) axis_inc(tc, tzmin, tmax, tzmax, kmin, kmax, tz0, ku, zdf, g.bz, g.dz, r.zs, g.Nz);
    else axis_dec(tc, tzmin, tmax, tzmax, kmin, kmax, tz0, ku, zdf, g.bz, g.dz, r.zs, g.Nz);
    Np = (kmax - kmin + 1) + (jmax - jmin + 1) + (imax - imin + 1);
    const float pt = (mn(mn(tz0, ty0), tx0) + tc) * 0.5f;
    ti = first_voxel(pt, xdf, g.dx, ax);
    tj = first_voxel(pt, ydf, g.dy, ay);
    tk = first_voxel(pt, zdf, g.dz, az);
    if (ti < 0 || (uint32_t)ti >= g.Nx || tj < 0 || (uint32_t)tj >= g.Ny || tk < 0 || (uint32_t)tk >= g.Nz) return true;
    txu = g.dx / fabsf(xdf);
    tyu = g.dy / fabsf(ydf);
    tzu = g.dz / fabsf(zdf);
    return false;
}

// LOR weight: PET 1/L * attenuation * norm * scatter * global factor (projector_functions.h:1817-1836,
// :1181-1211), CT 1. The attenuation factor multiplies in where the reference's backprojection applies
// its exp(jelppi) or a given per-LOR factor (:1825-1828); on axis-parallel LORs (perp) a given per-LOR
// factor comes after norm, as there (:1209-1210).
// MULTI: the geometry may have several rays per LOR (never in the direct-plane kernels).
// With several rays and an attenuation image each ray has its own factor, exp(jelppi) of its own walk
// (jelppi reset per ray, :1480, :1826, :2027).
template <bool MULTI = true>
__device__ __forceinline__ float make_weight(const SidArgs& g, int64_t lo, float L, bool ct, bool perp = false,
                                             int ray = 0) {
    if (ct) return 1.f;
    // 1 / (L * nRays) with several rays per LOR (:1819-1821, :1189-1190)
    float t = (MULTI && n_rays(g) > 1) ? 1.f / (L * (float)n_rays(g)) : 1.f / L;
    const bool late = perp && g.atten_lor;
    const float* atn = g.atten ? g.atten + (int64_t)ray * g.atten_ray_stride : nullptr;
    if (atn && !late) t *= atn[lo];
    if (g.norm) t *= g.norm[lo];
    if (atn && late) t *= atn[lo];
    if (g.scat) t *= g.scat[lo];
    t *= g.global_factor;
    return t;
}

__device__ __forceinline__ int perp_start(float b, float d, float dd, uint32_t N) {
    int idx = 0;
    float s = b - d + dd;
    for (uint32_t i = 0; i < N; i++) {
        if (s > 0.f) { idx = (int)i; break; }
        s += dd;
    }
    return idx;
}

// Traversal state of one LOR. Either an axis-parallel row/column walk (perp) or a Siddon walk.
struct Walk {
    float tx0, ty0, tz0, txu, tyu, tzu, tc, L;
    int32_t ti, tj, tk, ux, uy, uz;
    uint32_t pos;                  // PLANAR: LDS index of (ti, tj), kept incrementally (uint32 wrap-around)
    uint32_t sxs, sys;             // PLANAR: index steps of an x / y plane crossing;
                                   // otherwise the x / y strides of the walked image (1, Nx) or (Ny, 1)
    bool trans;                    // the walked image is y-fastest (transposed): y + x*Ny + z*Nx*Ny
    uint32_t left;                 // Siddon steps still to take (0 = finished)
    bool perp;
    uint32_t p_ind, p_stride, p_a1, p_a2, p_nrun, p_lim;
    uint32_t p_row, p_rstride;     // first voxel and stride of the whole image row the LOR runs along
    float p_dT1, p_dT2, p_d;
    float tofD;                    // TOF: signed distance of the walk's start from the LOR centre (TOF only)
};

// Set-up of LOR `lo` (projector_functions.h:1421-1840). Returns false when the LOR contributes
// nothing; otherwise fills the walk and the LOR weight.
// PLANAR: the caller guarantees zs == zd (a direct sinogram plane); only the in-plane branches
// are compiled and voxel indices are relative to the LOR's slice (tk * Nx * Ny omitted).
// allow_trans: an x-major LOR (|dx| > |dy| in voxel steps) walks the y-fastest image, where the
// neighbouring radial bins held by the lanes of a wave -- side by side along y for such LORs -- share
// cache lines (w.trans is set; the caller supplies the transposed image / accumulator).
template <bool PLANAR>
__device__ __forceinline__ bool siddon_setup(const SidArgs& g, int64_t lo, Walk& w, float& weight,
                                             bool allow_trans = false, int ray = 0) {
    Ray r;
    // the direct-plane kernels never see several rays per LOR or TOF (such geometries are not planar)
    if (PLANAR) fetch_ray_base(g, lo, r);
    else fetch_ray(g, lo, r, ray);
    const float ydf = r.yd - r.ys, xdf = r.xd - r.xs, zdf = r.zd - r.zs;
    if ((ydf == 0.f && xdf == 0.f) || isinf(ydf) || isinf(xdf)) return false;
    const bool ct = g.mode == MODE_CT_FLAT;
    const float L = sqrtf(xdf * xdf + ydf * ydf + zdf * zdf);
    w.L = L;

    // axis-parallel in-plane LOR (projector_functions.h:1508-1700); thresholds are double 1e-8
    if ((double)fabsf(zdf) < 1e-8 && ((double)fabsf(ydf) < 1e-8 || (double)fabsf(xdf) < 1e-8)) {
        const int32_t tk = (int32_t)(fabsf(r.zs - g.bz) / g.dz);
        if (tk < 0 || (uint32_t)tk >= g.Nz) return false;
        int a1 = 0, a2;
        float dT1, dT2, db, dd, ddb, dd2;
        uint32_t Nm, Nrow, Nrun, stride;
        bool along_x;   // the LOR runs along x (a row of constant y), else along y
        if ((double)fabsf(ydf) < 1e-8 && r.yd <= g.bmaxy && r.yd >= g.by) {
            along_x = true;
            a2 = (int)g.Nx - 1; dT1 = g.dx; dT2 = g.dx; db = g.by; dd = r.yd; ddb = g.dy;
            float e1, e2;
            if (r.xs > r.xd) { e1 = g.bx - r.xd; e2 = g.bx + (float)g.Nx * g.dx - r.xs; }
            else { e1 = g.bx - r.xs; e2 = g.bx + (float)g.Nx * g.dx - r.xd; }
            for (int k = 0; k < (int)g.Nx; k++) {
                if (e1 >= 0.f) { a1 = k; dT1 = (k == 0) ? g.dx : mn(e1, g.dx); break; }
                e1 += g.dx;
            }
            for (int k = (int)g.Nx - 1; k >= a1; k--) {
                if (e2 <= 0.f) { a2 = k; dT2 = (k == (int)g.Nx - 1) ? g.dx : mn(-e2, g.dx); break; }
                e2 -= g.dx;
            }
            dd2 = g.dx;
            // the row start uses Ny as the row stride, as the reference CPU code does (projector_functions.h:1570-1573, :1176);
            // planar (slice-relative) indices require Nx == Ny and use the LDS row pitch
            Nm = PLANAR ? g.pitch : g.Ny; Nrow = g.Ny; Nrun = g.Nx; stride = 1u;
        } else if ((double)fabsf(xdf) < 1e-8 && r.xd <= g.bmaxx && r.xd >= g.bx) {
            along_x = false;
            a2 = (int)g.Ny - 1; dT1 = g.dx; dT2 = g.dx; db = g.bx; dd = r.xd;
            float e1, e2;
            if (r.ys > r.yd) { e1 = g.by - r.yd; e2 = g.by + (float)g.Ny * g.dy - r.ys; }
            else { e1 = g.by - r.ys; e2 = g.by + (float)g.Ny * g.dy - r.yd; }
            for (int k = 0; k < (int)g.Ny; k++) {
                if (e1 >= 0.f) { a1 = k; dT1 = (k == 0) ? g.dy : mn(e1, g.dy); break; }
                e1 += g.dy;
            }
            for (int k = (int)g.Ny - 1; k >= a1; k--) {
                if (e2 <= 0.f) { a2 = k; dT2 = (k == (int)g.Ny - 1) ? g.dy : mn(-e2, g.dy); break; }
                e2 -= g.dy;
            }
            dd2 = g.dy; ddb = g.dx;
            Nm = 1u; Nrow = g.Nx; Nrun = g.Ny; stride = PLANAR ? g.pitch : g.Nx;
        } else {
            return false;
        }
        weight = make_weight<!PLANAR>(g, lo, L, ct, true, ray);
        uint32_t apu = 0;
        {
            float s = db - dd + ddb;
            for (uint32_t i = 0; i < Nrow; i++) {
                if (s > 0.f) { apu = i; break; }
                s += ddb;
            }
        }
        w.perp = true;
        w.trans = false;
        if (!PLANAR) {   // the whole row, for the attenuation line integral (k_atten_factor)
            const uint32_t zoff = (uint32_t)tk * g.Nyx;
            if (along_x) { w.p_row = apu * g.Nx + zoff; w.p_rstride = 1u; }
            else { w.p_row = apu + zoff; w.p_rstride = stride; }
        }
        w.p_ind = apu * Nm + (PLANAR ? 0u : (uint32_t)tk * Nrow * Nrun) + stride * (uint32_t)a1;
        w.p_stride = stride;
        w.p_a1 = (uint32_t)a1; w.p_a2 = (uint32_t)a2; w.p_nrun = Nrun;
        // With Nx != Ny the reference's row stride (Ny) can point past the image; such voxels are dropped.
        w.p_lim = PLANAR ? g.Ny * g.pitch : g.Nyx * g.Nz;
        w.p_dT1 = dT1; w.p_dT2 = dT2; w.p_d = dd2;
        w.left = 0;
        // half the row, signed by the (near-zero) difference across the row (projector_functions.h:1641-1645)
        if (!PLANAR && g.tof_bins > 1) w.tofD = (dd2 * (float)Nrun) / copysignf(2.f, along_x ? ydf : xdf);
        return true;
    }
    w.perp = false;
    int32_t ti = 0, tj = 0, tk = 0, ux = 0, uy = 0, uz = 0;
    float txu = 0.f, tyu = 0.f, tzu = 0.f, tc = 0.f, tx0 = 1e8f, ty0 = 1e8f, tz0 = 1e8f;
    uint32_t Np = 0;
    bool skip;
    if (PLANAR || fabsf(zdf) < 1e-8f) {
        tk = (int32_t)(fabsf(r.zs - g.bz) / g.dz);
        if (tk < 0 || (uint32_t)tk >= g.Nz) return false;
        skip = preloop2(g.bx, g.by, xdf, ydf, g.bmaxx, g.bmaxy, g.dx, g.dy, g.Nx, g.Ny, ti, tj, txu, tyu, Np,
                        r.ys, r.xs, r.yd, r.xd, tc, ux, uy, tx0, ty0);
    } else if (fabsf(ydf) < 1e-8f) {
        tj = perp_start(g.by, r.yd, g.dy, g.Ny);
        skip = preloop2(g.bx, g.bz, xdf, zdf, g.bmaxx, g.bmaxz, g.dx, g.dz, g.Nx, g.Nz, ti, tk, txu, tzu, Np,
                        r.zs, r.xs, r.zd, r.xd, tc, ux, uz, tx0, tz0);
        if (r.yd > g.bmaxy || r.yd < g.by) skip = true;
    } else if (fabsf(xdf) < 1e-8f) {
        ti = perp_start(g.bx, r.xd, g.dx, g.Nx);
        skip = preloop2(g.by, g.bz, ydf, zdf, g.bmaxy, g.bmaxz, g.dy, g.dz, g.Ny, g.Nz, tj, tk, tyu, tzu, Np,
                        r.zs, r.ys, r.zd, r.yd, tc, uy, uz, ty0, tz0);
        if (r.xd > g.bmaxx || r.xd < g.bx) skip = true;
    } else {
        skip = preloop3(g, r, xdf, ydf, zdf, ti, tj, tk, tyu, txu, tzu, Np, tc, ux, uy, uz, tx0, ty0, tz0);
    }
    if (skip) return false;
    weight = make_weight<!PLANAR>(g, lo, L, ct, false, ray);
    if (!PLANAR && g.tof_bins > 1) {   // TOFDis (projector_functions.h:588-594, :1741-1742)
        const float xI = xdf * tc, yI = ydf * tc, zI = zdf * tc;
        w.tofD = sqrtf(xI * xI + yI * yI + zI * zI) - L / 2.f;
    }
    w.tx0 = tx0; w.ty0 = ty0; w.tz0 = tz0; w.txu = txu; w.tyu = tyu; w.tzu = tzu; w.tc = tc;
    w.ti = ti; w.tj = tj; w.tk = tk; w.ux = ux; w.uy = uy; w.uz = uz;
    if (PLANAR) {
        w.trans = false;
        w.pos = (uint32_t)tj * g.pitch + (uint32_t)ti;
        w.sxs = (uint32_t)ux;
        w.sys = (uint32_t)(uy * (int32_t)g.pitch);
    } else {
        w.trans = allow_trans && txu > 0.f && (tyu == 0.f || g.dx * tyu > g.dy * txu);
        w.sxs = w.trans ? g.Ny : 1u;
        w.sys = w.trans ? 1u : g.Nx;
    }
    w.left = Np;
    return true;
}

// One Siddon step (projector_functions.h:1842-1998) without branches: the voxel the ray is in,
// the length inside it (0 when it contributes nothing or the walk is over), then the advance to
// the nearest plane crossing (z, then y, then x). Same float operations as the reference.
// RING: the caller guarantees that every crossing parameter of the walk lies in [0, 1] (both
// endpoints of the LOR outside the image, with margin); the reference's length expression then
// reduces to (tn - tc) * L, the same float result without its range selections.
// STRIDED (not PLANAR): the index uses the walk's per-LOR strides (siddon_setup's allow_trans);
// otherwise the x-fastest image's tj * Nx + ti + tk * Nx*Ny with the grid's uniform strides.
template <bool PLANAR, bool RING = false, bool STRIDED = false>
__device__ __forceinline__ void walk_step(const SidArgs& g, Walk& w, uint32_t& ind, float& len) {
    // bitwise & / | on the predicates keep the step free of short-circuit control flow
    const bool alive = w.left > 0u;
    if (PLANAR) ind = w.pos;
    else if (STRIDED) ind = (uint32_t)w.ti * w.sxs + (uint32_t)w.tj * w.sys + (uint32_t)w.tk * g.Nyx;
    else ind = (uint32_t)w.tj * g.Nx + (uint32_t)w.ti + (uint32_t)w.tk * g.Nyx;
    const bool zmin = !PLANAR && ((w.tz0 < w.ty0) & (w.tz0 < w.tx0));
    const bool ymin = (!zmin) & (w.ty0 < w.tx0);
    const float tn = zmin ? w.tz0 : (ymin ? w.ty0 : w.tx0);
    const float tc = w.tc;
    if (RING) {
        len = alive ? (tn - tc) * w.L : 0.f;
    } else {
        const bool tn_in = (tn >= 0.f) & (tn <= 1.f);
        const bool tc_in = (tc >= 0.f) & (tc <= 1.f);
        // operand of the length product, as chosen by the reference's branches
        const float d_nc = tn - tc, d_1c = 1.f - tc;
        const float o_in = (tc < 0.f) ? tn : d_nc;
        const float o_out = (tn > 1.f) ? d_1c : d_nc;
        const float l = (tn_in ? o_in : o_out) * w.L;
        len = (alive & (tn_in | tc_in)) ? l : 0.f;
    }
    const bool sx = alive & (!zmin) & (!ymin), sy = alive & ymin, sz = alive & zmin;
    w.tc = alive ? tn : tc;
    w.ti += sx ? w.ux : 0;
    w.tx0 = sx ? w.tx0 + w.txu : w.tx0;
    w.tj += sy ? w.uy : 0;
    w.ty0 = sy ? w.ty0 + w.tyu : w.ty0;
    if (PLANAR) w.pos += sx ? w.sxs : (sy ? w.sys : 0u);
    if (!PLANAR) {
        w.tk += sz ? w.uz : 0;
        w.tz0 = sz ? w.tz0 + w.tzu : w.tz0;
    }
    // a negative index wraps to >= N as uint32, so one unsigned compare per axis checks both ends
    bool out = ((uint32_t)w.ti >= g.Nx) | ((uint32_t)w.tj >= g.Ny);
    if (!PLANAR) out = out | ((uint32_t)w.tk >= g.Nz);
    w.left = (alive & !out) ? w.left - 1u : 0u;
}

// ---- time of flight (projector_functions.h:559-594, 632-657) ---------------------------------------
// A voxel's TOF weight for bin c is a 4-point trapezoid of a Gaussian over the voxel's stretch of the
// LOR, from the distance D of its entry point from the LOR centre towards the far end. As in the
// reference, the Gaussian's exponential is taken in double precision (its -0.5 is a double literal)
// and its product with (0.3989422804014327f / sigma) rounded to float.
constexpr uint32_t kTofMax = 64;   // TOF bins a geometry may have

__device__ __forceinline__ float tof_pdf(const SidArgs& g, float x, float mu) {
    const float a = (x - mu) / g.tof_sigma;
    return (float)((double)g.tof_pf * exp(-0.5 * (double)a * (double)a));
}

// TOFWeight (projector_functions.h:565-573) with TRAPZ_BINS = 4
__device__ __forceinline__ float tof_weight(const SidArgs& g, float elem, float D, float DD, float mu, float dX) {
    float out = tof_pdf(g, D, mu);
    const float dXs = copysignf(dX, DD);
    out += tof_pdf(g, D - dXs * 1.f, mu) * 2.f;
    out += tof_pdf(g, D - dXs * 2.f, mu) * 2.f;
    out += tof_pdf(g, D - copysignf(elem, DD), mu);
    return out;
}

// TOF walk of LOR `lo`: visit(voxel, length, D, DD, esum) for each voxel with length > 0 in traversal
// order, D the TOF distance at the voxel and esum the length the reference's weight sum (TOFLoop) uses:
// the length itself on a Siddon walk, the full voxel size on an axis-parallel row (whose end voxels
// have partial lengths, :1664-1673). D moves by each step's length (:1990-1991; rows: the voxel size,
// :1684-1685), also across steps of length 0. Returns false when the LOR contributes nothing.
template <typename Visit>
__device__ __forceinline__ bool siddon_tof_trace(const SidArgs& g, int64_t lo, float& weight, Visit&& visit,
                                                 int ray = 0) {
    Walk w;
    if (!siddon_setup<false>(g, lo, w, weight, false, ray)) return false;
    float D = w.tofD;
    const float DD = D;
    if (w.perp) {
        uint32_t ind = w.p_ind;
        for (uint32_t ii = w.p_a1; ii <= w.p_a2; ii++) {
            float din = w.p_d;
            if (w.p_a1 > 0 && ii == w.p_a1) din = w.p_dT1;
            else if (w.p_a2 < w.p_nrun - 1 && ii == w.p_a2) din = w.p_dT2;
            if (ind < w.p_lim) visit(ind, din, D, DD, w.p_d);
            ind += w.p_stride;
            D -= copysignf(w.p_d, DD);
        }
        return true;
    }
    while (w.left > 0u) {
        uint32_t ind;
        float len;
        walk_step<false>(g, w, ind, len);
        if (len > 0.f) visit(ind, len, D, DD, len);
        D -= copysignf(len, DD);
    }
    return true;
}

// Axis-parallel walk: visit(voxel, length) for each voxel in order.
template <typename Visit>
__device__ __forceinline__ void perp_walk(const Walk& w, Visit&& visit) {
    uint32_t ind = w.p_ind;
    for (uint32_t ii = w.p_a1; ii <= w.p_a2; ii++) {
        float din = w.p_d;
        if (w.p_a1 > 0 && ii == w.p_a1) din = w.p_dT1;
        else if (w.p_a2 < w.p_nrun - 1 && ii == w.p_a2) din = w.p_dT2;
        if (ind < w.p_lim) visit(ind, din);
        ind += w.p_stride;
    }
}

// Trace LOR `lo`: visit(voxel, length) for each voxel with length > 0 in traversal order.
// Returns false when the LOR contributes nothing; `weight` is the LOR weight.
template <bool PLANAR = false, typename Visit>
__device__ __forceinline__ bool siddon_trace(const SidArgs& g, int64_t lo, float& weight, Visit&& visit, int ray = 0) {
    Walk w;
    if (!siddon_setup<PLANAR>(g, lo, w, weight, false, ray)) return false;
    if (w.perp) {
        perp_walk(w, visit);
        return true;
    }
    while (w.left > 0u) {
        uint32_t ind;
        float len;
        walk_step<PLANAR>(g, w, ind, len);
        if (len > 0.f) visit(ind, len);
    }
    return true;
}

// Forward-projection gather of LOR `lo`: sum of len * src[voxel] in traversal order (the
// reference's accumulation order), with kBatch steps traced before their loads are consumed so
// kBatch gathers are in flight at once. Returns false when the LOR contributes nothing.
constexpr int kBatch = 8;

template <bool PLANAR, bool RING = false>
__device__ __forceinline__ bool siddon_gather(const SidArgs& g, int64_t lo, const float* __restrict__ src, float& ax,
                                              float& weight, int ray = 0) {
    ax = 0.f;
    Walk w;
    if (!siddon_setup<PLANAR>(g, lo, w, weight, false, ray)) return false;
    if (w.perp) {
        perp_walk(w, [&](uint32_t ind, float len) { ax += len * src[ind]; });
        return true;
    }
    while (w.left > 0u) {
        uint32_t ind[kBatch];
        float len[kBatch], v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) walk_step<PLANAR, RING>(g, w, ind[k], len[k]);
#pragma unroll
        for (int k = 0; k < kBatch; k++) v[k] = src[len[k] > 0.f ? ind[k] : 0u];
#pragma unroll
        for (int k = 0; k < kBatch; k++) ax = (len[k] > 0.f) ? ax + len[k] * v[k] : ax;
    }
    return true;
}

// Backprojection walk of LOR `lo`: visit(voxel, length) for each voxel with length > 0, kBatch steps
// traced per loop trip (one loop test per kBatch steps). Returns false when the LOR contributes nothing.
template <bool PLANAR, bool RING = false, typename Visit>
__device__ __forceinline__ bool siddon_scatter(const SidArgs& g, int64_t lo, float& weight, Visit&& visit) {
    Walk w;
    if (!siddon_setup<PLANAR>(g, lo, w, weight)) return false;
    if (w.perp) {
        perp_walk(w, visit);
        return true;
    }
    while (w.left > 0u) {
        uint32_t ind[kBatch];
        float len[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) walk_step<PLANAR, RING>(g, w, ind[k], len[k]);
#pragma unroll
        for (int k = 0; k < kBatch; k++)
            if (len[k] > 0.f) visit(ind[k], len[k]);
    }
    return true;
}

// Backprojection walk that lets x-major LORs accumulate into a y-fastest (transposed) image:
// visit(ind, len, trans) for each voxel with len > 0, ind in the layout named by trans (see
// siddon_setup's allow_trans). Returns false when the LOR contributes nothing.
template <typename Visit>
__device__ __forceinline__ bool siddon_trace_split(const SidArgs& g, int64_t lo, float& weight, Visit&& visit,
                                                   int ray = 0) {
    Walk w;
    if (!siddon_setup<false>(g, lo, w, weight, true, ray)) return false;
    if (w.perp) {
        perp_walk(w, [&](uint32_t ind, float len) { visit(ind, len, false); });
        return true;
    }
    const bool t = w.trans;
    while (w.left > 0u) {
        uint32_t ind;
        float len;
        walk_step<false, false, true>(g, w, ind, len);
        if (len > 0.f) visit(ind, len, t);
    }
    return true;
}

// ---- fixed-point LDS accumulation of the direct-plane (planar) kernels ----------------------------
// A bound B >= 0 satisfies B < 2^(exp_bound(B)) (finite B; exp_bound(0) = -126).
__device__ __forceinline__ int32_t exp_bound(float B) {
    return (int32_t)((__float_as_uint(B) >> 23) & 0xffu) - 126;
}

// Scale 2^F of a workgroup's fixed-point sums and its inverse: every term |t| <= Bt maps below 2^50
// (the conversion in lds_add_fixed needs |t 2^F| < 2^51) and every sum |s| <= Bs below 2^62.
__device__ __forceinline__ void fixed_scale(float Bt, float Bs, double& up, double& down) {
    const int32_t F = min(50 - exp_bound(Bt), 62 - exp_bound(Bs));   // in [-79, 176]
    up = __longlong_as_double((long long)(1023 + F) << 52);
    down = __longlong_as_double((long long)(1023 - F) << 52);
}

// Maximum of v over the workgroup (red: TPB / 64 floats of LDS). Contains __syncthreads.
template <int TPB>
__device__ __forceinline__ float block_max(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float m = red[0];
    for (int i = 1; i < TPB / 64; i++) m = fmaxf(m, red[i]);
    return m;
}

// Adds round(t 2^F) to a 64-bit LDS accumulator (ds_add_u64). The double fma t * 2^F + 1.5 * 2^52
// rounds t 2^F to an integer k (|k| < 2^51 keeps the exponent at 52), so the bits of the result
// less those of 1.5 * 2^52 are k in two's complement: two VALU operations and an integer subtract
// instead of a float -> int64 conversion sequence.
__device__ __forceinline__ void lds_add_fixed(unsigned long long* p, float t, double up) {
    const double q = __builtin_fma((double)t, up, 0x1.8p52);
    atomicAdd(p, (unsigned long long)(__double_as_longlong(q) - 0x4338000000000000ll));
}

// The float value of a fixed-point sum.
__device__ __forceinline__ float fixed_value(unsigned long long v, double down) {
    return (float)((double)(long long)v * down);
}

}  // namespace omg
