// omega_oracle.cpp -- CPU ORACLE (test infrastructure only).
//
// A plain C++17/OpenMP restatement of OMEGA's CPU reconstruction path for the
// projector hot path. It exists to CHECK the HIP library and to give the timed
// CPU baseline; it is never linked into, loaded by, or called from the product
// library (omega_amd/). Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it.
//
// What each part restates (reference file:line, villekf/OMEGA v2.3.0):
//   siddon_lor<T>        improved Siddon, projector type 1, nRays=1, no TOF:
//                        projectorType123Implementation4 source/cpp/projector_functions.h:1311-2063
//                        (index decode :1355-1408, detector fetch :357-498,
//                         perpendicular branch :1508-1700 with perpendicular_elements :1166-1222,
//                         pre-loops siddon_pre_loop_2D/3D :788-970 with d_g_s_precomp/s_g_d_precomp :723-778,
//                         traversal :1842-1998, LOR weight temp :1817-1836)
//   rdp_gradient<T>      RDPKernel source/cpp/RDP.h:4-78 (zero outside the volume)
//   ratio / em / osl     computeForwardStep source/cpp/computeForwardProject.h:20-21,94,400-401;
//                        EM algorithms.h:12-15; OSL-OSEM subiterStep.h:219-231; OSEM subiterStep.h:111-121;
//                        positivity reconstructionAF.h:1461-1465; sensitivity clamp functions.hpp:2981-2996
//
// Pinning (what checks this oracle, beyond retyping it):
//   * golden vectors written by the reference headers compiled as oracle/_ref
//     (tests/golden/make_golden.py) -- bit-exact Siddon FP values and index sets;
//   * brute-force double-precision plane-crossing Siddon on tiny grids (tests);
//   * closed forms: sum of lengths == analytic chord of the ray through the box.
//
// Arithmetic follows the reference expression by expression (same operation
// order, no FMA contraction: build with -ffp-contract=off) so float results are
// bit-identical to the reference's CPU build.
#include <cstdint>
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// Projection data modes.
enum Mode : int32_t { PET_SINO = 0, CT_FLAT = 1, LIST6 = 2 };

template <typename T>
struct Geo {
    uint32_t Nx, Ny, Nz;
    T dx, dy, dz, bx, by, bz;
    uint32_t size_x, size_y;   // detector rows (radial bins) and columns (angles)
    int32_t mode;
    int64_t plane_offset;      // first plane/projection of the subset in the (subset-ordered) z table
    T dPitchZ;                 // CT: axial detector pitch (used when ct_pitch == 0)
    int32_t ct_pitch;          // CT: z table holds 6 floats per projection
    T global_factor;
    const T* norm;             // optional per-measurement normalisation (subset-local), may be null
};

// float -> uint32 the way g++/x86-64 converts (cvttss2si to 64 bit, keep the low word).
template <typename T>
static inline uint32_t to_u32(T v) { return (uint32_t)(int64_t)v; }

template <typename T>
struct Ray { T xs, ys, zs, xd, yd, zd; };

template <typename T>
static inline void fetch_ray(const Geo<T>& g, int64_t lo, const T* x, const T* z, Ray<T>& r) {
    if (g.mode == LIST6) {
        r.xs = x[lo * 6]; r.ys = x[lo * 6 + 1]; r.zs = x[lo * 6 + 2];
        r.xd = x[lo * 6 + 3]; r.yd = x[lo * 6 + 4]; r.zd = x[lo * 6 + 5];
        return;
    }
    const int64_t per = (int64_t)g.size_x * (int64_t)g.size_y;
    int64_t iz = lo / per;
    const int64_t rem = lo - iz * per;
    const int64_t ix = rem % g.size_x;
    const int64_t iy = rem / g.size_x;
    iz += g.plane_offset;
    if (g.mode == PET_SINO) {
        const int64_t id = (ix + iy * (int64_t)g.size_x) * 4;
        r.xs = x[id]; r.ys = x[id + 1]; r.xd = x[id + 2]; r.yd = x[id + 3];
        r.zs = z[iz * 2]; r.zd = z[iz * 2 + 1];
    } else {
        const int64_t id = iz * 6;
        r.xs = x[id]; r.ys = x[id + 1]; r.zs = x[id + 2];
        r.xd = x[id + 3]; r.yd = x[id + 4]; r.zd = x[id + 5];
        // detector-pixel offsets, computed in double then rounded once (projector_functions.h:480-481)
        const T i1 = (T)((double)(T)ix - (double)(T)g.size_x / 2. + .5);
        const T i2 = (T)((double)(T)iy - (double)(T)g.size_y / 2. + .5);
        if (g.ct_pitch) {
            r.xd += z[id] * i1 + z[id + 3] * i2;
            r.yd += z[id + 1] * i1 + z[id + 4] * i2;
            r.zd += z[id + 2] * i1 + z[id + 5] * i2;
        } else {
            r.xd += z[iz * 2] * i1;
            r.yd += z[iz * 2 + 1] * i1;
            r.zd += g.dPitchZ * i2;
        }
    }
}

// Entry/exit voxel bookkeeping for one axis, source-to-detector increasing (projector_functions.h:723-749).
template <typename T>
static inline void axis_increasing(T tc, T tmin_axis, T tmax, T tmax_axis, uint32_t& vmin, uint32_t& vmax,
                                   T& t0, int32_t& step, T diff, T b, T d, T s, uint32_t N) {
    if (tc == tmin_axis) vmin = 1u;
    else { const T p = s + tc * diff; vmin = to_u32(std::ceil((p - b) / d)); }
    if (tmax == tmax_axis) vmax = N;
    else { const T p = s + tmax * diff; vmax = to_u32(std::floor((p - b) / d)); }
    t0 += (T)vmin * d / diff;
    step = 1;
}

// Same, source greater than detector along the axis (projector_functions.h:752-778).
template <typename T>
static inline void axis_decreasing(T tc, T tmin_axis, T tmax, T tmax_axis, uint32_t& vmin, uint32_t& vmax,
                                   T& t0, int32_t& step, T diff, T b, T d, T s, uint32_t N) {
    if (tc == tmin_axis) vmax = N - 1u;
    else { const T p = s + tc * diff; vmax = to_u32(std::floor((p - b) / d)); }
    if (tmax == tmax_axis) vmin = 0u;
    else { const T p = s + tmax * diff; vmin = to_u32(std::ceil((p - b) / d)); }
    t0 += (T)vmax * d / diff;
    step = -1;
}

template <typename T>
static inline int32_t first_voxel(T pt, T diff, T d, T a) { return (int32_t)std::floor((pt * diff - a) / d); }

// 2-D pre-loop over axes (1,2) (projector_functions.h:788-870). Returns true when the ray misses.
template <typename T>
static bool preloop2(T b1, T b2, T diff1, T diff2, T max1, T max2, T d1, T d2, uint32_t N1, uint32_t N2,
                     int32_t& i1, int32_t& i2, T& t1u, T& t2u, uint32_t& Np, T s2, T s1, T e2, T e1,
                     T& tc, int32_t& u1, int32_t& u2, T& t10, T& t20) {
    const T a1 = b1 - s1, a2 = b2 - s2;
    t10 = a1 / diff1;
    t20 = a2 / diff2;
    const T t1b = (max1 - s1) / diff1, t2b = (max2 - s2) / diff2;
    const T t1min = std::min(t10, t1b), t1max = std::max(t10, t1b);
    const T t2min = std::min(t20, t2b), t2max = std::max(t20, t2b);
    tc = std::max(t1min, t2min);
    const T tmax = std::min(t1max, t2max);
    if (tc >= tmax) return true;
    uint32_t v1min = 0, v1max = 0, v2min = 0, v2max = 0;
    if (s1 < e1) axis_increasing(tc, t1min, tmax, t1max, v1min, v1max, t10, u1, diff1, b1, d1, s1, N1);
    else axis_decreasing(tc, t1min, tmax, t1max, v1min, v1max, t10, u1, diff1, b1, d1, s1, N1);
    if (s2 < e2) axis_increasing(tc, t2min, tmax, t2max, v2min, v2max, t20, u2, diff2, b2, d2, s2, N2);
    else axis_decreasing(tc, t2min, tmax, t2max, v2min, v2max, t20, u2, diff2, b2, d2, s2, N2);
    Np = v1max + 1u + v2max + 1u - v1min - v2min;
    const T pt = (T)((double)(std::min(t10, t20) + tc) / 2.);
    i1 = first_voxel(pt, diff1, d1, a1);
    i2 = first_voxel(pt, diff2, d2, a2);
    if (i1 < 0 || (uint32_t)i1 >= N1 || i2 < 0 || (uint32_t)i2 >= N2) return true;
    t1u = d1 / std::fabs(diff1);
    t2u = d2 / std::fabs(diff2);
    return false;
}

// 3-D pre-loop (projector_functions.h:874-970).
template <typename T>
static bool preloop3(const Geo<T>& g, const Ray<T>& r, T xdf, T ydf, T zdf, T mx, T my, T mz,
                     int32_t& ti, int32_t& tj, int32_t& tk, T& tyu, T& txu, T& tzu, uint32_t& Np,
                     T& tc, int32_t& iu, int32_t& ju, int32_t& ku, T& tx0, T& ty0, T& tz0) {
    const T ax = g.bx - r.xs, ay = g.by - r.ys, az = g.bz - r.zs;
    tx0 = ax / xdf; ty0 = ay / ydf; tz0 = az / zdf;
    const T txb = (mx - r.xs) / xdf, tyb = (my - r.ys) / ydf, tzb = (mz - r.zs) / zdf;
    const T txmin = std::min(tx0, txb), txmax = std::max(tx0, txb);
    const T tymin = std::min(ty0, tyb), tymax = std::max(ty0, tyb);
    const T tzmin = std::min(tz0, tzb), tzmax = std::max(tz0, tzb);
    tc = std::max(std::max(txmin, tzmin), tymin);
    const T tmax = std::min(std::min(txmax, tzmax), tymax);
    if (tc >= tmax) return true;
    uint32_t imin = 0, imax = 0, jmin = 0, jmax = 0, kmin = 0, kmax = 0;
    if (r.xs < r.xd) axis_increasing(tc, txmin, tmax, txmax, imin, imax, tx0, iu, xdf, g.bx, g.dx, r.xs, g.Nx);
    else axis_decreasing(tc, txmin, tmax, txmax, imin, imax, tx0, iu, xdf, g.bx, g.dx, r.xs, g.Nx);
    if (r.ys < r.yd) axis_increasing(tc, tymin, tmax, tymax, jmin, jmax, ty0, ju, ydf, g.by, g.dy, r.ys, g.Ny);
    else axis_decreasing(tc, tymin, tmax, tymax, jmin, jmax, ty0, ju, ydf, g.by, g.dy, r.ys, g.Ny);
    if (r.zs < r.zd) axis_increasing(tc, tzmin, tmax, tzmax, kmin, kmax, tz0, ku, zdf, g.bz, g.dz, r.zs, g.Nz);
    else axis_decreasing(tc, tzmin, tmax, tzmax, kmin, kmax, tz0, ku, zdf, g.bz, g.dz, r.zs, g.Nz);
    Np = (kmax - kmin + 1) + (jmax - jmin + 1) + (imax - imin + 1);
    const T pt = (T)((double)(std::min(std::min(tz0, ty0), tx0) + tc) / 2.);
    ti = first_voxel(pt, xdf, g.dx, ax);
    tj = first_voxel(pt, ydf, g.dy, ay);
    tk = first_voxel(pt, zdf, g.dz, az);
    if (ti < 0 || (uint32_t)ti >= g.Nx || tj < 0 || (uint32_t)tj >= g.Ny || tk < 0 || (uint32_t)tk >= g.Nz) return true;
    txu = g.dx / std::fabs(xdf);
    tyu = g.dy / std::fabs(ydf);
    tzu = g.dz / std::fabs(zdf);
    return false;
}

// First voxel index whose far edge passes the coordinate (projector_functions.h:670-681).
template <typename T>
static inline int perp_start(T b, T d, T dd, uint32_t N) {
    int idx = 0;
    T s = b - d + dd;
    for (uint32_t i = 0; i < N; i++) {
        if (s > (T)0.) { idx = (int)i; break; }
        s += dd;
    }
    return idx;
}

// Traces LOR `lo` and calls visit(voxel_index, length) for every voxel with length > 0, in
// traversal order. Returns false if the LOR is skipped (no contribution). `w` receives the
// LOR weight (1/L * norm * global_factor for PET, 1 for CT).
template <typename T, typename F>
static bool siddon_lor(const Geo<T>& g, int64_t lo, const T* x, const T* z, T& w, F&& visit) {
    Ray<T> r;
    fetch_ray(g, lo, x, z, r);
    T ydf = r.yd - r.ys, xdf = r.xd - r.xs, zdf = r.zd - r.zs;
    if ((ydf == (T)0. && xdf == (T)0. && zdf == (T)0.) || (ydf == (T)0. && xdf == (T)0.) || std::isinf(ydf) || std::isinf(xdf))
        return false;
    const bool ct = g.mode == CT_FLAT;
    const uint32_t Nyx = g.Ny * g.Nx;
    const T bmaxz = g.bz + (T)g.Nz * g.dz;
    const T bmaxy = (T)g.Ny * g.dy + g.by;
    const T bmaxx = (T)g.Nx * g.dx + g.bx;
    const T L = std::sqrt(xdf * xdf + ydf * ydf + zdf * zdf);
    const T lnorm = g.norm ? g.norm[lo] : (T)0.;

    if ((double)std::fabs(zdf) < 1e-8 && ((double)std::fabs(ydf) < 1e-8 || (double)std::fabs(xdf) < 1e-8)) {
        // Axis-parallel in-plane LOR: straight row/column walk with partial end voxels.
        int32_t tk = (int32_t)(std::fabs(r.zs - g.bz) / g.dz);
        if (tk < 0 || (uint32_t)tk >= g.Nz) return false;
        int a1 = 0, a2 = 0;
        T dT1, dT2, db, dd, ddb, dd2;
        uint32_t N, Nrow, Nrun, stride;
        if ((double)std::fabs(ydf) < 1e-8 && r.yd <= bmaxy && r.yd >= g.by) {
            a1 = 0; a2 = (int)g.Nx - 1; dT1 = g.dx; dT2 = g.dx;
            db = g.by; dd = r.yd; ddb = g.dy;
            T e1, e2;
            if (r.xs > r.xd) { e1 = g.bx - r.xd; e2 = g.bx + (T)g.Nx * g.dx - r.xs; }
            else { e1 = g.bx - r.xs; e2 = g.bx + (T)g.Nx * g.dx - r.xd; }
            for (int k = 0; k < (int)g.Nx; k++) {
                if (e1 >= (T)0.) { a1 = k; dT1 = (k == 0) ? g.dx : std::min(e1, g.dx); break; }
                e1 += g.dx;
            }
            for (int k = (int)g.Nx - 1; k >= a1; k--) {
                if (e2 <= (T)0.) { a2 = k; dT2 = (k == (int)g.Nx - 1) ? g.dx : std::min(-e2, g.dx); break; }
                e2 -= g.dx;
            }
            dd2 = g.dx;
            // row index is multiplied by Ny here, as the reference does (projector_functions.h:1570-1573 -> :1176)
            N = g.Ny; Nrow = g.Ny; Nrun = g.Nx; stride = 1u;
        } else if ((double)std::fabs(xdf) < 1e-8 && r.xd <= bmaxx && r.xd >= g.bx) {
            a1 = 0; a2 = (int)g.Ny - 1; dT1 = g.dx; dT2 = g.dx;
            db = g.bx; dd = r.xd;
            T e1, e2;
            if (r.ys > r.yd) { e1 = g.by - r.yd; e2 = g.by + (T)g.Ny * g.dy - r.ys; }
            else { e1 = g.by - r.ys; e2 = g.by + (T)g.Ny * g.dy - r.yd; }
            for (int k = 0; k < (int)g.Ny; k++) {
                if (e1 >= (T)0.) { a1 = k; dT1 = (k == 0) ? g.dy : std::min(e1, g.dy); break; }
                e1 += g.dy;
            }
            for (int k = (int)g.Ny - 1; k >= a1; k--) {
                if (e2 <= (T)0.) { a2 = k; dT2 = (k == (int)g.Ny - 1) ? g.dy : std::min(-e2, g.dy); break; }
                e2 -= g.dy;
            }
            dd2 = g.dy; ddb = g.dx;
            N = 1u; Nrow = g.Nx; Nrun = g.Ny; stride = g.Nx;
        } else {
            return false;
        }
        // perpendicular_elements (projector_functions.h:1166-1222)
        uint32_t apu = 0;
        {
            T s = db - dd + ddb;
            for (int32_t i = 0; (uint32_t)i < Nrow; i++) {
                if (s > (T)0.) { apu = (uint32_t)i; break; }
                s += ddb;
            }
        }
        uint32_t ind = apu * N + (uint32_t)tk * Nrow * Nrun;
        if (!ct) {
            T t = (T)1. / L;
            if (g.norm) t *= lnorm;
            t *= g.global_factor;
            w = t;
        } else {
            w = (T)1.;
        }
        for (uint32_t ii = (uint32_t)a1; ii <= (uint32_t)a2; ii++) {
            T din = dd2;
            if (ii == (uint32_t)a1) ind += stride * ii;
            if (a1 > 0 && ii == (uint32_t)a1) din = dT1;
            else if ((uint32_t)a2 < Nrun - 1 && ii == (uint32_t)a2) din = dT2;
            // with Nx != Ny the reference's row stride can index past the image (undefined there); dropped here
            if (ind < Nyx * g.Nz) visit(ind, din);
            ind += stride;
        }
        return true;
    }

    int32_t ti = 0, tj = 0, tk = 0, ux = 0, uy = 0, uz = 0;
    T txu = 0, tyu = 0, tzu = 0, tc = 0, tx0 = (T)1e8, ty0 = (T)1e8, tz0 = (T)1e8;
    uint32_t Np = 0;
    bool skip;
    // the general branch compares against a float threshold (T)1e-8 (projector_functions.h:1708,1716,1724)
    if (std::fabs(zdf) < (T)1e-8) {
        tk = (int32_t)(std::fabs(r.zs - g.bz) / g.dz);
        if (tk < 0 || (uint32_t)tk >= g.Nz) return false;
        skip = preloop2(g.bx, g.by, xdf, ydf, bmaxx, bmaxy, g.dx, g.dy, g.Nx, g.Ny, ti, tj, txu, tyu, Np,
                        r.ys, r.xs, r.yd, r.xd, tc, ux, uy, tx0, ty0);
    } else if (std::fabs(ydf) < (T)1e-8) {
        tj = perp_start(g.by, r.yd, g.dy, g.Ny);
        skip = preloop2(g.bx, g.bz, xdf, zdf, bmaxx, bmaxz, g.dx, g.dz, g.Nx, g.Nz, ti, tk, txu, tzu, Np,
                        r.zs, r.xs, r.zd, r.xd, tc, ux, uz, tx0, tz0);
        if (r.yd > bmaxy || r.yd < g.by) skip = true;
    } else if (std::fabs(xdf) < (T)1e-8) {
        ti = perp_start(g.bx, r.xd, g.dx, g.Nx);
        skip = preloop2(g.by, g.bz, ydf, zdf, bmaxy, bmaxz, g.dy, g.dz, g.Ny, g.Nz, tj, tk, tyu, tzu, Np,
                        r.zs, r.ys, r.zd, r.yd, tc, uy, uz, ty0, tz0);
        if (r.xd > bmaxx || r.xd < g.bx) skip = true;
    } else {
        skip = preloop3(g, r, xdf, ydf, zdf, bmaxx, bmaxy, bmaxz, ti, tj, tk, tyu, txu, tzu, Np, tc, ux, uy, uz, tx0, ty0, tz0);
    }
    if (skip) return false;
    if (!ct) {
        T t = (T)1. / (L * (T)1);
        if (g.norm) t *= lnorm;
        t *= g.global_factor;
        w = t;
    } else {
        w = (T)1.;
    }
    for (uint32_t ii = 0; ii < Np; ii++) {
        const uint32_t ind = (uint32_t)tj * g.Nx + (uint32_t)ti + (uint32_t)tk * Nyx;
        // pick the nearest plane crossing: z, then y, then x (projector_functions.h:1857-1922)
        int axis;
        T tn;
        if (tz0 < ty0 && tz0 < tx0) { axis = 2; tn = tz0; }
        else if (ty0 < tx0) { axis = 1; tn = ty0; }
        else { axis = 0; tn = tx0; }
        T len = (T)0.;
        if (tn >= (T)0. && tn <= (T)1.) len = (tc < (T)0.) ? tn * L : (tn - tc) * L;
        else if (tc >= (T)0. && tc <= (T)1.) len = (tn > (T)1.) ? ((T)1. - tc) * L : (tn - tc) * L;
        tc = tn;
        if (axis == 2) { tk += uz; tz0 += tzu; }
        else if (axis == 1) { tj += uy; ty0 += tyu; }
        else { ti += ux; tx0 += txu; }
        if (len > (T)0.) visit(ind, len);
        if (ti < 0 || (uint32_t)ti >= g.Nx || tj < 0 || (uint32_t)tj >= g.Ny || tk < 0 || (uint32_t)tk >= g.Nz) break;
    }
    return true;
}

static inline void set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// y[lo] = w * sum(len * img[ind]) for lo in [0, nMeas)
template <typename T>
static void siddon_fp(const Geo<T>& g, int64_t nMeas, const T* x, const T* z, const T* img, T* y, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t lo = 0; lo < nMeas; lo++) {
        T ax = (T)0.;
        T w = (T)0.;
        const bool hit = siddon_lor(g, lo, x, z, w, [&](uint32_t ind, T len) { ax += len * img[ind]; });
        y[lo] = hit ? ((g.mode == CT_FLAT) ? ax : ax * w) : (T)0.;
    }
}

// rhs[v] += sum_i (len*w) * yin[i]; sens[v] += sum_i len*w (sens may be null)
template <typename T>
static void siddon_bp(const Geo<T>& g, int64_t nMeas, const T* x, const T* z, const T* yin, T* rhs, T* sens, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t lo = 0; lo < nMeas; lo++) {
        const T yi = yin[lo];
        T w = (T)0.;
        siddon_lor(g, lo, x, z, w, [&](uint32_t ind, T len) {
            const T lw = len * w;
            const T c = yi * lw;
#pragma omp atomic
            rhs[ind] += c;
            if (sens) {
#pragma omp atomic
                sens[ind] += lw;
            }
        });
    }
}

// RDPKernel restatement (RDP.h:4-78): grad[n] += beta * (sum over 6 neighbours, zero outside).
template <typename T>
static void rdp_gradient(T* grad, const T* u, int Nx, int Ny, int Nz, T gamma, T epps, T beta, int nthreads) {
    set_threads(nthreads);
    const int64_t Nxy = (int64_t)Nx * Ny, N = Nxy * Nz;
#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t n = 0; n < N; n++) {
        const int64_t zz = n / Nxy, yy = (n - zz * Nxy) / Nx, xx = n - zz * Nxy - yy * Nx;
        const T uj = u[n];
        T nb[6] = {0, 0, 0, 0, 0, 0};
        if (xx < Nx - 1) nb[0] = u[n + 1];
        if (xx > 0) nb[1] = u[n - 1];
        if (yy < Ny - 1) nb[2] = u[n + Nx];
        if (yy > 0) nb[3] = u[n - Nx];
        if (zz < Nz - 1) nb[4] = u[n + Nxy];
        if (zz > 0) nb[5] = u[n - Nxy];
        // Two lanes (+ neighbour, - neighbour) summed over axes x, y, z in that order, as float2 arithmetic does.
        T lane[2];
        for (int s = 0; s < 2; s++) {
            T acc = (T)0.;
            for (int a = 0; a < 3; a++) {
                const T uk = nb[2 * a + s];
                const T du = uj - uk;
                const T den = uj + uk + gamma * std::fabs(du);
                const T term = du * (gamma * std::fabs(du) + uj + (T)3.f * uk + epps * epps) / (den * den + epps);
                acc = (a == 0) ? term : acc + term;
            }
            lane[s] = std::isnan(acc) ? (T)0. : acc;
        }
        grad[n] += beta * (lane[0] + lane[1]);
    }
}

}  // namespace oracle

using namespace oracle;

template <typename T>
static Geo<T> make_geo(const uint32_t* n3, const T* d3, const T* b3, uint32_t size_x, uint32_t size_y, int32_t mode,
                       int64_t plane_offset, T dPitchZ, int32_t ct_pitch, T global_factor, const T* norm) {
    Geo<T> g;
    g.Nx = n3[0]; g.Ny = n3[1]; g.Nz = n3[2];
    g.dx = d3[0]; g.dy = d3[1]; g.dz = d3[2];
    g.bx = b3[0]; g.by = b3[1]; g.bz = b3[2];
    g.size_x = size_x; g.size_y = size_y; g.mode = mode; g.plane_offset = plane_offset;
    g.dPitchZ = dPitchZ; g.ct_pitch = ct_pitch; g.global_factor = global_factor; g.norm = norm;
    return g;
}

#define ORACLE_SIDDON_API(T, SUF)                                                                                     \
    extern "C" int oracle_siddon_fp_##SUF(const uint32_t* n3, const T* d3, const T* b3, uint32_t size_x,             \
                                          uint32_t size_y, int32_t mode, int64_t plane_offset, T dPitchZ,            \
                                          int32_t ct_pitch, T global_factor, const T* norm, int64_t nMeas,           \
                                          const T* x, const T* z, const T* img, T* y, int nthreads) {                \
        Geo<T> g = make_geo<T>(n3, d3, b3, size_x, size_y, mode, plane_offset, dPitchZ, ct_pitch, global_factor, norm); \
        siddon_fp(g, nMeas, x, z, img, y, nthreads);                                                                   \
        return 0;                                                                                                      \
    }                                                                                                                  \
    extern "C" int oracle_siddon_bp_##SUF(const uint32_t* n3, const T* d3, const T* b3, uint32_t size_x,             \
                                          uint32_t size_y, int32_t mode, int64_t plane_offset, T dPitchZ,            \
                                          int32_t ct_pitch, T global_factor, const T* norm, int64_t nMeas,           \
                                          const T* x, const T* z, const T* yin, T* rhs, T* sens, int nthreads) {      \
        Geo<T> g = make_geo<T>(n3, d3, b3, size_x, size_y, mode, plane_offset, dPitchZ, ct_pitch, global_factor, norm); \
        siddon_bp(g, nMeas, x, z, yin, rhs, sens, nthreads);                                                          \
        return 0;                                                                                                      \
    }                                                                                                                  \
    extern "C" int64_t oracle_siddon_trace_##SUF(const uint32_t* n3, const T* d3, const T* b3, uint32_t size_x,      \
                                                 uint32_t size_y, int32_t mode, int64_t plane_offset, T dPitchZ,     \
                                                 int32_t ct_pitch, T global_factor, const T* norm, int64_t nMeas,    \
                                                 const T* x, const T* z, int64_t* counts, int64_t cap,               \
                                                 uint32_t* idx_out, T* len_out, T* w_out) {                          \
        Geo<T> g = make_geo<T>(n3, d3, b3, size_x, size_y, mode, plane_offset, dPitchZ, ct_pitch, global_factor, norm); \
        int64_t total = 0;                                                                                             \
        for (int64_t lo = 0; lo < nMeas; lo++) {                                                                       \
            int64_t c = 0;                                                                                             \
            T w = (T)0.;                                                                                               \
            siddon_lor(g, lo, x, z, w, [&](uint32_t ind, T len) {                                                      \
                if (idx_out && total < cap) { idx_out[total] = ind; len_out[total] = len; }                            \
                total++; c++;                                                                                          \
            });                                                                                                        \
            if (counts) counts[lo] = c;                                                                                \
            if (w_out) w_out[lo] = w;                                                                                  \
        }                                                                                                              \
        return total;                                                                                                  \
    }                                                                                                                  \
    extern "C" int oracle_rdp_gradient_##SUF(T* grad, const T* u, int Nx, int Ny, int Nz, T gamma, T epps, T beta,   \
                                             int nthreads) {                                                           \
        rdp_gradient(grad, u, Nx, Ny, Nz, gamma, epps, beta, nthreads);                                               \
        return 0;                                                                                                      \
    }

ORACLE_SIDDON_API(float, f32)
ORACLE_SIDDON_API(double, f64)

// Elementwise OSEM pieces, FP32, in the reference's operation order.
// ratio: yhat += epps (non-CT); r = y / yhat; NaN/Inf -> epps  (computeForwardProject.h:20-21,94,400-401)
extern "C" void oracle_ratio_f32(const float* y, float* yhat_to_r, int64_t n, float epps) {
    for (int64_t i = 0; i < n; i++) {
        float a = yhat_to_r[i] + epps;
        float r = y[i] / a;
        if (std::isnan(r) || std::isinf(r)) r = epps;
        yhat_to_r[i] = r;
    }
}
// sensitivity clamp: S < epps -> epps (functions.hpp:2996)
extern "C" void oracle_sens_clamp_f32(float* s, int64_t n, float epps) {
    for (int64_t i = 0; i < n; i++) if (s[i] < epps) s[i] = epps;
}
// OSEM: x = x / S * (rhs + epps) (subiterStep.h:120, algorithms.h:14); then positivity x<epps -> epps
extern "C" void oracle_osem_update_f32(float* x, const float* S, const float* rhs, int64_t n, float epps, int positivity) {
    for (int64_t i = 0; i < n; i++) {
        float v = x[i] / S[i] * (rhs[i] + epps);
        if (positivity && v < epps) v = epps;
        x[i] = v;
    }
}
// OSL-OSEM: x = x / (S + dU) * rhs (subiterStep.h:223-226); then positivity
extern "C" void oracle_osl_update_f32(float* x, const float* S, const float* dU, const float* rhs, int64_t n, float epps, int positivity) {
    for (int64_t i = 0; i < n; i++) {
        float v = x[i] / (S[i] + dU[i]) * rhs[i];
        if (positivity && v < epps) v = epps;
        x[i] = v;
    }
}
