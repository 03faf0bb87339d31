// orth.hip -- orthogonal distance-based (projector type 2) and volume-of-intersection (type 3) projectors
// for gfx950.
//
// Semantics are those of the reference's GPU kernel, projectorType123.cl built with -DORTH and
// -DCRYSTXY / -DCRYSTZ (no TOF, masks or multi-ray), because the reference's CPU version
// of this projector is one-sided (projector_functions.h:1121-1123,1145-1147):
//   * a Siddon walk (GPU set-up: siddon_pre_loop_2D/3D, general_opencl_functions.h:1825-1957) fixes the
//     slab axis and, at every crossing of a slab boundary, the row/plane the ray leaves the slab through
//     (projectorType123.cl:991-1161);
//   * each slab is swept outwards from that row (and plane, with the axial sweep) while the voxel centre's
//     distance d to the ray stays below 0.99 * width; each voxel gets w = 1 - d / width
//     (orthDistance3D / orthogonalHelper3D, opencl_functions_orth3D.h:85-416);
//   * slabs before the first and after the last crossing are swept until one yields no voxel
//     (projectorType123.cl:1055-1128,1235-1313);
//   * rays parallel to x or y walk their row of slabs directly (projectorType123.cl:513-791).
// FP: y = temp * sum w x; BP: rhs += r * (w * temp), sens += w * temp, temp = atten * norm * scat * global_factor,
// atten the LOR's attenuation factor exp(-sum len mu) over its Siddon walk (projectorType123.cl:1022-1048,
// :1220-1233: the reference's -DORTH -DATN integrates mu along the same Siddon walk).
// Projector 3 (-DVOL -DTOTLENGTH) sweeps the same tube of radius R = width with orl = L: a voxel at axis
// distance d weighs Vmax inside bmin, else its volume of intersection from the look-up table, and is cut
// at d > bmax; temp carries 1 / TotV, TotV = L R^2 / pi (opencl_functions_orth3D.h:132-140,
// projectorType123.cl:502-505,641-643,952).
//
// One lane follows one LOR. FMAD of the reference is fmaf here; everything else is evaluated as written
// (the library is built with -ffp-contract=off and IEEE division/sqrt), so a forward projection is
// bit-identical to the oracle's restatement of the same kernel.
#include <type_traits>

#include "geom.hpp"

#include <algorithm>
#include <vector>

namespace omg {

constexpr int kOrthBlock = 256;
constexpr float kOrthThr = 0.01f;
constexpr double kCosLo = 0.7071068921727742;    // cos(0.785398f), in double
constexpr double kCosHi = -0.7071035822349067;   // cos(2.35619f), in double

struct OrthArgs {
    SidArgs a;
    float width;        // kernel tube width in mm (tube_width_z); the tube radius for projector 3
    int32_t cxy, cz;    // transaxial row sweep, axial sweep
    const float* vol;   // projector 3: volume-of-intersection table (null: projector 2)
    uint32_t vol_len;
    float bmin, bmax, vmax;
};

struct F3 { float x, y, z; };

__device__ __forceinline__ uint32_t f2u_ru(float v) {
    if (!(v > 0.f)) return 0u;
    const float c = ceilf(v);
    return c >= 4294967296.f ? 4294967295u : (uint32_t)c;
}
__device__ __forceinline__ uint32_t f2u_rd(float v) {
    if (!(v > 0.f)) return 0u;
    const float c = floorf(v);
    return c >= 4294967296.f ? 4294967295u : (uint32_t)c;
}

__device__ __forceinline__ void o_dgs(float tmin, float t_min, float tmax, float t_max, uint32_t& vmin, uint32_t& vmax,
                                      float& t0, int& vu, float diff, float b, float d, float s, uint32_t N) {
    if (tmin == t_min) vmin = 1u;
    else { const float pt = s + tmin * diff; vmin = f2u_ru((pt - b) / d); }
    if (tmax == t_max) vmax = N;
    else { const float pt = s + tmax * diff; vmax = f2u_rd((pt - b) / d); }
    t0 += ((float)vmin * d) / diff;
    vu = 1;
}
__device__ __forceinline__ void o_sgd(float tmin, float t_min, float tmax, float t_max, uint32_t& vmin, uint32_t& vmax,
                                      float& t0, int& vu, float diff, float b, float d, float s, uint32_t N) {
    if (tmin == t_min) vmax = N - 1u;
    else { const float pt = s + tmin * diff; vmax = f2u_rd((pt - b) / d); }
    if (tmax == t_max) vmin = 0u;
    else { const float pt = s + tmax * diff; vmin = f2u_ru((pt - b) / d); }
    t0 += ((float)vmax * d) / diff;
    vu = -1;
}

__device__ __forceinline__ int o_perp_start(float b, float d, float dd, uint32_t N) {
    int t = (int)floorf((d - b) / dd);
    if (t < 0 || t >= (int)N) t = 0;
    return t;
}

__device__ __forceinline__ bool o_pre2(float b1, float b2, float diff1, float diff2, float max1, float max2, float d1,
                                       float d2, uint32_t N1, uint32_t N2, int& temp1, int& temp2, float& t1u,
                                       float& t2u, uint32_t& Np, float ys, float xs, float yd, float xd, float& tc,
                                       int& u1, int& u2, float& t10, float& t20, bool& xy) {
    const float apu_tx = b1 - xs, apu_ty = b2 - ys;
    t10 = apu_tx / diff1;
    t20 = apu_ty / diff2;
    const float txback = (max1 - xs) / diff1, tyback = (max2 - ys) / diff2;
    const float txmin = fminf(t10, txback), txmax = fmaxf(t10, txback);
    const float tymin = fminf(t20, tyback), tymax = fmaxf(t20, tyback);
    tc = fmaxf(txmin, tymin);
    const float tmax = fminf(txmax, tymax);
    xy = (tc == t10 || tc == txback);
    if (tc >= tmax) return true;
    uint32_t imin, imax, jmin, jmax;
    if (xs < xd) o_dgs(tc, txmin, tmax, txmax, imin, imax, t10, u1, diff1, b1, d1, xs, N1);
    else o_sgd(tc, txmin, tmax, txmax, imin, imax, t10, u1, diff1, b1, d1, xs, N1);
    if (ys < yd) o_dgs(tc, tymin, tmax, tymax, jmin, jmax, t20, u2, diff2, b2, d2, ys, N2);
    else o_sgd(tc, tymin, tmax, tymax, jmin, jmax, t20, u2, diff2, b2, d2, ys, N2);
    Np = imax + 1u + jmax + 1u - imin - jmin;
    const float pt = (fminf(t10, t20) + tc) / 2.f;
    temp1 = (int)((pt * diff1 - apu_tx) / d1);
    temp2 = (int)((pt * diff2 - apu_ty) / d2);
    t1u = d1 / fabsf(diff1);
    t2u = d2 / fabsf(diff2);
    return temp1 < 0 || temp2 < 0 || temp1 >= (int)N1 || temp2 >= (int)N2;
}

__device__ __forceinline__ bool o_pre3(const SidArgs& g, F3 s, F3 d, F3 diff, int& ti, int& tj, int& tk, float& txu,
                                       float& tyu, float& tzu, uint32_t& Np, float& tc, int& ux, int& uy, int& uz,
                                       float& tx0, float& ty0, float& tz0, bool& xy) {
    const F3 apuT{g.bx - s.x, g.by - s.y, g.bz - s.z};
    const F3 t0{apuT.x / diff.x, apuT.y / diff.y, apuT.z / diff.z};
    const F3 tB{(g.bmaxx - s.x) / diff.x, (g.bmaxy - s.y) / diff.y, (g.bmaxz - s.z) / diff.z};
    const F3 tMin{fminf(t0.x, tB.x), fminf(t0.y, tB.y), fminf(t0.z, tB.z)};
    const F3 tMax{fmaxf(t0.x, tB.x), fmaxf(t0.y, tB.y), fmaxf(t0.z, tB.z)};
    tc = fmaxf(fmaxf(tMin.x, tMin.z), tMin.y);
    const float tmax = fminf(fminf(tMax.x, tMax.z), tMax.y);
    tx0 = t0.x; ty0 = t0.y; tz0 = t0.z;
    {
        const float p = d.x - s.x, pY = d.y - s.y;
        const float cx = p / sqrtf(p * p + pY * pY);
        xy = ((double)cx > kCosLo && cx < 1.f) || ((double)cx < kCosHi);
    }
    if (tc >= tmax) return true;
    uint32_t imin, imax, jmin, jmax, kmin, kmax;
    if (s.x < d.x) o_dgs(tc, tMin.x, tmax, tMax.x, imin, imax, tx0, ux, diff.x, g.bx, g.dx, s.x, g.Nx);
    else o_sgd(tc, tMin.x, tmax, tMax.x, imin, imax, tx0, ux, diff.x, g.bx, g.dx, s.x, g.Nx);
    if (s.y < d.y) o_dgs(tc, tMin.y, tmax, tMax.y, jmin, jmax, ty0, uy, diff.y, g.by, g.dy, s.y, g.Ny);
    else o_sgd(tc, tMin.y, tmax, tMax.y, jmin, jmax, ty0, uy, diff.y, g.by, g.dy, s.y, g.Ny);
    if (s.z < d.z) o_dgs(tc, tMin.z, tmax, tMax.z, kmin, kmax, tz0, uz, diff.z, g.bz, g.dz, s.z, g.Nz);
    else o_sgd(tc, tMin.z, tmax, tMax.z, kmin, kmax, tz0, uz, diff.z, g.bz, g.dz, s.z, g.Nz);
    Np = (kmax - kmin + 1u) + (jmax - jmin + 1u) + (imax - imin + 1u);
    const float pt = (fminf(fminf(tz0, ty0), tx0) + tc) / 2.f;
    ti = (int)fminf(fmaxf(__builtin_fmaf(pt, diff.x, -apuT.x) / g.dx, 0.f), (float)(g.Nx - 1));
    tj = (int)fminf(fmaxf(__builtin_fmaf(pt, diff.y, -apuT.y) / g.dy, 0.f), (float)(g.Ny - 1));
    tk = (int)fminf(fmaxf(__builtin_fmaf(pt, diff.z, -apuT.z) / g.dz, 0.f), (float)(g.Nz - 1));
    txu = g.dx / fabsf(diff.x);
    tyu = g.dy / fabsf(diff.y);
    tzu = g.dz / fabsf(diff.z);
    return false;
}

// Slab geometry shared by every sweep of one LOR (after the frame swap: "x" is the slab axis).
struct Tube {
    F3 s, diff;
    float b2, d2, bz, dz, orl;
    uint32_t Nxy, N1, N2p, N3p, Nz;
    bool cxy, cz;
    static constexpr bool kVol = false;
};

// Projector 3's tube: the look-up table on top. A type of its own keeps these fields and the table
// branch out of projector 2's kernels (their registers and occupancy stay as they were).
struct VolTube : Tube {
    static constexpr bool kVol = true;
    const float* V;
    uint32_t vlen;
    float bmin, bmax, vmax;
};

// One voxel: true when it lies outside the cut-off (the row sweep stops there).
template <class T, class V>
__device__ __forceinline__ bool o_voxel(const T& t, int tempi, int uu, int zz, const F3& l, float cy, V& visit) {
    const float x0 = __builtin_fmaf(-1.f, t.s.y, cy);
    const float y1 = __builtin_fmaf(t.diff.z, x0, -l.y);
    const float z1 = __builtin_fmaf(-t.diff.x, x0, l.z);
    const float norm2 = __builtin_fmaf(l.x, l.x, __builtin_fmaf(y1, y1, z1 * z1));
    float w;
    if constexpr (T::kVol) {
        // volume of intersection (opencl_functions_orth3D.h:132-140): orl = L, so w is the axis distance;
        // past the table's end (which the reference leaves open) the voxel counts 0
        if (norm2 > t.bmax * t.bmax * t.orl * t.orl) return true;
        w = sqrtf(norm2) / t.orl;
        if (w < t.bmin) {
            w = t.vmax;
        } else {
            const uint32_t k = __float2uint_rn((w - t.bmin) * 1e3f);
            w = k < t.vlen ? t.V[k] : 0.f;
        }
    } else {
        if (norm2 >= (1.f - kOrthThr) * (1.f - kOrthThr) * t.orl * t.orl) return true;
        w = 1.f - sqrtf(norm2) / t.orl;
    }
    const int64_t ind = (int64_t)((uint32_t)tempi * t.N2p + (uint32_t)uu * t.N3p) + (int64_t)zz * (int64_t)t.Nxy;
    visit(ind, w);
    return false;
}

// Rows of one z plane. `recompute`: the upward rows take their centre from the row index (first z loop);
// otherwise the centre accumulates from the start row (second z loop), as in the reference.
template <class T, class V>
__device__ __forceinline__ void o_rows(const T& t, int tempi, int temp2, int zz, const F3& l, float centerY0,
                                       bool recompute, int& uu, int& uu1, int& uu2, bool& brk, V& visit) {
    const int maxXY = (int)t.N1;
    if (t.cxy) {
        float cy = centerY0;
        for (uu1 = max(temp2, 0); uu1 < maxXY; uu1++) {
            if (recompute) cy = t.b2 + (float)uu1 * t.d2 + t.d2 / 2.f;
            brk = o_voxel(t, tempi, uu1, zz, l, cy, visit);
            if (brk) break;
            uu++;
            cy += t.d2;
        }
        cy = centerY0 - t.d2;
        for (uu2 = min(temp2, maxXY) - 1; uu2 >= 0; uu2--) {
            brk = o_voxel(t, tempi, uu2, zz, l, cy, visit);
            if (brk) break;
            uu++;
            cy -= t.d2;
        }
    } else {
        uu1 = temp2;
        const float cy = recompute ? t.b2 + (float)uu1 * t.d2 + t.d2 / 2.f : centerY0;
        brk = o_voxel(t, tempi, uu1, zz, l, cy, visit);
        uu2 = temp2 - 1;
    }
}

// orthDistance3D: the tube's voxels in slab `tempi` (centre cx), rows from temp2, planes from tempk.
// Inlined into every caller, as is orth_lor: as real calls they kept the tube in a stack frame in scratch
// memory (312-328 B per lane in the one-lane-per-LOR kernels, 3 waves per SIMD); inlined, those kernels
// use no scratch and ~95 VGPRs (5 waves per SIMD).
template <class T, class V>
__device__ __forceinline__ int o_sweep(const T& t, int tempi, float cx, int temp2, int tempk, bool preStep, int& k, int ku,
                       V& visit) {
    int uu = 0, uu1 = 0, uu2 = 0;
    bool brk = false;
    F3 l;
    const float v0 = cx - t.s.x;
    l.z = t.diff.y * v0;
    const float apu1 = t.diff.z * v0;
    const int maxZ = (int)t.Nz;
    const float centerY0 = t.b2 + (float)temp2 * t.d2 + t.d2 / 2.f;
    const float centerZ0 = t.bz + (float)tempk * t.dz + t.dz / 2.f;
    if (!t.cz) {
        const float cz = t.bz + (float)tempk * t.dz + t.dz / 2.f;
        const float z0 = cz - t.s.z;
        l.x = t.diff.x * z0 - apu1;
        l.y = t.diff.y * z0;
        o_rows(t, tempi, temp2, tempk, l, centerY0, true, uu, uu1, uu2, brk, visit);
        return uu;
    }
    int zz;
    for (zz = max(tempk, 0); zz < maxZ; zz++) {
        const float cz = t.bz + (float)zz * t.dz + t.dz / 2.f;
        const float z0 = cz - t.s.z;
        l.x = t.diff.x * z0 - apu1;
        l.y = t.diff.y * z0;
        o_rows(t, tempi, temp2, zz, l, centerY0, true, uu, uu1, uu2, brk, visit);
        if (uu1 == temp2 && uu2 == temp2 - 1 && brk) break;
    }
    k = zz - 1;
    float cz = centerZ0 - t.dz;
    for (zz = min(tempk, maxZ) - 1; zz >= 0; zz--) {
        const float z0 = cz - t.s.z;
        l.x = t.diff.x * z0 - apu1;
        l.y = t.diff.y * z0;
        o_rows(t, tempi, temp2, zz, l, centerY0, false, uu, uu1, uu2, brk, visit);
        if (uu1 == temp2 && uu2 == temp2 - 1 && brk) break;
        cz -= t.dz;
    }
    if (preStep) k = (ku < 0) ? max(k, zz) - 1 : min(k, zz) + 1;
    return uu;
}

// One LOR: visit(ind, w) for every voxel of its tube. Returns the LOR weight temp (0: ray skipped); with
// temp_out, temp is also stored there before the first visit (for callers whose visit applies it).
// No time of flight: nothing to do per stretch of the centre ray.
struct NoSeg {
    __device__ void operator()(float, float, float) const {}
};

// The OpenCL built-in sign(): +-1, the argument itself for +-0, 0 for NaN.
__device__ __forceinline__ float cl_sign(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : (v == v ? v : 0.f)); }

// One LOR: visit(ind, w) for every voxel of its tube; returns the LOR weight (0 for a skipped ray). With TOF,
// seg(elem, D, DDsign) announces the centre ray's current stretch before the planes swept for it
// (projectorType123.cl:653-686 axis-parallel, :845-847, :1022-1054, :1215-1220 otherwise): axis-parallel rays
// start at D = (d N / 2) sign(diff_y) after the x/y swap and step by the voxel size (first / last partial
// voxels: dT1 / dT2); other rays start at |diff tc| - L / 2 and step by each stretch
// (min(t0, 1) - max(tc, 0)) L; the planes swept after the walk keep the last stretch's weights.
template <bool VOL, bool TOF, class V, class S>
__device__ __forceinline__ float orth_lor_t(const OrthArgs& oa, int64_t lo, V&& visit, float* temp_out, S&& seg) {
    const SidArgs& g = oa.a;
    Ray r;
    fetch_ray(g, lo, r);
    std::conditional_t<VOL, VolTube, Tube> t;
    t.s = F3{r.xs, r.ys, r.zs};
    const F3 d{r.xd, r.yd, r.zd};
    t.diff = F3{d.x - t.s.x, d.y - t.s.y, d.z - t.s.z};
    F3& s = t.s;
    F3& diff = t.diff;
    if ((diff.x == 0.f && diff.y == 0.f) || !isfinite(diff.x) || !isfinite(diff.y) || !isfinite(diff.z)) return 0.f;
    const float L = sqrtf(diff.x * diff.x + diff.y * diff.y + diff.z * diff.z);
    constexpr bool vol = VOL;
    t.orl = vol ? L : L * oa.width;
    if constexpr (VOL) {
        t.V = oa.vol;
        t.vlen = oa.vol_len;
        t.bmin = oa.bmin;
        t.bmax = oa.bmax;
        t.vmax = oa.vmax;
    }
    // corrections in the reference kernel's order: factors made from an attenuation image first (-DATN,
    // projectorType123.cl:975-976); given per-LOR factors (-DATNM) after the global factor (:986-988), or
    // before it on the axis-parallel path (perpendicular_elements, general_opencl_functions.h:1722-1734)
    auto chain = [&](float w, bool perp) {
        if (g.atten && !g.atten_lor) w *= g.atten[lo];
        if (g.norm) w *= g.norm[lo];
        if (g.scat) w *= g.scat[lo];
        if (g.atten && g.atten_lor && perp) w *= g.atten[lo];
        w *= g.global_factor;
        if (g.atten && g.atten_lor && !perp) w *= g.atten[lo];
        return w;
    };
    // projector 3: TotV = L * (1/pi) * R^2 (projectorType123.cl:505); an axis-parallel ray's weight is the
    // corrections' product times 1 / TotV (:641-643), any other ray's starts from 1 / TotV (:952)
    const float inv_v = vol ? 1.f / (L * (float)M_1_PI * oa.width * oa.width) : 1.f;
    const float w1 = chain(1.f, true);
    t.Nxy = g.Nx * g.Ny;
    t.Nz = g.Nz;
    t.bz = g.bz;
    t.dz = g.dz;
    t.cxy = oa.cxy != 0;
    t.cz = oa.cz != 0;
    int tempk_b = 0;

    if (fabsf(diff.z) < 1e-6f && (fabsf(diff.y) < 1e-6f || fabsf(diff.x) < 1e-6f)) {
        const int tempk = (int)(fabsf(s.z - g.bz) / g.dz);
        if (tempk < 0 || tempk >= (int)g.Nz) return 0.f;
        int apuX1, apuX2, indO;
        float b1, d1, dT1 = 0.f, dT2 = 0.f;
        uint32_t Nrun = 0;
        auto span = [&](float sa, float da, float ba, float dd, uint32_t N) {
            apuX1 = 0;
            apuX2 = (int)N - 1;
            if constexpr (TOF) {
                dT1 = dd;
                dT2 = dd;
                Nrun = N;
            }
            float dist1, dist2;
            if (sa > da) { dist1 = ba - da; dist2 = ba + (float)N * dd - sa; }
            else { dist1 = ba - sa; dist2 = ba + (float)N * dd - da; }
            if (dist1 < 0.f) {
                const int kk = (int)ceilf(-dist1 / dd);
                if (kk < (int)N) {
                    apuX1 = kk;
                    if constexpr (TOF) dT1 = fminf(dist1 + (float)kk * dd, dd);
                }
            }
            if (dist2 > 0.f) {
                const int mm = (int)ceilf(dist2 / dd);
                const int kk = (int)N - 1 - mm;
                if (kk >= apuX1) {
                    apuX2 = kk;
                    if constexpr (TOF) dT2 = fminf((float)mm * dd - dist2, dd);
                }
            }
        };
        if (fabsf(diff.y) < 1e-6f && d.y <= g.bmaxy && d.y >= g.by && s.y <= g.bmaxy && s.y >= g.by) {
            span(s.x, d.x, g.bx, g.dx, g.Nx);
            b1 = g.bx; t.b2 = g.by; d1 = g.dx; t.d2 = g.dy;
            t.N1 = g.Nx; t.N2p = 1u; t.N3p = g.Ny;
            indO = o_perp_start(g.by, d.y, g.dy, g.Ny);
        } else if (fabsf(diff.x) < 1e-6f && d.x <= g.bmaxx && d.x >= g.bx && s.x <= g.bmaxx && s.x >= g.bx) {
            span(s.y, d.y, g.by, g.dy, g.Ny);
            b1 = g.by; t.b2 = g.bx; d1 = g.dy; t.d2 = g.dx;
            const float sb = s.x; s.x = s.y; s.y = sb;
            const float db = diff.x; diff.x = diff.y; diff.y = db;
            t.N1 = g.Ny; t.N2p = g.Nx; t.N3p = 1u;
            indO = o_perp_start(g.bx, d.x, g.dx, g.Nx);
        } else {
            return 0.f;
        }
        const float w = vol ? w1 * inv_v : w1;
        if (temp_out) *temp_out = w;
        float D = 0.f, DDsign = 0.f;
        if constexpr (TOF) {
            D = ((d1 * (float)Nrun) / 2.f) * cl_sign(diff.y);
            DDsign = cl_sign(D);
        }
        for (int ii = apuX1; ii <= apuX2; ii++) {
            if constexpr (TOF) {
                float d_in = d1;
                if (apuX1 > 0 && ii == apuX1) d_in = dT1;
                else if (apuX2 < (int)Nrun - 1 && ii == apuX2) d_in = dT2;
                seg(d_in, D, DDsign);
            }
            const float cx = b1 + d1 * (float)ii + d1 / 2.f;
            o_sweep(t, ii, cx, indO, tempk, false, tempk_b, 0, visit);
            if constexpr (TOF) D -= d1 * DDsign;
        }
        return w;
    }

    float txu = 0.f, tyu = 0.f, tzu = 0.f, tc = 0.f, tx0 = 1e8f, ty0 = 1e8f, tz0 = 1e8f;
    bool skip = false, XY = true;
    int tempi = 0, tempj = 0, tempk = 0, ux = 0, uy = 0, uz = 0;
    uint32_t Np = 0u;
    if (fabsf(diff.z) < 1e-6f) {
        tempk = (int)(fabsf(s.z - g.bz) / g.dz);
        if (tempk < 0 || tempk >= (int)g.Nz) return 0.f;
        skip = o_pre2(g.bx, g.by, diff.x, diff.y, g.bmaxx, g.bmaxy, g.dx, g.dy, g.Nx, g.Ny, tempi, tempj, txu, tyu, Np,
                      s.y, s.x, d.y, d.x, tc, ux, uy, tx0, ty0, XY);
    } else if (fabsf(diff.y) < 1e-6f) {
        tempj = o_perp_start(g.by, d.y, g.dy, g.Ny);
        skip = o_pre2(g.bx, g.bz, diff.x, diff.z, g.bmaxx, g.bmaxz, g.dx, g.dz, g.Nx, g.Nz, tempi, tempk, txu, tzu, Np,
                      s.z, s.x, d.z, d.x, tc, ux, uz, tx0, tz0, XY);
        XY = true;
        if (d.y > g.bmaxy || d.y < g.by) skip = true;
    } else if (fabsf(diff.x) < 1e-6f) {
        tempi = o_perp_start(g.bx, d.x, g.dx, g.Nx);
        skip = o_pre2(g.by, g.bz, diff.y, diff.z, g.bmaxy, g.bmaxz, g.dy, g.dz, g.Ny, g.Nz, tempj, tempk, tyu, tzu, Np,
                      s.z, s.y, d.z, d.y, tc, uy, uz, ty0, tz0, XY);
        XY = false;
        if (d.x > g.bmaxx || d.x < g.bx) skip = true;
    } else {
        skip = o_pre3(g, s, d, diff, tempi, tempj, tempk, txu, tyu, tzu, Np, tc, ux, uy, uz, tx0, ty0, tz0, XY);
    }
    if (skip) return 0.f;
    float D = 0.f, DDsign = 0.f, tcv = tc;
    if constexpr (TOF) {   // TOFDis with the unswapped difference vector
        const float ax = diff.x * tc, ay = diff.y * tc, az = diff.z * tc;
        D = sqrtf(ax * ax + ay * ay + az * az) - L / 2.f;
        DDsign = cl_sign(D);
    }

    float b1, d1;
    uint32_t NNx;
    if (!XY) {
        b1 = g.by; t.b2 = g.bx; d1 = g.dy; t.d2 = g.dx;
        int ib = tempi; tempi = tempj; tempj = ib;
        float fb = s.x; s.x = s.y; s.y = fb;
        ib = ux; ux = uy; uy = ib;
        fb = tx0; tx0 = ty0; ty0 = fb;
        fb = txu; txu = tyu; tyu = fb;
        fb = diff.x; diff.x = diff.y; diff.y = fb;
        NNx = g.Ny;
        t.N1 = g.Nx; t.N2p = g.Nx; t.N3p = 1u;
    } else {
        b1 = g.bx; t.b2 = g.by; d1 = g.dx; t.d2 = g.dy;
        NNx = g.Nx;
        t.N1 = g.Ny; t.N2p = 1u; t.N3p = g.Nx;
    }
    const float w = chain(inv_v, false);
    if (temp_out) *temp_out = w;
    int tempi_a = tempi, tempj_a = tempj, tempk_a = tempk;
    bool swept = false;
    int tempiOld = 0;
    for (uint32_t ii = 0u; ii < Np; ii++) {
        const int li = tempi, lj = tempj, lk = tempk;
        const float tx0_a = tx0, ty0_a = ty0, tz0_a = tz0;
        tempi_a = tempi; tempj_a = tempj; tempk_a = tempk;
        float t0;
        if (tz0 < ty0 && tz0 < tx0) { t0 = tz0; tempk += uz; tz0 += tzu; }
        else if (ty0 < tx0) { t0 = ty0; tempj += uy; ty0 += tyu; }
        else { t0 = tx0; tempi += ux; tx0 += txu; }
        float elem = 0.f;
        if constexpr (TOF) {   // compute_element, general_opencl_functions.h:1809-1818
            const bool pass = (t0 >= 0.f && t0 <= 1.f) || (tcv >= 0.f && tcv <= 1.f);
            if (pass) elem = (fminf(t0, 1.f) - fmaxf(tcv, 0.f)) * L;
            tcv = t0;
            seg(elem, D, DDsign);
        }
        (void)t0;
        if (ii == 0u) {
            tempk_b = tempk_a;
            if (ux >= 0) {
                float cx = b1 + d1 * (float)(tempi_a - 1) + d1 / 2.f;
                for (int kk = tempi_a - 1; kk >= 0; kk--) {
                    if (o_sweep(t, kk, cx, tempj_a, tempk_b, true, tempk_b, uz, visit) == 0) break;
                    cx -= d1;
                }
            } else {
                float cx = b1 + d1 * (float)(tempi_a + 1) + d1 / 2.f;
                for (int kk = tempi_a + 1; kk < (int)NNx; kk++) {
                    if (o_sweep(t, kk, cx, tempj_a, tempk_b, true, tempk_b, uz, visit) == 0) break;
                    cx += d1;
                }
            }
        }
        if (tz0_a >= tx0_a && ty0_a >= tx0_a) {
            const float cx = b1 + d1 * (float)li + d1 / 2.f;
            o_sweep(t, li, cx, lj, lk, false, tempk_b, uz, visit);
            tempiOld = tempi_a;
            swept = true;
        }
        if constexpr (TOF) D -= elem * DDsign;
        if (tempi < 0 || tempi >= (int)g.Nx || tempj < 0 || tempj >= (int)g.Ny || tempk < 0 || tempk >= (int)g.Nz)
            break;
    }
    if (ux < 0) {
        if (!swept || tempiOld != tempi_a) tempi_a++;
        for (int ii = tempi_a - 1; ii >= 0; ii--) {
            const float cx = b1 + d1 * (float)ii + d1 / 2.f;
            if (o_sweep(t, ii, cx, tempj_a, tempk_a, false, tempk_b, uz, visit) == 0) break;
        }
    } else {
        if (!swept || tempiOld != tempi_a) tempi_a--;
        float cx = b1 + d1 * (float)(tempi_a + 1) + d1 / 2.f;
        for (int ii = tempi_a + 1; ii < (int)NNx; ii++) {
            if (o_sweep(t, ii, cx, tempj_a, tempk_a, false, tempk_b, uz, visit) == 0) break;
            cx += d1;
        }
    }
    return w;
}

template <bool VOL, class V>
__device__ __forceinline__ float orth_lor(const OrthArgs& oa, int64_t lo, V&& visit, float* temp_out = nullptr) {
    return orth_lor_t<VOL, false>(oa, lo, visit, temp_out, NoSeg{});
}

template <bool VOL>
__global__ __launch_bounds__(kOrthBlock) void k_orth_fp(OrthArgs oa, const float* __restrict__ img,
                                                        float* __restrict__ y) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    float ax = 0.f;
    const float temp = orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) { ax += w * img[ind]; });
    y[lo] = ax * temp;
}

template <bool VOL>
__global__ __launch_bounds__(kOrthBlock) void k_orth_bp(OrthArgs oa, const float* __restrict__ yin,
                                                        float* __restrict__ rhs, float* __restrict__ sens) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    const float r = yin[lo];
    if (r == 0.f && sens == nullptr) return;
    float temp = 0.f;   // the LOR weight, set by orth_lor before its first visit
    if (sens) {
        orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) {
            const float e = w * temp;
            unsafeAtomicAdd(rhs + ind, r * e);
            unsafeAtomicAdd(sens + ind, e);
        }, &temp);
    } else {
        orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) { unsafeAtomicAdd(rhs + ind, r * (w * temp)); }, &temp);
    }
}

// FP, ratio r = y / (A x + eps) (NaN/Inf -> eps), BP of r; optional sensitivity.
template <bool VOL>
__global__ __launch_bounds__(kOrthBlock) void k_orth_osem(OrthArgs oa, const float* __restrict__ img,
                                                          const float* __restrict__ meas, float* __restrict__ rhs,
                                                          float* __restrict__ sens, float epps) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    float ax = 0.f;
    const float temp = orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) { ax += w * img[ind]; });
    if (temp == 0.f) return;
    float r = meas[lo] / ratio_den(oa.a, lo, ax * temp, epps);
    if (isnan(r) || isinf(r)) r = epps;
    if (r == 0.f && sens == nullptr) return;
    if (sens) {
        orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) {
            const float e = w * temp;
            unsafeAtomicAdd(rhs + ind, r * e);
            unsafeAtomicAdd(sens + ind, e);
        });
    } else {
        orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) { unsafeAtomicAdd(rhs + ind, r * (w * temp)); });
    }
}

// ---- Time of flight (projector 3; projectorType123.cl -DTOF with -DORTH -DVOL) ----
// One lane per LOR, one wave per workgroup; a lane's bins (accumulators, measurements or ratios) and the
// current stretch's bin weights live in LDS, laid out [bin][lane].
constexpr int kOrthTofBlock = 64;

// The bin weights of one centre-ray stretch (TOFLoop / TOFWeight, general_opencl_functions.h:1074-1116 with
// -DUSEMAD, TRAPZ_BINS 4): normPDF(x) = (0.3989423f / sigma) exp(-0.5 ((x - mu) / sigma)^2), weight_b =
// (p(D) + 2 p(D - dX s) + 2 p(D - 2 dX s) + p(D - elem s)) dX with dX = elem / 3, s = DDsign, and their sum
// floored at epps. The walk announces every stretch; the weights are evaluated only when a tube voxel of the
// stretch's planes needs them. Before the first stretch the weights are zero.
//
// Every tube voxel of a stretch's planes takes the same weights, so the bins are visited once per stretch,
// not per voxel: the forward projection sums the stretch's (w x[v]) / sum and adds weight_b times that sum to
// bin b when the stretch changes (the reference adds weight_b ((w x[v]) / sum) per voxel: the same terms,
// summed in another order); the backprojection forms sum_b weight_b r_b and sum_b weight_b once per stretch.
struct OrthTof {
    const SidArgs& a;
    float* wl;
    float inv, pf, sum, elem, D, sgn;
    float part;      // FP: this stretch's sum of (w x[v]) / sum, not yet added to the bins
    float s1, sr;    // BP: sum_b weight_b and sum_b weight_b r_b of the current stretch
    bool dirty;
    __device__ OrthTof(const SidArgs& args, float* w)
        : a(args), wl(w), inv(1.f / args.tof_sigma), pf(0.3989423f * (1.f / args.tof_sigma)), sum(1.f), elem(0.f),
          D(0.f), sgn(0.f), part(0.f), s1(0.f), sr(0.f), dirty(false) {
        for (uint32_t b = 0; b < a.tof_bins; b++) wl[b * kOrthTofBlock] = 0.f;
    }
    __device__ void set(float e, float d, float s) {
        elem = e;
        D = d;
        sgn = s;
        dirty = true;
    }
    __device__ float pdf(float x, float mu) const {
        const float q = (x - mu) * inv;
        return pf * expf(-0.5f * q * q);
    }
    __device__ float weight(uint32_t b) const {
        const float dX = elem / 3.f;
        const float dXs = dX * sgn;
        const float mu = a.tof_c[b];
        float out = pdf(D, mu);
        out += pdf(fmaf(-dXs, 1.f, D), mu) * 2.f;
        out += pdf(fmaf(-dXs, 2.f, D), mu) * 2.f;
        out += pdf(fmaf(-elem, sgn, D), mu);
        return out * dX;
    }
    // FP: add the pending stretch sum to the bins with the weights it was gathered under
    __device__ void flush(float* ax) {
        if (part != 0.f)
            for (uint32_t b = 0; b < a.tof_bins; b++) ax[b * kOrthTofBlock] += wl[b * kOrthTofBlock] * part;
        part = 0.f;
    }
    __device__ void ready_fp(float* ax) {
        if (!dirty) return;
        dirty = false;
        flush(ax);
        float s = 0.f;
        for (uint32_t b = 0; b < a.tof_bins; b++) {
            const float v = weight(b);
            wl[b * kOrthTofBlock] = v;
            s += v;
        }
        sum = s < a.tof_epps ? a.tof_epps : s;
    }
    __device__ void ready_bp(const float* r) {
        if (!dirty) return;
        dirty = false;
        float s = 0.f, q = 0.f;
        for (uint32_t b = 0; b < a.tof_bins; b++) {
            const float v = weight(b);
            s += v;
            q += v * r[b * kOrthTofBlock];
        }
        s1 = s;
        sr = q;
        sum = s < a.tof_epps ? a.tof_epps : s;
    }
};

// FP (denominator, general_opencl_functions.h:1178-1206): ax_b += weight_b ((w x[v]) / sum); y_b = ax_b temp.
template <bool VOL>
__device__ __forceinline__ float orth_tof_gather(const OrthArgs& oa, int64_t lo, const float* __restrict__ img,
                                                 float* ax, OrthTof& t) {
    for (uint32_t b = 0; b < oa.a.tof_bins; b++) ax[b * kOrthTofBlock] = 0.f;
    const float temp = orth_lor_t<VOL, true>(oa, lo, [&](int64_t ind, float w) {
        t.ready_fp(ax);
        t.part += (w * img[ind]) / t.sum;
    }, nullptr, [&](float e, float D, float sg) { t.set(e, D, sg); });
    t.flush(ax);
    return temp;
}

// BP (rhs, :1210-1245): rhs[v] += sum_b ((w temp) / sum) weight_b r_b; sens[v] += sum_b ((w temp) / sum) weight_b.
template <bool VOL>
__device__ __forceinline__ void orth_tof_scatter(const OrthArgs& oa, int64_t lo, const float* r, OrthTof& t,
                                                 float* __restrict__ rhs, float* __restrict__ sens) {
    float temp = 0.f;   // the LOR weight, set before the first visit
    orth_lor_t<VOL, true>(oa, lo, [&](int64_t ind, float w) {
        t.ready_bp(r);
        const float dX = (w * temp) / t.sum;
        unsafeAtomicAdd(rhs + ind, dX * t.sr);
        if (sens) unsafeAtomicAdd(sens + ind, dX * t.s1);
    }, &temp, [&](float e, float D, float sg) { t.set(e, D, sg); });
}

template <bool VOL>
__global__ __launch_bounds__(kOrthTofBlock) void k_orth_tof_fp(OrthArgs oa, const float* __restrict__ img,
                                                               float* __restrict__ y) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kOrthTofBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    float* ax = lds + threadIdx.x;
    OrthTof t(oa.a, lds + oa.a.tof_bins * kOrthTofBlock + threadIdx.x);
    const float temp = orth_tof_gather<VOL>(oa, lo, img, ax, t);
    for (uint32_t b = 0; b < oa.a.tof_bins; b++) y[lo + (int64_t)b * oa.a.nmeas] = ax[b * kOrthTofBlock] * temp;
}

template <bool VOL>
__global__ __launch_bounds__(kOrthTofBlock) void k_orth_tof_bp(OrthArgs oa, const float* __restrict__ yin,
                                                               float* __restrict__ rhs, float* __restrict__ sens) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kOrthTofBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    float* yv = lds + threadIdx.x;
    bool any = false;
    for (uint32_t b = 0; b < oa.a.tof_bins; b++) {
        const float v = yin[lo + (int64_t)b * oa.a.nmeas];
        yv[b * kOrthTofBlock] = v;
        any |= v != 0.f;
    }
    if (!any && sens == nullptr) return;   // adds of +0 leave rhs unchanged
    OrthTof t(oa.a, lds + oa.a.tof_bins * kOrthTofBlock + threadIdx.x);
    orth_tof_scatter<VOL>(oa, lo, yv, t, rhs, sens);
}

// Fused data step: yhat_b = ax_b temp; r_b = meas_b / ((yhat_b + epps) + add); NaN/Inf -> epps; BP of r.
template <bool VOL>
__global__ __launch_bounds__(kOrthTofBlock) void k_orth_tof_osem(OrthArgs oa, const float* __restrict__ img,
                                                                 const float* __restrict__ meas,
                                                                 float* __restrict__ rhs, float* __restrict__ sens,
                                                                 float epps) {
    extern __shared__ float lds[];
    const int64_t lo = (int64_t)blockIdx.x * kOrthTofBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    float* ax = lds + threadIdx.x;
    float* wl = lds + oa.a.tof_bins * kOrthTofBlock + threadIdx.x;
    OrthTof t(oa.a, wl);
    const float temp = orth_tof_gather<VOL>(oa, lo, img, ax, t);
    if (temp == 0.f) return;
    for (uint32_t b = 0; b < oa.a.tof_bins; b++) {
        float r = meas[lo + (int64_t)b * oa.a.nmeas] / ratio_den(oa.a, lo, ax[b * kOrthTofBlock] * temp, epps);
        if (isnan(r) || isinf(r)) r = epps;
        ax[b * kOrthTofBlock] = r;
    }
    OrthTof t2(oa.a, wl);
    orth_tof_scatter<VOL>(oa, lo, ax, t2, rhs, sens);
}

// Per-LOR voxel lists: counts only (offsets == null) or indices and tube weights at offsets[lo].
template <bool VOL>
__global__ __launch_bounds__(kOrthBlock) void k_orth_trace(OrthArgs oa, int32_t* __restrict__ counts,
                                                           const int64_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ idx, float* __restrict__ wout) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    int32_t c = 0;
    if (offsets) {
        int64_t o = offsets[lo];
        orth_lor<VOL>(oa, lo, [&](int64_t ind, float w) { idx[o] = (uint32_t)ind; wout[o] = w; o++; c++; });
    } else {
        orth_lor<VOL>(oa, lo, [&](int64_t, float) { c++; });
    }
    counts[lo] = c;
}

// ---- direct-plane (planar) kernels ---------------------------------------------------------------
// For direct-plane sinograms (zs == zd) whose tube cannot reach a neighbouring slice -- transaxial sweep
// only, or an axial sweep whose neighbouring slice centres all lie farther than the 0.99 * width
// cut-off from the plane (checked on the host, orth_prepare) -- every tube voxel of a LOR lies in the
// LOR's slice tk = (int)(|zs - bz| / dz), the slice of orth_lor's in-plane branch. The Siddon planar
// design then applies: a workgroup owns (plane, chunk of LORs); the FP gathers from the slice staged
// in LDS (same visit order: bit-identical to k_orth_fp); the BP accumulates t = r * (w * temp) in a
// 64-bit fixed-point LDS slice (the scale bound uses w <= 1 and, as margin, up to four visits of a
// voxel per LOR) and stores a partial slice, reduced by the shared planar reduce / update.
// Tube lists kept in device memory (see k_tube_fill): lane lo's k-th visit at base[lo / 64] + 64 k + lo % 64.
struct TubeView {
    const uint32_t* ind;
    const float* w;
    const int64_t* base;
    const int32_t* count;
    const float* temp;   // the LOR weight orth_lor returns (0: nothing visited)
    int64_t n;
};

// LISTS: the tubes come from the kept lists (tv) instead of being traced: the same pairs in the same order,
// so every sum is the traced kernel's, bit for bit.
template <bool LISTS, class V>
__device__ __forceinline__ float planar_tube(const OrthArgs& oa, const TubeView& tv, int64_t lo, V&& visit) {
    if constexpr (LISTS) {
        const int32_t c = tv.count[lo];
        const int64_t p = tv.base[lo >> 6] + (lo & 63);
        for (int32_t k = 0; k < c; k++) visit((int64_t)tv.ind[p + 64 * k], tv.w[p + 64 * k]);
        return tv.temp[lo];
    } else {
        return orth_lor<false>(oa, lo, visit);
    }
}

template <int TPB, bool FP, bool BP, bool SENS, bool LISTS>
__global__ __launch_bounds__(TPB) void k_orth_planar(OrthArgs oa, TubeView tv, uint32_t n_planes, uint32_t chunk,
                                                     const float* __restrict__ img, const float* __restrict__ yin,
                                                     float* __restrict__ yout, float* __restrict__ rhs,
                                                     float* __restrict__ sens, float epps) {
    extern __shared__ float lds[];
    const SidArgs& a = oa.a;
    const uint32_t plane = blockIdx.x % n_planes, grp = blockIdx.x / n_planes;
    const uint32_t per = a.size_x * a.size_y;
    const uint32_t l0 = grp * chunk;
    if (l0 >= per) return;
    const uint32_t l1 = min(per, l0 + chunk);
    const uint32_t Nyx = a.Nyx;
    const int64_t base = (int64_t)plane * per;
    const float zs = a.z[(a.plane_offset + plane) * 2];
    const int32_t tk = (int32_t)(fabsf(zs - a.bz) / a.dz);
    if (tk < 0 || (uint32_t)tk >= a.Nz) {   // every LOR of this plane misses the image
        if (FP && !BP)
            for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) yout[base + l] = 0.f;
        return;
    }
    const int64_t sbase = (int64_t)tk * Nyx;
    // LDS: [slice: Nyx floats (FP) / Nyx int64 (BP)][R: chunk floats][red: TPB / 64 floats]
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds);
    float* R = lds + (BP ? 2 : 1) * Nyx;
    float* red = R + chunk;
    float mr = 0.f, ms = 0.f;   // this lane's max |r temp| and max temp
    if (FP) {
        const float4* src = reinterpret_cast<const float4*>(img + sbase);
        float4* dst = reinterpret_cast<float4*>(lds);
        for (uint32_t i = threadIdx.x; i < Nyx / 4; i += TPB) dst[i] = src[i];
        __syncthreads();
        for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
            const int64_t lo = base + l;
            float ax = 0.f;
            const float temp = planar_tube<LISTS>(oa, tv, lo, [&](int64_t ind, float w) {
                const uint64_t j = (uint64_t)(ind - sbase);
                if (j < Nyx) ax += w * lds[j];
            });
            if (!BP) {
                yout[lo] = ax * temp;
            } else {
                float r = 0.f;
                if (temp != 0.f) {
                    r = yin[lo] / ratio_den(a, lo, ax * temp, epps);
                    if (isnan(r) || isinf(r)) r = epps;
                    mr = fmaxf(mr, fabsf(r * temp));
                    ms = fmaxf(ms, fabsf(temp));
                }
                R[l - l0] = r;
            }
        }
    } else if (BP) {
        for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
            const int64_t lo = base + l;
            float temp = 1.f;
            if (a.atten) temp *= a.atten[lo];
            if (a.norm) temp *= a.norm[lo];
            if (a.scat) temp *= a.scat[lo];
            temp *= a.global_factor;
            mr = fmaxf(mr, fabsf(yin[lo] * temp));
            ms = fmaxf(ms, fabsf(temp));
        }
    }
    if (BP) {
        // |r (w temp)| <= |r temp| (w <= 1); a voxel's sum is bounded, with a margin of four visits per
        // LOR, by 4 * (LORs of the workgroup) * that
        const float nl4 = 4.f * (float)(l1 - l0);
        double up_r, down_r, up_s = 1.0, down_s = 1.0;
        const float br = block_max<TPB>(mr, red) * (1.f + 0x1p-10f);
        fixed_scale(br, br * nl4, up_r, down_r);
        if (SENS) {
            __syncthreads();
            const float bs = block_max<TPB>(ms, red) * (1.f + 0x1p-10f);
            fixed_scale(bs, bs * nl4, up_s, down_s);
        }
        for (int pass = 0; pass < (SENS ? 2 : 1); pass++) {
            for (uint32_t i = threadIdx.x; i < Nyx; i += TPB) acc[i] = 0ull;
            __syncthreads();
            for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
                const int64_t lo = base + l;
                const float r = FP ? R[l - l0] : yin[lo];
                if (pass == 0 && r == 0.f) continue;
                float temp = 1.f;
                if (a.atten) temp *= a.atten[lo];
                if (a.norm) temp *= a.norm[lo];
                if (a.scat) temp *= a.scat[lo];
                temp *= a.global_factor;
                if (pass == 0) {
                    planar_tube<LISTS>(oa, tv, lo, [&](int64_t ind, float w) {
                        const uint64_t j = (uint64_t)(ind - sbase);
                        if (j < Nyx) lds_add_fixed(acc + j, r * (w * temp), up_r);
                    });
                } else {
                    planar_tube<LISTS>(oa, tv, lo, [&](int64_t ind, float w) {
                        const uint64_t j = (uint64_t)(ind - sbase);
                        if (j < Nyx) lds_add_fixed(acc + j, w * temp, up_s);
                    });
                }
            }
            __syncthreads();
            float* part = (pass == 0 ? rhs : sens) + (size_t)blockIdx.x * Nyx;
            const double down = pass == 0 ? down_r : down_s;
            for (uint32_t i = threadIdx.x; i < Nyx; i += TPB) part[i] = fixed_value(acc[i], down);
            if (SENS && pass == 0) __syncthreads();
        }
    }
}

// ---- projector 3 on direct-plane sinograms: fused FP -> ratio -> BP with LDS slice accumulators ----
// The volume tube reaches the slices next to the LOR's own (a voxel counts out to bmax = 2.77 mm with the
// defaults): 64% of a 2-D plane subset's visits land in slice tk, 18% in each neighbour. A workgroup
// owns (plane, chunk of LORs), all in slice tk:
//   * FP: slice tk staged in LDS, the others read from img, summed in the visit order (bit-identical to
//     k_orth_osem's forward projection); r and the LOR weight are kept per LOR in LDS;
//   * BP: slices tk and tk + 1 accumulate in LDS (float ds_add: ~0.33 lanes / clock / CU, still ~8x the
//     scattered global float atomics), slice tk - 1 takes global atomics; one coalesced flush of the
//     two slices (non-zero entries) into rhs at the end.
// r * (w * temp) is the value k_orth_osem adds, so rhs differs from it only in the order of the float
// additions. No sensitivity image (the first iteration keeps k_orth_osem).
template <int TPB>
__global__ __launch_bounds__(TPB) void k_orth_vol_planar(OrthArgs oa, uint32_t n_planes, uint32_t chunk,
                                                         const float* __restrict__ img, const float* __restrict__ meas,
                                                         float* __restrict__ rhs, float epps) {
    extern __shared__ float lds[];
    const SidArgs& a = oa.a;
    const uint32_t plane = blockIdx.x % n_planes, grp = blockIdx.x / n_planes;
    const uint32_t per = a.size_x * a.size_y;
    const uint32_t l0 = grp * chunk;
    if (l0 >= per) return;
    const uint32_t l1 = min(per, l0 + chunk);
    const uint32_t Nyx = a.Nyx;
    const int64_t base = (int64_t)plane * per;
    const float zs = a.z[(a.plane_offset + plane) * 2];
    const int32_t tk = (int32_t)(fabsf(zs - a.bz) / a.dz);
    if (tk < 0 || (uint32_t)tk >= a.Nz) return;   // every LOR of the plane misses the image
    const int64_t sbase = (int64_t)tk * Nyx;
    const uint32_t nacc = ((uint32_t)tk + 1u < a.Nz) ? 2u * Nyx : Nyx;   // slices tk (and tk + 1) in LDS
    // LDS: [2 Nyx floats: slice tk for the FP, then the accumulators][R: chunk floats][T: chunk floats]
    float* R = lds + 2 * Nyx;
    float* T = R + chunk;
    {
        const float4* src = reinterpret_cast<const float4*>(img + sbase);
        float4* dst = reinterpret_cast<float4*>(lds);
        for (uint32_t i = threadIdx.x; i < Nyx / 4; i += TPB) dst[i] = src[i];
    }
    __syncthreads();
    for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
        const int64_t lo = base + l;
        float ax = 0.f;
        const float temp = orth_lor<true>(oa, lo, [&](int64_t ind, float w) {
            const uint64_t j = (uint64_t)(ind - sbase);
            ax += w * (j < Nyx ? lds[j] : img[ind]);
        });
        float r = 0.f;   // 0: nothing to backproject (a LOR that misses the image, or y = 0)
        if (temp != 0.f) {
            r = meas[lo] / ratio_den(a, lo, ax * temp, epps);
            if (isnan(r) || isinf(r)) r = epps;
        }
        R[l - l0] = r;
        T[l - l0] = temp;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nacc; i += TPB) lds[i] = 0.f;
    __syncthreads();
    for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
        const float r = R[l - l0];
        if (r == 0.f) continue;
        const float temp = T[l - l0];
        orth_lor<true>(oa, base + l, [&](int64_t ind, float w) {
            const float v = r * (w * temp);
            const uint64_t j = (uint64_t)(ind - sbase);
            if (j < nacc) __hip_atomic_fetch_add(lds + j, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else unsafeAtomicAdd(rhs + ind, v);
        });
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nacc; i += TPB) {
        const float v = lds[i];
        if (v != 0.f) unsafeAtomicAdd(rhs + sbase + i, v);
    }
}

// ---- projectors 2 and 3: tube lists kept in device memory ------------------------------------------
// A LOR's tube (the voxels it visits and their distance or table weights) depends on the geometry only,
// and tracing it costs ~290 VALU instructions per visit (projector 3). The lists of a measurement range are traced once and kept:
// lane lo's k-th visit at base[lo / 64] + 64 k + lo % 64 (the 64 LORs of a block side by side, so each
// load of a wave is one coalesced row), the block padded to its longest list. The projections then
// stream them: the same (voxel, weight) pairs in the same order, so the forward projection is
// bit-identical to the tracing kernels' and the backprojection adds the same values.

// The walk is the counting pass's (same code, no contraction), so it visits count[lo] voxels; the writes
// are bounded by that count all the same.
template <bool VOL>
__global__ __launch_bounds__(kOrthBlock) void k_tube_fill(OrthArgs oa, const int64_t* __restrict__ base,
                                                          const int32_t* __restrict__ count,
                                                          uint32_t* __restrict__ ind, float* __restrict__ w,
                                                          float* __restrict__ temp) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= oa.a.nmeas) return;
    const int64_t p = base[lo >> 6] + (lo & 63);
    const int32_t c = count[lo];
    int32_t k = 0;
    temp[lo] = orth_lor<VOL>(oa, lo, [&](int64_t i, float wt) {
        if (k < c) {
            ind[p + 64 * (int64_t)k] = (uint32_t)i;
            w[p + 64 * (int64_t)k] = wt;
        }
        k++;
    });
}

// sum_k w_k img[ind_k] in visit order
__device__ __forceinline__ float tube_gather(const TubeView& t, int64_t lo, const float* __restrict__ img) {
    const int32_t c = t.count[lo];
    const int64_t p = t.base[lo >> 6] + (lo & 63);
    float ax = 0.f;
    for (int32_t k = 0; k < c; k++) ax += t.w[p + 64 * k] * img[t.ind[p + 64 * k]];
    return ax;
}

__global__ __launch_bounds__(kOrthBlock) void k_tube_fp(TubeView t, const float* __restrict__ img,
                                                        float* __restrict__ y) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= t.n) return;
    y[lo] = tube_gather(t, lo, img) * t.temp[lo];
}

// The backprojection of r over one LOR: rhs += r (w temp), sens += w temp (k_orth_bp's values).
__device__ __forceinline__ void tube_scatter(const TubeView& t, int64_t lo, float r, float temp, float* __restrict__ rhs,
                                             float* __restrict__ sens) {
    const int32_t c = t.count[lo];
    const int64_t p = t.base[lo >> 6] + (lo & 63);
    for (int32_t k = 0; k < c; k++) {
        const float e = t.w[p + 64 * k] * temp;
        const uint32_t i = t.ind[p + 64 * k];
        unsafeAtomicAdd(rhs + i, r * e);
        if (sens) unsafeAtomicAdd(sens + i, e);
    }
}

__global__ __launch_bounds__(kOrthBlock) void k_tube_bp(TubeView t, const float* __restrict__ yin,
                                                        float* __restrict__ rhs, float* __restrict__ sens) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= t.n) return;
    const float r = yin[lo];
    if (r == 0.f && sens == nullptr) return;
    tube_scatter(t, lo, r, t.temp[lo], rhs, sens);
}

// k_orth_osem from the lists.
__global__ __launch_bounds__(kOrthBlock) void k_tube_osem(TubeView t, SidArgs a, const float* __restrict__ img,
                                                          const float* __restrict__ meas, float* __restrict__ rhs,
                                                          float* __restrict__ sens, float epps) {
    const int64_t lo = (int64_t)blockIdx.x * kOrthBlock + threadIdx.x;
    if (lo >= t.n) return;
    const float temp = t.temp[lo];
    if (temp == 0.f) return;
    const float ax = tube_gather(t, lo, img);
    float r = meas[lo] / ratio_den(a, lo, ax * temp, epps);
    if (isnan(r) || isinf(r)) r = epps;
    if (r == 0.f && sens == nullptr) return;
    tube_scatter(t, lo, r, temp, rhs, sens);
}

// k_orth_vol_planar from the lists: slice tk staged in LDS for the gathers, slices tk and tk + 1
// accumulated in LDS, slice tk - 1 through global atomics. SENS: the sensitivity image instead, the sum of
// w temp over every LOR that visits the image (no forward projection; `rhs` receives it).
template <int TPB, bool SENS>
__global__ __launch_bounds__(TPB) void k_tube_vol_planar(TubeView t, SidArgs a, uint32_t n_planes, uint32_t chunk,
                                                         const float* __restrict__ img, const float* __restrict__ meas,
                                                         float* __restrict__ rhs, float epps) {
    extern __shared__ float lds[];
    const uint32_t plane = blockIdx.x % n_planes, grp = blockIdx.x / n_planes;
    const uint32_t per = a.size_x * a.size_y;
    const uint32_t l0 = grp * chunk;
    if (l0 >= per) return;
    const uint32_t l1 = min(per, l0 + chunk);
    const uint32_t Nyx = a.Nyx;
    const int64_t base = (int64_t)plane * per;
    const float zs = a.z[(a.plane_offset + plane) * 2];
    const int32_t tk = (int32_t)(fabsf(zs - a.bz) / a.dz);
    if (tk < 0 || (uint32_t)tk >= a.Nz) return;
    const int64_t sbase = (int64_t)tk * Nyx;
    const uint32_t nacc = ((uint32_t)tk + 1u < a.Nz) ? 2u * Nyx : Nyx;
    float* R = lds + 2 * Nyx;
    if (!SENS) {
        const float4* src = reinterpret_cast<const float4*>(img + sbase);
        float4* dst = reinterpret_cast<float4*>(lds);
        for (uint32_t i = threadIdx.x; i < Nyx / 4; i += TPB) dst[i] = src[i];
    }
    __syncthreads();
    for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
        const int64_t lo = base + l;
        const float temp = t.temp[lo];
        float r = 0.f;
        if (SENS) {
            r = temp != 0.f ? 1.f : 0.f;
        } else if (temp != 0.f) {
            const int32_t c = t.count[lo];
            const int64_t p = t.base[lo >> 6] + (lo & 63);
            float ax = 0.f;
            for (int32_t k = 0; k < c; k++) {
                const uint32_t i = t.ind[p + 64 * k];
                const uint64_t j = (uint64_t)((int64_t)i - sbase);
                ax += t.w[p + 64 * k] * (j < Nyx ? lds[j] : img[i]);
            }
            r = meas[lo] / ratio_den(a, lo, ax * temp, epps);
            if (isnan(r) || isinf(r)) r = epps;
        }
        R[l - l0] = r;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nacc; i += TPB) lds[i] = 0.f;
    __syncthreads();
    for (uint32_t l = l0 + threadIdx.x; l < l1; l += TPB) {
        const float r = R[l - l0];
        if (r == 0.f) continue;
        const int64_t lo = base + l;
        const float temp = t.temp[lo];
        const int32_t c = t.count[lo];
        const int64_t p = t.base[lo >> 6] + (lo & 63);
        for (int32_t k = 0; k < c; k++) {
            const uint32_t i = t.ind[p + 64 * k];
            const float e = t.w[p + 64 * k] * temp;
            const float v = SENS ? e : r * e;
            const uint64_t j = (uint64_t)((int64_t)i - sbase);
            if (j < nacc) __hip_atomic_fetch_add(lds + j, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else unsafeAtomicAdd(rhs + i, v);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nacc; i += TPB) {
        const float v = lds[i];
        if (v != 0.f) unsafeAtomicAdd(rhs + sbase + i, v);
    }
}

static OrthArgs make_orth_args(const omg_geom* g, int64_t first, int64_t n) {
    OrthArgs oa;
    oa.a = make_sid_args(g, first, n);
    // b + N*d in double, rounded once (init.py:727); the Siddon path keeps the CPU form.
    oa.a.bmaxx = (float)((double)g->d.bx + (double)g->d.Nx * (double)g->d.dx);
    oa.a.bmaxy = (float)((double)g->d.by + (double)g->d.Ny * (double)g->d.dy);
    oa.a.bmaxz = (float)((double)g->d.bz + (double)g->d.Nz * (double)g->d.dz);
    oa.width = g->d.orth_width;
    oa.cxy = (g->d.orth_flags & 1) != 0;
    oa.cz = (g->d.orth_flags & 2) != 0;
    oa.vol = g->vol_dev;
    oa.vol_len = (uint32_t)g->d.vol_len;
    oa.bmin = g->d.vol_bmin;
    oa.bmax = g->d.vol_bmax;
    oa.vmax = g->d.vol_vmax;
    return oa;
}

}  // namespace omg

using namespace omg;

// The planar kernels apply when the sinogram is direct-plane, the slice fits the LDS accumulator and,
// with the axial sweep, the centres of the slices next to every plane's slice lie at least
// 0.99 * width * (1 + 1e-3) from the plane, so no voxel of another slice passes the tube cut-off
// (a voxel's distance to an in-plane LOR is at least its axial offset).
static bool orth_planar_check(const omg_geom* g) {
    const omg_geom_desc& d = g->d;
    if (!g->planar || d.mode != OMG_MODE_PET_SINO || d.Nx % 4 != 0) return false;
    const size_t Nyx = (size_t)d.Nx * d.Ny;
    if (Nyx * 8 + sizeof(float) * (kPlanarTPB + kPlanarTPB / 64) > 160 * 1024) return false;
    if (!(d.orth_flags & 2)) return true;
    const double cut = (1.0 - (double)kOrthThr) * (double)d.orth_width * (1.0 + 1e-3);
    for (uint32_t p = 0; p < d.n_proj; p++) {
        const float zs = g->z_host[2 * (size_t)p];
        const int32_t tk = (int32_t)(fabsf(zs - d.bz) / d.dz);
        if (tk < 0 || (uint32_t)tk >= d.Nz) continue;
        for (int k = tk - 1; k <= tk + 1; k += 2) {
            if (k < 0 || (uint32_t)k >= d.Nz) continue;
            const float c = d.bz + (float)k * d.dz + d.dz / 2.f;   // the kernel's slice centre
            if (std::fabs((double)c - (double)zs) < cut) return false;
        }
    }
    return true;
}

struct TubeEntry;
static int tube_lists(const omg_geom* gc, int64_t first, int64_t n, hipStream_t s, const TubeEntry** out);
static TubeView tube_view(const TubeEntry* e);

template <bool FP, bool BP, bool SENS>
static int orth_launch_planar(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* yin,
                              float* yout, float* rhs, float* sens, float epps, hipStream_t s,
                              const float* add = nullptr) {
    const PlanarShape sh = planar_shape(g, n);
    const uint32_t Nyx = g->d.Nx * g->d.Ny;
    const size_t lds = BP ? (size_t)Nyx * 8 + sizeof(float) * ((size_t)sh.chunk + kPlanarTPB / 64)
                          : sizeof(float) * (size_t)Nyx;
    const TubeEntry* e = nullptr;   // the kept tube lists of the range, when there are
    if (int rc = tube_lists(g, first, n, s, &e)) return rc;
    const TubeView tv = e ? tube_view(e) : TubeView{};
    auto kern = e ? k_orth_planar<kPlanarTPB, FP, BP, SENS, true> : k_orth_planar<kPlanarTPB, FP, BP, SENS, false>;
    static size_t attr_bytes[2] = {0, 0};
    if (lds > attr_bytes[e ? 1 : 0]) {
        OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_bytes[e ? 1 : 0] = lds;
    }
    OrthArgs oa = make_orth_args(g, first, n);
    oa.a.add = add;
    float* part = nullptr;
    float* part2 = nullptr;
    if (BP) {
        const int rc = scratch_reserve(const_cast<omg_geom*>(g), (size_t)sh.blocks * Nyx * (SENS ? 2 : 1));
        if (rc) return rc;
        part = scratch_ptr(g);
        part2 = SENS ? scratch_ptr(g) + (size_t)sh.blocks * Nyx : nullptr;
    }
    hipEvent_t stop;
    timing_begin(s, &stop);
    hipLaunchKernelGGL(kern, dim3(sh.blocks), dim3(kPlanarTPB), lds, s, oa, tv, sh.n_planes, sh.chunk, img, yin, yout,
                       part, part2, epps);
    timing_end(s, stop);
    OMG_CHECK_LAUNCH();
    if (BP && rhs) return planar_reduce_partials(g, first, n, rhs, SENS ? sens : nullptr, s);
    return OMG_OK;
}

// Fused planar FP -> ratio -> BP of [first, first + n) into the partial slices of scratch_ptr(g) (the caller
// applies planar_update). Returns OMG_ERR_ARG when the planar kernels do not apply.
int orth_planar_fused(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas,
                      int32_t compute_sens, float epps, hipStream_t s, const float* add) {
    OMG_REQUIRE(g->orth_planar, "the orthogonal planar kernels do not apply to this geometry");
    if (compute_sens)
        return orth_launch_planar<true, true, true>(g, first, n, img, meas, nullptr, nullptr, nullptr, epps, s, add);
    return orth_launch_planar<true, true, false>(g, first, n, img, meas, nullptr, nullptr, nullptr, epps, s, add);
}

int orth_prepare(omg_geom* g) {
    const omg_geom_desc& d = g->d;
    OMG_REQUIRE(d.mode == OMG_MODE_PET_SINO || d.mode == OMG_MODE_LIST6,
                "projector_type %d is provided for PET sinogram and list-mode data", d.projector_type);
    OMG_REQUIRE(d.Nx == d.Ny, "projector_type %d needs Nx == Ny (the reference kernel's axis-parallel branch "
                              "indexes rows with the other axis' extent)", d.projector_type);
    OMG_REQUIRE(d.orth_width > 0.f && std::isfinite(d.orth_width),
                "orth_width must be a positive tube width (projector 2) or radius (projector 3)");
    OMG_REQUIRE((d.orth_flags & ~3) == 0, "orth_flags holds bit 0 (transaxial sweep) and bit 1 (axial sweep) only");
    // volume weights exceed 1, outside the planar kernels' fixed-point bound: one lane per LOR, and the
    // float-accumulator slice kernel for the fused step on direct-plane sinograms
    g->orth_planar = d.projector_type == OMG_PROJ_ORTH && orth_planar_check(g);
    g->vol_planar = d.projector_type == OMG_PROJ_VOLUME && g->planar && d.mode == OMG_MODE_PET_SINO && d.Nx % 4 == 0;
    g->tube_planar = g->planar && d.mode == OMG_MODE_PET_SINO && d.Nx % 4 == 0;
    return OMG_OK;
}

// ---- the tube-list cache -----------------------------------------------------------------------------
struct TubeEntry {
    int64_t first = 0, n = 0;
    bool cached = false;   // false: over the budget (kept so the range is not counted again)
    uint32_t* ind = nullptr;
    float* w = nullptr;
    int64_t* base = nullptr;
    int32_t* count = nullptr;
    float* temp = nullptr;
    size_t bytes = 0;
};
struct TubeCache {
    std::vector<TubeEntry*> e;
    size_t bytes = 0;
};

static void tube_free(TubeEntry* e) {
    hipFree(e->ind);
    hipFree(e->w);
    hipFree(e->base);
    hipFree(e->count);
    hipFree(e->temp);
    delete e;
}

static size_t tube_budget(const omg_geom* g) {
    const int32_t r = g->d.reserved[4];
    if (r < 0) return 0;
    return (r == 0 ? (size_t)16384 : (size_t)r) << 20;
}

// *out: the lists of [first, first + n) of a projector-2 or -3 geometry, built on the first call for the range
// (two tracing passes: counts, then the lists); nullptr when they do not apply (budget, a no-allocation
// call, stream capture, a one-lane-per-LOR routing) or would not fit the budget.
static int tube_lists(const omg_geom* gc, int64_t first, int64_t n, hipStream_t s, const TubeEntry** out) {
    *out = nullptr;
    omg_geom* g = const_cast<omg_geom*>(gc);
    if (!is_tube(g) || g->force_generic || fixed_region_active() || n <= 0) return OMG_OK;
    const size_t budget = tube_budget(g);
    if (budget == 0) return OMG_OK;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return OMG_OK;
    if (!g->tubes) g->tubes = new TubeCache();
    for (TubeEntry* e : g->tubes->e)
        if (e->first == first && e->n == n) {
            *out = e->cached ? e : nullptr;
            return OMG_OK;
        }
    TubeEntry* e = new TubeEntry();
    e->first = first;
    e->n = n;
    g->tubes->e.push_back(e);
    const OrthArgs oa = make_orth_args(g, first, n);
    const int64_t blocks = (n + 63) / 64;
    if (hipMalloc(reinterpret_cast<void**>(&e->count), (size_t)n * sizeof(int32_t)) != hipSuccess) {
        set_error("could not allocate %lld tube counts", (long long)n);
        return OMG_ERR_HIP;
    }
    const bool vol = oa.vol != nullptr;
    hipLaunchKernelGGL(vol ? k_orth_trace<true> : k_orth_trace<false>, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock),
                       0, s, oa, e->count, nullptr, nullptr, nullptr);
    OMG_CHECK_LAUNCH();
    std::vector<int32_t> cnt((size_t)n);
    OMG_CHECK_HIP(hipMemcpyAsync(cnt.data(), e->count, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OMG_CHECK_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> base((size_t)blocks);
    int64_t total = 0;
    for (int64_t b = 0; b < blocks; b++) {
        int32_t m = 0;
        for (int64_t l = b * 64; l < std::min(n, b * 64 + 64); l++) m = std::max(m, cnt[(size_t)l]);
        base[(size_t)b] = total;
        total += 64 * (int64_t)m;
    }
    e->bytes = (size_t)total * 8 + (size_t)blocks * 8 + (size_t)n * 8;
    if (g->tubes->bytes + e->bytes > budget) {   // keep tracing this range: the lists would not fit
        hipFree(e->count);
        e->count = nullptr;
        e->bytes = 0;
        return OMG_OK;
    }
    bool ok = hipMalloc(reinterpret_cast<void**>(&e->ind), (size_t)std::max<int64_t>(total, 1) * 4) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&e->w), (size_t)std::max<int64_t>(total, 1) * 4) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&e->base), (size_t)blocks * 8) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&e->temp), (size_t)n * 4) == hipSuccess;
    if (!ok) {   // device memory short of the budget: trace instead
        (void)hipGetLastError();
        hipFree(e->ind); hipFree(e->w); hipFree(e->base); hipFree(e->temp); hipFree(e->count);
        e->ind = nullptr; e->w = nullptr; e->base = nullptr; e->temp = nullptr; e->count = nullptr;
        e->bytes = 0;
        return OMG_OK;
    }
    OMG_CHECK_HIP(hipMemcpyAsync(e->base, base.data(), (size_t)blocks * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(vol ? k_tube_fill<true> : k_tube_fill<false>, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0,
                       s, oa, e->base, e->count, e->ind, e->w, e->temp);
    OMG_CHECK_LAUNCH();
    OMG_CHECK_HIP(hipStreamSynchronize(s));   // base (a host vector) must outlive the copy
    e->cached = true;
    g->tubes->bytes += e->bytes;
    *out = e;
    return OMG_OK;
}

static TubeView tube_view(const TubeEntry* e) {
    return TubeView{e->ind, e->w, e->base, e->count, e->temp, e->n};
}

// k_orth_vol_planar's launch, or OMG_ERR_UNSUPPORTED when its LDS would not fit.
static int vol_launch_planar(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas,
                             float* rhs, float epps, hipStream_t s, const float* add) {
    const PlanarShape sh = planar_shape(g, n);
    const size_t Nyx = (size_t)g->d.Nx * g->d.Ny;
    const size_t lds = sizeof(float) * (2 * Nyx + 2 * (size_t)sh.chunk);
    if (lds > 160 * 1024) return OMG_ERR_UNSUPPORTED;
    if (sh.blocks == 0) return OMG_OK;
    auto kern = k_orth_vol_planar<kPlanarTPB>;
    static size_t attr = 0;
    if (lds > attr) {
        OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = lds;
    }
    OrthArgs oa = make_orth_args(g, first, n);
    oa.a.add = add;
    hipLaunchKernelGGL(kern, dim3(sh.blocks), dim3(kPlanarTPB), lds, s, oa, sh.n_planes, sh.chunk, img, meas, rhs,
                       epps);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

void orth_release(omg_geom* g) {
    if (!g->tubes) return;
    for (TubeEntry* e : g->tubes->e) tube_free(e);
    delete g->tubes;
    g->tubes = nullptr;
}

// The tube-list form of k_orth_vol_planar (its LDS: 2 Nyx + chunk floats); with sens, a second launch
// accumulates the sensitivity image the same way.
static int tube_launch_planar(const omg_geom* g, const TubeEntry* e, int64_t first, int64_t n, const float* img,
                              const float* meas, float* rhs, float* sens, float epps, hipStream_t s, const float* add) {
    const PlanarShape sh = planar_shape(g, n);
    const size_t Nyx = (size_t)g->d.Nx * g->d.Ny;
    const size_t lds = sizeof(float) * (2 * Nyx + (size_t)sh.chunk);
    if (lds > 160 * 1024) return OMG_ERR_UNSUPPORTED;
    if (sh.blocks == 0) return OMG_OK;
    auto kern = k_tube_vol_planar<kPlanarTPB, false>;
    auto kern_s = k_tube_vol_planar<kPlanarTPB, true>;
    static size_t attr = 0;
    if (lds > attr) {
        OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern_s),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = lds;
    }
    SidArgs a = make_sid_args(g, first, n);
    a.add = add;
    hipLaunchKernelGGL(kern, dim3(sh.blocks), dim3(kPlanarTPB), lds, s, tube_view(e), a, sh.n_planes, sh.chunk, img,
                       meas, rhs, epps);
    OMG_CHECK_LAUNCH();
    if (sens) {
        hipLaunchKernelGGL(kern_s, dim3(sh.blocks), dim3(kPlanarTPB), lds, s, tube_view(e), a, sh.n_planes, sh.chunk,
                           img, meas, sens, epps);
        OMG_CHECK_LAUNCH();
    }
    return OMG_OK;
}

// LDS of the TOF kernels: [bins: accumulators / measurements / ratios][bins: weights] x 64 lanes.
static size_t orth_tof_lds(const omg_geom* g) { return 2 * (size_t)g->d.tof_bins * kOrthTofBlock * sizeof(float); }

int orth_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s) {
    if (g->d.tof_bins > 1) {
        const OrthArgs oa = make_orth_args(g, first, n);
        auto* kern = oa.vol ? k_orth_tof_fp<true> : k_orth_tof_fp<false>;
        hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthTofBlock)), dim3(kOrthTofBlock), orth_tof_lds(g), s, oa, img,
                           y);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    if (g->orth_planar && !g->force_generic)
        return orth_launch_planar<true, false, false>(g, first, n, img, nullptr, y, nullptr, nullptr, 0.f, s);
    const TubeEntry* e = nullptr;
    if (int rc = tube_lists(g, first, n, s, &e)) return rc;
    if (e) {
        hipLaunchKernelGGL(k_tube_fp, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, tube_view(e), img, y);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    const OrthArgs oa = make_orth_args(g, first, n);
    auto* kern = oa.vol ? k_orth_fp<true> : k_orth_fp<false>;
    hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, oa, img, y);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int orth_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens,
                  hipStream_t s) {
    if (g->d.tof_bins > 1) {
        const OrthArgs oa = make_orth_args(g, first, n);
        auto* kern = oa.vol ? k_orth_tof_bp<true> : k_orth_tof_bp<false>;
        hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthTofBlock)), dim3(kOrthTofBlock), orth_tof_lds(g), s, oa, yin,
                           rhs, sens);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    if (g->orth_planar && !g->force_generic) {
        if (sens) return orth_launch_planar<false, true, true>(g, first, n, nullptr, yin, nullptr, rhs, sens, 0.f, s);
        return orth_launch_planar<false, true, false>(g, first, n, nullptr, yin, nullptr, rhs, nullptr, 0.f, s);
    }
    const TubeEntry* e = nullptr;
    if (int rc = tube_lists(g, first, n, s, &e)) return rc;
    if (e) {
        hipLaunchKernelGGL(k_tube_bp, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, tube_view(e), yin, rhs,
                           sens);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    const OrthArgs oa = make_orth_args(g, first, n);
    auto* kern = oa.vol ? k_orth_bp<true> : k_orth_bp<false>;
    hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, oa, yin, rhs, sens);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int orth_osem(const omg_geom* g, int64_t first, int64_t n, const float* img, const float* meas, float* rhs,
              float* sens, float epps, hipStream_t s, const float* add) {
    if (g->d.tof_bins > 1) {
        OrthArgs oa = make_orth_args(g, first, n);
        oa.a.add = add;
        auto* kern = oa.vol ? k_orth_tof_osem<true> : k_orth_tof_osem<false>;
        hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthTofBlock)), dim3(kOrthTofBlock), orth_tof_lds(g), s, oa, img,
                           meas, rhs, sens, epps);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    if (g->orth_planar && !g->force_generic) {
        if (sens) return orth_launch_planar<true, true, true>(g, first, n, img, meas, nullptr, rhs, sens, epps, s, add);
        return orth_launch_planar<true, true, false>(g, first, n, img, meas, nullptr, rhs, nullptr, epps, s, add);
    }
    const TubeEntry* e = nullptr;
    if (int rc = tube_lists(g, first, n, s, &e)) return rc;
    if (e) {
        if (g->tube_planar) {
            const int rc = tube_launch_planar(g, e, first, n, img, meas, rhs, sens, epps, s, add);
            if (rc != OMG_ERR_UNSUPPORTED) return rc;
        }
        SidArgs a = make_sid_args(g, first, n);
        a.add = add;
        hipLaunchKernelGGL(k_tube_osem, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, tube_view(e), a, img,
                           meas, rhs, sens, epps);
        OMG_CHECK_LAUNCH();
        return OMG_OK;
    }
    if (g->vol_planar && !sens && !g->force_generic) {
        const int rc = vol_launch_planar(g, first, n, img, meas, rhs, epps, s, add);
        if (rc != OMG_ERR_UNSUPPORTED) return rc;
    }
    OrthArgs oa = make_orth_args(g, first, n);
    oa.a.add = add;
    auto* kern = oa.vol ? k_orth_osem<true> : k_orth_osem<false>;
    hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, oa, img, meas, rhs, sens,
                       epps);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int orth_trace(const omg_geom* g, int64_t first, int64_t n, int32_t* counts, const int64_t* offsets, uint32_t* idx,
               float* w, hipStream_t s) {
    const OrthArgs oa = make_orth_args(g, first, n);
    auto* kern = oa.vol ? k_orth_trace<true> : k_orth_trace<false>;
    hipLaunchKernelGGL(kern, dim3(grid_for(n, kOrthBlock)), dim3(kOrthBlock), 0, s, oa, counts, offsets, idx,
                       w);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}
