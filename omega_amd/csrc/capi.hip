// capi.hip -- geometry handles, error text and version of libomega_hip.so.
#include "geom.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

namespace omg {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_range(const omg_geom* g, int64_t first, int64_t n) {
    OMG_REQUIRE(first >= 0 && n >= 0, "negative measurement range (first=%lld, n=%lld)", (long long)first, (long long)n);
    if (g->d.mode == OMG_MODE_LIST6) {
        OMG_REQUIRE((first + n) * 6 <= g->d.x_len, "list-mode range [%lld, %lld) exceeds the %lld LORs in the table",
                    (long long)first, (long long)(first + n), (long long)(g->d.x_len / 6));
        return OMG_OK;
    }
    if (g->index_dev) {
        OMG_REQUIRE(first + n <= g->d.meas_index_len, "measurement range [%lld, %lld) exceeds the %lld entries of "
                    "the subset index table", (long long)first, (long long)(first + n), (long long)g->d.meas_index_len);
        OMG_REQUIRE(n <= 0xFFFFFFFFll, "one call handles at most 2^32-1 measurements");
        return OMG_OK;
    }
    const int64_t per = (int64_t)g->d.size_x * g->d.size_y;
    OMG_REQUIRE(per > 0, "detector size is zero");
    OMG_REQUIRE(n % per == 0, "measurement count %lld is not a whole number of %lld-bin planes", (long long)n,
                (long long)per);
    OMG_REQUIRE(n <= 0xFFFFFFFFll, "one call handles at most 2^32-1 measurements");
    OMG_REQUIRE(first + n / per <= (int64_t)g->d.n_proj, "planes [%lld, %lld) exceed the %u planes in the tables",
                (long long)first, (long long)(first + n / per), g->d.n_proj);
    return OMG_OK;
}

static thread_local const FixedRegion* t_fixed = nullptr;

FixedRegionGuard::FixedRegionGuard(const FixedRegion* r) { t_fixed = r; }
FixedRegionGuard::~FixedRegionGuard() { t_fixed = nullptr; }
bool fixed_region_active() { return t_fixed != nullptr; }

// Floats of the work part of a fixed region: [dU: N][rhs: N][slice map: Nz int32], rounded up to 64
// floats (256 B) so the scratch part that follows keeps the region's alignment.
static size_t fixed_work_floats(const omg_geom* g) {
    const size_t w = 2 * (size_t)g->d.Nx * g->d.Ny * g->d.Nz + g->d.Nz;
    return (w + 63) & ~(size_t)63;
}

static int fixed_need(const omg_geom* g, size_t floats, const char* what) {
    const size_t need = fixed_work_floats(g) + floats;
    OMG_REQUIRE(need <= t_fixed->floats, "this call needs %zu floats of work memory (%s), more than the %zu floats "
                "(%zu MiB) of the no-allocation scratch", need, what, t_fixed->floats,
                t_fixed->floats * sizeof(float) >> 20);
    return OMG_OK;
}

float* scratch_ptr(const omg_geom* g) { return t_fixed ? t_fixed->base + fixed_work_floats(g) : g->scratch; }
float* work_ptr(const omg_geom* g) { return t_fixed ? t_fixed->base : g->work; }
float* q_ptr(const omg_geom* g) { return t_fixed ? t_fixed->base + fixed_work_floats(g) : g->vbp_q; }

int scratch_reserve(omg_geom* g, size_t floats) {
    if (t_fixed) return fixed_need(g, floats, "partial images");
    if (g->scratch_floats >= floats) return OMG_OK;
    if (g->scratch) {
        OMG_CHECK_HIP(hipDeviceSynchronize());
        OMG_CHECK_HIP(hipFree(g->scratch));
        g->scratch = nullptr;
        g->scratch_floats = 0;
    }
    OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g->scratch), floats * sizeof(float)));
    g->scratch_floats = floats;
    return OMG_OK;
}

int work_reserve(omg_geom* g, uint32_t N) {
    if (t_fixed) return fixed_need(g, 0, "image work buffers");
    const size_t need = 2 * (size_t)N + g->d.Nz;
    if (g->work_floats >= need) return OMG_OK;
    if (g->work) {
        OMG_CHECK_HIP(hipDeviceSynchronize());
        OMG_CHECK_HIP(hipFree(g->work));
        g->work = nullptr;
        g->work_floats = 0;
    }
    OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g->work), need * sizeof(float)));
    g->work_floats = need;
    return OMG_OK;
}

int q_reserve(omg_geom* g, size_t floats) {
    if (t_fixed) return fixed_need(g, floats, "per-LOR q buffers");
    if (g->vbp_q_floats >= floats) return OMG_OK;
    if (g->vbp_q) {
        OMG_CHECK_HIP(hipDeviceSynchronize());
        OMG_CHECK_HIP(hipFree(g->vbp_q));
        g->vbp_q = nullptr;
        g->vbp_q_floats = 0;
    }
    OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g->vbp_q), floats * sizeof(float)));
    g->vbp_q_floats = floats;
    return OMG_OK;
}

}  // namespace omg

using namespace omg;

extern "C" const char* omg_version(void) { return "omega_hip 0.1 (gfx950)"; }

extern "C" const char* omg_last_error(void) { return g_err; }

static int upload(const float* src, int64_t n, float** dst) {
    *dst = nullptr;
    if (!src || n <= 0) return OMG_OK;
    OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(dst), (size_t)n * sizeof(float)));
    OMG_CHECK_HIP(hipMemcpy(*dst, src, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    return OMG_OK;
}

// Measurements the tables describe: the length of a per-measurement table (norm, LOR attenuation).
static int64_t total_measurements(const omg_geom_desc& d) {
    if (d.mode == OMG_MODE_LIST6) return d.x_len / 6;
    if (d.meas_index) return d.meas_index_len;
    return (int64_t)d.size_x * d.size_y * d.n_proj;
}

// Attenuation image -> one factor per measurement (g->atten_dev), computed once on the device; with
// several rays per LOR one factor per ray and measurement, ray-major.
static int attenuation_prepare(omg_geom* g, const float* mu_host, int64_t n_total) {
    const int64_t N = (int64_t)g->d.Nx * g->d.Ny * g->d.Nz;
    const int64_t rays = (int64_t)std::max(g->d.rays_xy, 1u) * std::max(g->d.rays_z, 1u);
    g->atten_ray_stride = rays > 1 ? n_total : 0;
    float* mu = nullptr;
    int rc = upload(mu_host, N, &mu);
    if (!rc && hipMalloc(reinterpret_cast<void**>(&g->atten_dev), (size_t)(n_total * rays) * sizeof(float)) !=
                   hipSuccess) {
        set_error("could not allocate %lld attenuation factors", (long long)(n_total * rays));
        rc = OMG_ERR_HIP;
    }
    if (!rc) rc = g->d.projector_type == OMG_PROJ_INTERP ? interp_atten_factors(g, mu, g->atten_dev, n_total, nullptr)
                                                         : atten_factors(g, mu, g->atten_dev, n_total, nullptr);
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) {
        set_error("attenuation factor computation failed");
        rc = OMG_ERR_HIP;
    }
    hipFree(mu);
    return rc;
}

int bdd_prepare(omg_geom* g);   // BDD per-geometry set-up (bdd.hip)
int vbp_prepare(omg_geom* g, const float* xt);  // voxel-driven planar backprojection set-up (vbp.hip)
int orth_prepare(omg_geom* g);  // orthogonal projector set-up (orth.hip)

// True when every LOR of a sinogram's transaxial table [xs, ys, xd, yd] either misses the image or
// enters it after t = 1e-3 and leaves it before t = 1 - 1e-3 of the segment source -> detector
// (slab clipping in double against the float image box). Every crossing parameter of such a walk
// then lies in [0, 1], also after the float accumulation of up to Nx + Ny crossings.
static bool ring_table(const omg_geom_desc& d, const float* xt, float bmaxx, float bmaxy) {
    const int64_t n = (int64_t)d.size_x * d.size_y;
    const double lo[2] = {d.bx, d.by}, hi[2] = {bmaxx, bmaxy};
    for (int64_t i = 0; i < n; i++) {
        const double s[2] = {xt[4 * i], xt[4 * i + 1]}, e[2] = {xt[4 * i + 2], xt[4 * i + 3]};
        double t0 = 0.0, t1 = 1.0;
        bool miss = false;
        for (int k = 0; k < 2; k++) {
            const double dd = e[k] - s[k];
            if (dd == 0.0) {
                if (s[k] < lo[k] || s[k] > hi[k]) miss = true;
                continue;
            }
            double ta = (lo[k] - s[k]) / dd, tb = (hi[k] - s[k]) / dd;
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
        }
        if (miss || t0 >= t1) continue;
        if (t0 < 1e-3 || t1 > 1.0 - 1e-3) return false;
    }
    return true;
}

extern "C" int omg_geom_create(const omg_geom_desc* d, omg_geom** out) {
    OMG_REQUIRE(d && out, "null descriptor or output pointer");
    OMG_REQUIRE(d->struct_size == sizeof(omg_geom_desc), "omg_geom_desc size mismatch: caller %u, library %zu",
                d->struct_size, sizeof(omg_geom_desc));
    OMG_REQUIRE(d->Nx > 0 && d->Ny > 0 && d->Nz > 0, "empty image grid");
    OMG_REQUIRE((uint64_t)d->Nx * d->Ny * d->Nz < (1ull << 32), "image grid exceeds 2^32 voxels");
    OMG_REQUIRE(d->dx > 0.f && d->dy > 0.f && d->dz > 0.f, "voxel sizes must be positive");
    OMG_REQUIRE(d->mode == OMG_MODE_PET_SINO || d->mode == OMG_MODE_CT_FLAT || d->mode == OMG_MODE_LIST6,
                "unknown data mode %d", d->mode);
    OMG_REQUIRE(d->projector_type == OMG_PROJ_SIDDON || d->projector_type == OMG_PROJ_ORTH ||
                    d->projector_type == OMG_PROJ_VOLUME || d->projector_type == OMG_PROJ_INTERP ||
                    d->projector_type == OMG_PROJ_BDD,
                "projector_type %d is not provided (1 Siddon, 2 orthogonal, 3 volume of intersection, 4 interpolation, "
                "5 branchless distance-driven)",
                d->projector_type);
    OMG_REQUIRE(d->x && d->x_len > 0, "missing x coordinate table");
    if (d->mode == OMG_MODE_PET_SINO) {
        OMG_REQUIRE(d->x_len >= (int64_t)d->size_x * d->size_y * 4, "x table holds %lld floats, need 4*%u*%u",
                    (long long)d->x_len, d->size_x, d->size_y);
        OMG_REQUIRE(d->z && d->z_len >= (int64_t)d->n_proj * 2, "z table holds %lld floats, need 2*%u planes",
                    (long long)d->z_len, d->n_proj);
    } else if (d->mode == OMG_MODE_CT_FLAT) {
        OMG_REQUIRE(d->x_len >= (int64_t)d->n_proj * 6, "x table holds %lld floats, need 6*%u projections",
                    (long long)d->x_len, d->n_proj);
        OMG_REQUIRE(d->z && d->z_len >= (int64_t)d->n_proj * (d->ct_pitch ? 6 : 2), "z (uV) table too short");
    } else {
        OMG_REQUIRE(d->x_len % 6 == 0, "list-mode x table length must be a multiple of 6");
    }
    if (d->projector_type == OMG_PROJ_BDD)
        OMG_REQUIRE(d->mode == OMG_MODE_CT_FLAT, "projector_type 5 (BDD) supports CT flat-panel data only");
    if (d->projector_type == OMG_PROJ_INTERP) {
        OMG_REQUIRE(std::isfinite(d->interp_dl), "interp_dl must be finite");
        if (d->mode == OMG_MODE_CT_FLAT)
            OMG_REQUIRE(std::isfinite(d->interp_kerroin) && d->interp_kerroin > 0.f && d->dPitchX > 0.f &&
                            d->dPitchY > 0.f,
                        "projector_type 4 on CT data needs interp_kerroin > 0 and the detector pitches dPitchX, dPitchY");
    }
    if (d->projector_type == OMG_PROJ_VOLUME) {
        OMG_REQUIRE(d->vol_V && d->vol_len > 0 && d->vol_len <= 0x7FFFFFFFll,
                    "projector_type 3 needs the volume table vol_V (1 to 2^31 - 1 entries), got %lld",
                    (long long)d->vol_len);
        OMG_REQUIRE(std::isfinite(d->vol_bmin) && d->vol_bmin >= 0.f && std::isfinite(d->vol_bmax) &&
                        d->vol_bmax > d->vol_bmin && std::isfinite(d->vol_vmax) && d->vol_vmax > 0.f,
                    "volume table scalars need 0 <= vol_bmin < vol_bmax and vol_vmax > 0 (got %g, %g, %g)",
                    (double)d->vol_bmin, (double)d->vol_bmax, (double)d->vol_vmax);
    }
    if (d->meas_index) {
        OMG_REQUIRE(d->mode == OMG_MODE_PET_SINO, "meas_index (subset types 1-5) is for PET sinogram data");
        OMG_REQUIRE(d->meas_index_len > 0 && d->meas_index_len <= 0xFFFFFFFFll,
                    "meas_index length %lld outside [1, 2^32)", (long long)d->meas_index_len);
        const uint64_t bins = (uint64_t)d->size_x * d->size_y * d->n_proj;
        for (int64_t i = 0; i < d->meas_index_len; i++)
            OMG_REQUIRE((uint64_t)d->meas_index[i] < bins, "meas_index[%lld] = %u is outside the %llu sinogram bins",
                        (long long)i, d->meas_index[i], (unsigned long long)bins);
    }

    const int64_t n_total = total_measurements(*d);
    if (d->norm)
        OMG_REQUIRE(d->norm_len >= n_total, "norm holds %lld floats, the data %lld measurements",
                    (long long)d->norm_len, (long long)n_total);
    if (d->scat) {
        OMG_REQUIRE(d->mode != OMG_MODE_CT_FLAT, "the multiplicative (scatter) correction applies to PET data, not CT");
        OMG_REQUIRE(d->scat_len >= n_total, "scat holds %lld floats, the data %lld measurements",
                    (long long)d->scat_len, (long long)n_total);
    }
    if (d->tof_bins > 1) {
        OMG_REQUIRE(d->tof_bins <= kTofMax, "tof_bins %u exceeds %u", d->tof_bins, kTofMax);
        OMG_REQUIRE((d->projector_type == OMG_PROJ_SIDDON || d->projector_type == OMG_PROJ_VOLUME ||
                     d->projector_type == OMG_PROJ_INTERP) && d->mode != OMG_MODE_CT_FLAT,
                    "time of flight is provided for the improved Siddon (projector_type 1), volume-of-intersection "
                    "(projector_type 3) and interpolation (projector_type 4) projectors on PET data");
        OMG_REQUIRE(d->tof_sigma > 0.f && std::isfinite(d->tof_sigma), "tof_sigma must be a positive standard deviation");
        OMG_REQUIRE(d->tof_centers != nullptr, "tof_bins %u without tof_centers", d->tof_bins);
        for (uint32_t b = 0; b < d->tof_bins; b++)
            OMG_REQUIRE(std::isfinite(d->tof_centers[b]), "tof_centers[%u] is not finite", b);
        OMG_REQUIRE(d->tof_epps >= 0.f, "tof_epps must be >= 0");
    }
    const uint64_t rays = (uint64_t)std::max(d->rays_xy, 1u) * std::max(d->rays_z, 1u);
    if (rays > 1) {
        OMG_REQUIRE(rays <= 256, "%llu rays per LOR exceeds 256", (unsigned long long)rays);
        OMG_REQUIRE((d->projector_type == OMG_PROJ_SIDDON || d->projector_type == OMG_PROJ_INTERP) &&
                        d->mode != OMG_MODE_CT_FLAT,
                    "several rays per LOR are provided for the improved Siddon (projector_type 1) and interpolation "
                    "(projector_type 4) projectors on PET data");
        OMG_REQUIRE(std::isfinite(d->cr_xy) && std::isfinite(d->cr_z), "cr_xy and cr_z must be finite");
    }
    OMG_REQUIRE(d->atten_mode >= 0 && d->atten_mode <= 2, "atten_mode %d is not 0 (off), 1 (image) or 2 (per LOR)",
                d->atten_mode);
    if (d->atten_mode) {
        OMG_REQUIRE(d->mode != OMG_MODE_CT_FLAT, "attenuation correction applies to PET data, not CT");
        OMG_REQUIRE(d->atten != nullptr, "atten_mode %d without an atten table", d->atten_mode);
        if (d->atten_mode == 1)
            OMG_REQUIRE(d->atten_len == (int64_t)d->Nx * d->Ny * d->Nz,
                        "attenuation image holds %lld floats, the image grid %llu voxels", (long long)d->atten_len,
                        (unsigned long long)d->Nx * d->Ny * d->Nz);
        else
            OMG_REQUIRE(d->atten_len >= n_total, "attenuation table holds %lld floats, the data %lld measurements",
                        (long long)d->atten_len, (long long)n_total);
    }

    omg_geom* g = new (std::nothrow) omg_geom();
    OMG_REQUIRE(g != nullptr, "out of host memory");
    g->d = *d;
    g->force_generic = d->reserved[0] != 0;
    g->planar_groups = d->reserved[1];
    hipGetDevice(&g->device);
    // far FOV edges in the reference's float arithmetic (projector_functions.h:1331-1333)
    g->bmaxz = d->bz + (float)d->Nz * d->dz;
    g->bmaxy = (float)d->Ny * d->dy + d->by;
    g->bmaxx = (float)d->Nx * d->dx + d->bx;
    int rc = upload(d->x, d->x_len, &g->x_dev);
    if (!rc) rc = upload(d->z, d->z_len, &g->z_dev);
    if (!rc) rc = upload(d->norm, d->norm_len, &g->norm_dev);
    if (!rc && d->meas_index) {
        const size_t bytes = (size_t)d->meas_index_len * sizeof(uint32_t);
        if (hipMalloc(reinterpret_cast<void**>(&g->index_dev), bytes) != hipSuccess ||
            hipMemcpy(g->index_dev, d->meas_index, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("could not copy the %lld-entry subset index table to the device", (long long)d->meas_index_len);
            rc = OMG_ERR_HIP;
        }
    }
    if (!rc && d->scat) rc = upload(d->scat, n_total, &g->scat_dev);
    if (!rc && d->tof_bins > 1) rc = upload(d->tof_centers, d->tof_bins, &g->tof_c);
    if (!rc && d->atten_mode == 2) rc = upload(d->atten, n_total, &g->atten_dev);
    if (!rc && d->atten_mode == 1) rc = attenuation_prepare(g, d->atten, n_total);
    if (!rc && d->mode == OMG_MODE_PET_SINO) {
        g->z_host.assign(d->z, d->z + d->z_len);
        g->planar = d->meas_index == nullptr && d->tof_bins <= 1 && rays <= 1;
        for (uint32_t p = 0; p < d->n_proj && g->planar; p++) g->planar = d->z[2 * p] == d->z[2 * p + 1];
    }
    if (!rc && g->planar) g->ring = ring_table(*d, d->x, g->bmaxx, g->bmaxy);
    if (!rc && d->mode == OMG_MODE_PET_SINO) rc = vbp_prepare(g, d->x);
    if (!rc && d->projector_type == OMG_PROJ_BDD) rc = bdd_prepare(g);
    if (!rc && d->projector_type == OMG_PROJ_VOLUME) rc = upload(d->vol_V, d->vol_len, &g->vol_dev);
    if (!rc && (d->projector_type == OMG_PROJ_ORTH || d->projector_type == OMG_PROJ_VOLUME)) rc = orth_prepare(g);
    g->d.x = nullptr; g->d.z = nullptr; g->d.norm = nullptr; g->d.meas_index = nullptr; g->d.atten = nullptr;
    g->d.scat = nullptr; g->d.tof_centers = nullptr; g->d.vol_V = nullptr;
    if (rc) {
        omg_geom_destroy(g);
        return rc;
    }
    *out = g;
    return OMG_OK;
}

void bdd_release(omg_geom* g);
void orth_release(omg_geom* g);
void vbp_release(omg_geom* g);

extern "C" int omg_geom_destroy(omg_geom* g) {
    if (!g) return OMG_OK;
    hipFree(g->x_dev);
    hipFree(g->z_dev);
    hipFree(g->norm_dev);
    hipFree(g->atten_dev);
    hipFree(g->scat_dev);
    hipFree(g->vol_dev);
    hipFree(g->tof_c);
    hipFree(g->index_dev);
    hipFree(g->frames_t);
    hipFree(g->scratch);
    hipFree(g->work);
    bdd_release(g);
    orth_release(g);
    vbp_release(g);
    delete g;
    return OMG_OK;
}
