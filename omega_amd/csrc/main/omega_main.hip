// omega_main.hip -- the built-in reconstruction entry point omegaMain (include/omega_main.h).
//
// Host driver of the reference's reconstructionAF (source/cpp/reconstructionAF.h:56-1568) for the
// algorithms and data this library provides, built on libomega_hip's C ABI (geometry handle,
// forward / backward projection, RDP gradient, the OSEM / OSL / PDHG update kernels, dot products)
// plus the few elementwise kernels below that only this driver uses (the measurement-domain
// forward step of each algorithm, PoissonUpdate, the BSREM MAP step, the image preconditioner).
//
// Compiled three times with OMG_MAIN_MEAS_TYPE = float, uint16_t, uint8_t into
// libomega_matrixfree{,_uint16,_uint8}.so, each linked against libomega_hip.so.
//
// Provided: PET sinograms (whole-plane subsets, subset types 8-11, and measurement-level subsets
// through xy_index / z_index, types 1-7), list-mode PET (event coordinates in x) and flat-panel CT;
// projector types 1-5 with the corrections libomega_hip accepts (normalisation, attenuation image
// or factors, additional correction, randoms, TOF, several rays); algorithms OSEM, OSL-OSEM, BSREM,
// MBSREM, PKMA, PDHG (least-squares data term) and CGLS; the RDP prior; image preconditioners
// 0 and 1. Everything else the block can select is refused with a message before any device work.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/omega_hip.h"
#include "../../../include/omega_main.h"

typedef OMG_MAIN_MEAS_TYPE MeasT;

// The block is read with the reference's 1-byte packing; these pin a few offsets of that layout
// (the full layout is checked from Python against the reference front end's field list).
static_assert(sizeof(bool) == 1, "bool members are one byte");
static_assert(offsetof(inputStruct, n_rays_transaxial) == 3, "packed layout");
static_assert(offsetof(inputStruct, projector_type) == 7, "packed layout");
static_assert(offsetof(inputStruct, nProjections) == 231, "packed layout");
static_assert(offsetof(inputStruct, tau) == 247, "packed layout");
static_assert(offsetof(inputStruct, OSEM) == 498, "packed layout");
static_assert(offsetof(inputStruct, mDim) == 550, "packed layout");
static_assert(offsetof(inputStruct, x) == 686, "packed layout");
static_assert(sizeof(inputStruct) == 1346, "packed layout");

namespace {

constexpr int kB = 256;

struct MainError : std::runtime_error {
    explicit MainError(const std::string& m) : std::runtime_error(m) {}
};

[[noreturn]] void fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw MainError(buf);
}

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) fail("%s failed: %s", what, hipGetErrorString(e));
}

void omg_ok(int rc, const char* what) {
    if (rc != OMG_OK) fail("%s failed: %s", what, omg_last_error());
}

// Device array of floats (or doubles) owned by the driver.
template <typename T>
struct Dev {
    T* p = nullptr;
    size_t n = 0;
    Dev() = default;
    explicit Dev(size_t count) { alloc(count); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    Dev(Dev&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    Dev& operator=(Dev&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~Dev() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) hip_ok(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count) {
        if (count != n) alloc(count);
        if (count) hip_ok(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void download(T* h, size_t count) const {
        if (count) hip_ok(hipMemcpy(h, p, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
};

unsigned blocks(int64_t n) {
    int64_t b = (n + kB - 1) / kB;
    if (b > 65536) b = 65536;
    return (unsigned)(b > 0 ? b : 1);
}

#define GRID_LOOP(i, n) for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kB)

__global__ void k_fill(float* a, int64_t n, float v) {
    GRID_LOOP(i, n) a[i] = v;
}

// The measurement-domain step of computeForwardStep (source/cpp/computeForwardProject.h:5-409) for the
// Poisson algorithms, in place on v = A_k x (TOF: bins blocks of n_lor values). Emission data:
// v = v + epps, then v = v + add[LOR] (randoms; :20-21,66-69); CT: v < 0 -> 0. Then by `kind`:
//   0 EM (OSEM, OSL):        PET y / v                        CT exp(-v)                     (:76-95)
//   1 BSREM / MBSREM:        PET y / v - 1                    CT exp(-v) flat - y            (:97-125,153-185)
//   2 PKMA:                  PET 1 - y / v                    CT y - exp(-v) flat            (:126-152)
// and NaN / Inf -> epps (:400-401). The exponential is evaluated in double and rounded once.
__global__ void k_forward_step(int kind, int ct, const float* __restrict__ y, float* __restrict__ v,
                               const float* __restrict__ add, int64_t n, int64_t n_lor, float epps, float flat) {
    GRID_LOOP(i, n) {
        float a = v[i];
        float r;
        if (!ct) {
            a = a + epps;
            if (add) a = a + add[i % n_lor];
            const float q = y[i] / a;
            r = kind == 0 ? q : (kind == 1 ? q - 1.f : 1.f - q);
        } else {
            if (a < 0.f) a = 0.f;
            const float e = (float)exp(-(double)a);
            r = kind == 0 ? e : (kind == 1 ? e * flat - y[i] : y[i] - e * flat);
        }
        if (isnan(r) || isinf(r)) r = epps;
        v[i] = r;
    }
}

// PDHG dual step, least-squares data term (computeForwardProject.h:227-316): res = v - y;
// p' = (p + sigma res) / (1 + sigma); v = p' - p with subsets (:377-381), p' without; p = p'; NaN/Inf -> epps.
__global__ void k_pdhg_dual(const float* __restrict__ y, float* __restrict__ v, float* __restrict__ p, int64_t n,
                            float sigma, int subsets, float epps) {
    GRID_LOOP(i, n) {
        const float res = v[i] - y[i];
        const float pn = (p[i] + sigma * res) / (1.f + sigma);
        float out = subsets ? pn - p[i] : pn;
        if (isnan(out) || isinf(out)) out = epps;
        p[i] = pn;
        v[i] = out;
    }
}

// The start of every emission forward step (computeForwardProject.h:20-21,66-69): v = v + epps, then
// v = v + add[LOR] (randoms).
__global__ void k_shift(float* __restrict__ v, const float* __restrict__ add, int64_t n, int64_t n_lor, float epps) {
    GRID_LOOP(i, n) {
        float a = v[i] + epps;
        if (add) a = a + add[i % n_lor];
        v[i] = a;
    }
}

__global__ void k_sub(float* __restrict__ v, const float* __restrict__ y, int64_t n) {
    GRID_LOOP(i, n) v[i] = v[i] - y[i];
}

// CGLS forward step (computeForwardProject.h:197-206): v = r - alpha v; r = v; NaN/Inf -> epps.
__global__ void k_cgls_r(float* __restrict__ v, float* __restrict__ r, int64_t n, float alpha, float epps) {
    GRID_LOOP(i, n) {
        float o = r[i] - alpha * v[i];
        r[i] = o;
        if (isnan(o) || isinf(o)) o = epps;
        v[i] = o;
    }
}

// OSEM / OSL-OSEM on CT data (subiterStep.h:117-118,223-224): x = x / (S [+ dU]) * (flat * rhs).
__global__ void k_ct_em(float* __restrict__ x, const float* __restrict__ S, const float* __restrict__ dU,
                        const float* __restrict__ rhs, int64_t n, float flat) {
    GRID_LOOP(i, n) {
        const float den = dU ? S[i] + dU[i] : S[i];
        x[i] = x[i] / den * (flat * rhs[i]);
    }
}

// Image-based preconditioners 0 and 1 (functions.hpp:2071-2081): rhs /= D, or rhs *= im / D with im the
// image, for MBSREM the image mirrored at U / 2 (algorithms.h:35-40: im >= U / 2 -> U - im).
__global__ void k_precond(float* __restrict__ rhs, const float* __restrict__ im, const float* __restrict__ D,
                          int64_t n, int type, int mirror, float U) {
    GRID_LOOP(i, n) {
        if (type == 0) {
            rhs[i] = rhs[i] / D[i];
        } else {
            float u = im[i];
            if (mirror && u >= U / 2.f) u = U - u;
            rhs[i] = rhs[i] * (u / D[i]);
        }
    }
}

// PoissonUpdate (source/opencl/auxKernels.cl:1997-2026): 0 PKMA, 1 MBSREM, 2 BSREM.
__global__ void k_poisson_update(float* __restrict__ im, const float* __restrict__ rhs, int64_t n, float lambda,
                                 float epps, float alpha, int pos, int variant) {
    GRID_LOOP(i, n) {
        const float old = im[i];
        float v;
        if (variant == 0) v = old - lambda * rhs[i];
        else if (variant == 1) v = old + lambda * rhs[i];
        else v = old + lambda * rhs[i] * old;
        if (pos) v = fmaxf(epps, v);
        if (variant == 0) {
            im[i] = (1.f - alpha) * old + alpha * v;
        } else {
            if (variant == 1 && v >= alpha) v = alpha - epps;
            im[i] = v;
        }
    }
}

// MAP step of BSREM at the end of an iteration (algorithms.h:128-134): im -= lam im dU; im < epps -> epps.
__global__ void k_map_step(float* __restrict__ im, const float* __restrict__ dU, int64_t n, float lam, float epps) {
    GRID_LOOP(i, n) {
        float v = im[i];
        v -= lam * v * dU[i];
        if (v < epps) v = epps;
        im[i] = v;
    }
}

__global__ void k_clamp(float* __restrict__ x, int64_t n, float epps) {
    GRID_LOOP(i, n) if (x[i] < epps) x[i] = epps;
}

// The full sensitivity image D (reconstructionAF.h:808-817): D /= subsets (when > 1); D == 0 -> 1.
__global__ void k_d_finish(float* __restrict__ D, int64_t n, float subsets) {
    GRID_LOOP(i, n) {
        float v = D[i];
        if (subsets > 1.f) v = v / subsets;
        if (v == 0.f) v = 1.f;
        D[i] = v;
    }
}

enum class Algo { OSEM, OSL, BSREM, MBSREM, PKMA, PDHG, CGLS };

const char* algo_name(Algo a) {
    switch (a) {
        case Algo::OSEM: return "OSEM";
        case Algo::OSL: return "OSL-OSEM";
        case Algo::BSREM: return "BSREM";
        case Algo::MBSREM: return "MBSREM";
        case Algo::PKMA: return "PKMA";
        case Algo::PDHG: return "PDHG";
        case Algo::CGLS: return "CGLS";
    }
    return "?";
}

// Refuses every switch of the block this driver does not provide.
Algo check_supported(const inputStruct& o) {
    struct Flag { bool on; const char* name; };
    const Flag algos_off[] = {
        {o.LSQR, "LSQR"}, {o.SART, "SART"}, {o.FISTA, "FISTA"}, {o.FISTAL1, "FISTAL1"}, {o.MRAMLA, "MRAMLA"},
        {o.RAMLA, "RAMLA"}, {o.ROSEM, "ROSEM"}, {o.RBI, "RBI"}, {o.DRAMA, "DRAMA"}, {o.COSEM, "COSEM"},
        {o.ECOSEM, "ECOSEM"}, {o.ACOSEM, "ACOSEM"}, {o.ROSEM_MAP, "ROSEM_MAP"}, {o.OSL_RBI, "OSL_RBI"},
        {o.OSL_COSEM, "OSL_COSEM"}, {o.SPS, "SPS"}, {o.PDHGKL, "PDHGKL"}, {o.PDHGL1, "PDHGL1"}, {o.PDDY, "PDDY"},
        {o.CV, "CV"}, {o.POCS, "POCS"}, {o.FDK, "FDK"}, {o.SAGA, "SAGA"}, {o.BB, "BB"}};
    for (const Flag& f : algos_off)
        if (f.on) fail("the %s algorithm is not provided by this library", f.name);
    const Flag priors_off[] = {
        {o.MRP, "MRP"}, {o.quad, "quadratic"}, {o.Huber, "Huber"}, {o.L, "L-filter"}, {o.FMH, "FMH"},
        {o.weighted_mean, "weighted mean"}, {o.TV, "TV"}, {o.hyperbolic, "hyperbolic"}, {o.AD, "AD"},
        {o.APLS, "APLS"}, {o.TGV, "TGV"}, {o.NLM, "NLM"}, {o.GGMRF, "GGMRF"}, {o.ProxTV, "ProxTV"},
        {o.ProxRDP, "ProxRDP"}, {o.ProxNLM, "ProxNLM"}, {o.temporalTV, "temporal TV"},
        {o.temporal_smoothness, "temporal smoothness"}, {o.custom, "custom"}};
    for (const Flag& f : priors_off)
        if (f.on) fail("the %s prior is not provided by this library (RDP is)", f.name);
    const Flag feats_off[] = {
        {o.use_psf, "PSF blurring (use_psf)"}, {o.SPECT, "SPECT data"}, {o.largeDim, "largeDim reconstruction"},
        {o.FISTA_acceleration, "FISTA acceleration"}, {o.meanFP, "meanFP"}, {o.meanBP, "meanBP"},
        {o.useMaskFP, "forward projection masks"}, {o.useMaskBP, "backprojection masks"},
        {o.useMultiResolutionVolumes, "multi-resolution volumes"}, {o.deblurring, "deblurring"},
        {o.useEFOV, "extended FOV"}, {o.offsetCorrection, "offset correction"},
        {o.relaxationScaling, "relaxation scaling"},
        {o.computeRelaxationParameters, "computed relaxation parameters"},
        {o.RDPIncludeCorners, "RDP with corners"}, {o.RDP_use_anatomical, "RDP with a reference image"},
        {o.compute_sensitivity_image, "list-mode sensitivity image from the full sinogram"},
        {o.useIndexBasedReconstruction, "index-based list-mode data"},
        {o.stochasticSubsetSelection, "stochastic subset selection"}, {o.useHelical, "helical CT"},
        {o.use_raw_data != 0, "raw detector-pair data"}};
    for (const Flag& f : feats_off)
        if (f.on) fail("%s is not provided by this library", f.name);
    if (o.Nt > 1) fail("dynamic data (Nt = %u) is not provided by omegaMain here", o.Nt);
    if (o.nMultiVolumes > 0) fail("multi-volume reconstruction is not provided by this library");
    if (o.nLayers > 1) fail("multiple crystal layers are not provided by this library");
    if (o.PDAdaptiveType != 0) fail("adaptive PDHG step sizes (PDAdaptiveType %u) are not provided", o.PDAdaptiveType);
    if (o.precondTypeMeas && (o.precondTypeMeas[0] || o.precondTypeMeas[1]))
        fail("measurement-based preconditioners are not provided by this library");
    if (o.precondTypeImage)
        for (int t = 2; t < 7; t++)
            if (o.precondTypeImage[t]) fail("image-based preconditioner type %d is not provided (types 0 and 1 are)", t);

    int n = 0;
    Algo a = Algo::OSEM;
    if (o.OSEM) { a = Algo::OSEM; n++; }
    if (o.OSL_OSEM) { a = Algo::OSL; n++; }
    if (o.BSREM) { a = Algo::BSREM; n++; }
    if (o.MBSREM) { a = Algo::MBSREM; n++; }
    if (o.PKMA) { a = Algo::PKMA; n++; }
    if (o.PDHG) { a = Algo::PDHG; n++; }
    if (o.CGLS) { a = Algo::CGLS; n++; }
    if (n != 1) fail("exactly one reconstruction algorithm must be selected (%d are)", n);
    const bool prec = o.precondTypeImage && (o.precondTypeImage[0] || o.precondTypeImage[1]);
    if (prec && !(a == Algo::MBSREM || a == Algo::PKMA || a == Algo::PDHG))
        fail("image-based preconditioners apply to MBSREM, PKMA and PDHG here");
    if (o.RDP && (a == Algo::OSEM || a == Algo::PDHG || a == Algo::CGLS))
        fail("the RDP prior is used with OSL-OSEM, BSREM, MBSREM and PKMA here, not with %s", algo_name(a));
    if (o.CT && o.randoms_correction) fail("randoms correction of CT data is not provided by this library");
    if (o.listmode && o.TOF) fail("TOF list-mode data is not provided by this library");
    if (a == Algo::MBSREM && o.randoms_correction)
        fail("MBSREM with randoms correction (the epsilon_mramla branch) is not provided by this library");
    if (a == Algo::CGLS && o.subsets > 1) fail("CGLS runs without subsets (subsets = %u)", o.subsets);
    return a;
}

// Per-subset data ranges in the library's call units (first: planes / projections / measurements /
// events, n: measurements).
struct SubsetRange { int64_t first, n; };

class Driver {
public:
    Driver(const inputStruct& o, const MeasT* sino, float* out, float* fp, float* resid)
        : o_(o), sino_(sino), out_(out), fp_out_(fp), resid_(resid) {}
    ~Driver() {
        if (g_) omg_geom_destroy(g_);
    }
    void run();

private:
    const inputStruct& o_;
    const MeasT* sino_;
    float* out_;
    float* fp_out_;
    float* resid_;
    Algo algo_ = Algo::OSEM;
    omg_geom* g_ = nullptr;
    int64_t N_ = 0;
    uint32_t subsets_ = 1;
    int64_t bins_ = 1;
    std::vector<SubsetRange> ranges_;
    std::vector<uint32_t> meas_index_;
    std::vector<Dev<float>> y_;     // per subset: measurements (bins blocks of n)
    std::vector<Dev<float>> add_;   // per subset: randoms (one per LOR) or empty
    std::vector<Dev<float>> S_;     // per subset: sensitivity (EM algorithms)
    Dev<float> x_, rhs_, v_, dU_, D_, tmp_;
    Dev<double> acc_;
    std::vector<Dev<float>> p_;     // PDHG duals per subset
    Dev<float> u_;                  // PDHG uCP
    Dev<float> r_, f_;              // CGLS residual and solution
    float gammaCGLS_ = 0.f, alphaCGLS_ = 0.f;
    size_t saved_ = 0;              // images written to out_
    uint32_t ee_ = 0;               // saveNiter cursor
    size_t fp_pos_ = 0;

    void make_geometry();
    void load_data();
    void compute_D();
    double dot(const float* a, const float* b, int64_t n);
    void prior(float* grad, bool fresh);
    void subset_step(uint32_t iter, uint32_t k);
    void end_of_iteration(uint32_t iter);
    void save_image(const float* dev);
    bool computeReg(int64_t counter, int64_t counterMax) const {
        return o_.regEveryIter <= 1 || counter == 0 || counter == counterMax || (counter % o_.regEveryIter) == 0;
    }
    int64_t n_lor(uint32_t k) const { return ranges_[k].n; }
};

void Driver::make_geometry() {
    const inputStruct& o = o_;
    if (!o.Nx || !o.Ny || !o.Nz || !o.dx || !o.dy || !o.dz || !o.bx || !o.by || !o.bz)
        fail("the image size (Nx, Ny, Nz), voxel size (dx, dy, dz) and FOV corner (bx, by, bz) arrays are required");
    if (!o.x || !o.pituus) fail("the coordinate table x and the subset boundaries pituus are required");
    int pt = (int)o.projector_type;
    if (pt > 10) {
        if (pt / 10 != pt % 10) fail("separate forward / backward projector types (%d) are not provided", pt);
        pt %= 10;
    }
    if (pt < 1 || pt > 5) fail("projector_type %u is not provided (types 1-5 are)", o.projector_type);

    omg_geom_desc d;
    std::memset(&d, 0, sizeof(d));
    d.struct_size = sizeof(d);
    d.projector_type = pt;
    d.mode = o.CT ? OMG_MODE_CT_FLAT : (o.listmode ? OMG_MODE_LIST6 : OMG_MODE_PET_SINO);
    d.Nx = o.Nx[0]; d.Ny = o.Ny[0]; d.Nz = o.Nz[0];
    d.dx = o.dx[0]; d.dy = o.dy[0]; d.dz = o.dz[0];
    d.bx = o.bx[0]; d.by = o.by[0]; d.bz = o.bz[0];
    N_ = (int64_t)d.Nx * d.Ny * d.Nz;
    d.size_x = o.listmode ? 1 : o.nRowsD;
    d.size_y = o.listmode ? 1 : o.nColsD;
    d.n_proj = (uint32_t)o.nProjections;
    d.global_factor = o.global_factor;
    d.x = o.x; d.x_len = (int64_t)o.sizeX;
    if (o.z && o.sizeZ) { d.z = o.z; d.z_len = (int64_t)o.sizeZ; }
    static const float zero2[2] = {0.f, 0.f};
    if (!d.z) { d.z = zero2; d.z_len = 2; }
    if (o.CT) {
        // copyStruct's CT branch (libHeader.h:956-959): the panel pitches are dPitchX / dPitchY
        d.ct_pitch = o.pitch ? 1 : 0;
        d.dPitchZ = o.dPitchY;
        d.dPitchX = o.dPitchX;
        d.dPitchY = o.dPitchY;
    } else {
        // PET (libHeader.h:969-972): the crystal pitches cr_p (transaxial) and cr_pz (axial) spread the rays
        const uint32_t rx = o.n_rays_transaxial > 1 ? o.n_rays_transaxial : 1;
        const uint32_t rz = o.n_rays_axial > 1 ? o.n_rays_axial : 1;
        if (rx * rz > 1) {
            d.rays_xy = rx; d.rays_z = rz;
            d.cr_xy = o.cr_p; d.cr_z = o.cr_pz;
        }
    }
    if (o.normalization_correction) {
        if (!o.norm || !o.sizeNorm) fail("normalization_correction needs norm");
        d.norm = o.norm; d.norm_len = (int64_t)o.sizeNorm;
    }
    if (o.additionalCorrection && !o.CT) {
        if (!o.corrVector || !o.sizeScat) fail("additionalCorrection needs corrVector");
        d.scat = o.corrVector; d.scat_len = (int64_t)o.sizeScat;
    }
    if (o.attenuation_correction && !o.CT) {
        if (!o.atten || !o.sizeAtten) fail("attenuation_correction needs atten");
        d.atten_mode = o.CTAttenuation ? 1 : 2;
        d.atten = o.atten; d.atten_len = (int64_t)o.sizeAtten;
    }
    if (o.TOF && o.TOF_bins > 1) {
        if (!o.TOFCenter) fail("TOF needs TOFCenter");
        d.tof_bins = (uint32_t)o.TOF_bins;
        d.tof_sigma = o.sigma_x;
        d.tof_epps = o.epps;
        d.tof_centers = o.TOFCenter;
        bins_ = o.TOF_bins;
    }
    if (pt == 2) {
        d.orth_width = o.tube_width_z;
        d.orth_flags = (o.orthTransaxial ? 1 : 0) | (o.orthAxial ? 2 : 0);
    } else if (pt == 3) {
        if (!o.V || !o.sizeV) fail("projector_type 3 needs the volume table V");
        d.orth_width = o.tube_radius;
        d.orth_flags = 3;
        d.vol_V = o.V; d.vol_len = (int64_t)o.sizeV;
        d.vol_bmin = o.bmin; d.vol_bmax = o.bmax; d.vol_vmax = o.Vmax;
    } else if (pt == 4) {
        d.interp_dl = o.dL;
        d.interp_kerroin = (o.CT && o.kerroin4) ? o.kerroin4[0] : 0.f;
    }
    // subsets and their ranges
    subsets_ = o.subsets > 1 ? o.subsets : 1;
    const bool meas_subsets = subsets_ > 1 && o.subsetType < 8 && !o.listmode && !o.CT;
    const int64_t per = o.listmode ? 1 : (int64_t)o.nRowsD * o.nColsD;
    if (meas_subsets) {
        if (!o.xy_index || !o.z_index || o.sizeXYind != o.sizeZind)
            fail("subset types 1-7 need xy_index and z_index of equal length");
        if ((int64_t)o.sizeX != 4 * per) fail("x holds %llu floats, expected 4 per sinogram bin (%lld)",
                                             (unsigned long long)o.sizeX, (long long)(4 * per));
        meas_index_.resize(o.sizeXYind);
        for (uint64_t i = 0; i < o.sizeXYind; i++)
            meas_index_[i] = (uint32_t)o.z_index[i] * (uint32_t)per + o.xy_index[i];
        d.meas_index = meas_index_.data();
        d.meas_index_len = (int64_t)meas_index_.size();
    }
    ranges_.resize(subsets_);
    for (uint32_t k = 0; k < subsets_; k++) {
        const int64_t a = o.pituus[k], b = o.pituus[k + 1];
        if (b < a) fail("pituus is not increasing at subset %u", k);
        ranges_[k].first = a;
        ranges_[k].n = (meas_subsets || o.listmode) ? (b - a) : (b - a) * per;
    }
    omg_ok(omg_geom_create(&d, &g_), "omg_geom_create");
}

// Measurement data of each subset to the device: bins blocks of the subset's measurements
// (reconstructionAF.h:1131-1140: Sin[pituus * per + kokoNonTOF * bin]); randoms one per LOR.
void Driver::load_data() {
    const inputStruct& o = o_;
    const bool meas_subsets = subsets_ > 1 && o.subsetType < 8 && !o.listmode && !o.CT;
    const int64_t per = (o.listmode || meas_subsets) ? 1 : (int64_t)o.nRowsD * o.nColsD;
    const int64_t koko = (int64_t)(o.measElem / (uint64_t)bins_);   // measurements per TOF bin
    y_.resize(subsets_);
    add_.resize(subsets_);
    std::vector<float> h;
    for (uint32_t k = 0; k < subsets_; k++) {
        const int64_t n = ranges_[k].n, off = ranges_[k].first * per;
        if (off + n > koko) fail("subset %u reaches measurement %lld of %lld", k, (long long)(off + n), (long long)koko);
        h.resize((size_t)(n * bins_));
        for (int64_t b = 0; b < bins_; b++)
            for (int64_t i = 0; i < n; i++) h[(size_t)(b * n + i)] = (float)sino_[b * koko + off + i];
        y_[k].upload(h.data(), h.size());
        if (o.randoms_correction) {
            if (!o.randoms) fail("randoms_correction needs randoms");
            add_[k].upload(o.randoms + off, (size_t)n);
        }
    }
}

double Driver::dot(const float* a, const float* b, int64_t n) {
    omg_ok(omg_dot(a, b, n, acc_.p, nullptr), "omg_dot");
    double h = 0.0;
    acc_.download(&h, 1);
    return h;
}

// The RDP gradient of the current image accumulated into grad (applySpatialPrior, priors.h:361-460;
// RDPKernel, auxKernels.cl:664-873): beta * dR/dx with the image read as a texture (clamp to edge)
// when useImages is set, as a buffer (zero outside) otherwise. fresh: grad is zeroed first (vec.dU).
void Driver::prior(float* grad, bool fresh) {
    if (fresh) {
        hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, grad, N_, 0.f);
        hip_ok(hipGetLastError(), "fill");
    }
    if (!o_.RDP) return;
    omg_ok(omg_rdp_gradient(x_.p, grad, o_.Nx[0], o_.Ny[0], o_.Nz[0], o_.RDP_gamma, o_.epps, o_.beta,
                            o_.useImages ? 1 : 0, nullptr), "omg_rdp_gradient");
}

// D = (sum_k A_k^T 1) / subsets, D == 0 -> 1 (reconstructionAF.h:687-822; computeD for image
// preconditioners 0-2, libHeader.h:1084-1088).
void Driver::compute_D() {
    D_.alloc((size_t)N_);
    hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, D_.p, N_, 0.f);
    for (uint32_t k = 0; k < subsets_; k++) {
        const int64_t m = ranges_[k].n * bins_;
        tmp_.alloc((size_t)m);
        hipLaunchKernelGGL(k_fill, dim3(blocks(m)), dim3(kB), 0, nullptr, tmp_.p, m, 1.f);
        omg_ok(omg_backward_project(g_, ranges_[k].first, ranges_[k].n, tmp_.p, D_.p, nullptr, nullptr),
               "omg_backward_project (D)");
    }
    hipLaunchKernelGGL(k_d_finish, dim3(blocks(N_)), dim3(kB), 0, nullptr, D_.p, N_, (float)subsets_);
    hip_ok(hipGetLastError(), "D");
}

void Driver::save_image(const float* dev) {
    hip_ok(hipMemcpy(out_ + saved_ * (size_t)N_, dev, (size_t)N_ * sizeof(float), hipMemcpyDeviceToHost),
           "hipMemcpy (output)");
    saved_++;
}

// One sub-iteration of the main loop (reconstructionAF.h:1073-1466).
void Driver::subset_step(uint32_t iter, uint32_t k) {
    const inputStruct& o = o_;
    const int64_t first = ranges_[k].first, n = ranges_[k].n, m = n * bins_;
    const float epps = o.epps;
    const int64_t kk = (int64_t)iter * subsets_ + k;
    const bool ct = o.CT;

    // CGLS initialisation in the first sub-iteration (functions.hpp:2315-2340): r = y, f = x0,
    // p = A^T y (held in x_), gamma = <p, p>
    if (algo_ == Algo::CGLS && iter == 0) {
        r_.alloc((size_t)m);
        hip_ok(hipMemcpy(r_.p, y_[k].p, (size_t)m * sizeof(float), hipMemcpyDeviceToDevice), "copy");
        f_.alloc((size_t)N_);
        hip_ok(hipMemcpy(f_.p, x_.p, (size_t)N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
        hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, rhs_.p, N_, 0.f);
        omg_ok(omg_backward_project(g_, first, n, y_[k].p, rhs_.p, nullptr, nullptr), "omg_backward_project");
        hip_ok(hipMemcpy(x_.p, rhs_.p, (size_t)N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
        gammaCGLS_ = (float)dot(rhs_.p, rhs_.p, N_);
    }

    // forward projection
    v_.alloc((size_t)m);
    omg_ok(omg_forward_project(g_, first, n, x_.p, v_.p, nullptr), "omg_forward_project");
    if (o.storeFP && fp_out_) {
        v_.download(fp_out_ + fp_pos_, (size_t)m);
        fp_pos_ += (size_t)m;
    }

    // measurement-domain step
    bool want_sens = false;
    const float* bp_in = v_.p;
    switch (algo_) {
        case Algo::OSEM:
        case Algo::OSL:
            if (ct) {
                // CT: S_k = A_k^T y_k in the first iteration (reconstructionAF.h:1224-1260)
                if (iter == 0) {
                    S_[k].alloc((size_t)N_);
                    hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, S_[k].p, N_, 0.f);
                    omg_ok(omg_backward_project(g_, first, n, y_[k].p, S_[k].p, nullptr, nullptr),
                           "omg_backward_project (S)");
                    omg_ok(omg_sens_clamp(S_[k].p, N_, epps, nullptr), "omg_sens_clamp");
                }
            } else if (iter == 0) {
                want_sens = true;
            }
            hipLaunchKernelGGL(k_forward_step, dim3(blocks(m)), dim3(kB), 0, nullptr, 0, ct ? 1 : 0, y_[k].p, v_.p,
                               add_[k].p, m, n, epps, o.flat);
            break;
        case Algo::BSREM:
        case Algo::MBSREM:
            hipLaunchKernelGGL(k_forward_step, dim3(blocks(m)), dim3(kB), 0, nullptr, 1, ct ? 1 : 0, y_[k].p, v_.p,
                               add_[k].p, m, n, epps, o.flat);
            break;
        case Algo::PKMA:
            hipLaunchKernelGGL(k_forward_step, dim3(blocks(m)), dim3(kB), 0, nullptr, 2, ct ? 1 : 0, y_[k].p, v_.p,
                               add_[k].p, m, n, epps, o.flat);
            break;
        case Algo::PDHG: {
            if (!ct)
                hipLaunchKernelGGL(k_shift, dim3(blocks(m)), dim3(kB), 0, nullptr, v_.p, add_[k].p, m, n, epps);
            if (o.storeResidual && resid_) {
                // residual[kk] = |Ax - y|^2 / 2 + |p|^2 / 2 + <p, y> (computeForwardProject.h:251-252,301-304)
                tmp_.alloc((size_t)m);
                hip_ok(hipMemcpy(tmp_.p, v_.p, (size_t)m * sizeof(float), hipMemcpyDeviceToDevice), "copy");
                hipLaunchKernelGGL(k_sub, dim3(blocks(m)), dim3(kB), 0, nullptr, tmp_.p, y_[k].p, m);
                const float nr = (float)std::sqrt(dot(tmp_.p, tmp_.p, m));
                float res = nr * nr * .5f;
                const float np = (float)std::sqrt(dot(p_[k].p, p_[k].p, m));
                const float py = (float)dot(p_[k].p, y_[k].p, m);
                res = (float)((double)res + ((double)(np * np) * .5 + (double)py));
                resid_[kk] = res;
            }
            const float sigma = o.sigmaCP ? o.sigmaCP[0] : 1.f;
            hipLaunchKernelGGL(k_pdhg_dual, dim3(blocks(m)), dim3(kB), 0, nullptr, y_[k].p, v_.p, p_[k].p, m, sigma,
                               subsets_ > 1 ? 1 : 0, epps);
            break;
        }
        case Algo::CGLS: {
            if (!ct)
                hipLaunchKernelGGL(k_shift, dim3(blocks(m)), dim3(kB), 0, nullptr, v_.p, add_[k].p, m, n, epps);
            const float nq = (float)std::sqrt(dot(v_.p, v_.p, m));
            alphaCGLS_ = gammaCGLS_ / (nq * nq);
            hipLaunchKernelGGL(k_cgls_r, dim3(blocks(m)), dim3(kB), 0, nullptr, v_.p, r_.p, m, alphaCGLS_, epps);
            break;
        }
    }
    hip_ok(hipGetLastError(), "forward step");

    // backprojection (with the subset's sensitivity image in the first EM iteration)
    hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, rhs_.p, N_, 0.f);
    float* sens = nullptr;
    if (want_sens) {
        S_[k].alloc((size_t)N_);
        hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, S_[k].p, N_, 0.f);
        sens = S_[k].p;
    }
    omg_ok(omg_backward_project(g_, first, n, bp_in, rhs_.p, sens, nullptr), "omg_backward_project");
    if (sens) omg_ok(omg_sens_clamp(sens, N_, epps, nullptr), "omg_sens_clamp");

    // image update (computeOSEstimates, subiterStep.h:5-377)
    const int64_t counterMax = (int64_t)subsets_ * o.Niter - 1;
    const bool reg = computeReg(kk, counterMax);
    const int pos_now = (o.enforcePositivity && subsets_ > 1) ? 1 : 0;
    switch (algo_) {
        case Algo::OSEM:
            if (ct) {
                hipLaunchKernelGGL(k_ct_em, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, S_[k].p, nullptr, rhs_.p,
                                   N_, o.flat);
                if (pos_now) hipLaunchKernelGGL(k_clamp, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, N_, epps);
            } else {
                omg_ok(omg_osem_update(x_.p, S_[k].p, rhs_.p, N_, epps, pos_now, 0, nullptr), "omg_osem_update");
            }
            break;
        case Algo::OSL:
            if (reg) prior(dU_.p, true);
            if (ct) {
                hipLaunchKernelGGL(k_ct_em, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, S_[k].p, dU_.p, rhs_.p,
                                   N_, o.flat);
                if (pos_now) hipLaunchKernelGGL(k_clamp, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, N_, epps);
            } else {
                omg_ok(omg_osl_update(x_.p, S_[k].p, dU_.p, rhs_.p, N_, epps, pos_now, 0, nullptr), "omg_osl_update");
            }
            break;
        case Algo::BSREM:
            // the prior enters at the end of the iteration (iterStep.h:13-35)
            hipLaunchKernelGGL(k_poisson_update, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, rhs_.p, N_,
                               o.lambdaN[iter], epps, 1.f, o.enforcePositivity ? 1 : 0, 2);
            if (pos_now) hipLaunchKernelGGL(k_clamp, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, N_, epps);
            break;
        case Algo::MBSREM:
        case Algo::PKMA: {
            if (reg) prior(rhs_.p, false);   // dU is rhs for these (priors.h:365-370)
            if (D_.p) {
                const int type = o.precondTypeImage[0] ? 0 : 1;
                hipLaunchKernelGGL(k_precond, dim3(blocks(N_)), dim3(kB), 0, nullptr, rhs_.p, x_.p, D_.p, N_, type,
                                   algo_ == Algo::MBSREM ? 1 : 0, o.U);
            }
            if (algo_ == Algo::MBSREM)
                hipLaunchKernelGGL(k_poisson_update, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, rhs_.p, N_,
                                   o.lambdaN[iter], epps, o.U, o.enforcePositivity ? 1 : 0, 1);
            else
                hipLaunchKernelGGL(k_poisson_update, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, rhs_.p, N_,
                                   o.lambdaN[iter], epps, o.alpha_PKMA[kk], o.enforcePositivity ? 1 : 0, 0);
            if (pos_now) hipLaunchKernelGGL(k_clamp, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, N_, epps);
            break;
        }
        case Algo::PDHG: {
            const float tau = o.tauCP[0];
            if (subsets_ > 1) {
                // PDHG1 (algorithms.h:306-327): u += rhs; rhs = u + subsets theta_kk rhs
                omg_ok(omg_axpby(u_.p, rhs_.p, N_, 1.f, 1.f, nullptr), "omg_axpby");
                omg_ok(omg_axpby(rhs_.p, u_.p, N_, 1.f, (float)subsets_ * o.thetaCP[kk], nullptr), "omg_axpby");
            }
            if (D_.p) {
                const int type = o.precondTypeImage[0] ? 0 : 1;
                hipLaunchKernelGGL(k_precond, dim3(blocks(N_)), dim3(kB), 0, nullptr, rhs_.p, x_.p, D_.p, N_, type, 0,
                                   0.f);
            }
            omg_ok(omg_pdhg_update(x_.p, rhs_.p, u_.p, N_, epps, subsets_ > 1 ? 1.f : o.thetaCP[kk], tau,
                                   o.enforcePositivity ? 1 : 0, subsets_ > 1 ? 1 : 0, nullptr), "omg_pdhg_update");
            break;
        }
        case Algo::CGLS: {
            // algorithms.h:243-269: gamma' = <s, s>; f += alpha p; p = s + (gamma' / gamma) p, or x = f at the end
            const float gamma_new = (float)dot(rhs_.p, rhs_.p, N_);
            const float beta = gamma_new / gammaCGLS_;
            omg_ok(omg_axpby(f_.p, x_.p, N_, alphaCGLS_, 1.f, nullptr), "omg_axpby");
            if (iter == o.Niter - 1)
                hip_ok(hipMemcpy(x_.p, f_.p, (size_t)N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
            else
                omg_ok(omg_axpby(x_.p, rhs_.p, N_, 1.f, beta, nullptr), "omg_axpby");
            gammaCGLS_ = gamma_new;
            break;
        }
    }
    hip_ok(hipGetLastError(), "update");
}

// After the subsets of an iteration (reconstructionAF.h:1494-1508; computeOSEstimatesIter, iterStep.h).
void Driver::end_of_iteration(uint32_t iter) {
    const inputStruct& o = o_;
    if (algo_ == Algo::BSREM) {
        if (computeReg(iter, (int64_t)o.Niter - 1)) prior(dU_.p, true);
        hipLaunchKernelGGL(k_map_step, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, dU_.p, N_, o.lambdaN[iter],
                           o.epps);
        hip_ok(hipGetLastError(), "MAP step");
    }
    const uint32_t mid = (uint32_t)o.nIterSaved;
    const bool save = o.save_iter || (mid > 0 && (iter == o.Niter - 1 || (ee_ < mid && o.saveNiter[ee_] == iter)));
    if (save) {
        if (o.save_iter && iter == 0) {
            std::memcpy(out_ + saved_ * (size_t)N_, o.x0, (size_t)N_ * sizeof(float));
            saved_++;
        }
        save_image(x_.p);
        ee_++;
    }
    const bool cp = algo_ == Algo::PDHG;
    if (o.enforcePositivity && subsets_ == 1 && !cp) {
        hipLaunchKernelGGL(k_clamp, dim3(blocks(N_)), dim3(kB), 0, nullptr, x_.p, N_, o.epps);
        hip_ok(hipGetLastError(), "clamp");
    }
}

void Driver::run() {
    const inputStruct& o = o_;
    algo_ = check_supported(o);
    if (!sino_ || !out_) fail("the measurement data and the output buffer are required");
    if (o.Niter == 0) fail("Niter must be positive");
    if (!o.x0) fail("the initial value x0 is required");
    if ((algo_ == Algo::BSREM || algo_ == Algo::MBSREM || algo_ == Algo::PKMA) && !o.lambdaN)
        fail("%s needs the relaxation parameters lambdaN", algo_name(algo_));
    if (algo_ == Algo::PKMA && !o.alpha_PKMA) fail("PKMA needs alpha_PKMA");
    if (algo_ == Algo::PDHG && (!o.tauCP || !o.thetaCP || !o.sigmaCP))
        fail("PDHG needs tauCP, sigmaCP and thetaCP");
    if (algo_ == Algo::PDHG && !(o.tauCP[0] > 0.f))
        fail("PDHG needs a positive primal step tauCP (the power method is not run here)");
    if (o.nIterSaved > 0 && !o.saveNiter) fail("nIterSaved > 0 needs saveNiter");

    int ndev = 0;
    hip_ok(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if ((int)o.deviceNum >= ndev) fail("deviceNum %u is not one of the %d GPUs", o.deviceNum, ndev);
    hip_ok(hipSetDevice((int)o.deviceNum), "hipSetDevice");

    make_geometry();
    load_data();
    acc_.alloc(1);
    x_.upload(o.x0, (size_t)N_);
    rhs_.alloc((size_t)N_);
    dU_.alloc((size_t)N_);
    hipLaunchKernelGGL(k_fill, dim3(blocks(N_)), dim3(kB), 0, nullptr, dU_.p, N_, 0.f);
    S_.resize(subsets_);
    const bool need_D = o.precondTypeImage && (o.precondTypeImage[0] || o.precondTypeImage[1]);
    if (need_D) compute_D();
    if (algo_ == Algo::PDHG) {
        p_.resize(subsets_);
        for (uint32_t k = 0; k < subsets_; k++) {
            const int64_t m = ranges_[k].n * bins_;
            p_[k].alloc((size_t)m);
            hipLaunchKernelGGL(k_fill, dim3(blocks(m)), dim3(kB), 0, nullptr, p_[k].p, m, 0.f);
        }
        // uCP starts as a copy of the initial image (functions.hpp:2373-2375)
        u_.alloc((size_t)N_);
        hip_ok(hipMemcpy(u_.p, x_.p, (size_t)N_ * sizeof(float), hipMemcpyDeviceToDevice), "copy");
    }
    for (uint32_t iter = 0; iter < o.Niter; iter++) {
        for (uint32_t k = 0; k < subsets_; k++) {
            subset_step(iter, k);
            if (o.verbose > 0 && subsets_ > 1) std::printf("Sub-iteration %u complete\n", k + 1);
        }
        end_of_iteration(iter);
        if (o.verbose > 0) std::printf("Iteration %u complete\n", iter + 1);
    }
    if (!(o.save_iter || o.nIterSaved > 0)) save_image(x_.p);
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

}  // namespace

extern "C" __attribute__((visibility("default"))) int omegaMain(inputStruct options, const char* header_directory,
                                                                const MeasT* Sino, float* outputPtr, float* FPptr,
                                                                float* residual) {
    (void)header_directory;   // kernels are compiled ahead of time; no run-time sources are read
    try {
        Driver d(options, Sino, outputPtr, FPptr, residual);
        d.run();
    } catch (const std::exception& e) {
        std::printf("omegaMain: %s\nReconstruction failed!\n", e.what());
        std::fflush(stdout);
        return -1;
    }
    std::fflush(stdout);
    return 0;
}
