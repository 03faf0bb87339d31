// ref_driver.cpp -- C-ABI shim around the REFERENCE's own CPU headers (test infrastructure only).
//
// Compiled by oracle/Makefile against the unmodified headers where they lie
// (source/cpp/projector_functions.h and source/cpp/RDP.h of villekf/OMEGA), into
// oracle/_ref/libomega_ref.so. It is used to validate the oracle restatement and
// to write golden vectors; nothing in the product links or loads it.
//
// The parameter block is filled the way ProjectorClassCPU::forwardProjection /
// backwardProjection do (source/cpp/ProjectorClassCPU.h:125-235), except that the
// subset's first plane (not its length) is passed as paramStruct::nMeas, i.e. the
// cumulative-start semantics of m-files/projectorClass.m:1009.
#include "projector_functions.h"
#include "RDP.h"

template <typename T>
static void fill(paramStruct<T>& p, const uint32_t* n3, const T* d3, const T* b3, uint32_t size_x, uint32_t size_y,
                 int32_t mode, int64_t plane_offset, T dPitchZ, int32_t ct_pitch, T global_factor, T* norm,
                 T* atten, int32_t atten_mode, T* scat, uint32_t tof_bins, T tof_sigma, T* tof_c, T tof_epps,
                 uint32_t rays_xy, uint32_t rays_z, T cr_xy, T cr_z, uint32_t nProjections) {
    p.Nx = n3[0]; p.Ny = n3[1]; p.Nz = n3[2];
    p.dx = d3[0]; p.dy = d3[1]; p.dz = d3[2];
    p.bx = b3[0]; p.by = b3[1]; p.bz = b3[2];
    p.size_x = size_x; p.size_y = size_y;
    p.projType = 1;
    p.subsetType = 8;
    p.subsets = plane_offset > 0 ? 2u : 1u;   // > 1 enables the plane offset branch (projector_functions.h:1387-1389)
    p.currentSubset = 0;
    p.nMeas = (uint64_t)plane_offset;
    p.listMode = (mode == 2) ? 1 : 0;
    p.raw = false;
    p.nRays2D = (uint16_t)(rays_xy > 1 ? rays_xy : 1);
    p.nRays3D = (uint16_t)(rays_z > 1 ? rays_z : 1);
    p.TOF = tof_bins > 1;
    p.nBins = tof_bins > 1 ? tof_bins : 1u;
    p.sigma_x = tof_sigma;
    p.TOFCenters = tof_c;
    p.globalFactor = global_factor;
    p.normalizationCorrection = norm != nullptr;
    p.normCoefs = norm;
    // attenuation: mode 1 image (CTAttenuation), mode 2 per-measurement factors
    p.attenuationCorrection = atten != nullptr && atten_mode != 0;
    p.CTAttenuation = atten_mode != 2;
    p.atten = atten;
    p.scatterCorrectionMult = scat != nullptr;
    p.scatterCoefs = scat;
    p.dPitchZ = dPitchZ; p.dPitchXY = dPitchZ;
    if (rays_xy > 1 || rays_z > 1) { p.dPitchXY = cr_xy; p.dPitchZ = cr_z; }
    p.pitch = ct_pitch != 0;
    p.nProjections = nProjections;
    p.epps = tof_bins > 1 ? tof_epps : (T)1e-5;
}

#define REF_API(T, SUF)                                                                                               \
    extern "C" int ref_siddon_##SUF(const uint32_t* n3, const T* d3, const T* b3, uint32_t size_x, uint32_t size_y,  \
                                    int32_t mode, int64_t plane_offset, T dPitchZ, int32_t ct_pitch, T global_factor, \
                                    T* norm, T* atten, int32_t atten_mode, T* scat, uint32_t tof_bins, T tof_sigma,  \
                                    T* tof_c, T tof_epps, uint32_t rays_xy, uint32_t rays_z, T cr_xy, T cr_z,        \
                                    uint32_t nProjections,                                                            \
                                    int64_t nMeas,                                                                    \
                                    const T* x, const T* z, T* input, T* output, T* sens, int fp, int nthreads) {     \
        paramStruct<T> p;                                                                                              \
        fill(p, n3, d3, b3, size_x, size_y, mode, plane_offset, dPitchZ, ct_pitch, global_factor, norm, atten,        \
             atten_mode, scat, tof_bins, tof_sigma, tof_c, tof_epps, rays_xy, rays_z, cr_xy, cr_z, nProjections);      \
        p.noSensImage = (sens == nullptr);                                                                             \
        projectorType123Implementation4<T>(p, nMeas, output, x, z, input, mode == 1, false, (uint8_t)fp, sens,        \
                                           nullptr, (uint32_t)nthreads);                                               \
        return 0;                                                                                                      \
    }

REF_API(float, f32)
REF_API(double, f64)

extern "C" int ref_rdp_f32(float* grad, const float* u, int Nx, int Ny, int Nz, float gamma, float epps, float beta) {
    RDPKernel<float>(grad, u, Nx, Ny, Nz, Nx, Ny, Nz, gamma, epps, beta);
    return 0;
}
