/* omega_main.h -- the built-in reconstruction entry point `omegaMain` and its parameter block.
 *
 * Drop-in for the reference's compiled reconstruction library (source/cpp/omega_maincpp.cpp:7-12,
 * declared at source/cpp/libHeader.h:1466-1472), which the Python front end loads by measurement type
 * (source/Python/omegatomo/reconstruction/recomain.py:652-700) and calls as
 *     omegaMain(param, headerDir, SinM, output, FPOutput, residual)
 * with `param` a ctypes Structure passed by value (_pack_ = 1, proj.py:1959-2282).
 *
 * Three libraries export the same symbol, one per measurement element type:
 *     libomega_matrixfree.so        float     (CUDA_matrixfree_lib / OpenCL_matrixfree_lib)
 *     libomega_matrixfree_uint16.so uint16_t  (..._uint16_lib)
 *     libomega_matrixfree_uint8.so  uint8_t   (..._uint8_lib)
 * Build-time macro OMG_MAIN_MEAS_TYPE selects the declaration below (default float).
 *
 * The block's layout is a binary interface fixed by the reference: the same members, in the same
 * order, with the same types and 1-byte packing as source/cpp/libHeader.h:11-534 (packing from the
 * #pragma pack(1) in source/cpp/functions.hpp:44 that precedes it). Defaults of the C++ initialisers
 * are not part of the interface (the Python caller sets every member) and are not repeated here.
 */
#ifndef OMEGA_MAIN_H
#define OMEGA_MAIN_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#pragma pack(push, 1)
typedef struct inputStruct {
    /* data kind and projector */
    uint8_t use_raw_data, listmode;
    int8_t verbose;
    uint16_t n_rays_transaxial, n_rays_axial;
    uint32_t projector_type, attenuation_correction, additionalCorrection, normalization_correction,
        randoms_correction;
    /* sizes */
    uint32_t nColsD, nRowsD, Nang, Ndist, subsets, det_per_ring, rings;
    uint32_t NxOrig, NyOrig, NzOrig, NxPrior, NyPrior, NzPrior;
    uint32_t Niter, Nt, subsetType, nMultiVolumes, nLayers, PDAdaptiveType, powerIterations, deblur_iterations,
        gradInitIter, gradLastIter, filteringIterations, mean_type;
    uint32_t Ndx, Ndy, Ndz, Nlx, Nly, Nlz, g_dim_x, g_dim_y, g_dim_z;
    int32_t localSizeX, localSizeY, localSizeZ, regEveryIter;
    uint32_t NiterAD, inffi, Nf, deviceNum, platform, derivativeType, TVtype, FluxType, DiffusionType,
        POCS_NgradIter, maskFPZ, maskBPZ, FISTAType;
    int64_t nProjections, TOF_bins;
    /* scalars */
    float tau, helicalRadius, tube_radius, epps, sigma_x, tube_width_z, tube_width_xy, bmin, bmax, Vmax,
        global_factor, dL, flat, U, h_ACOSEM, dPitchX, dPitchY, cr_p, cr_pz, NLMsigma, NLAdaptiveConstant, w_sum,
        KAD, TimeStepAD, RDP_gamma, huber_delta, gradV1, gradV2, alpha0TGV, alpha1TGV, GGMRF_p, GGMRF_q, GGMRF_c,
        beta, beta_temporal, T, dSizeXBP, dSizeZBP, TVsmoothing, temporalTVsmoothing, C, SATVPhi, eta,
        APLSsmoothing, hyperbolicDelta, sourceToCRot, POCS_alpha, POCS_rMax, POCS_alphaRed, POCSepps;
    /* switches */
    bool use_psf, TOF, pitch, SPECT, PET, CT, largeDim, loadTOF, storeResidual, FISTA_acceleration, meanFP, meanBP,
        useMaskFP, useMaskBP, orthTransaxial, orthAxial, enforcePositivity, useMultiResolutionVolumes, save_iter,
        deblurring, useMAD, useImages, useEFOV, CTAttenuation, offsetCorrection, relaxationScaling,
        computeRelaxationParameters, storeFP, use2DTGV, med_no_norm, NLM_MRP, NLTV, NLRD, NLLange, NLGGMRF,
        NLM_use_anatomical, NLAdaptive, TV_use_anatomical, RDPIncludeCorners, RDP_use_anatomical, useL2Ball,
        saveSens, use_64bit_atomics, use_32bit_atomics, compute_sensitivity_image, useFDKWeights,
        useIndexBasedReconstruction, stochasticSubsetSelection, useTotLength, useParallelBeam, useHelical;
    /* algorithms */
    bool OSEM, LSQR, CGLS, SART, FISTA, FISTAL1, MRAMLA, RAMLA, ROSEM, RBI, DRAMA, COSEM, ECOSEM, ACOSEM, OSL_OSEM,
        MBSREM, BSREM, ROSEM_MAP, OSL_RBI, OSL_COSEM, PKMA, SPS, PDHG, PDHGKL, PDHGL1, PDDY, CV, POCS, FDK, SAGA, BB;
    /* priors */
    bool MRP, quad, Huber, L, FMH, weighted_mean, TV, hyperbolic, AD, APLS, TGV, NLM, RDP, GGMRF, ProxTV, ProxRDP,
        ProxNLM, temporalTV, temporal_smoothness, MAP, custom;
    /* element counts of the arrays below */
    uint64_t mDim, nIterSaved, sizeScat, eFOV, sizeX, sizeZ, sizeAtten, sizeNorm, sizePSF, sizeXYind, sizeZind,
        xCenterSize, yCenterSize, zCenterSize, sizeV, measElem;
    int64_t seed;
    /* arrays (host memory owned by the caller) */
    float *x, *z, *uV, *dx, *dy, *dz, *bx, *by, *bz, *atten, *norm;
    int64_t* pituus;
    uint32_t* xy_index;
    uint16_t* z_index;
    float *x_center, *y_center, *z_center, *V, *gaussPSF, *gaussianNLM;
    uint32_t *saveNiter, *Nx, *Ny, *Nz;
    float *randoms, *corrVector, *x0, *offsetVal, *dScaleX4, *dScaleY4, *dScaleZ4, *dSizeX, *dSizeY, *dScaleX,
        *dScaleY, *dScaleZ, *kerroin4, *lam_drama;
    uint8_t *maskFP, *maskBP, *eFOVIndices, *maskPrior, *TOFIndices;
    float *angles, *swivelAngles;
    int32_t *blurPlanes, *blurPlanes2;
    float* gFilter;
    uint64_t* gFSize;
    bool *precondTypeImage, *precondTypeMeas;
    float *referenceImage, *filterIm, *filter, *filter2, *Ffilter, *s, *weights_quad, *weights_huber,
        *weighted_weights, *APLS_ref_image, *lambdaN, *lambdaFiltered, *alpha_PKMA, *alphaPrecond, *tauCP, *tauCPFilt,
        *sigmaCP, *sigma2CP, *thetaCP, *TOFCenter, *TV_ref;
    uint16_t *trIndices, *axIndices;
    /* SPECT */
    float *rayShiftsDetector, *rayShiftsSource;
    float coneOfResponseStdCoeffA, coneOfResponseStdCoeffB, coneOfResponseStdCoeffC, totalFOVxmin, totalFOVymin,
        totalFOVzmin, totalFOVxmax, totalFOVymax, totalFOVzmax;
    float *NLM_ref, *RDP_ref;
} inputStruct;
#pragma pack(pop)

#ifndef OMG_MAIN_MEAS_TYPE
#define OMG_MAIN_MEAS_TYPE float
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the reconstruction the block selects on the GPU `deviceNum` and writes, all in host memory:
 *   outputPtr: the final image (Nx*Ny*Nz floats, x fastest); with save_iter, x0 and then every iteration's
 *              image ((Niter + 1) images); with nIterSaved > 0, the iterations listed in saveNiter and then
 *              the last one (source/cpp/iterStep.h:39-62, reconstructionAF.h:1446-1458);
 *   FPptr:     with storeFP, each sub-iteration's forward projection, iteration-major
 *              (reconstructionAF.h:1199-1210);
 *   residual:  with storeResidual and PDHG, the primal-dual value per sub-iteration
 *              (computeForwardProject.h:251-252,301-304).
 * Sino holds measElem values in the layout the Python front end passes (subset order).
 * Returns 0 on success. A configuration outside what this library provides, or a failure, prints a
 * message and "Reconstruction failed!" (omega_maincpp.cpp:177-178) and returns a nonzero value (the
 * reference returns 0 in both cases; its Python caller ignores the value). */
int omegaMain(inputStruct options, const char* header_directory, const OMG_MAIN_MEAS_TYPE* Sino, float* outputPtr,
              float* FPptr, float* residual);

#ifdef __cplusplus
}
#endif
#endif /* OMEGA_MAIN_H */
