/* omega_hip.h -- C ABI of the MI355X (gfx950) projector library libomega_hip.so.
 *
 * Plain C: fixed-width integers, float pointers, an opaque geometry handle and
 * a HIP stream passed as void* (0 = the null stream). Every device pointer is
 * borrowed from the caller (e.g. torch's data_ptr()); the library never frees
 * caller memory. Geometry handles own their device copies of the coordinate
 * tables. Calls are asynchronous on the given stream. Every entry point returns
 * 0 on success or a negative OMG_ERR_* code; omg_last_error() gives a message
 * (thread-local).
 *
 * Each entry point names the reference interface it replaces (villekf/OMEGA,
 * paths relative to the repository root).
 */
#ifndef OMEGA_HIP_H
#define OMEGA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMG_OK 0
#define OMG_ERR_ARG (-1)          /* invalid argument (size, null pointer, unsupported combination) */
#define OMG_ERR_HIP (-2)          /* HIP runtime error (allocation, launch, copy) */
#define OMG_ERR_UNSUPPORTED (-3)  /* valid in the reference but not provided by this library */

/* Projection data layouts (source/cpp/projector_functions.h:357-498). */
#define OMG_MODE_PET_SINO 0   /* x: 4 floats per (radial, angle) bin; z: 2 floats per plane */
#define OMG_MODE_CT_FLAT 1    /* x: 6 floats per projection; z: uV (2) or pitch vectors (6) per projection */
#define OMG_MODE_LIST6 2      /* x: 6 floats per LOR */

/* Projector types (source/Python/omegatomo/projector/proj.py:242, projector_type). */
#define OMG_PROJ_SIDDON 1
#define OMG_PROJ_ORTH 2
#define OMG_PROJ_VOLUME 3
#define OMG_PROJ_INTERP 4
#define OMG_PROJ_BDD 5

typedef struct omg_geom omg_geom;

/* Geometry description. Mirrors the scalars and tables the reference binds into
 * paramStruct<T> (source/cpp/projector_functions.h:223-347) and the kernel
 * arguments of projectorType123 (source/opencl/projectorType123.cl:81-267). */
typedef struct omg_geom_desc {
    uint32_t struct_size;      /* sizeof(omg_geom_desc) as seen by the caller */
    int32_t projector_type;    /* OMG_PROJ_* */
    int32_t mode;              /* OMG_MODE_* */
    uint32_t Nx, Ny, Nz;       /* voxels */
    float dx, dy, dz;          /* voxel size (mm) */
    float bx, by, bz;          /* FOV lower corner (mm) */
    uint32_t size_x;           /* nRowsD: radial bins / detector rows */
    uint32_t size_y;           /* nColsD: angles / detector columns */
    uint32_t n_proj;           /* planes (PET) or projections (CT) held by the tables */
    int32_t ct_pitch;          /* CT: z holds 6 floats per projection */
    float dPitchZ;             /* CT: axial detector pitch when ct_pitch == 0 */
    float global_factor;       /* PET global correction factor */
    const float* x; int64_t x_len;        /* host coordinate tables, copied at create */
    const float* z; int64_t z_len;
    const float* norm; int64_t norm_len;  /* optional per-measurement normalisation (full data order) */
    /* orthogonal distance-based projector (projector_type 2) */
    float orth_width;          /* kernel tube width in mm (the reference passes tube_width_z, init.py:652) */
    int32_t orth_flags;        /* bit 0: transaxial row sweep (-DCRYSTXY, tube_width_xy > 0);
                                  bit 1: axial plane sweep (-DCRYSTZ, tube_width_z > 0) (init.py:286-290) */
    /* branchless distance-driven projector (projector_type 5); CT mode only */
    float dPitchX;             /* transaxial detector pitch */
    float dPitchY;             /* axial detector pitch */
    /* launch tuning: reserved[0] != 0 routes all Siddon calls through the one-lane-per-LOR kernels;
     * reserved[1] > 0 sets the workgroups per plane of the direct-plane (LDS) kernels; reserved[2] == 1
     * opts the fused subset iteration into the voxel-driven backprojection on parallel-beam ring
     * sinograms (closed-form chord lengths: within 5e-7 of an FP64 Siddon backprojection, but not the
     * reference's FP32 walk lengths, which are themselves ~1e-5 from FP64); geometries that do not
     * qualify keep the LDS fixed-point planar kernel; reserved[3] != 0 keeps the one-lane-per-LOR kernels
     * on PET sinograms in the x-fastest image only (by default LORs closer to the x axis accumulate
     * their backprojection into y-fastest buffers, where the lanes of a wave share cache lines);
     * reserved[4] is the device-memory budget in MiB of projectors 2 and 3's tube lists (each
     * measurement range's voxel indices and table weights, traced once and then streamed by the
     * projections; 0: 16384 MiB, < 0: off, every call traces the tubes).
     * Zero = automatic. */
    int32_t reserved[8];
    /* measurement-level subsets of PET sinograms (subset types 1-5: indexMaker's `index`,
     * source/Python/omegatomo/projector/indices.py:51-123, read by getIndex,
     * source/opencl/general_opencl_functions.h:913-941): optional host table of sinogram bins, one per
     * measurement in subset order, copied at create. With it, PET sinogram calls take `first` and `n`
     * in measurements of this table instead of planes: measurement i of a call is sinogram bin
     * meas_index[first + i] (radial bin fastest, then angle, then plane), and `norm` is in the same
     * (subset) order. The direct-plane kernels do not apply; frame batching is not offered. */
    const uint32_t* meas_index; int64_t meas_index_len;
    /* attenuation correction of PET data (projectorClass.attenuation_correction, CTAttenuation, vaimennus:
     * source/Python/omegatomo/projector/proj.py:85,296,358; kernel options -DATN / -DATNM, init.py:318-321).
     * atten_mode 0: off. 1: `atten` is the attenuation image (Nx*Ny*Nz floats, 1/mm, x fastest); the
     * library computes each measurement's factor exp(-sum len * mu) over its Siddon walk once, here (the
     * reference integrates it inside every projection, projector_functions.h:1754-1767,1938-1939). 2:
     * `atten` holds one factor per measurement, in the order of `norm`. The factor multiplies the LOR
     * weight of every projector and of the fused steps. */
    int32_t atten_mode; const float* atten; int64_t atten_len;
    /* optional multiplicative correction per measurement (projectorClass.additionalCorrection / corrVector,
     * kernel option -DSCATTER, init.py:324,592-595; paramStruct::scatterCoefs, projector_functions.h:1436-1440),
     * in the order of `norm`; it multiplies the LOR weight after norm. PET data only. */
    const float* scat; int64_t scat_len;
    /* time of flight (projectorClass.TOF_bins, TOF_FWHM, TOF_width, TOFCenter; paramStruct TOF, nBins, sigma_x,
     * TOFCenters, projector_functions.h:306-313,559-657; for projectors 3 and 4 the GPU kernels' TOFLoop,
     * projectorType123.cl -DTOF -DVOL, projectorType4.cl -DTOF). tof_bins > 1 turns it on (at most 64;
     * projectors 1, 3 and 4, PET data): every measurement array of a call then holds tof_bins blocks of n_meas
     * values, bin b of measurement i at i + b * n_meas (projector_functions.h:2030-2035); the additive term of the ratio
     * stays one value per measurement, added to every bin. tof_sigma: the TOF standard deviation (mm);
     * tof_centers: the bins' centre offsets along the LOR (mm, tof_bins floats, copied at create);
     * tof_epps: the floor of a voxel's weight sum (the reference's epps). */
    uint32_t tof_bins; float tof_sigma; float tof_epps; const float* tof_centers;
    /* several rays per LOR (projectorClass.n_rays_transaxial / n_rays_axial, -DN_RAYS2D / -DN_RAYS3D,
     * init.py:357-360; paramStruct nRays2D, nRays3D, projector_functions.h:262-264,1444-1466): rays_xy x
     * rays_z rays (0 or 1: one), their endpoints spread over cr_xy (transaxially, on x and y alike) and
     * cr_z (axially) (paramStruct dPitchXY, dPitchZ; the operator API passes dPitchX and dPitchY); each
     * ray weighs 1 / (L nRays), the forward projection sums them (per TOF bin with TOF). Siddon and
     * interpolation (projectorType4.cl -DN_RAYS), PET data.
     * With an attenuation image each ray has its own factor, exp of its own walk's sum (the library keeps
     * rays x measurements factors). */
    uint32_t rays_xy, rays_z; float cr_xy, cr_z;
    /* volume-of-intersection projector (projector_type 3; the reference's -DORTH -DVOL kernel,
     * opencl_functions_orth3D.h:132-140, projectorType123.cl:502-505,641-643,952-953): the tube of
     * radius orth_width (tube_radius, init.py:653) sweeps like projector 2 (orth_flags); a voxel at axis
     * distance d counts vol_vmax when d < vol_bmin, vol_V[rne((d - vol_bmin) * 1000)] when
     * d <= vol_bmax (0 past the table's end), nothing beyond; every LOR weighs 1 / (L R^2 / pi).
     * vol_V, vol_bmin, vol_bmax, vol_vmax: computeVoxelVolumes (coordinates.py:190-205), copied at create. */
    const float* vol_V; int64_t vol_len; float vol_bmin, vol_bmax, vol_vmax;
    /* interpolation projector (projector_type 4; projectorType4Forward / projectorType4Backward,
     * source/opencl/projectorType4.cl:84,718): samples every interp_dl mm along the ray (interp_dl <= 0: one
     * sample per voxel plane of the ray's dominant axis, the Joseph variant), trilinear in the image; CT data
     * are backprojected voxel by voxel, each panel sampled bilinearly and weighted by L^3 / |v|^2 *
     * interp_kerroin, interp_kerroin = (dx dy dz) / (dPitchX dPitchY sourceToDetector)
     * (computeProjectorScalingValues, coordinates.py:241-262). */
    float interp_dl; float interp_kerroin;
} omg_geom_desc;

/* Library/version info and error text. */
const char* omg_version(void);
const char* omg_last_error(void);

/* Geometry handles: replaces the per-projector set-up of projectorClass.initProj /
 * initProjector (source/Python/omegatomo/projector/proj.py:1913, init.py:39) and the
 * device buffers ProjectorClass::createBuffers makes (source/cpp/ProjectorClass.h). */
int omg_geom_create(const omg_geom_desc* desc, omg_geom** out);
int omg_geom_destroy(omg_geom* g);

/* Forward projection y = A_k x for measurements [first, first + n_meas) of the
 * subset-ordered data: first is the subset's first plane/projection (list mode: LOR).
 * y is overwritten. Replaces projectorClass.forwardProject / projfunctions.forwardProjection
 * (proj.py:1923, projfunctions.py:86) and projectorType123Implementation4(fp=1)
 * (source/cpp/projector_functions.h:1311). */
int omg_forward_project(const omg_geom* g, int64_t first, int64_t n_meas, const float* img, float* y, void* stream);

/* Backprojection rhs += A_k^T y; if sens != NULL also sens += A_k^T 1. rhs/sens are
 * accumulated (the caller zeroes them). Replaces projectorClass.backwardProject
 * (proj.py:1929, projfunctions.py:735) and projectorType123Implementation4(fp=2)
 * (projector_functions.h:1311, rhs :633-657). */
int omg_backward_project(const omg_geom* g, int64_t first, int64_t n_meas, const float* y, float* rhs, float* sens,
                         void* stream);

/* One OSEM subset data step, fused: yhat = A_k x; r = meas / (yhat + epps) (NaN/Inf -> epps);
 * rhs += A_k^T r; sens += A_k^T 1 when sens != NULL. PET modes, projector_type 1 or 2. Replaces the
 * forwardProjectionAFOpenCL -> computeForwardStep -> backwardProjectionAFOpenCL sequence
 * of reconstructionAF (source/cpp/reconstructionAF.h:1089-1264, computeForwardProject.h:20-21,94,400-401). */
int omg_osem_data_step(const omg_geom* g, int64_t first, int64_t n_meas, const float* img, const float* meas,
                       float* rhs, float* sens, float epps, void* stream);

/* One OSEM (beta == 0) or OSL-OSEM with the RDP prior (beta != 0) subset iteration on image x, in place:
 * the body of the reference's subset loop (source/cpp/reconstructionAF.h:1073-1466): yhat = A_k x;
 * r = meas / (yhat + epps) (NaN/Inf -> epps); rhs = A_k^T r; if compute_sens, S = A_k^T 1 with
 * S < epps -> epps (stored to S; else S is read); OSEM x = x / S * (rhs + epps) (subiterStep.h:111-121)
 * or OSL x = x / (S + dU) * rhs with dU = beta * grad RDP(x; gamma, epps), zero boundary
 * (subiterStep.h:219-231, RDP.h:4-78); positivity x < epps -> epps when positivity != 0.
 * Emission data, projector_type 1 or 2. Uses library-owned work buffers of the geometry handle:
 * calls on one handle must not overlap. */
int omg_osem_subset_iteration(const omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas, float* S,
                              int32_t compute_sens, float epps, float beta, float gamma, int32_t positivity,
                              void* stream);

/* The data step and the subset iteration with the randoms / scatter term of the ratio: the denominator
 * becomes (A_k x + epps) + additive[i] (computeForwardProject.h:20-21, 66-69: input += epps; input +=
 * randomsData), additive holding one float per measurement of the call in the order of meas, or NULL
 * (the plain entry points above). Not offered for the batched frames. */
int omg_osem_data_step_ex(const omg_geom* g, int64_t first, int64_t n_meas, const float* img, const float* meas,
                          const float* additive, float* rhs, float* sens, float epps, void* stream);
int omg_osem_subset_iteration_ex(const omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas,
                                 const float* additive, float* S, int32_t compute_sens, float epps, float beta,
                                 float gamma, int32_t positivity, void* stream);

/* One OSEM (beta == 0) or OSL-OSEM with the RDP prior (beta != 0) subset iteration of transmission (CT) data on
 * image x, in place, with any projector of the geometry: if compute_sens, S = A_k^T meas with S < epps -> epps
 * (stored; else S is read); yhat = A_k x, yhat < 0 -> 0; r = exp(-yhat); rhs = A_k^T r; OSEM
 * x = x / S * (flat * rhs), OSL x = x / (S + dU) * (flat * rhs) with dU = beta * grad RDP(x; gamma, epps), zero
 * boundary; positivity x < epps -> epps when positivity != 0. meas: the measured counts of the subset's
 * projections; flat: the blank-scan (flat-field) value. Replaces the CT branch of reconstructionAF for OSEM /
 * OSL-OSEM without randoms (source/cpp/reconstructionAF.h:322-327,1224-1260; computeForwardProject.h:76-90;
 * subiterStep.h:111-121,219-231). Uses work buffers of the geometry handle: calls on one handle must not
 * overlap. */
int omg_ct_osem_subset_iteration(const omg_geom* g, int64_t first, int64_t n_meas, float* x, const float* meas,
                                 float* S, int32_t compute_sens, float flat, float epps, float beta, float gamma,
                                 int32_t positivity, void* stream);

/* omg_osem_subset_iteration_ex without allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). The same kernels in the same order, with every work buffer (OSL gradient, rhs image, slice map,
 * the direct-plane kernels' per-workgroup partial slices, the voxel-driven path's per-LOR buffers) carved
 * out of one library-owned device array of 64 MiB (16,777,216 floats) that each GPU holds until the process
 * ends. A call that needs more (2 N + Nz floats for an image of N voxels and Nz slices, plus, on the
 * direct-plane path, 2 x planes x workgroups-per-plane x Nx Ny floats) fails with OMG_ERR_ARG before any
 * launch. Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other threads, other handles) race; the caller serialises them. The current device must be the
 * geometry's. Nothing in the array has to start at zero. */
int omg_osem_subset_iteration_noalloc(const omg_geom* g, int64_t first, int64_t n_meas, float* x,
                                      const float* meas, const float* additive, float* S, int32_t compute_sens,
                                      float epps, float beta, float gamma, int32_t positivity, void* stream);

/* Kernel timing of the projector launches with HIP events, the counterpart of the reference's
 * verbose >= 3 event timers (source/cpp/ProjectorClass.h:3594-3599). omg_timing_enable(1) records
 * every following projector kernel launch (Siddon FP/BP/fused kernels) until omg_timing_enable(0);
 * omg_timing_collect waits for the recorded launches, returns their summed duration (ms) and count,
 * and clears the record. Process-wide, not thread-safe. */
int omg_timing_enable(int32_t on);
int omg_timing_collect(double* total_ms, int64_t* launches);

/* Per-LOR voxel lists of the ray tracer (verification of the traversal itself).
 * counts[i] = number of voxels visited for measurement i. If offsets != NULL (exclusive prefix sums of
 * counts), voxel indices are written to idx and per-voxel weights to len: the intersection length for
 * projector_type 1 (the index/element lists of improved_siddon_precomputed,
 * source/cpp/improved_siddon_precomputed.cpp:32-251), the tube weight 1 - d/width in visiting order for
 * projector_type 2 (orthogonalHelper3D, source/opencl/opencl_functions_orth3D.h:85-193). */
int omg_siddon_trace(const omg_geom* g, int64_t first, int64_t n_meas, int32_t* counts, const int64_t* offsets,
                     uint32_t* idx, float* len, void* stream);

/* ---- Batched dynamic-study OSEM (all frames of a study per launch) ----------------------------------
 * The reference runs the frames of a dynamic study one after another inside every subset
 * (source/cpp/reconstructionAF.h:1089-1453, per-frame data :429-436). These entry points advance all
 * n_frames frames through a subset step together, tracing each LOR once per tile of frames.
 * Layouts (frame-major): image stack x[n_frames][N]; measurements meas[n_frames][...] with frame stride
 * meas_stride, each frame in the subset order of a single-frame run, meas pointing at frame 0's first LOR
 * of the subset; rhs[n_frames][N] is a work stack that must be zero on the first call and is left zero by
 * every step. S is the subset's frame-independent sensitivity image (omg_frames_sensitivity).
 * Siddon projector, emission data. */

/* S = max(A_k^T 1, epps) for the subset's measurements [first, first + n_meas) (reconstructionAF.h:356-381,
 * functions.hpp:2981-2996). */
int omg_frames_sensitivity(const omg_geom* g, int64_t first, int64_t n_meas, float* S, float epps, void* stream);

/* yhat[f * yhat_stride + i] = (A_k x_f)_i for the subset's measurements (Siddon FP of every frame). */
int omg_frames_forward(const omg_geom* g, int64_t first, int64_t n_meas, int32_t n_frames, const float* x, float* yhat,
                       int64_t yhat_stride, void* stream);
/* out[f] = sum_i y[f][i] log(yhat[f][i]) - yhat[f][i] (terms with y = 0: -yhat), FP64, [n_frames][n] stacks. */
int omg_frames_loglik(const float* y, const float* yhat, int64_t n, int32_t n_frames, double* out, void* stream);

/* One OSEM subset step of every frame: yhat_f = A_k x_f; r_f = meas_f / (yhat_f + epps) (NaN/Inf -> epps);
 * x_f = x_f / S * (A_k^T r_f + epps) (subiterStep.h:111-121), or with beta != 0 the OSL-RDP step
 * x_f = x_f / (S + dU_f) * A_k^T r_f, dU_f = beta * grad RDP(x_f; gamma, epps) with zero boundary
 * (subiterStep.h:219-231, RDP.h:4-78), dU a work stack [n_frames][N]; positivity x < epps -> epps.
 * Launch shape: rays_per_block in {64,128,256} lanes (one LOR each), frames_per_tile in {1,2,4,8} frames per
 * lane (n_frames must be a multiple), grid_blocks persistent workgroups (0: one resident wave). */
int omg_frames_osem_step(const omg_geom* g, int64_t first, int64_t n_meas, int32_t n_frames, float* x,
                         const float* meas, int64_t meas_stride, const float* S, float* rhs, float* dU, float beta,
                         float gamma, float epps, int32_t positivity, int32_t rays_per_block, int32_t frames_per_tile,
                         int32_t grid_blocks, void* stream);

/* A captured iteration: n_subsets subset steps (subset k: planes/LORs from first[k], n_meas[k] measurements at
 * meas + meas_offset[k] (frame stride meas_stride), sensitivity S + k * N) recorded once in a hipGraph and replayed by
 * omg_graph_launch. The pointers are baked into the graph: keep the buffers alive and in place. */
typedef struct omg_graph omg_graph;
int omg_frames_graph_create(const omg_geom* g, int32_t n_subsets, const int64_t* first, const int64_t* n_meas,
                            const int64_t* meas_offset, int32_t n_frames, float* x, const float* meas,
                            int64_t meas_stride, const float* S, float* rhs, float* dU, float beta, float gamma,
                            float epps,
                            int32_t positivity, int32_t rays_per_block, int32_t frames_per_tile, int32_t grid_blocks,
                            omg_graph** out);
int omg_graph_launch(omg_graph* h, void* stream);
int omg_graph_destroy(omg_graph* h);

/* ---- 3-D U-Net warm start (no counterpart in the reference) -------------------------------------------
 * Kernels of the learned warm start for OSEM-RDP. BF16 tensors are uint16 bit patterns. Activations are
 * channels-last [D][H][W][C] (C a multiple of 16); conv weights [27][Cout][Cin], tap = (dz*3 + dy)*3 + dx;
 * 3x3x3 convolutions with zero padding 1 on the matrix cores (BF16 in, FP32 accumulation). */

/* Y = relu?(conv(X) + bias) with X = [X1 | X2] (C1 + C2 input channels; X2 may be null with C2 = 0);
 * Cout in {16, 32, 48, 64, 96}; bias may be null. Also the data gradient: dX = conv(dY) with the weights
 * from omg_conv3d_flipT and no bias or ReLU. */
int omg_conv3d_forward(const uint16_t* X1, int32_t C1, const uint16_t* X2, int32_t C2, const uint16_t* W,
                       const float* bias, uint16_t* Y, int32_t Cout, int32_t D, int32_t H, int32_t Wd, int32_t relu,
                       void* stream);
/* dW[tap][co][ci] += sum_v dY[v][co] * X[v + off(tap)][ci] (FP32 dW). */
int omg_conv3d_wgrad(const uint16_t* dY, int32_t Cout, const uint16_t* X1, int32_t C1, const uint16_t* X2,
                     int32_t C2, float* dW, int32_t D, int32_t H, int32_t Wd, void* stream);
/* Wf[tap][ci][co] = W[26 - tap][co][ci]. */
int omg_conv3d_flipT(const uint16_t* W, uint16_t* Wf, int32_t Cout, int32_t Cin, void* stream);
/* db[c] += sum_v dY[v][c], C in {16, 32, 64}. */
int omg_bias_grad(const uint16_t* dY, float* db, int32_t V, int32_t C, void* stream);
/* dY = 0 where Y <= 0. */
int omg_relu_backward(const uint16_t* Y, uint16_t* dY, int64_t n, void* stream);
/* 2x2x2 average pooling of the fine grid D x H x W and its adjoint ((+)= into dX). */
int omg_pool_forward(const uint16_t* X, uint16_t* Y, int32_t D, int32_t H, int32_t W, int32_t C, void* stream);
int omg_pool_backward(const uint16_t* dY, uint16_t* dX, int32_t D, int32_t H, int32_t W, int32_t C,
                      int32_t accumulate, void* stream);
/* Nearest x2 upsampling onto the fine grid D x H x W; its adjoint reads channels [off, off + C) of a
 * fine-grid tensor with Cs channels and writes the coarse grid D x H x W. */
int omg_upsample_forward(const uint16_t* X, uint16_t* Y, int32_t D, int32_t H, int32_t W, int32_t C, void* stream);
int omg_upsample_backward(const uint16_t* dY, int32_t Cs, int32_t off, uint16_t* dX, int32_t D, int32_t H, int32_t W,
                          int32_t C, void* stream);
/* dst[v][c] (+)= src[v][off + c] for a source with Cs channels. */
int omg_take_channels(const uint16_t* src, int32_t Cs, int32_t off, uint16_t* dst, int64_t V, int32_t C,
                      int32_t accumulate, void* stream);
/* Output layer out[v] = base[v] + b + sum_c w[c] X[v][c] (FP32) and its backward. */
int omg_head_forward(const uint16_t* X, const float* w, const float* b, const float* base, float* out, int32_t V,
                     int32_t C, void* stream);
int omg_head_backward(const uint16_t* X, const float* w, const float* g, uint16_t* dX, float* dw, float* db, int32_t V,
                      int32_t C, void* stream);
/* Masked MSE: loss += sum (mask (out - t))^2 * inv_n (FP64), g = 2 mask (out - t) * inv_n. */
int omg_mse(const float* out, const float* t, const float* mask, float* g, double* loss, int32_t V, float inv_n,
            void* stream);
/* Adam step `step` (1-based) on FP32 parameters; q (optional) receives the BF16 copy. */
int omg_adam(float* p, const float* g, float* m, float* v, uint16_t* q, int32_t n, float lr, float b1, float b2,
             float eps, int32_t step, void* stream);
int omg_to_bf16(const float* src, uint16_t* dst, int64_t n, void* stream);
/* Network input: img[Nz][H][W] * scale into channel 0 of a 16-channel BF16 volume with D >= Nz planes. */
int omg_embed(const float* img, uint16_t* X, int32_t D, int32_t H, int32_t W, int32_t Nz, float scale, void* stream);

/* Relative difference prior gradient: grad += beta * dR/du, 6-neighbourhood.
 * boundary 0: zero outside the volume (CPU RDPKernel, source/cpp/RDP.h:4-78);
 * boundary 1: clamp to edge (GPU RDPKernel with images, source/opencl/auxKernels.cl:664-873). */
int omg_rdp_gradient(const float* u, float* grad, uint32_t Nx, uint32_t Ny, uint32_t Nz, float gamma, float epps,
                     float beta, int32_t boundary, void* stream);

/* Elementwise image/measurement updates (FP32). */
/* r = y / (yhat + epps), NaN/Inf -> epps (computeForwardProject.h:20-21,94,400-401); in place on yhat_r. */
int omg_ratio(const float* y, float* yhat_r, int64_t n, float epps, void* stream);
/* s = max-clamp: s < epps -> epps (functions.hpp:2996). */
int omg_sens_clamp(float* s, int64_t n, float epps, void* stream);
/* OSEM: x = x / S * (rhs + epps); positivity x < epps -> epps; optionally zero rhs afterwards
 * (algorithms.h:12-15, subiterStep.h:111-121, reconstructionAF.h:1461-1465). */
int omg_osem_update(float* x, const float* S, float* rhs, int64_t n, float epps, int32_t positivity,
                    int32_t clear_rhs, void* stream);
/* OSL-OSEM: x = x / (S + dU) * rhs; positivity; optionally zero rhs (subiterStep.h:219-231). */
int omg_osl_update(float* x, const float* S, const float* dU, float* rhs, int64_t n, float epps,
                   int32_t positivity, int32_t clear_rhs, void* stream);
/* PDHG image update (auxKernels.cl:2034-2059, PDHGUpdate). subsets != 0: im = max(epps, im - tau*rhs).
 * subsets == 0: u_new = max(epps, u - tau*rhs); im = u_new + theta*(u_new - u); u = u_new. */
int omg_pdhg_update(float* im, const float* rhs, float* u, int64_t n, float epps, float theta, float tau,
                    int32_t positivity, int32_t subsets, void* stream);
/* PDHG dual step for the least-squares data term (computeForwardProject.h:227-316):
 * p = (p + sigma * (Ax - y)) / (1 + sigma); in place on p, Ax read from ax. */
int omg_pdhg_dual_l2(float* p, const float* ax, const float* y, int64_t n, float sigma, void* stream);

/* Block-reduced dot products (double accumulation) for CGLS/PDHG scalars:
 * out[0] = sum a*b. out is a device double. */
int omg_dot(const float* a, const float* b, int64_t n, double* out, void* stream);
/* y = a*x + b*y. */
int omg_axpby(float* y, const float* x, int64_t n, float a, float b, void* stream);

/* Objectives (absent in the reference, which never evaluates them; reconstructionAF only
 * ascends them). Both accumulate in double into the device scalar *out.
 * Poisson log-likelihood L = sum_i y_i log(yhat_i) - yhat_i (terms with y_i = 0 are -yhat_i). */
int omg_poisson_loglik(const float* y, const float* yhat, int64_t n, double* out, void* stream);
/* RDP value R(u) = beta * [ sum over interior 6-neighbour pairs {j,k} of phi(u_j, u_k)
 * + sum over voxels j and their neighbours outside the volume of phi(u_j, 0) ],
 * phi(a, b) = (a - b)^2 / (a + b + gamma |a - b|). dR/du is the eps -> 0 limit of
 * omg_rdp_gradient with boundary 0 (RDPKernel, source/cpp/RDP.h:4-78). */
int omg_rdp_value(const float* u, uint32_t Nx, uint32_t Ny, uint32_t Nz, float gamma, float beta, double* out,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OMEGA_HIP_H */
