/*
 * vf_diffusion.h -- the reference's diffusion C interface, served by libvfd.so.
 *
 * Drop-in replacement for src/Diffusion/convolutionFFT_common.h: the same
 * types (fComplex, c_ctx, pad_t), the same extern "C" names and argument
 * meaning, backed by the MI355X kernels of include/vfd.h.  Differences, all
 * on the error side: a failing call returns false (bool functions) and leaves
 * a message in vfd_last_error() instead of exiting the process; plans are
 * library handles (vfdPlan3d) instead of cuFFT plans, and fftDiffuse3D keeps
 * its workspace between calls instead of allocating per call.
 *
 * Every field and kernel buffer lives on one GPU: gpu_id[] / dev_id[] are 0.
 */
#ifndef VF_DIFFUSION_H
#define VF_DIFFUSION_H

#include <stdbool.h>
#include <stddef.h>

#include "vfd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* convolutionFFT_common.h:15 */
#ifndef VFD_NO_UINT_TYPEDEF
typedef unsigned int uint;
#endif

/* convolutionFFT_common.h:17-25: interleaved FP32 complex */
typedef struct {
    float x;
    float y;
} fComplex;

/* src/enums.h:35-40 */
#ifndef VFD_NO_PAD_T
typedef enum { pClampToBorder = 0, pRightWall = 1, pMirror = 2, pConstantVF = 3 } pad_t;
#endif

/* cuFFT's plan handle type (an int); here a libvfd plan id. */
#ifndef VFD_NO_CUFFTHANDLE_TYPEDEF
typedef int cufftHandle;
#endif
enum { VFD_R2C = 0x2a, VFD_C2R = 0x2c };

/* convolutionFFT_common.h:292-303.  One GPU holds every field. */
#define N_CHEM 8
#define N_GPU 1
#define KCoeffsH 3
#define KCoeffsW 3
#define KCoeffsD 3

/* convolutionFFT_common.h:305-346: identical field order and types. */
typedef struct CCTX {
    float dt;
    float dx2;
    int kernelRadius;
    int niter;

    int KD;
    int KH;
    int KW;

    int KZ;
    int KX;
    int KY;

    int DD;
    int DH;
    int DW;

    int FFTD;
    int FFTH;
    int FFTW;

    int windowD;
    int windowH;
    int windowW;

    float* h_ibuffs[N_CHEM];
    float* h_obuffs[N_CHEM];

    int gpu_id[N_CHEM];
    int dev_id[N_CHEM];

    float* d_data[N_CHEM];
    float* h_data[N_CHEM];

    fComplex* h_kernelspectrum[N_CHEM];
    fComplex* d_kernelspectrum_h[N_CHEM];
} c_ctx;

/* convolutionFFT_common.h:32-41 */
static inline int iDivUp(int a, int b) { return (a % b != 0) ? (a / b + 1) : (a / b); }
static inline int iAlignUp(int a, int b) { return (a % b != 0) ? (a - a % b + b) : a; }

/* convolutionFFT_common.h:384 (chem ownership); all zeros on one GPU */
extern int gpu_id[N_CHEM];

/* ---- sizes and transfers: Diffusion.cu:39-122, convolutionFFT_common.h:350-352 ---- */
int snapTransformSize(int dataSize);
void H2D(int ic, c_ctx* chem_cctx, int np);
void D2H(int ic, c_ctx* chem_cctx, int np);
void reportMemUsageGPU(void);

/* ---- kernel precompute: Diffusion.cu:988-1619, convolutionFFT_common.h:389-412 ----
 * computeKernel3D builds the reference's stencil heat-kernel window in FP64 on
 * the device; computeKernelSpectrum3D forms its R2C spectrum on the chem FFT box
 * (FP64 transform, stored as fComplex). */
bool computeKernel3DBatch(int kernelRadius, float lambda[N_CHEM], float gamma[N_CHEM], float dt,
                          c_ctx kern_cctx, c_ctx chem_cctx);
bool computeKernel3D(int kernelRadius, float lambda, float gamma, float dt, c_ctx cctx,
                     cufftHandle fftPlanFwd, cufftHandle fftPlanInv, short int ic);
bool computeKernelSpectrum3D(c_ctx kernel_cctx, c_ctx chem_cctx, short int ic);

/* ---- the hot path: Diffusion.cu:1622-1934, convolutionFFT_common.h:414-424 ----
 * d_Data [DD][DH][DW] is diffused in place with the constant-VF pad (interior
 * 0<z<DD-1, 0<y<DH-1, 0<x<epiBoundary, baseChem elsewhere) and the R2C-layout
 * spectrum d_KernelSpectrum on the FFTD x FFTH x FFTW box.  C has no default
 * arguments; C++ callers may omit iter. */
#ifdef __cplusplus
bool fftDiffuse3D(float* d_Data, fComplex* d_KernelSpectrum, cufftHandle fftPlanFwd,
                  cufftHandle fftPlanInv, c_ctx cctx, int epiBoundary, float baseChem, int chemIndex,
                  int iter = 0);
#else
bool fftDiffuse3D(float* d_Data, fComplex* d_KernelSpectrum, cufftHandle fftPlanFwd,
                  cufftHandle fftPlanInv, c_ctx cctx, int epiBoundary, float baseChem, int chemIndex,
                  int iter);
#endif

/* ---- batched per-tick form of the chem loop of Diffuser::diffuseChemGPU
 * (Diffuser.cpp:300-421): for ic < nchem, h_ibuffs[ic] -> d_data[ic] -> the
 * fftDiffuse3D operator (constant-VF pad with baseChem[ic] and epiBoundary,
 * spectrum d_kernelspectrum_h[ic]) -> h_obuffs[ic].  Copies in, diffusion and
 * copies out run on three streams, so one chem's copies overlap its
 * neighbours' diffusion; distinct d_data[ic] buffers give the most overlap.
 * Synchronous on return; false (and vfd_last_error) on bad arguments. ---- */
bool vfdDiffuseChems(c_ctx cctx, int nchem, int epiBoundary, const float* baseChem);

/* ---- one ABM tick of the 8 chems with the levels resident on the device:
 * Diffuser::diffuseChemGPU (Diffuser.cpp:221-470) followed by
 * WHChemical::update with gpuUpdate (WHChemical.cpp:395-450) in one call, with
 * only the agents' deltas crossing PCIe towards the device.  d_data[ic] holds
 * chem ic's levels (pChem) from the previous tick and keeps the new ones; the
 * caller loads it once (H2D) before the first tick, one buffer per chem.
 * Per tick: h_dchem[ic] -> device; the fftDiffuse3D operator on d_data[ic]
 * with p = d + t * decay[ic] fused into its last pass (FP32 product rounded
 * before the add) and the device delta cleared; new levels -> h_ibuffs[ic]
 * (pChem); h_dchem[ic] zeroed on the host; the ECM fragment fields
 * (fragmnt_col, fragmnt_ela of WHWorld::worldECM, packed) -> device; the
 * NEU, MAC, FIB gradients -> h_grads[0..2]; compensated totals -> totals[8].
 * Bit-identical to vfdDiffuseChems followed by the host update.  Synchronous
 * on return; false (and vfd_last_error) on bad arguments. ---- */
bool vfdTickChems(c_ctx cctx, int epiBoundary, const float* baseChem, float* const* h_dchem, const float* decay,
                  const float* h_frag_col, const float* h_frag_ela, float* const* h_grads, double* totals);

/* ---- vfdTickChems split at the agents' step, the go() split of
 * woundHealingWorld.cpp:2013-2055 (diffusion on one OpenMP thread while the
 * agents run on another): vfdTickChemsBegin queues the diffusion of every
 * chem's resident levels d_data[ic] (in place: they become tChem) and returns
 * without waiting, so the device diffuses while the agents write dChem and
 * read pChem on the host; vfdTickChemsEnd, after the agents, does the rest of
 * vfdTickChems (deltas and fragments in, p = d + t * decay, gradients and
 * totals, levels and gradients out, host dChem cleared) and returns when all
 * is done.  The numbers equal vfdTickChems'.  Between the two calls no other
 * diffusion entry point of this interface runs (they return false). ---- */
/* ---- frees what the interface keeps between calls (the per-box workspaces of
 * fftDiffuse3D / vfdDiffuseChems / vfdTickChems, the resident tick's arrays,
 * the pipeline's streams and events); later calls make them again.  0, or a
 * negative vfd error code (a pending vfdTickChemsBegin is one).  Called by
 * diffusion_helper::deallocConvCtxBuffers. ---- */
int vfdReleaseWorkspaces(void);

bool vfdTickChemsBegin(c_ctx cctx, int epiBoundary, const float* baseChem);
bool vfdTickChemsEnd(c_ctx cctx, float* const* h_dchem, const float* decay, const float* h_frag_col,
                     const float* h_frag_ela, float* const* h_grads, double* totals);

/* ---- memory for the diffusion_helper / Diffuser layer and WHChemical: stands
 * in for the cudaGetDeviceCount / cudaSetDevice / cudaMalloc / cudaFree /
 * cudaMallocHost / cudaFreeHost calls of DiffusionHelper.cpp:6-26, 160-225,
 * 464-588 and WHChemical.cpp:198-206.  Pointer functions return NULL on
 * failure; int functions return 0 or a negative vfd error code. ---- */
int vfdDeviceCount(void);
int vfdSetDevice(int dev);
void* vfdMalloc(size_t bytes);
void vfdFree(void* d_ptr);
void* vfdHostAlloc(size_t bytes); /* pinned (page-locked) host memory */
void vfdHostFree(void* h_ptr);
int vfdMemcpy(void* dst, const void* src, size_t bytes); /* any direction, synchronous */

/* ---- plan handles standing in for cufftPlan3d / cufftDestroy (Diffuser.cpp:273-274,
 * 429-430).  type is VFD_R2C or VFD_C2R; a plan only records the box. ---- */
int vfdPlan3d(cufftHandle* plan, int nz, int ny, int nx, int type);
int vfdDestroy(cufftHandle plan);

/* ---- device helpers: convolutionFFT.cu (3D) ---- */
void extractCenter3D(float* d_Dst, float* d_Src, int dataD, int dataH, int dataW, int windowD,
                     int windowH, int windowW);                                  /* .cu:106-136 */
void unpadResult3D(float* d_Dst, float* d_Src, int dataD, int dataH, int dataW, int fftD, int fftH,
                   int fftW);                                                     /* .cu:138-165 */
void padKernel3D(float* d_Dst, float* d_Src, int fftD, int fftH, int fftW, int kernelD, int kernelH,
                 int kernelW, int kernelZ, int kernelY, int kernelX);             /* .cu:318-351 */
void padDataClampToBorder3D(float* d_PaddedData, float* d_Data, int fftD, int fftH, int fftW,
                            int dataD, int dataH, int dataW, int kernelD, int kernelH, int kernelW,
                            int kernelZ, int kernelY, int kernelX);               /* .cu:442-481 */
void padDataConstantVF3D(float* d_Dst, float* d_Src, int fftD, int fftH, int fftW, int dataD,
                         int dataH, int dataW, int kernelD, int kernelH, int kernelW, int kernelZ,
                         int kernelY, int kernelX, int epiBoundary, float baseChem); /* .cu:356-399 */
void modulateAndNormalize3D(fComplex* d_Dst, fComplex* d_Src, int fftD, int fftH, int fftW,
                            int padding);                                         /* .cu:562-581 */
void modulate3D(fComplex* d_Dst, fComplex* d_Src, int fftD, int fftH, int fftW, int padding);
                                                                                  /* .cu:537-556 */
void complexPower(fComplex* d_Dst, int fftD, int fftH, int fftW, int padding, int n);
                                                                                  /* .cu:512-530 */

/* ---- the 2D operator (convolutionFFT_common.h:356-380, the !MODEL_3D build) ---- */
/* Diffusion.cu:303-675: the 5-point stencil's kernel after kernelRadius
 * convolution rounds on the FFT box; the centred windowH x windowW window
 * into d_Window. */
bool computeKernel(float* d_Window, int kernelRadius, float lambda, float gamma, float dt, c_ctx cctx);
/* Diffusion.cu:677-771: the packed spectrum (the C2C transform of the padded
 * kernel read as FFTH x FFTW/2 complex) of the chem_cctx KH x KW window. */
bool computeKernelSpectrum(fComplex* d_KernelSpectrum, float* d_Kernel, c_ctx kernel_cctx, c_ctx chem_cctx);
/* Diffusion.cu:773-907: constant-VF pad, convolution with the packed
 * spectrum's kernel, crop; in place on d_Data (DH x DW). */
bool fftDiffuse2D(float* d_Data, fComplex* d_KernelSpectrum0, c_ctx cctx, int epiBoundary, float baseChem);

/* ---- device helpers: convolutionFFT.cu (2D, the !MODEL_3D build) ---- */
void padData2D(pad_t pt, float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW,
               int kernelH, int kernelW, int kernelY, int kernelX, int epiBoundary, float baseChem);
                                                                       /* Diffusion.cu:227-300 */
void extractCenter(float* d_Dst, float* d_Src, int dataH, int dataW, int windowH, int windowW);
void unpadResult(float* d_Dst, float* d_Src, int dataH, int dataW, int fftH, int fftW);
void padKernel(float* d_PaddedKernel, float* d_Kernel, int fftH, int fftW, int kernelH, int kernelW,
               int kernelY, int kernelX);
void padDataConstantVF(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW,
                       int kernelH, int kernelW, int kernelY, int kernelX, int epiBoundary,
                       float baseChem);
void padDataRightWall(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW,
                      int kernelH, int kernelW, int kernelY, int kernelX);
void padDataMirror(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW,
                   int kernelH, int kernelW, int kernelY, int kernelX);
void padDataClampToBorder(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH,
                          int dataW, int kernelH, int kernelW, int kernelY, int kernelX);
void modulateAndNormalize(fComplex* d_Dst, fComplex* d_Src, int fftH, int fftW, int padding);
/* convolutionFFT_gold.cpp:22-58: host FP64 direct 2D clamp-to-border convolution. */
void convolutionClampToBorderCPU(float* h_Result, float* h_Data, float* h_Kernel, int dataH, int dataW,
                                 int kernelH, int kernelW, int kernelY, int kernelX);

/* ---- additions (no counterpart in the reference) ---- */
/* 1: cpow keeps the reference's atan(b/a) phase (default); 0: principal power. */
void vfdSetKernelPowerQuirk(int on);

#ifdef __cplusplus
}
#endif

#endif /* VF_DIFFUSION_H */
