/*
 * DiffusionHelper.h -- the reference's diffusion_helper namespace
 * (src/Diffusion/DiffusionHelper.h:30-43), same signatures, served by the
 * MI355X layer (libvfd.so).  The implementation is
 * vf-version_7_3_amd/cxx/DiffusionHelper.cpp, compiled into the ABM's build
 * beside its own WHChemical.h, as the reference's DiffusionHelper.cpp is.
 *
 * Differences, all on the ownership and error side:
 *   - every chem gets its own device field buffer d_data[ic] and device
 *     spectrum d_kernelspectrum_h[ic] (resident; no per-tick spectrum upload),
 *     and h_kernelspectrum[ic] stays NULL (spectra never travel to the host);
 *   - deallocConvCtxBuffers frees only what this layer allocated, each buffer
 *     once, and never the WHChemical arrays that h_ibuffs / h_obuffs alias;
 *   - initChemDimensions reports bad dimensions through vfd_last_error() and
 *     leaves the context unchanged instead of exiting the process.
 */
#ifndef VFD_DIFFUSIONHELPER_H_
#define VFD_DIFFUSIONHELPER_H_

#include "vf_diffusion.h"

class WHChemical;

namespace diffusion_helper {

void findGPUs();
void initKernelDimensions(c_ctx* kernel_cctx);
void allocKernelBuffers(c_ctx* kernel_cctx);
void deallocConvCtxBuffers(c_ctx* ctx);
void initChemDimensions(c_ctx* chem_cctx, c_ctx kernel_cctx, int nx, int ny, int nz);
void allocChemBuffers(c_ctx* chem_cctx, WHChemical* WHWorldChem, int nx, int ny, int nz);
void prepAndComputeKernelSpectrum(c_ctx* chem_cctx, c_ctx kernel_cctx, float** h_dWindow);
void prepKernel(c_ctx* chem_cctx, c_ctx kernel_context, float* lambda, float* gamma, int numChem);
void printContext(c_ctx c);

}  // namespace diffusion_helper

#endif /* VFD_DIFFUSIONHELPER_H_ */
