// DiffusionHelper.cpp -- diffusion_helper (src/Diffusion/DiffusionHelper.h:30-43)
// on the MI355X layer.  Compiled into the ABM's build (it needs the ABM's
// WHChemical.h for the pChem / tChem accessors), linked against libvfd.so;
// it calls only the C interface of include/vf_diffusion.h.
//
// Context contents follow the reference (DiffusionHelper.cpp:29-463) for the 3D
// model: the same kernel and chem dimensions, window sizes, centres and FFT
// boxes, so fftDiffuse3D / computeKernel3DBatch see the same c_ctx.  Buffers
// differ in ownership only (see include/DiffusionHelper.h).
#include "DiffusionHelper.h"

#include <cstdio>
#include <iostream>
#include <mutex>
#include <set>

#include "WHChemical.h"

namespace {

// buffers this layer allocated (device and pinned host), so that
// deallocConvCtxBuffers frees each one once and nothing it does not own
std::mutex g_own_mu;
std::set<void*> g_dev_owned, g_host_owned;

void* dev_alloc(size_t bytes) {
    void* p = vfdMalloc(bytes);
    if (!p) {
        fprintf(stderr, "diffusion_helper: device allocation of %zu bytes failed: %s\n", bytes, vfd_last_error());
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_own_mu);
    g_dev_owned.insert(p);
    return p;
}

void* host_alloc(size_t bytes) {
    void* p = vfdHostAlloc(bytes);
    if (!p) {
        fprintf(stderr, "diffusion_helper: pinned allocation of %zu bytes failed: %s\n", bytes, vfd_last_error());
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_own_mu);
    g_host_owned.insert(p);
    return p;
}

void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_own_mu);
    if (g_dev_owned.erase(p)) vfdFree(p);
    else if (g_host_owned.erase(p)) vfdHostFree(p);
}

void assign_gpus(c_ctx* c) {
    // chem ic runs on GPU ic % N_GPU (DiffusionHelper.cpp:34-41); one GPU here
    for (int ic = 0; ic < N_CHEM; ic++) {
        c->gpu_id[ic] = ic % N_GPU;
        c->dev_id[ic] = c->gpu_id[ic];
    }
}

}  // namespace

namespace diffusion_helper {

// DiffusionHelper.cpp:6-26
void findGPUs() {
    const int total = vfdDeviceCount();
    if (N_GPU > total) printf("Error: Only have %d GPUs. Require %d.\n", total, N_GPU);
}

// DiffusionHelper.cpp:29-163 (MODEL_3D): kernel-build dimensions.  The kernel
// is niter = totalSec/nSec rounds of nSec/dt stencil steps; its domain is
// 2*(kernelRadius-1)+1 wide, its FFT box snapTransformSize(domain + 2), and
// the window kept for the chems (kernelRadius/2)+1 wide (147 and a 616 box
// for HUMAN_VF, domain/2 for RAT_VF).
void initKernelDimensions(c_ctx* kernel_cctx) {
    assign_gpus(kernel_cctx);
    const int totalSec = 1800;
#ifdef RAT_VF
    const int nSec = 200;
#else
    const int nSec = 300;
#endif
    const int k_niter = totalSec / nSec;
    const float dt = kernel_cctx->dt;
    if (totalSec % nSec) printf("ERROR: totalSec (%d) NOT integer multiple of nSec (%d)\n", totalSec, nSec);
#ifdef RAT_VF
    const int kernelRadius = nSec;
#else
    const int kernelRadius = (int)(nSec / dt);
#endif
    const int outKernel = (kernelRadius - 1) * 2 + 1;
#if defined(HUMAN_VF)
    const int kfft = 616, window = 147;
    const int kfftD = kfft, kfftH = kfft, kfftW = kfft;
    const int windowD = window, windowH = window, windowW = window;
#elif defined(RAT_VF)
    const int kfftD = snapTransformSize(outKernel + KCoeffsD - 1);
    const int kfftH = snapTransformSize(outKernel + KCoeffsH - 1);
    const int kfftW = snapTransformSize(outKernel + KCoeffsW - 1);
    const int windowD = outKernel / 2, windowH = outKernel / 2, windowW = outKernel / 2;
#else
    const int kfftD = snapTransformSize(outKernel + KCoeffsD - 1);
    const int kfftH = snapTransformSize(outKernel + KCoeffsH - 1);
    const int kfftW = snapTransformSize(outKernel + KCoeffsW - 1);
    const int windowD = kernelRadius / 2 + 1, windowH = kernelRadius / 2 + 1, windowW = kernelRadius / 2 + 1;
#endif
    kernel_cctx->niter = k_niter;
    kernel_cctx->dt = dt;
    kernel_cctx->kernelRadius = kernelRadius;
    kernel_cctx->DD = outKernel;
    kernel_cctx->DH = outKernel;
    kernel_cctx->DW = outKernel;
    kernel_cctx->KD = KCoeffsD;
    kernel_cctx->KH = KCoeffsH;
    kernel_cctx->KW = KCoeffsW;
    kernel_cctx->KZ = 1;
    kernel_cctx->KX = 1;
    kernel_cctx->KY = 1;
    kernel_cctx->FFTD = kfftD;
    kernel_cctx->FFTH = kfftH;
    kernel_cctx->FFTW = kfftW;
    kernel_cctx->windowD = windowD;
    kernel_cctx->windowH = windowH;
    kernel_cctx->windowW = windowW;
}

// DiffusionHelper.cpp:166-240 (MODEL_3D): one host and one device window per
// GPU, shared by that GPU's chems.  computeKernel3D builds its FP64 work
// arrays itself, so the kernel context carries no spectrum buffer.
void allocKernelBuffers(c_ctx* kernel_cctx) {
    const size_t wbytes = (size_t)kernel_cctx->windowD * kernel_cctx->windowH * kernel_cctx->windowW * sizeof(float);
    for (int ic = 0; ic < N_CHEM; ic++) {
        kernel_cctx->h_ibuffs[ic] = nullptr;
        kernel_cctx->h_obuffs[ic] = nullptr;
        kernel_cctx->h_kernelspectrum[ic] = nullptr;
        kernel_cctx->d_kernelspectrum_h[ic] = nullptr;
    }
    for (int ig = 0; ig < N_GPU; ig++) {
        vfdSetDevice(kernel_cctx->dev_id[ig]);
        kernel_cctx->h_data[ig] = (float*)host_alloc(wbytes);
        kernel_cctx->d_data[ig] = (float*)dev_alloc(wbytes);
    }
    for (int ic = N_GPU; ic < N_CHEM; ic++) {
        const int ig = kernel_cctx->gpu_id[ic];
        kernel_cctx->h_data[ic] = kernel_cctx->h_data[ig];
        kernel_cctx->d_data[ic] = kernel_cctx->d_data[ig];
    }
}

// DiffusionHelper.cpp:243-316: every pointer slot is reset; a buffer is freed
// once, and only if this layer allocated it.
void deallocConvCtxBuffers(c_ctx* ctx) {
    if (!ctx) return;
    for (int ic = 0; ic < N_CHEM; ic++) {
        void* slots[] = {ctx->h_ibuffs[ic], ctx->h_obuffs[ic], ctx->h_kernelspectrum[ic],
                         ctx->d_kernelspectrum_h[ic], ctx->d_data[ic], ctx->h_data[ic]};
        for (void* p : slots) release(p);
        ctx->h_ibuffs[ic] = nullptr;
        ctx->h_obuffs[ic] = nullptr;
        ctx->h_kernelspectrum[ic] = nullptr;
        ctx->d_kernelspectrum_h[ic] = nullptr;
        ctx->d_data[ic] = nullptr;
        ctx->h_data[ic] = nullptr;
    }
    // the layer's own workspaces (made again if diffusion continues)
    if (vfdReleaseWorkspaces() < 0) fprintf(stderr, "diffusion_helper: %s\n", vfd_last_error());
}

// DiffusionHelper.cpp:319-378
void printContext(c_ctx c) {
    using std::cout;
    using std::endl;
    cout << "dt" << "\t" << c.dt << endl;
    cout << "dx2" << "\t" << c.dx2 << endl;
    cout << "kernelRadius" << "\t" << c.kernelRadius << endl;
    cout << "niter" << "\t" << c.niter << endl << endl;
    cout << "KD" << "\t" << c.KD << endl << "KH" << "\t" << c.KH << endl << "KW" << "\t" << c.KW << endl << endl;
    cout << "KZ" << "\t" << c.KZ << endl << "KX" << "\t" << c.KX << endl << "KY" << "\t" << c.KY << endl << endl;
    cout << "DD" << "\t" << c.DD << endl << "DH" << "\t" << c.DH << endl << "DW" << "\t" << c.DW << endl << endl;
    cout << "FFTD" << "\t" << c.FFTD << endl << "FFTH" << "\t" << c.FFTH << endl << "FFTW" << "\t" << c.FFTW
         << endl << endl;
    cout << "windowD" << "\t" << c.windowD << endl << "windowH" << "\t" << c.windowH << endl << "windowW" << "\t"
         << c.windowW << endl << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "h_ibuffs[" << i << "]" << (void*)c.h_ibuffs[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "h_obuffs[" << i << "]" << (void*)c.h_obuffs[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "gpu_id[" << i << "]" << c.gpu_id[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "dev_id[" << i << "]" << c.dev_id[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "d_data[" << i << "]" << (void*)c.d_data[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "h_data[" << i << "]" << (void*)c.h_data[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++) cout << "h_kernelspectrum[" << i << "]" << (void*)c.h_kernelspectrum[i] << endl;
    cout << endl;
    for (int i = 0; i < N_CHEM; i++)
        cout << "d_kernelspectrum_h[" << i << "]" << (void*)c.d_kernelspectrum_h[i] << endl;
    cout << endl;
}

// DiffusionHelper.cpp:381-461 (MODEL_3D): chem FFT box snapTransformSize(n +
// window - 1) per axis (1152 x 1536 x 256 for HUMAN_VF), checked to hold the
// linear convolution; window centre from the window size.  The centre uses
// the reference's crossed even-size flags (KZ from x, KX from y, KY from z),
// which agree for the cubic odd windows every configuration builds.
void initChemDimensions(c_ctx* chem_cctx, c_ctx kernel_cctx, int nx, int ny, int nz) {
    if (!chem_cctx || nx < 1 || ny < 1 || nz < 1) {
        fprintf(stderr, "InitializeChemGPU Error: bad arguments\n");
        return;
    }
    const int windowW = kernel_cctx.windowW, windowH = kernel_cctx.windowH, windowD = kernel_cctx.windowD;
    const int chemH = ny, chemW = nx, chemD = nz;
#ifdef HUMAN_VF
    const int cfftD = 1152, cfftW = 256, cfftH = 1536;
#else
    const int cfftD = snapTransformSize(chemD + windowD - 1);
    const int cfftW = snapTransformSize(chemW + windowW - 1);
    const int cfftH = snapTransformSize(chemH + windowH - 1);
#endif
    bool bad = false;
    if (cfftD - windowD + 1 < chemD) {
        printf("InitializeChemGPU Error: bad dimensions\n\tcfftD: %d\twindowD: %d\tchemD: %d\n", cfftD, windowD, chemD);
        bad = true;
    }
    if (cfftH - windowH + 1 < chemH) {
        printf("InitializeChemGPU Error: bad dimensions\n\tcfftH: %d\twindowH: %d\tchemH: %d\n", cfftH, windowH, chemH);
        bad = true;
    }
    if (cfftW - windowW + 1 < chemW) {
        printf("InitializeChemGPU Error: bad dimensions\n\tcfftW: %d\twindowW: %d\tchemW: %d\n", cfftW, windowW, chemW);
        bad = true;
    }
    if (bad) return;  // the reference exits here; the context is left unchanged
    assign_gpus(chem_cctx);
    chem_cctx->DD = chemD;
    chem_cctx->DH = chemH;
    chem_cctx->DW = chemW;
    chem_cctx->KD = kernel_cctx.windowD;
    chem_cctx->KH = kernel_cctx.windowH;
    chem_cctx->KW = kernel_cctx.windowW;
    const int isEvenX = (kernel_cctx.windowW % 2) ? 0 : 1;
    const int isEvenY = (kernel_cctx.windowH % 2) ? 0 : 1;
    const int isEvenZ = (kernel_cctx.windowD % 2) ? 0 : 1;
    chem_cctx->KZ = (kernel_cctx.windowD / 2) - isEvenX;
    chem_cctx->KX = (kernel_cctx.windowW / 2) - isEvenY;
    chem_cctx->KY = (kernel_cctx.windowH / 2) - isEvenZ;
    chem_cctx->FFTD = cfftD;
    chem_cctx->FFTH = cfftH;
    chem_cctx->FFTW = cfftW;
}

// DiffusionHelper.cpp:464-647 (MODEL_3D): host buffers alias the WHChemical
// arrays (pChem in, tChem out), or pinned staging buffers under OPT_CHEM;
// every chem gets its own resident device field and spectrum.
void allocChemBuffers(c_ctx* chem_cctx, WHChemical* WHWorldChem, int nx, int ny, int nz) {
    const size_t fftbytes = (size_t)chem_cctx->FFTD * chem_cctx->FFTH * (chem_cctx->FFTW / 2 + 1) * sizeof(fComplex);
    const size_t databytes = (size_t)nx * ny * nz * sizeof(float);
    for (int ic = 0; ic < N_CHEM; ic++) {
        chem_cctx->h_data[ic] = nullptr;
        chem_cctx->h_kernelspectrum[ic] = nullptr;
#ifdef OPT_CHEM
        chem_cctx->h_obuffs[ic] = (float*)host_alloc(databytes);
        chem_cctx->h_ibuffs[ic] = (float*)host_alloc(databytes);
#else
        chem_cctx->h_obuffs[ic] = WHWorldChem->getTchemPtr(ic);
        chem_cctx->h_ibuffs[ic] = WHWorldChem->getPchemPtr(ic);
#endif
        vfdSetDevice(chem_cctx->dev_id[ic]);
        chem_cctx->d_data[ic] = (float*)dev_alloc(databytes);
        chem_cctx->d_kernelspectrum_h[ic] = (fComplex*)dev_alloc(fftbytes);
    }
}

// DiffusionHelper.cpp:650-693: in the 3D model the spectra are formed by
// computeKernel3DBatch (prepKernel); nothing to do here.
void prepAndComputeKernelSpectrum(c_ctx* chem_cctx, c_ctx kernel_cctx, float** h_dWindow) {
    (void)chem_cctx;
    (void)kernel_cctx;
    (void)h_dWindow;
}

// DiffusionHelper.cpp:696-794 (MODEL_3D): kernel windows and chem spectra for
// every chem, built on the device (computeKernel3DBatch).
void prepKernel(c_ctx* chem_cctx, c_ctx kernel_context, float* lambda, float* gamma, int numChem) {
    (void)numChem;
    if (!computeKernel3DBatch(kernel_context.kernelRadius, lambda, gamma, kernel_context.dt, kernel_context,
                              *chem_cctx))
        fprintf(stderr, "prepKernel: %s\n", vfd_last_error());
}

}  // namespace diffusion_helper
