// diffuser_check.cpp -- drives the C++ diffusion surface (diffusion_helper +
// Diffuser, include/DiffusionHelper.h, include/Diffuser.h) the way the ABM
// does (woundHealingWorld.cpp:1009-1049 at init, :2403-2417 per tick), on a
// test WHChemical, and hands the results to the parity tests through one
// extern "C" function.
#include <cstring>
#include <vector>

#include "Diffuser.h"
#include "DiffusionHelper.h"
#include "WHChemical.h"

WHChemical::WHChemical(int nx, int ny, int nz) : np(nx * ny * nz) {
    for (int ic = 0; ic < 8; ++ic) {
        pChem[ic] = (float*)vfdHostAlloc((size_t)np * sizeof(float));
        tChem[ic] = (float*)vfdHostAlloc((size_t)np * sizeof(float));
    }
    pbaseline = new float[8]();
}

WHChemical::~WHChemical() {
    for (int ic = 0; ic < 8; ++ic) {
        vfdHostFree(pChem[ic]);
        vfdHostFree(tChem[ic]);
    }
    delete[] pbaseline;
}

// info[0..12] = kernel FFT box, kernel window, kernelRadius, niter, chem FFT
// box (D,H,W), chem window (KD,KH,KW), centre (KZ,KY,KX).  Between ticks the
// diffused field becomes the next input (the ABM's update with no agents and
// no extra decay).  spectra (optional) receives the 8 chem spectra; windows
// (optional) the 8 kernel-window slots (windowD x windowH x windowW each; the
// kernel context holds one window buffer per GPU, which every chem's build
// reuses, DiffusionHelper.cpp:166-240); rows (optional) the first spectrum
// row (kz = ky = 0, FFTW/2 + 1 values) of every chem.
extern "C" int diffuser_check(int nx, int ny, int nz, float dt, const float* lambda, const float* gamma,
                              const float* baseline, int rightWall, int ticks, float* fields, fComplex* spectra,
                              float* windows, fComplex* rows, int* info) {
    c_ctx* kc = new c_ctx();
    c_ctx* cc = new c_ctx();
    std::memset(kc, 0, sizeof(c_ctx));
    std::memset(cc, 0, sizeof(c_ctx));
    kc->dt = dt;
    cc->dt = dt;
    diffusion_helper::findGPUs();
    diffusion_helper::initKernelDimensions(kc);
    diffusion_helper::allocKernelBuffers(kc);
    diffusion_helper::initChemDimensions(cc, *kc, nx, ny, nz);
    int rc = cc->FFTD > 0 ? 0 : -1;
    WHChemical chem(nx, ny, nz);
    if (!rc) {
        for (int ic = 0; ic < 8; ++ic) {
            std::memcpy(chem.pChem[ic], fields + (size_t)ic * chem.np, (size_t)chem.np * sizeof(float));
            chem.pbaseline[ic] = baseline[ic];
        }
        diffusion_helper::allocChemBuffers(cc, &chem, nx, ny, nz);
        std::vector<float> lam(lambda, lambda + 8), gam(gamma, gamma + 8);
        diffusion_helper::prepKernel(cc, *kc, lam.data(), gam.data(), 8);
        const int out[13] = {kc->FFTD, kc->windowD, kc->kernelRadius, kc->niter, cc->FFTD, cc->FFTH, cc->FFTW,
                             cc->KD,   cc->KH,      cc->KW,           cc->KZ,    cc->KY,  cc->KX};
        std::memcpy(info, out, sizeof out);
        const size_t sn = (size_t)cc->FFTD * cc->FFTH * (cc->FFTW / 2 + 1);
        for (int ic = 0; ic < 8 && spectra && !rc; ++ic)
            rc = vfdMemcpy(spectra + ic * sn, cc->d_kernelspectrum_h[ic], sn * sizeof(fComplex));
        const size_t wn = (size_t)kc->windowD * kc->windowH * kc->windowW;
        for (int ic = 0; ic < 8 && windows && !rc; ++ic)
            rc = vfdMemcpy(windows + ic * wn, kc->d_data[ic], wn * sizeof(float));
        const size_t rn = (size_t)cc->FFTW / 2 + 1;
        for (int ic = 0; ic < 8 && rows && !rc; ++ic)
            rc = vfdMemcpy(rows + ic * rn, cc->d_kernelspectrum_h[ic], rn * sizeof(fComplex));
        Diffuser d(8, cc, kc, &chem);
        for (int t = 0; t < ticks && !rc; ++t) {
            d.diffuseChemGPU(rightWall);
            for (int ic = 0; ic < 8; ++ic)
                std::memcpy(chem.pChem[ic], chem.tChem[ic], (size_t)chem.np * sizeof(float));
        }
        for (int ic = 0; ic < 8; ++ic)
            std::memcpy(fields + (size_t)ic * chem.np, chem.pChem[ic], (size_t)chem.np * sizeof(float));
    }
    diffusion_helper::deallocConvCtxBuffers(cc);
    diffusion_helper::deallocConvCtxBuffers(kc);
    delete cc;
    delete kc;
    return rc;
}
