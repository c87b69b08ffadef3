// Diffuser.cpp -- the reference's Diffuser (src/Diffusion/Diffuser.cpp:16-470)
// on the MI355X layer.  Compiled into the ABM's build beside its WHChemical.h
// and linked against libvfd.so.
//
// The reference's diffuseChemGPU creates two cuFFT plans per GPU every tick,
// then for each chem copies pChem to the device, runs fftDiffuse3D and copies
// the result back into tChem, one chem at a time (Diffuser.cpp:264-431).  Here
// one call, vfdDiffuseChems, does the same work for every chem with the
// copies of one chem overlapped with the diffusion of the others; the plan
// and workspace for the chem box persist between ticks.
#include "Diffuser.h"

#include <cstdio>

#include "WHChemical.h"

Diffuser::Diffuser()
    : nBaseChem(8), chem_cctx(nullptr), kernel_cctx(nullptr), WHWorldChem(nullptr), iter(0) {}

Diffuser::Diffuser(int nBaseChem, c_ctx* cc, c_ctx* kc, WHChemical* WHWorldChem)
    : nBaseChem(nBaseChem), chem_cctx(cc), kernel_cctx(kc), WHWorldChem(WHWorldChem), iter(0) {}

Diffuser::~Diffuser() {}

// One tick of every chem (Diffuser.cpp:221-470): h_ibuffs[ic] -> diffused with
// the constant-VF pad (baseline WHChemical::pbaseline[ic], epithelium at
// rightWallIndex) -> h_obuffs[ic].
void Diffuser::diffuseChemGPU(int rightWallIndex) {
    if (!chem_cctx || !WHWorldChem) {
        fprintf(stderr, "Diffuser::diffuseChemGPU: no context\n");
        return;
    }
    const int n = nBaseChem < N_CHEM ? nBaseChem : N_CHEM;
    if (!vfdDiffuseChems(*chem_cctx, n, rightWallIndex, WHWorldChem->pbaseline))
        fprintf(stderr, "Diffuser::diffuseChemGPU: %s\n", vfd_last_error());
    this->iter++;
}
