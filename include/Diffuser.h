/*
 * Diffuser.h -- the reference's Diffuser class (src/Diffusion/Diffuser.h:24-70),
 * same public interface, served by the MI355X layer.  The implementation is
 * vf-version_7_3_amd/cxx/Diffuser.cpp, compiled into the ABM's build beside
 * its WHChemical.h.
 *
 * diffuseChemGPU(rightWallIndex) advances every chem one tick: h_ibuffs[ic]
 * (WHChemical pChem) -> device -> constant-VF-padded FFT convolution with the
 * chem's resident spectrum -> h_obuffs[ic] (WHChemical tChem), with the copies
 * of one chem overlapped with the diffusion of the others (vfdDiffuseChems).
 * The reference's explicit-stencil diffuseChemCPU is not part of this layer.
 */
#ifndef VFD_DIFFUSER_H_
#define VFD_DIFFUSER_H_

#include "DiffusionHelper.h"
#include "vf_diffusion.h"

class WHChemical;

class Diffuser {
public:
    Diffuser();
    Diffuser(int nBaseChem, c_ctx* cc, c_ctx* kc, WHChemical* WHWorldChem);
    virtual ~Diffuser();

    void diffuseChemGPU(int rightWallIndex);

private:
    int nBaseChem;
    c_ctx* chem_cctx;
    c_ctx* kernel_cctx;
    WHChemical* WHWorldChem;
    int iter;
};

#endif /* VFD_DIFFUSER_H_ */
