// WHChemical.h -- test stand-in for the ABM's WHChemical
// (src/FieldVariable/Usr_FieldVariables/WHChemical.h:81-82, 144-175) with only
// what the diffusion layer touches: the pChem / tChem field arrays (pinned, as
// WHChemical.cpp:198-206 allocates them), their accessors, pbaseline and np.
#ifndef VFD_TEST_WHCHEMICAL_H_
#define VFD_TEST_WHCHEMICAL_H_

class WHChemical {
public:
    WHChemical(int nx, int ny, int nz);
    ~WHChemical();
    float* getPchemPtr(int ic) { return pChem[ic]; }
    float* getTchemPtr(int ic) { return tChem[ic]; }

    float* tChem[8];
    float* pChem[8];
    float* pbaseline;
    int np;
};

#endif
