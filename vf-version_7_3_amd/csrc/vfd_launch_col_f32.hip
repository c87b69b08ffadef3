// Instantiates the col passes for float.
#define VFD_REAL float
#define VFD_COL_TU
#include "vfd_launch_impl.hpp"

namespace vfd {
template <int RM>
static hipError_t zdirect_go(const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                             hipStream_t st) {
    dim3 grid((unsigned)((cd.ncols + 63) / 64), (unsigned)nplanes, (unsigned)nfields);
    hipLaunchKernelGGL((zdirect_kernel<RM>), grid, dim3(64), 0, st, fields, cd);
    return hipGetLastError();
}

hipError_t launch_zdirect(int rm, const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                          hipStream_t st) {
    switch (rm) {
        case 8: return zdirect_go<8>(fields, nfields, cd, nplanes, st);
        case 12: return zdirect_go<12>(fields, nfields, cd, nplanes, st);
        case 16: return zdirect_go<16>(fields, nfields, cd, nplanes, st);
        default: return hipErrorInvalidValue;
    }
}
}  // namespace vfd
