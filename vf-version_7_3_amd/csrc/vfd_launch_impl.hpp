// vfd_launch_impl.hpp -- runtime-length dispatch of the templated passes.
// Included by vfd_launch_{row,col}_{f32,f64}.hip with VFD_REAL set.
#pragma once
#include "vfd_passes.hpp"
#include "vfd_launch.h"

// Transform lengths compiled into the library: 2^a 3^b (b <= 2) up to 2304,
// 5 * 2^a up to 1280,
// plus 616 = 8*7*11 and 308 (the production kernel-build box and its half).
#define VFD_LENGTHS(X)                                                                         \
    X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(3) X(6) X(12)      \
    X(24) X(48) X(96) X(192) X(384) X(768) X(1536) X(9) X(18) X(36) X(72) X(144) X(288) X(576) \
    X(1152) X(2304) X(616) X(308) X(5) X(10) X(20) X(40) X(80) X(160) X(320) X(640) X(1280)

namespace vfd {

#ifdef VFD_ROW_TU
template <>
hipError_t launch_xfwd<VFD_REAL>(int L, const FieldBatch& fields, int nfields, const GridDesc& g,
                                 const void* tw, hipStream_t st) {
    const long rows = (long)g.Ny * g.Nz;
    switch (L) {
#define VFD_CASE(Q)                                                                           \
    case Q: {                                                                                 \
        dim3 grid((unsigned)((rows + RowCfg<Q>::G - 1) / RowCfg<Q>::G), 1, (unsigned)nfields); \
        hipLaunchKernelGGL((xfwd_kernel<VFD_REAL, Q>), grid, dim3(RowCfg<Q>::NT), 0, st,     \
                           fields, g, (const cx<VFD_REAL>*)tw);                                \
        return hipGetLastError();                                                             \
    }
        VFD_LENGTHS(VFD_CASE)
#undef VFD_CASE
        default:
            return hipErrorInvalidValue;
    }
}

template <>
hipError_t launch_xinv<VFD_REAL>(int L, const FieldBatch& fields, int nfields, const GridDesc& g,
                                 const void* tw, hipStream_t st) {
    const long rows = (long)g.Ny * g.Nz;
    switch (L) {
#define VFD_CASE(Q)                                                                           \
    case Q: {                                                                                 \
        dim3 grid((unsigned)((rows + RowCfg<Q>::G - 1) / RowCfg<Q>::G), 1, (unsigned)nfields); \
        hipLaunchKernelGGL((xinv_kernel<VFD_REAL, Q>), grid, dim3(RowCfg<Q>::NT), 0, st,     \
                           fields, g, (const cx<VFD_REAL>*)tw);                                \
        return hipGetLastError();                                                             \
    }
        VFD_LENGTHS(VFD_CASE)
#undef VFD_CASE
        default:
            return hipErrorInvalidValue;
    }
}
#endif

#ifdef VFD_COL_TU
template <typename T, int Q, int MODE, int RS>
static hipError_t col_go(const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                         const void* tw, hipStream_t st) {
    using C = ColCfg<T, Q>;
    dim3 grid((unsigned)((cd.ncols + C::TC - 1) / C::TC), (unsigned)nplanes, (unsigned)nfields);
    hipLaunchKernelGGL((col_kernel<T, Q, MODE, RS>), grid, dim3(C::NT), 0, st, fields, cd,
                       (const cx<T>*)tw);
    return hipGetLastError();
}

template <typename T, int Q>
static hipError_t col_mode(const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                           const void* tw, hipStream_t st) {
    const int key = cd.mode * 2 + (cd.rs - 1);
    switch (key) {
        case COL_FWD * 2 + 0: return col_go<T, Q, COL_FWD, 1>(fields, nfields, cd, nplanes, tw, st);
        case COL_FWD * 2 + 1: return col_go<T, Q, COL_FWD, 2>(fields, nfields, cd, nplanes, tw, st);
        case COL_INV * 2 + 0: return col_go<T, Q, COL_INV, 1>(fields, nfields, cd, nplanes, tw, st);
        case COL_INV * 2 + 1: return col_go<T, Q, COL_INV, 2>(fields, nfields, cd, nplanes, tw, st);
        case COL_FUSED * 2 + 0: return col_go<T, Q, COL_FUSED, 1>(fields, nfields, cd, nplanes, tw, st);
        case COL_FUSED * 2 + 1: return col_go<T, Q, COL_FUSED, 2>(fields, nfields, cd, nplanes, tw, st);
        default: return hipErrorInvalidValue;
    }
}

template <>
hipError_t launch_col<VFD_REAL>(int Q, const FieldBatch& fields, int nfields, const ColDesc& cd,
                                long nplanes, const void* tw, hipStream_t st) {
    switch (Q) {
#define VFD_CASE(QQ) \
    case QQ: return col_mode<VFD_REAL, QQ>(fields, nfields, cd, nplanes, tw, st);
        VFD_LENGTHS(VFD_CASE)
#undef VFD_CASE
        default:
            return hipErrorInvalidValue;
    }
}
#endif

}  // namespace vfd
