// vfd_launch_impl.hpp -- runtime-length dispatch of the templated passes.
// Included by vfd_launch_{row,col}_{f32,f64}.hip with VFD_REAL set.
#pragma once
#include "vfd_passes.hpp"
#include "vfd_launch.h"

// Transform lengths compiled into the library: 2^a 3^b (b <= 2) up to 2304,
// 5 * 2^a up to 1280, plus 616 = 8*7*11 and 308 (the production kernel-build
// box and its half).
#define VFD_COLD_LENGTHS(X)                                                                    \
    X(2) X(4) X(8) X(16) X(2048) X(3) X(6) X(12) X(24) X(48) X(96) X(9) X(18) X(36) X(72)      \
    X(144) X(288) X(576) X(2304) X(616) X(308) X(5) X(10) X(20) X(40) X(80) X(160)     \
    X(320) X(640) X(1280)
// lengths of the tissue-scale, ensemble, precision and production (1152 x 1536
// z-y box) configurations: every
// schedule variant is compiled for these and the autotuner picks one
#define VFD_HOT_LENGTHS(X) X(32) X(64) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1152) X(1536)

namespace vfd {

// resident work-groups per CU for a kernel (callers cache it per instantiation)
template <typename K>
static int blocks_per_cu(K kernel, int nt) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, nt, 0) != hipSuccess || n <= 0) n = 1;
    return n;
}

[[maybe_unused]] static int num_cus() {
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
    }
    return ncu;
}

#ifdef VFD_ROW_TU
template <typename T, int Q, int V, bool FWD, bool VIS = false>
static hipError_t row_go(const FieldBatch& fields, int nfields, const GridDesc& g, const void* tw,
                         hipStream_t st) {
    using C = RowCfg<T, Q, V>;
    const long rows = (long)g.Ny * g.Nz;
    const long items = (rows + C::G - 1) / C::G * nfields;
    auto kern = FWD ? xfwd_kernel<T, Q, V> : xinv_kernel<T, Q, V, VIS>;
    // 4 rounds of resident work-groups: the rows whose waves drift apart
    // (wave-level stage sync) balance better over the CUs than with one round
    // (x passes 1.34 -> 1.31 ms at 1024x512x128, 8 fields), while a work-group
    // per row group restages the twiddle table for every 4 rows (1.65 ms)
    static const int per_cu = blocks_per_cu(kern, C::NT);  // this instantiation's
    const long cap = 4L * num_cus() * per_cu;
    const unsigned grid = (unsigned)(items < cap ? items : cap);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(C::NT), 0, st, fields, g, (const cx<T>*)tw, nfields);
    return hipGetLastError();
}

template <typename T, int Q, bool FWD>
static hipError_t row_var(int variant, const FieldBatch& fields, int nfields, const GridDesc& g,
                          const void* tw, hipStream_t st) {
    switch (variant) {
        case 0: return row_go<T, Q, 0, FWD>(fields, nfields, g, tw, st);
        case 1: return row_go<T, Q, 1, FWD>(fields, nfields, g, tw, st);
        case 2: return row_go<T, Q, 2, FWD>(fields, nfields, g, tw, st);
        case 3: return row_go<T, Q, 3, FWD>(fields, nfields, g, tw, st);
        default: return hipErrorInvalidValue;
    }
}

template <bool FWD>
static hipError_t row_dispatch(int L, int variant, const FieldBatch& fields, int nfields, const GridDesc& g,
                               const void* tw, hipStream_t st) {
    constexpr bool kVariants = sizeof(VFD_REAL) == 4;  // FP64 keeps the default schedule
    switch (L) {
#define VFD_HOT(Q)                                                                   \
    case Q:                                                                          \
        if constexpr (kVariants)                                                     \
            return row_var<VFD_REAL, Q, FWD>(variant, fields, nfields, g, tw, st);   \
        else                                                                         \
            return row_go<VFD_REAL, Q, 0, FWD>(fields, nfields, g, tw, st);
#define VFD_COLD(Q) \
    case Q: return row_go<VFD_REAL, Q, 0, FWD>(fields, nfields, g, tw, st);
        VFD_HOT_LENGTHS(VFD_HOT)
        VFD_COLD_LENGTHS(VFD_COLD)
#undef VFD_HOT
#undef VFD_COLD
        default:
            return hipErrorInvalidValue;
    }
}

template <>
hipError_t launch_xfwd<VFD_REAL>(int L, const FieldBatch& fields, int nfields, const GridDesc& g,
                                 const void* tw, hipStream_t st, int variant) {
    return row_dispatch<true>(L, variant, fields, nfields, g, tw, st);
}

// x inverse with the fused visualisation downsample: FP32, default schedule
static hipError_t xinv_vis(int L, const FieldBatch& fields, int nfields, const GridDesc& g, const void* tw,
                           hipStream_t st) {
    if constexpr (sizeof(VFD_REAL) == 4) {
        switch (L) {
#define VFD_ANY(Q) \
    case Q: return row_go<VFD_REAL, Q, 0, false, true>(fields, nfields, g, tw, st);
            VFD_HOT_LENGTHS(VFD_ANY)
            VFD_COLD_LENGTHS(VFD_ANY)
#undef VFD_ANY
            default:
                return hipErrorInvalidValue;
        }
    }
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_xinv<VFD_REAL>(int L, const FieldBatch& fields, int nfields, const GridDesc& g,
                                 const void* tw, hipStream_t st, int variant, bool vis) {
    if (vis) return xinv_vis(L, fields, nfields, g, tw, st);
    return row_dispatch<false>(L, variant, fields, nfields, g, tw, st);
}
#endif

#ifdef VFD_COL_TU
template <typename T, int Q, int MODE, int RS, int V, int KS = KS_REAL, bool FOLD = false>
static hipError_t col_go(const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                         const void* tw, hipStream_t st) {
    using C = ColCfg<T, Q, V, RS>;
    if constexpr (C::SH > 1) {
        // grouped 1-D launch (col_kernel): 8 segment groups per 8 * SH ids
        ColDesc g = cd;
        g.grid_segs = (unsigned)((cd.ncols + C::TC_LINE - 1) / C::TC_LINE);
        g.grid_planes = (unsigned)nplanes;
        g.grid_fields = (unsigned)nfields;
        const unsigned long groups = (unsigned long)g.grid_segs * nplanes * nfields;
        const unsigned long ids = (groups + 7) / 8 * 8 * C::SH;
        if (ids > 0x7fffffffu) return hipErrorInvalidConfiguration;
        hipLaunchKernelGGL((col_kernel<T, Q, MODE, RS, V, KS, FOLD>), dim3((unsigned)ids), dim3(C::NT), 0, st, fields,
                           g, (const cx<T>*)tw);
    } else {
        dim3 grid((unsigned)((cd.ncols + C::TC - 1) / C::TC), (unsigned)nplanes, (unsigned)nfields);
        hipLaunchKernelGGL((col_kernel<T, Q, MODE, RS, V, KS, FOLD>), grid, dim3(C::NT), 0, st, fields, cd,
                           (const cx<T>*)tw);
    }
    return hipGetLastError();
}

// fused pass: the spectrum format picks the instantiation
template <typename T, int Q, int RS, int V>
static hipError_t col_fused(const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                            const void* tw, hipStream_t st) {
    if (cd.kspec_sep) return col_go<T, Q, COL_FUSED, RS, V, KS_SEP>(fields, nfields, cd, nplanes, tw, st);
    if (cd.kspec_complex) return col_go<T, Q, COL_FUSED, RS, V, KS_COMPLEX>(fields, nfields, cd, nplanes, tw, st);
    return col_go<T, Q, COL_FUSED, RS, V, KS_REAL>(fields, nfields, cd, nplanes, tw, st);
}

template <typename T, int Q, int V>
static hipError_t col_mode(const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                           const void* tw, hipStream_t st) {
    if (cd.fold) {
        if constexpr (sizeof(T) == 4 && fold_supported(Q)) {
            if (cd.rs == 2 && cd.mode == COL_INV)
                return col_go<T, Q, COL_INV, 2, V, KS_REAL, true>(fields, nfields, cd, nplanes, tw, st);
        }
        return hipErrorInvalidValue;
    }
    const int key = cd.mode * 2 + (cd.rs - 1);
    switch (key) {
        case COL_FWD * 2 + 0: return col_go<T, Q, COL_FWD, 1, V>(fields, nfields, cd, nplanes, tw, st);
        case COL_FWD * 2 + 1: return col_go<T, Q, COL_FWD, 2, V>(fields, nfields, cd, nplanes, tw, st);
        case COL_INV * 2 + 0: return col_go<T, Q, COL_INV, 1, V>(fields, nfields, cd, nplanes, tw, st);
        case COL_INV * 2 + 1: return col_go<T, Q, COL_INV, 2, V>(fields, nfields, cd, nplanes, tw, st);
        case COL_FUSED * 2 + 0: return col_fused<T, Q, 1, V>(fields, nfields, cd, nplanes, tw, st);
        case COL_FUSED * 2 + 1: return col_fused<T, Q, 2, V>(fields, nfields, cd, nplanes, tw, st);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int Q>
static hipError_t col_var(int variant, const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                          const void* tw, hipStream_t st) {
    switch (variant) {
        case 0: return col_mode<T, Q, 0>(fields, nfields, cd, nplanes, tw, st);
        case 1: return col_mode<T, Q, 1>(fields, nfields, cd, nplanes, tw, st);
        case 2: return col_mode<T, Q, 2>(fields, nfields, cd, nplanes, tw, st);
        case 3: return col_mode<T, Q, 3>(fields, nfields, cd, nplanes, tw, st);
        default: return hipErrorInvalidValue;
    }
}

template <>
hipError_t launch_col<VFD_REAL>(int Q, const FieldBatch& fields, int nfields, const ColDesc& cd,
                                long nplanes, const void* tw, hipStream_t st, int variant) {
    constexpr bool kVariants = sizeof(VFD_REAL) == 4;  // FP64 keeps the default schedule
    switch (Q) {
#define VFD_HOT(QQ)                                                                     \
    case QQ:                                                                            \
        if constexpr (kVariants)                                                        \
            return col_var<VFD_REAL, QQ>(variant, fields, nfields, cd, nplanes, tw, st); \
        else                                                                            \
            return col_mode<VFD_REAL, QQ, 0>(fields, nfields, cd, nplanes, tw, st);
#define VFD_COLD(QQ) \
    case QQ: return col_mode<VFD_REAL, QQ, 0>(fields, nfields, cd, nplanes, tw, st);
        VFD_HOT_LENGTHS(VFD_HOT)
        VFD_COLD_LENGTHS(VFD_COLD)
#undef VFD_HOT
#undef VFD_COLD
        default:
            return hipErrorInvalidValue;
    }
}
#endif

}  // namespace vfd
