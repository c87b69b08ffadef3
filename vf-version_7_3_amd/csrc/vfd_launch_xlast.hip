// Instantiates the x-last order's z and x passes (FP32, every schedule variant)
// for the lengths of the tuned set; xlast_supported() names them.
#include "vfd_xlast.hpp"
#include "vfd_launch.h"

#define VFD_XLAST_LENGTHS(X) X(32) X(64) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1152) X(1536)

namespace vfd {

namespace {

template <typename K>
int occupancy(K kernel, int nt) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, nt, 0) != hipSuccess || n <= 0) n = 1;
    return n;
}

int cu_count() {
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
    }
    return ncu;
}

template <int Q, int V, bool FWD>
hipError_t zr_go(const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw, hipStream_t st) {
    using C = ColCfg<float, Q, V, 1>;
    XlDesc g = d;
    if constexpr (C::SH > 1) {
        g.grid_segs = (unsigned)((d.ncolp + C::TC_LINE - 1) / C::TC_LINE);
        g.grid_fields = (unsigned)nfields;
        const unsigned long groups = (unsigned long)g.grid_segs * nfields;
        const unsigned long ids = (groups + 7) / 8 * 8 * C::SH;
        if (ids > 0x7fffffffu) return hipErrorInvalidConfiguration;
        hipLaunchKernelGGL((zr_kernel<float, Q, V, FWD>), dim3((unsigned)ids), dim3(C::NT), 0, st, fields, g,
                           (const float2*)tw);
    } else {
        dim3 grid((unsigned)((d.ncolp + C::TC - 1) / C::TC), (unsigned)nfields);
        hipLaunchKernelGGL((zr_kernel<float, Q, V, FWD>), grid, dim3(C::NT), 0, st, fields, g, (const float2*)tw);
    }
    return hipGetLastError();
}

template <int P, int V>
hipError_t xc_go(const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw, hipStream_t st) {
    using C = XcCfg<float, P, V>;
    const long rows = (long)d.hz * d.py;
    const long items = (rows + C::G - 1) / C::G * nfields;
    static int per_cu = 0;
    if (!per_cu) per_cu = occupancy(xc_kernel<float, P, V>, C::NT);
    // 4 rounds of resident work-groups, as the x-first row passes
    const long cap = 4L * cu_count() * per_cu;
    const unsigned grid = (unsigned)(items < cap ? items : cap);
    hipLaunchKernelGGL((xc_kernel<float, P, V>), dim3(grid), dim3(C::NT), 0, st, fields, d, (const float2*)tw,
                       nfields);
    return hipGetLastError();
}

template <int Q, bool FWD>
hipError_t zr_var(int variant, const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw,
                  hipStream_t st) {
    switch (variant) {
        case 0: return zr_go<Q, 0, FWD>(fields, nfields, d, tw, st);
        case 1: return zr_go<Q, 1, FWD>(fields, nfields, d, tw, st);
        case 2: return zr_go<Q, 2, FWD>(fields, nfields, d, tw, st);
        case 3: return zr_go<Q, 3, FWD>(fields, nfields, d, tw, st);
        default: return hipErrorInvalidValue;
    }
}

template <int P>
hipError_t xc_var(int variant, const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw,
                  hipStream_t st) {
    switch (variant) {
        case 0: return xc_go<P, 0>(fields, nfields, d, tw, st);
        case 1: return xc_go<P, 1>(fields, nfields, d, tw, st);
        case 2: return xc_go<P, 2>(fields, nfields, d, tw, st);
        case 3: return xc_go<P, 3>(fields, nfields, d, tw, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool xlast_supported(int pz, int px) {
    bool z = false, x = false;
#define VFD_IS(Q) z = z || pz == Q; x = x || px == Q;
    VFD_XLAST_LENGTHS(VFD_IS)
#undef VFD_IS
    return z && x;
}

hipError_t launch_zr(bool fwd, int pz, const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw,
                     hipStream_t st, int variant) {
    switch (pz) {
#define VFD_Z(Q)                                                                          \
    case Q:                                                                               \
        return fwd ? zr_var<Q, true>(variant, fields, nfields, d, tw, st)                 \
                   : zr_var<Q, false>(variant, fields, nfields, d, tw, st);
        VFD_XLAST_LENGTHS(VFD_Z)
#undef VFD_Z
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t launch_xc(int px, const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw,
                     hipStream_t st, int variant) {
    switch (px) {
#define VFD_X(P) \
    case P: return xc_var<P>(variant, fields, nfields, d, tw, st);
        VFD_XLAST_LENGTHS(VFD_X)
#undef VFD_X
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace vfd
