// vfd_launch.h -- host entry points of the templated passes (one per real type).
#pragma once
#include <hip/hip_runtime.h>
#include "vfd_types.h"

namespace vfd {

// True when a transform of this length is compiled into the library.
bool length_supported(int q);
// Sub-lengths Q (= P/2) with the folded even/odd split of the FP32 inverse y
// pass compiled (col_kernel FOLD): long columns whose unsplit tiles are narrow.
constexpr bool fold_supported(int q) { return q == 512 || q == 768; }

// x-forward (pad + R2C) over Ny*Nz rows of nfields fields; L = Px/2.
// `variant` selects a compiled schedule (0..3, see vfd_passes.hpp); lengths
// without variants run schedule 0.
constexpr int kScheduleVariants = 4;
template <typename T>
hipError_t launch_xfwd(int L, const FieldBatch& fields, int nfields, const GridDesc& g,
                       const void* tw, hipStream_t st, int variant = 0);
// x-inverse (C2R + unpad).
template <typename T>
hipError_t launch_xinv(int L, const FieldBatch& fields, int nfields, const GridDesc& g,
                       const void* tw, hipStream_t st, int variant = 0, bool vis = false);
// column pass of sub-length Q (P = Q * cd.rs) over nplanes planes.
template <typename T>
hipError_t launch_col(int Q, const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                      const void* tw, hipStream_t st, int variant = 0);

// direct z pass (FP32, separable spectrum) for window radius R <= rm, rm one of
// 8, 12, 16 (kZDirectMaxR); cd.scale carries Pz.
hipError_t launch_zdirect(int rm, const FieldBatch& fields, int nfields, const ColDesc& cd, long nplanes,
                          hipStream_t st);

// x-last order (FP32): z R2C (fwd) / C2R (inverse) of column pairs with FFT
// length pz, and the fused x row pass of length px.  xlast_supported tells
// whether both lengths are compiled.
bool xlast_supported(int pz, int px);
hipError_t launch_zr(bool fwd, int pz, const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw,
                     hipStream_t st, int variant = 0);
hipError_t launch_xc(int px, const FieldBatch& fields, int nfields, const XlDesc& d, const void* tw,
                     hipStream_t st, int variant = 0);

}  // namespace vfd
