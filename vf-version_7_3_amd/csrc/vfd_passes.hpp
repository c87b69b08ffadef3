// vfd_passes.hpp -- the five HBM passes of one diffusion tick.
//
//   xfwd : pad + R2C along x   (data rows -> half-spectrum rows, natural order)
//   col  : C2C along y or z    (forward, inverse, or forward * K^ * inverse fused)
//   xinv : C2R along x + unpad (+ the constant-pad restore term)
//
// Work buffer ("work") of one field: complex [planes][Py][ld], plane = Py*ld.
// x-rows (z,y) live at work + z*plane + y*ld; the y pass runs over planes z,
// the z pass over ky with a row stride of one plane.  Every pass is in place
// per work-group (a block reads its whole strip before it writes any of it).
//
// Reference semantics: the operator of fftDiffuse3D (src/Diffusion/Diffusion.cu:
// 1622-1934): padDataConstantVF3D (convolutionFFT.cuh:590-638) or a zero pad,
// R2C (Diffusion.cu:1814), modulateAndNormalize (convolutionFFT.cuh:690-716),
// C2R (Diffusion.cu:1842), unpadResult3D (convolutionFFT.cuh:152-171).
#pragma once
#include "vfd_fft.hpp"
#include "vfd_types.h"

namespace vfd {

// ------------------------------------------------------ launch geometry
template <int Q> struct RowCfg {
    static constexpr int E0 = 8;
    static constexpr int TPF = (Q / E0) > 0 ? (Q / E0) : 1;
    static constexpr int G = (256 / TPF) > 0 ? (256 / TPF) : 1;  // rows per block
    static constexpr int NT = TPF * G;
    static constexpr int PITCH = Q + Q / 8 + 1;                  // padded LDS row
};

template <typename T, int Q> struct ColCfg {
    static constexpr int E0 = 8;
    static constexpr int TPF = (Q / E0) > 0 ? (Q / E0) : 1;
    static constexpr int LDS_CAP = 80 * 1024;
    static constexpr int tc_for() {
        int tc = 64;
        while (tc > 1 && (tc * TPF > 1024 || tc * Q * (int)sizeof(cx<T>) > LDS_CAP)) tc /= 2;
        return tc;
    }
    static constexpr int TC = tc_for();  // columns per block
    static constexpr int NT = TC * TPF;
};

// ------------------------------------------------------------ LDS views
template <typename T, int TC> struct ColLds {
    cx<T>* base;  // element n of this thread's column: base[n*TC]
    __device__ __forceinline__ cx<T>& at(int n) { return base[n * TC]; }
};
template <typename T> struct RowLds {
    cx<T>* base;  // row start
    __device__ __forceinline__ cx<T>& at(int n) { return base[n + (n >> 3)]; }
};

template <typename V> __device__ __forceinline__ V tw_load(const V* tw, long k) { return tw[k]; }

// ------------------------------------------------------------- x forward
// rows r = z*Ny + y (z < Nz, y < Ny).  Packed z[n] = u(2n) + i u(2n+1) of the
// padded row (length L = Px/2), length-L FFT, then the real-FFT post-process
// A_k = 1/2 (Z_k + Z*_{L-k}) - i/2 W_Px^k (Z_k - Z*_{L-k}),  k = 0..L.
template <typename T, int L>
__global__ void __launch_bounds__(RowCfg<L>::NT)
xfwd_kernel(const FieldBatch fb, GridDesc g, const cx<T>* __restrict__ tw) {
    using C = RowCfg<L>;
    using Sh = FftShape<T, L, C::TPF>;
    __shared__ cx<T> smem[C::G * C::PITCH];
    const FieldDesc& fd = fb.f[blockIdx.z];
    const int t = threadIdx.x % C::TPF;
    const int gr = threadIdx.x / C::TPF;
    const long row = (long)blockIdx.x * C::G + gr;
    const long nrows = (long)g.Ny * g.Nz;
    const bool live = row < nrows;
    const int z = live ? (int)(row / g.Ny) : 0;
    const int y = live ? (int)(row % g.Ny) : 0;
    const T* __restrict__ src = (const T*)fd.in + row * g.Nx;
    // constant-pad mode: interior 0<z<D-1, 0<y<H-1, 0<x<epi keeps u-b, the rest 0
    const bool cvf = g.pad_mode == PAD_CONSTANT_VF;
    const bool row_in = live && (!cvf || (z > 0 && z < g.Nz - 1 && y > 0 && y < g.Ny - 1));
    const int xhi = cvf ? (fd.epi < g.Nx ? fd.epi : g.Nx) : g.Nx;
    const int xlo = cvf ? 1 : 0;
    const T base = cvf ? (T)fd.base : T(0);

    cx<T> v[Sh::E];
#pragma unroll
    for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
        for (int r = 0; r < Sh::R0; ++r) {
            const int n = Sh::in_elem(t, b, r);
            cx<T> val = {T(0), T(0)};
            if (n >= 0 && row_in) {
                const int x0 = 2 * n, x1 = 2 * n + 1;
                if (x0 >= xlo && x0 < xhi) val.x = src[x0] - base;
                if (x1 >= xlo && x1 < xhi) val.y = src[x1] - base;
            }
            v[b * Sh::R0 + r] = val;
        }
    RowLds<T> lds{smem + gr * C::PITCH};
    fft_regs<T, L, C::TPF, false>(t, v, lds, tw, 2);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
        for (int r = 0; r < Sh::RL; ++r) {
            const int k = Sh::out_elem(t, b, r);
            if (k >= 0) lds.at(k) = v[b * Sh::RL + r];
        }
    __syncthreads();
    if (!live) return;
    cx<T>* __restrict__ dst = (cx<T>*)fd.work + (long)z * g.plane + (long)y * g.ld;
    for (int k = t; k <= L; k += C::TPF) {
        const cx<T> zk = lds.at(k == L ? 0 : k);
        const cx<T> zc = cconj(lds.at(k == 0 ? 0 : L - k));
        const cx<T> e = cscale(cadd(zk, zc), T(0.5));
        const cx<T> d = cscale(csub(zk, zc), T(0.5));
        const cx<T> wd = cmul(d, tw_load(tw, k));     // W_Px^k d
        dst[k] = cx<T>{e.x + wd.y, e.y - wd.x};       // e - i W d
    }
}

// ------------------------------------------------------------- x inverse
// Z_k = (A_k + A*_{L-k}) + i (A_k - A*_{L-k}) W_Px^-k, length-L inverse FFT,
// u(2n) = Re z_n, u(2n+1) = Im z_n (unnormalized C2R of length Px), crop to Nx,
// plus base * K^(0,0,0) in constant-pad mode.
template <typename T, int L>
__global__ void __launch_bounds__(RowCfg<L>::NT)
xinv_kernel(const FieldBatch fb, GridDesc g, const cx<T>* __restrict__ tw) {
    using C = RowCfg<L>;
    using Sh = FftShape<T, L, C::TPF>;
    __shared__ cx<T> smem[C::G * C::PITCH];
    const FieldDesc& fd = fb.f[blockIdx.z];
    const int t = threadIdx.x % C::TPF;
    const int gr = threadIdx.x / C::TPF;
    const long row = (long)blockIdx.x * C::G + gr;
    const long nrows = (long)g.Ny * g.Nz;
    const bool live = row < nrows;
    const int z = live ? (int)(row / g.Ny) : 0;
    const int y = live ? (int)(row % g.Ny) : 0;
    const cx<T>* __restrict__ srow = (const cx<T>*)fd.work + (long)z * g.plane + (long)y * g.ld;

    cx<T> v[Sh::E];
#pragma unroll
    for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
        for (int r = 0; r < Sh::R0; ++r) {
            const int k = Sh::in_elem(t, b, r);
            cx<T> val = {T(0), T(0)};
            if (k >= 0 && live) {
                const cx<T> ak = srow[k];
                const cx<T> ac = cconj(srow[L - k]);
                const cx<T> xe = cadd(ak, ac);
                const cx<T> xo = cmulc(csub(ak, ac), tw_load(tw, k));  // (A - A*) W^-k
                val = cx<T>{xe.x - xo.y, xe.y + xo.x};                  // Xe + i Xo
            }
            v[b * Sh::R0 + r] = val;
        }
    RowLds<T> lds{smem + gr * C::PITCH};
    fft_regs<T, L, C::TPF, true>(t, v, lds, tw, 2);
    if (!live) return;
    T addc = T(0);
    if (g.pad_mode == PAD_CONSTANT_VF) {
        const T dc = g.kspec_complex ? (T)((const cx<T>*)fd.kspec)[0].x : ((const T*)fd.kspec)[0];
        addc = (T)fd.base * dc;
    }
    T* __restrict__ dst = (T*)fd.out + row * g.Nx;
#pragma unroll
    for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
        for (int r = 0; r < Sh::RL; ++r) {
            const int n = Sh::out_elem(t, b, r);
            if (n >= 0) {
                const cx<T> val = v[b * Sh::RL + r];
                if (2 * n < g.Nx) dst[2 * n] = val.x + addc;
                if (2 * n + 1 < g.Nx) dst[2 * n + 1] = val.y + addc;
            }
        }
}

// ------------------------------------------------------------ column pass
// One strip = TC adjacent columns of one plane; a column is a line along y
// (row stride ld, planes z) or along z (row stride plane, planes ky).
// Length P = Q * RS, RS in {1, 2}.  With RS = 2 the zero-padded transform is
// split into its even and odd outputs X[2m+s] = FFT_Q(x[n] W_P^{sn})[m]
// (requires nin, nout <= Q), so the LDS strip holds Q rows instead of P.
//   COL_FWD  : nin rows in, P rows out (natural frequency order)
//   COL_INV  : P rows in, nout rows out
//   COL_FUSED: nin rows in, forward, x K^[k]*scale, inverse, nout rows out
template <typename T, int Q, int MODE, int RS>
__global__ void __launch_bounds__((ColCfg<T, Q>::NT))
col_kernel(const FieldBatch fb, ColDesc cd, const cx<T>* __restrict__ tw) {
    using C = ColCfg<T, Q>;
    using Sh = FftShape<T, Q, C::TPF>;
    __shared__ cx<T> smem[C::TC * Q];
    const FieldDesc& fd = fb.f[blockIdx.z];
    const int c = threadIdx.x % C::TC;
    const int t = threadIdx.x / C::TC;
    const long col = (long)blockIdx.x * C::TC + c;
    const bool live = col < cd.ncols;
    cx<T>* __restrict__ base = (cx<T>*)fd.work + (long)blockIdx.y * cd.pstride + col;
    ColLds<T, C::TC> lds{smem + c};
    constexpr int NI = Sh::NB0 * Sh::R0;
    constexpr int NO = Sh::NBL * Sh::RL;

    cx<T> x[MODE == COL_INV ? 1 : NI];   // saved inputs (forward modes)
    cx<T> acc[MODE == COL_FWD ? 1 : NO]; // accumulated outputs (inverse modes)
    cx<T> v[Sh::E];

    if constexpr (MODE != COL_INV) {
#pragma unroll
        for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
            for (int r = 0; r < Sh::R0; ++r) {
                const int n = Sh::in_elem(t, b, r);
                cx<T> val = {T(0), T(0)};
                if (n >= 0 && n < cd.nin && live) val = base[(long)n * cd.rstride];
                x[b * Sh::R0 + r] = val;
            }
    }
    if constexpr (MODE != COL_FWD) {
#pragma unroll
        for (int i = 0; i < NO; ++i) acc[i] = cx<T>{T(0), T(0)};
    }

#pragma unroll
    for (int s = 0; s < RS; ++s) {
        if constexpr (MODE != COL_INV) {
            // forward sub-transform s on the saved inputs
#pragma unroll
            for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
                for (int r = 0; r < Sh::R0; ++r) {
                    const int n = Sh::in_elem(t, b, r);
                    cx<T> val = x[b * Sh::R0 + r];
                    if (s && n >= 0) val = cmul(val, tw_load(tw, n));  // W_P^n
                    v[b * Sh::R0 + r] = val;
                }
            fft_regs<T, Q, C::TPF, false>(t, v, lds, tw, RS);
        }
        if constexpr (MODE == COL_FWD) {
#pragma unroll
            for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
                for (int r = 0; r < Sh::RL; ++r) {
                    const int m = Sh::out_elem(t, b, r);
                    if (m >= 0 && live) base[(long)(m * RS + s) * cd.rstride] = v[b * Sh::RL + r];
                }
        } else {
            if constexpr (MODE == COL_FUSED) {
                // spectral multiply by K^ (frequency k = m*RS + s along this axis)
                const long kbase = (long)blockIdx.y * cd.ks_plane + col;
#pragma unroll
                for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
                    for (int r = 0; r < Sh::RL; ++r) {
                        const int m = Sh::out_elem(t, b, r);
                        if (m >= 0) {
                            const long ki = kbase + (long)(m * RS + s) * cd.ks_row;
                            cx<T> val = v[b * Sh::RL + r];
                            if (live) {
                                if (cd.kspec_complex) {
                                    val = cmul(val, ((const cx<T>*)fd.kspec)[ki]);
                                } else {
                                    val = cscale(val, ((const T*)fd.kspec)[ki]);
                                }
                            }
                            v[b * Sh::RL + r] = cscale(val, (T)cd.scale);
                        }
                    }
                // re-distribute: last-stage ownership -> first-stage ownership
                __syncthreads();
#pragma unroll
                for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
                    for (int r = 0; r < Sh::RL; ++r) {
                        const int m = Sh::out_elem(t, b, r);
                        if (m >= 0) lds.at(m) = v[b * Sh::RL + r];
                    }
                __syncthreads();
#pragma unroll
                for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
                    for (int r = 0; r < Sh::R0; ++r) {
                        const int m = Sh::in_elem(t, b, r);
                        if (m >= 0) v[b * Sh::R0 + r] = lds.at(m);
                    }
            } else {
                // inverse-only: load sub-sequence X[m*RS + s]
#pragma unroll
                for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
                    for (int r = 0; r < Sh::R0; ++r) {
                        const int m = Sh::in_elem(t, b, r);
                        cx<T> val = {T(0), T(0)};
                        if (m >= 0 && live) val = base[(long)(m * RS + s) * cd.rstride];
                        v[b * Sh::R0 + r] = val;
                    }
            }
            fft_regs<T, Q, C::TPF, true>(t, v, lds, tw, RS);
#pragma unroll
            for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
                for (int r = 0; r < Sh::RL; ++r) {
                    const int n = Sh::out_elem(t, b, r);
                    if (n >= 0) {
                        cx<T> val = v[b * Sh::RL + r];
                        if (s) val = cmulc(val, tw_load(tw, n));  // W_P^-n
                        acc[b * Sh::RL + r] = cadd(acc[b * Sh::RL + r], val);
                    }
                }
        }
    }
    if constexpr (MODE != COL_FWD) {
        if (!live) return;
#pragma unroll
        for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
            for (int r = 0; r < Sh::RL; ++r) {
                const int n = Sh::out_elem(t, b, r);
                if (n >= 0 && n < cd.nout) base[(long)n * cd.rstride] = acc[b * Sh::RL + r];
            }
    }
}

}  // namespace vfd
