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
#include <type_traits>

#include "vfd_fft.hpp"
#include "vfd_types.h"

namespace vfd {

// ------------------------------------------------------ launch geometry
// Schedule variants (picked per pass by the autotuner):
//   V = 0: radix <= 16, 16 elements per thread, blocks <= 512 threads
//   V = 1: radix <= 8,   8 elements per thread, blocks <= 512 threads
//   V = 2: radix <= 8,   8 elements per thread, blocks <= 256 threads
//   V = 3: radix <= 16, 16 elements per thread, blocks <= 256 threads
constexpr int kNumVariants = 4;
constexpr int var_maxr(int v) { return (v == 1 || v == 2) ? 8 : 16; }
constexpr int var_ntmax(int v) { return v >= 2 ? 256 : 512; }

template <typename T, int Q, int V = 0> struct RowCfg {
    static constexpr int MAXR = var_maxr(V);
    static constexpr int TPF = tpf_for(Q, MAXR);
    static constexpr int NTT = var_ntmax(V) / 2;                    // target threads
    static constexpr int G = (NTT / TPF) > 0 ? (NTT / TPF) : 1;     // rows per group
    static constexpr int NT = TPF * G;
    static constexpr int PITCH = Q + Q / 32 + 1;                 // padded LDS row (RowLds)
    // each row's transform and LDS row belong to one wavefront: its stages
    // synchronize that wave alone (stage_sync)
    static constexpr bool WS = wave_local<TPF>();
    // x inverse: the real-FFT mirror pairs (A_k, A_{L-k}) are exchanged across
    // lanes (ds_bpermute) instead of loaded twice: FP32, wave-local rows, every
    // first-stage slot busy (MirrorSlots).  (The x forward pass's post-process
    // exchange done the same way held all 16 outputs across it: 162 -> 170
    // VGPRs, 3 -> 2 waves per SIMD, 1.33 -> 1.37 ms; it keeps the LDS exchange.)
    using Sh_ = FftShape<T, Q, TPF, false, MAXR>;
    static constexpr bool MIRROR = sizeof(T) == 4 && WS && (Q / Sh_::R0) % TPF == 0;
    // twiddle table W_Px (2Q entries) staged in LDS when it fits beside the rows
    static constexpr bool TW_LDS = (2 * Q + G * PITCH) * (int)sizeof(cx<T>) <= 64 * 1024;
    static constexpr int TW_N = TW_LDS ? 2 * Q : 1;
};

template <typename T, int Q, int V = 0, int RS = 2> struct ColCfg {
    static constexpr int MAXR = var_maxr(V);
    static constexpr int TPF = tpf_for(Q, MAXR);
    // (a 144 KB cap for P >= 1024, 8-column tiles of 1536- and 1152-long
    // columns at one work-group per CU, was slower than 4-column tiles at two:
    // 0.59 - 0.66 ms against 0.38 - 0.40 ms per production y pass)
    static constexpr int LDS_CAP = 72 * 1024;
    // columns per block: the variant's thread cap, but never narrower than one
    // 128-byte line of a row (narrower tiles fetch whole lines and use part)
    static constexpr int TC_LINE = 128 / (int)sizeof(cx<T>);
    static constexpr bool lds_fits(int tc) { return (tc * Q + Q * RS) * (int)sizeof(cx<T>) <= LDS_CAP; }
    static constexpr int tc_for() {
        int tc = 64;
        while (tc > 1 && (tc * TPF > var_ntmax(V) || !lds_fits(tc))) tc /= 2;
        while (tc < TC_LINE && lds_fits(2 * tc) && 2 * tc * TPF <= 1024) tc *= 2;
        return tc;
    }
    static constexpr int TC = tc_for();  // columns per block
    static constexpr int NT = TC * TPF;
    // tiles sharing one 128-byte row segment (narrow tiles of long columns)
    static constexpr int SH = TC < TC_LINE ? TC_LINE / TC : 1;
    // twiddle table W_P (P = Q * RS entries) staged in LDS
    static constexpr int TW_N = Q * RS;
};

// ------------------------------------------------------------ LDS views
template <typename T, int TC> struct ColLds {
    cx<T>* base;  // element n of this thread's column: base[n*TC]
    __device__ __forceinline__ cx<T>& at(int n) { return base[n * TC]; }
};
template <typename T> struct RowLds {
    cx<T>* base;  // row start
    // one pad element per 32: the Stockham stage stores (stride NS*R) and
    // loads (stride Q/R) of the row transforms then spread over the banks
    __device__ __forceinline__ cx<T>& at(int n) { return base[n + (n >> 5)]; }
};

// ------------------------------------------------------------- x forward
// rows r = z*Ny + y (z < Nz, y < Ny).  Packed z[n] = u(2n) + i u(2n+1) of the
// padded row (length L = Px/2), length-L FFT, then the real-FFT post-process
// A_k = 1/2 (Z_k + Z*_{L-k}) - i/2 W_Px^k (Z_k - Z*_{L-k}),  k = 0..L.
// Persistent: each work-group loops over groups of G rows of every field.
template <typename T, int L, int V>
__global__ void __launch_bounds__((RowCfg<T, L, V>::NT))
xfwd_kernel(const FieldBatch fb, GridDesc g, const cx<T>* __restrict__ twg, int nfields) {
    using C = RowCfg<T, L, V>;
    using Sh = FftShape<T, L, C::TPF, false, C::MAXR>;
    using B = BufOps<T>;
    using O = typename B::O;
    __shared__ cx<T> smem[C::G * C::PITCH];
    __shared__ cx<T> twl[C::TW_N];
    const cx<T>* tw = twg;
    if constexpr (C::TW_LDS) {
        stage_table(twl, twg, 2 * L);
        __syncthreads();
        tw = twl;
    }
    const int t = threadIdx.x % C::TPF;
    const int gr = threadIdx.x / C::TPF;
    const long nrows = (long)g.Ny * g.Nz;
    const long ngroups = (nrows + C::G - 1) / C::G;
    RowLds<T> lds{smem + gr * C::PITCH};
    const bool cvf = g.pad_mode == PAD_CONSTANT_VF;
    constexpr unsigned RB = sizeof(T), CB = sizeof(cx<T>);
    for (long item = blockIdx.x; item < ngroups * nfields; item += gridDim.x) {
        const int f = (int)(item / ngroups);
        const long row = (item % ngroups) * C::G + gr;
        const FieldDesc& fd = fb.f[f];
        const auto rin = B::make(fd.in, g.in_bytes);
        const auto rwork = B::make(fd.work, g.work_bytes);
        const bool live = row < nrows;
        const int z = live ? (int)(row / g.Ny) : 0;
        const int y = live ? (int)(row % g.Ny) : 0;
        const O rowoff = (O)(row * g.Nx) * RB;
        // constant-pad mode: interior 0<z<D-1, 0<y<H-1, 0<x<epi keeps u-b, the rest 0
        const bool row_in = live && (!cvf || (z > 0 && z < g.Nz - 1 && y > 0 && y < g.Ny - 1));
        const int xhi = cvf ? (fd.epi < g.Nx ? fd.epi : g.Nx) : g.Nx;
        const int xlo = cvf ? 1 : 0;
        const T base = cvf ? (T)fd.base : T(0);
        const bool vec = !cvf && (g.Nx % 2 == 0);

        cx<T> v[Sh::E];
#pragma unroll
        for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
            for (int r = 0; r < Sh::R0; ++r) {
                const int n = Sh::in_elem(t, b, r);
                const int x0 = 2 * n, x1 = 2 * n + 1;
                cx<T> val;
                if (vec) {
                    const bool ok = n >= 0 && row_in && x1 < g.Nx;
                    val = B::template ld_cx<kStream>(rin, ok ? rowoff + (O)x0 * RB : B::OOB);
                } else {
                    const bool ok0 = n >= 0 && row_in && x0 >= xlo && x0 < xhi;
                    const bool ok1 = n >= 0 && row_in && x1 >= xlo && x1 < xhi;
                    const T a0 = B::template ld_r<kStream>(rin, ok0 ? rowoff + (O)x0 * RB : B::OOB);
                    const T a1 = B::template ld_r<kStream>(rin, ok1 ? rowoff + (O)x1 * RB : B::OOB);
                    val = cx<T>{ok0 ? a0 - base : T(0), ok1 ? a1 - base : T(0)};
                }
                v[b * Sh::R0 + r] = val;
            }
        fft_regs<T, L, C::TPF, false, false, C::MAXR, C::WS>(t, v, lds, tw, 2);
        const O dst = (O)((long)z * g.plane + (long)y * g.ld) * CB;
        auto post = [&](int k, cx<T> zk, cx<T> zc) {
            const cx<T> e = cscale(cadd(zk, zc), T(0.5));
            const cx<T> d = cscale(csub(zk, zc), T(0.5));
            const cx<T> wd = cmul(d, tw[k]);  // W_Px^k d
            return cx<T>{e.x + wd.y, e.y - wd.x};  // e - i W d
        };
        stage_sync<C::WS>();
#pragma unroll
        for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
            for (int r = 0; r < Sh::RL; ++r) {
                const int k = Sh::out_elem(t, b, r);
                if (k >= 0) lds.at(k) = v[b * Sh::RL + r];
            }
        stage_sync<C::WS>();
        for (int k = t; k <= L; k += C::TPF) {
            const cx<T> zk = lds.at(k == L ? 0 : k);
            const cx<T> zc = cconj(lds.at(k == 0 ? 0 : L - k));
            B::template st_cx<kStream>(post(k, zk, zc), rwork, live ? dst + (O)k * CB : B::OOB);
        }
        // the next item's first LDS store happens after a barrier inside fft_regs
    }
}

// d + t * decay with the product rounded to FP32 before the add (no fused
// multiply-add), the level update of WHChemical::update
__device__ __forceinline__ float upd_level(float d, float t, float decay) {
#pragma clang fp contract(off)
    const float scaled = t * decay;
    return d + scaled;
}
__device__ __forceinline__ double upd_level(double d, double t, double decay) { return d + t * decay; }

// ------------------------------------------------------------- x inverse
// Z_k = (A_k + A*_{L-k}) + i (A_k - A*_{L-k}) W_Px^-k, length-L inverse FFT,
// u(2n) = Re z_n, u(2n+1) = Im z_n (unnormalized C2R of length Px), crop to Nx,
// plus base * K^(0,0,0) in constant-pad mode.
template <typename T, int L, int V, bool VIS = false>
__global__ void __launch_bounds__((RowCfg<T, L, V>::NT))
xinv_kernel(const FieldBatch fb, GridDesc g, const cx<T>* __restrict__ twg, int nfields) {
    using C = RowCfg<T, L, V>;
    using Sh = FftShape<T, L, C::TPF, false, C::MAXR>;
    using B = BufOps<T>;
    using O = typename B::O;
    __shared__ cx<T> smem[C::G * C::PITCH];
    __shared__ cx<T> twl[C::TW_N];
    const cx<T>* tw = twg;
    if constexpr (C::TW_LDS) {
        stage_table(twl, twg, 2 * L);
        __syncthreads();
        tw = twl;
    }
    const int t = threadIdx.x % C::TPF;
    const int gr = threadIdx.x / C::TPF;
    const long nrows = (long)g.Ny * g.Nz;
    const long ngroups = (nrows + C::G - 1) / C::G;
    RowLds<T> lds{smem + gr * C::PITCH};
    constexpr unsigned RB = sizeof(T), CB = sizeof(cx<T>);
    for (long item = blockIdx.x; item < ngroups * nfields; item += gridDim.x) {
        const int f = (int)(item / ngroups);
        const long row = (item % ngroups) * C::G + gr;
        const FieldDesc& fd = fb.f[f];
        const auto rwork = B::make(fd.work, g.work_bytes);
        const auto rout = B::make(fd.out, g.out_bytes);
        const bool live = row < nrows;
        const int z = live ? (int)(row / g.Ny) : 0;
        const int y = live ? (int)(row % g.Ny) : 0;
        const O srow = (O)((long)z * g.plane + (long)y * g.ld) * CB;

        cx<T> v[Sh::E];
        auto pre = [&](int k, cx<T> ak, cx<T> ac) {
            const cx<T> xe = cadd(ak, ac);
            const cx<T> xo = cmulc(csub(ak, ac), tw[k]);  // (A - A*) W^-k
            return cx<T>{xe.x - xo.y, xe.y + xo.x};       // Xe + i Xo
        };
        if constexpr (C::MIRROR) {
            // each A_k is loaded once; A_{L-k} comes from the mirror lane's
            // registers, A_L (the mirror of k = 0) by one more load on lane 0
            using M = MirrorSlots<Sh::NB0, Sh::R0>;
            const int maddr = mirror_lane_addr<C::TPF>(t);
            cx<T> a[Sh::NB0 * Sh::R0];
#pragma unroll
            for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
                for (int r = 0; r < Sh::R0; ++r)
                    a[b * Sh::R0 + r] = B::ld_cx(rwork, live ? srow + (O)Sh::in_elem(t, b, r) * CB : B::OOB);
            const cx<T> aL = B::ld_cx(rwork, live && t == 0 ? srow + (O)L * CB : B::OOB);
#pragma unroll
            for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
                for (int r = 0; r < Sh::R0; ++r) {
                    cx<T> m = lane_read(a[M::other(b, r)], maddr);
                    if (t == 0) m = (b == 0 && r == 0) ? aL : a[M::lane0(b, r)];
                    v[b * Sh::R0 + r] = pre(Sh::in_elem(t, b, r), a[b * Sh::R0 + r], cconj(m));
                }
        } else {
#pragma unroll
            for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
                for (int r = 0; r < Sh::R0; ++r) {
                    const int k = Sh::in_elem(t, b, r);
                    const bool ok = k >= 0 && live;
                    const cx<T> ak = B::ld_cx(rwork, ok ? srow + (O)k * CB : B::OOB);
                    const cx<T> ac = cconj(B::ld_cx(rwork, ok ? srow + (O)(L - k) * CB : B::OOB));
                    v[b * Sh::R0 + r] = pre(ok ? k : 0, ak, ac);
                }
        }
        fft_regs<T, L, C::TPF, true, false, C::MAXR, C::WS>(t, v, lds, tw, 2);
        T addc = T(0);
        if (g.pad_mode == PAD_CONSTANT_VF) {
            const auto rk = B::make(fd.kspec, g.kspec_bytes);
            const T dc = B::ld_r(rk, 0);  // Re K^(0,0,0) (real part first in both layouts)
            // base * sum(k): the complex spectrum holds sum(k) itself; the real
            // one sum(k) / (Px Py Pz), which kdc_scale undoes
            addc = g.kspec_complex ? (T)fd.base * dc : (T)((double)fd.base * ((double)dc * g.kdc_scale));
        }
        const O orow = (O)(row * g.Nx) * RB;
        const bool vec = (g.Nx % 2 == 0);
        // fused chem update (FP32): p = d + t * decay with the product rounded
        // before the add (WHChemical.cpp:425-428), and d cleared
        [[maybe_unused]] const bool upd = sizeof(T) == 4 && fd.delta != nullptr;
        [[maybe_unused]] const auto rdel = B::make(fd.delta, g.out_bytes);
        [[maybe_unused]] const T dec = (T)fd.decay;
        // fused visualisation downsample of t (FP32): sampleChem's every
        // vs-th voxel, min(t * scale, 1), into the dense [Nz/vs][Ny/vs][Nx/vs] array
        const bool vis = VIS && fd.vis != nullptr;
        const int vs = vis ? fd.vis_stride : 1;
        const int vD = g.Nz / vs, vH = g.Ny / vs, vW = g.Nx / vs;
        const bool vrow = vis && live && z % vs == 0 && y % vs == 0 && z / vs < vD && y / vs < vH;
        const auto rvis = B::make(fd.vis, vis ? (unsigned long)vD * vH * vW * RB : 0ul);
        const O vbase = vrow ? (O)(((long)(z / vs) * vH + y / vs) * vW) * RB : (O)0;
        const float vscale = fd.vis_scale;
        auto vis_put = [&](int x, T tv) {
            if (x % vs == 0 && x / vs < vW) {
                float s = (float)tv * vscale;
                if (s > 1.0f) s = 1.0f;
                B::st_r((T)s, rvis, vbase + (O)(x / vs) * RB);
            }
        };
#pragma unroll
        for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
            for (int r = 0; r < Sh::RL; ++r) {
                const int n = Sh::out_elem(t, b, r);
                const cx<T> val = v[b * Sh::RL + r];
                if (vec) {
                    const bool ok = n >= 0 && live && 2 * n + 1 < g.Nx;
                    const O off = ok ? orow + (O)(2 * n) * RB : B::OOB;
                    cx<T> o{val.x + addc, val.y + addc};
                    if constexpr (VIS) {
                        if (vrow && ok) {
                            vis_put(2 * n, o.x);
                            vis_put(2 * n + 1, o.y);
                        }
                    }
                    if constexpr (sizeof(T) == 4) {
                        if (upd) {
                            const cx<T> d = B::ld_cx(rdel, off);
                            o = cx<T>{upd_level(d.x, o.x, dec), upd_level(d.y, o.y, dec)};
                            B::template st_cx<kStream>(cx<T>{T(0), T(0)}, rdel, off);
                        }
                    }
                    B::template st_cx<kStream>(o, rout, off);
                } else {
                    const bool ok0 = n >= 0 && live && 2 * n < g.Nx;
                    const bool ok1 = n >= 0 && live && 2 * n + 1 < g.Nx;
                    const O off0 = ok0 ? orow + (O)(2 * n) * RB : B::OOB;
                    const O off1 = ok1 ? orow + (O)(2 * n + 1) * RB : B::OOB;
                    T o0 = val.x + addc, o1 = val.y + addc;
                    if constexpr (VIS) {
                        if (vrow && ok0) vis_put(2 * n, o0);
                        if (vrow && ok1) vis_put(2 * n + 1, o1);
                    }
                    if constexpr (sizeof(T) == 4) {
                        if (upd) {
                            o0 = upd_level(B::ld_r(rdel, off0), o0, dec);
                            o1 = upd_level(B::ld_r(rdel, off1), o1, dec);
                            B::template st_r<kStream>(T(0), rdel, off0);
                            B::template st_r<kStream>(T(0), rdel, off1);
                        }
                    }
                    B::template st_r<kStream>(o0, rout, off0);
                    B::template st_r<kStream>(o1, rout, off1);
                }
            }
        // the next item's first LDS store happens after a barrier inside fft_regs
    }
}

// v * k for a spectrum factor formed in FP64.  FP32 carries k as a pair of
// floats, k ~ kh + kl, so that the factor's own rounding (up to 6e-8 of it,
// the same on every tick) does not become a per-tick gain: the DC factor is
// the decay of the mass, and a gain of 6e-8 a tick is 6e-6 after 100 ticks.
__device__ __forceinline__ float2 kscale(float2 v, double k) {
    const float kh = (float)k, kl = (float)(k - (double)kh);
    return float2{fmaf(kh, v.x, kl * v.x), fmaf(kh, v.y, kl * v.y)};
}
__device__ __forceinline__ double2 kscale(double2 v, double k) { return double2{v.x * k, v.y * k}; }

// ------------------------------------------------------------ column pass
// One strip = TC adjacent columns of one plane; a column is a line along y
// (row stride ld, planes z) or along z (row stride plane, planes ky).
// Length P = Q * RS, RS in {1, 2}.  With RS = 2 the zero-padded transform is
// split into its even and odd outputs X[2m+s] = FFT_Q(x[n] W_P^{sn})[m]
// (requires nin, nout <= Q), so the LDS strip holds Q rows instead of P.
//   COL_FWD  : nin rows in, P rows out (natural frequency order)
//   COL_INV  : P rows in, nout rows out
//   COL_FUSED: nin rows in, forward, x K^[k]*scale, inverse, nout rows out.
//              The inverse runs the reversed radix plan, so it starts on the
//              elements the forward left in registers (no LDS redistribution).
// KS (fused mode): the kernel-spectrum format -- KS_REAL, KS_COMPLEX (R2C
// layout, scaled by cd.scale), or KS_SEP (FP64 tables, K^ formed here).
enum { KS_REAL = 0, KS_COMPLEX = 1, KS_SEP = 2 };

// Which z-pass form an FP32 field with a separable table takes, from the
// window radius R the table itself stores (tab[cd.sep_g]): 8, 12 or 16 -- the
// direct form compiled for radius bound 8 (R in [0, 8]), 12 (R in (8, 12]) or
// 16 (R in (12, 16]) -- or 0, the FFT form (R > 16, or no valid radius).
// Every z-pass launch of a separable batch covers all its fields and each
// work-group keeps only the fields of its own form, so the choice follows the
// table's contents wherever the table came from.
__device__ __forceinline__ int zform_of(double rd) {
    if (rd >= 0.0 && rd <= 8.0) return 8;
    if (rd > 8.0 && rd <= 12.0) return 12;
    if (rd > 12.0 && rd <= (double)kZDirectMaxR) return 16;
    return 0;  // also NaN
}
__device__ __forceinline__ int zform_of_table(const void* kspec, unsigned long kspec_bytes, int sep_g) {
    using BD = BufOps<double>;
    const auto tk = BD::make(kspec, kspec_bytes);
    return zform_of(BD::ld_r(tk, (typename BD::O)sep_g * 8u));
}

// Waves per SIMD the compiler must leave room for: the FP32 inverse pass of
// 512-thread work-groups (both sub-sequences' inputs loaded up front) needs 131
// VGPRs unconstrained, one work-group per CU; at 128 two fit
template <typename T, int MODE, int NT> constexpr int col_waves_per_eu() {
    return sizeof(T) == 4 && MODE == COL_INV && NT == 512 ? 4 : 1;
}

// FOLD (RS = 2, inverse-only mode): the output rows reach past Q = P/2 (a box
// less than twice the data, as the ABM's 1536 rows for 1390 cells), so the
// even/odd split folds them: with Y_s the inverse transform of the sub-sequence
// X[2k + s], row n takes Y_0[n] + W_P^-n Y_1[n] and row n + Q takes
// Y_0[n] - W_P^-n Y_1[n].  Half the LDS per column of the unsplit transform:
// twice the columns per tile.
template <typename T, int Q, int MODE, int RS, int V, int KS = KS_REAL, bool FOLD = false>
__global__ void __launch_bounds__((ColCfg<T, Q, V, RS>::NT))
__attribute__((amdgpu_waves_per_eu((col_waves_per_eu<T, MODE, ColCfg<T, Q, V, RS>::NT>()))))
col_kernel(const FieldBatch fb, ColDesc cd, const cx<T>* __restrict__ twg) {
    static_assert(!FOLD || (RS == 2 && MODE == COL_INV), "folded split: RS = 2, inverse-only mode");
    using C = ColCfg<T, Q, V, RS>;
    using ShF = FftShape<T, Q, C::TPF, false, C::MAXR>;                // forward / inverse-only plan
    using ShI = FftShape<T, Q, C::TPF, MODE == COL_FUSED, C::MAXR>;    // inverse plan
    using B = BufOps<T>;
    using O = typename B::O;
    constexpr int P = Q * RS;
    constexpr unsigned CB = sizeof(cx<T>), RB = sizeof(T);
    __shared__ cx<T> smem[C::TC * Q];
    __shared__ cx<T> tw[C::TW_N];
    constexpr bool SEP = MODE == COL_FUSED && KS == KS_SEP;
    __shared__ double gzs[SEP ? P : 1];  // separable: the gz table of the z frequencies
    // Block coordinates.  Narrow tiles (SH of them share each 128-byte row
    // segment) launch as a 1-D grid in which the SH tiles of one segment get
    // work-group ids 8 apart.  Ids are dealt round-robin to the 8 XCDs, so the
    // tiles sharing a line run side by side on one XCD and its L2 fetches the
    // line once.
    unsigned bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if constexpr (C::SH > 1) {
        const unsigned lin = blockIdx.x;
        const unsigned grp = (lin / (8u * C::SH)) * 8u + lin % 8u;
        const unsigned seg = grp % cd.grid_segs, rest = grp / cd.grid_segs;
        bx = seg * C::SH + (lin / 8u) % C::SH;
        by = rest % cd.grid_planes;
        bz = rest / cd.grid_planes;
        // whole-block exits, before any barrier: padding ids and tiles past the row end
        if (bz >= cd.grid_fields || (long)bx * C::TC >= cd.ncols) return;
    }
    const FieldDesc& fd = fb.f[bz];
    if constexpr (SEP && sizeof(T) == 4) {
        // FP32 separable tables of radius <= kZDirectMaxR take the direct z pass
        // (zdirect_kernel, launched over the same fields); whole-block exit
        if (zform_of_table(fd.kspec, cd.kspec_bytes, cd.sep_g) != 0) return;
    }
    const auto rwork = B::make(fd.work, cd.work_bytes);
    const int c = threadIdx.x % C::TC;
    const int t = threadIdx.x / C::TC;
    const long col = (long)bx * C::TC + c;
    const bool live = col < cd.ncols;
    const O base = (O)((long)by * cd.pstride + col) * CB;
    const O rstep = (O)cd.rstride * CB;
    ColLds<T, C::TC> lds{smem + c};
    // non-temporal loads only when a tile row spans whole 128-byte lines: a
    // narrower tile (long columns) shares each line with its neighbours, which
    // must still find it in the cache
    constexpr int NTP = C::TC * (int)CB >= 128 ? kStream : kCacheDefault;
    // stores: non-temporal for whole 128-byte row segments; a narrower segment
    // is written back through the L2, where the tiles sharing its line merge.
    // (Choosing the L2 path at run time for the last, mostly empty tile of a
    // 1025-column row measured 10-30 % slower on the y forward and z FFT
    // passes; the direct z pass, one column per lane, gains from it.)
    constexpr int NTS = C::TC * (int)CB >= 128 ? kStream : kCacheDefault;
    constexpr int NI = ShF::NB0 * ShF::R0;
    constexpr int NO = ShI::NBL * ShI::RL;
    constexpr int NK = ShF::NBL * ShF::RL;

    constexpr int NII = ShI::NB0 * ShI::R0;
    cx<T> x[MODE == COL_INV ? NII * RS : NI];   // saved inputs (all sub-sequences in INV mode)
    cx<T> acc[MODE == COL_FWD ? 1 : NO]; // accumulated outputs (inverse modes)
    [[maybe_unused]] cx<T> acq[FOLD && MODE == COL_INV ? NO : 1];  // folded: output rows n + Q
    cx<T> v[ShF::E > ShI::E ? ShF::E : ShI::E];

    if constexpr (MODE != COL_INV) {
#pragma unroll
        for (int b = 0; b < ShF::NB0; ++b)
#pragma unroll
            for (int r = 0; r < ShF::R0; ++r) {
                const int n = ShF::in_elem(t, b, r);
                const bool ok = n >= 0 && n < cd.nin && live;
                x[b * ShF::R0 + r] = B::template ld_cx<NTP>(rwork, ok ? base + (O)n * rstep : B::OOB);
            }
    }
    if constexpr (MODE == COL_INV) {
        // inverse-only: every sub-sequence X[m*RS + s] is loaded up front
#pragma unroll
        for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int b = 0; b < ShI::NB0; ++b)
#pragma unroll
                for (int r = 0; r < ShI::R0; ++r) {
                    const int m = ShI::in_elem(t, b, r);
                    const bool ok = m >= 0 && live;
                    x[s * NII + b * ShI::R0 + r] = B::template ld_cx<NTP>(rwork, ok ? base + (O)(m * RS + s) * rstep : B::OOB);
                }
    }
    if constexpr (MODE != COL_FWD) {
#pragma unroll
        for (int i = 0; i < NO; ++i) acc[i] = cx<T>{T(0), T(0)};
        if constexpr (FOLD) {
#pragma unroll
            for (int i = 0; i < NO; ++i) acq[i] = cx<T>{T(0), T(0)};
        }
    }

    // kernel spectrum for the fused mode, prefetched before any transform work:
    // real K^ stores the RS frequencies m*RS + s of one element side by side
    // (one load per element), complex K^ is the R2C layout (one load per s)
    constexpr bool STREAM_K = MODE == COL_FUSED && KS != KS_SEP;
    [[maybe_unused]] cx<T> kv[STREAM_K ? NK : 1];
    [[maybe_unused]] cx<T> kvc[STREAM_K && KS == KS_COMPLEX && RS > 1 ? NK : 1];
    [[maybe_unused]] double ka = 0.0;  // separable: (gx*gy)*scale of this column
    if constexpr (SEP) {
        // separable K^(kz) = (((gx*gy)*scale)*gz[kz]) rounded to T: the numbers
        // the 3D real layout stores (gauss_product_kernel); gz staged in LDS
        using BD = BufOps<double>;
        using OD = typename BD::O;
        const auto tk = BD::make(fd.kspec, cd.kspec_bytes);
        const double sc = BD::ld_r(tk, OD(0));
        const double gyv = BD::ld_r(tk, (OD)(cd.sep_gy + by) * 8u);
        const double gxv = BD::ld_r(tk, live ? (OD)(1 + col) * 8u : BD::OOB);
        ka = (gxv * gyv) * sc;
        for (int i = threadIdx.x; i < P; i += C::NT) gzs[i] = BD::ld_r(tk, (OD)(cd.sep_gz + i) * 8u);
    } else if constexpr (MODE == COL_FUSED) {
        const auto rk = B::make(fd.kspec, cd.kspec_bytes);
        if constexpr (KS == KS_COMPLEX) {
            const O kbase = (O)((long)by * cd.ks_plane + col) * CB;
            const O kstep = (O)cd.ks_row * CB;
#pragma unroll
            for (int b = 0; b < ShF::NBL; ++b)
#pragma unroll
                for (int r = 0; r < ShF::RL; ++r) {
                    const int m = ShF::out_elem(t, b, r);
                    const bool ok = m >= 0 && live;
                    kv[b * ShF::RL + r] = B::ld_cx(rk, ok ? kbase + (O)(m * RS) * kstep : B::OOB);
                    if constexpr (RS > 1)
                        kvc[b * ShF::RL + r] = B::ld_cx(rk, ok ? kbase + (O)(m * RS + 1) * kstep : B::OOB);
                }
        } else {
            static_assert(KS == KS_REAL, "streamed spectrum format");
            const O kbase = (O)((long)by * cd.ks_plane + col * RS) * RB;
            const O kstep = (O)cd.ks_row * RB;
#pragma unroll
            for (int b = 0; b < ShF::NBL; ++b)
#pragma unroll
                for (int r = 0; r < ShF::RL; ++r) {
                    const int m = ShF::out_elem(t, b, r);
                    const bool ok = m >= 0 && live;
                    const O off = ok ? kbase + (O)m * kstep : B::OOB;
                    if constexpr (RS > 1) kv[b * ShF::RL + r] = B::ld_cx(rk, off);  // (K^[2m], K^[2m+1])
                    else kv[b * ShF::RL + r] = cx<T>{B::ld_r(rk, off), T(0)};
                }
        }
    }

    // twiddle table: loads issued after the data loads, one latency for both
    stage_table_unrolled<C::TW_N, C::NT>(tw, twg, P);
    __syncthreads();

#pragma unroll
    for (int s = 0; s < RS; ++s) {
        if constexpr (MODE == COL_FUSED) {
#pragma unroll
            for (int b = 0; b < ShF::NB0; ++b)
#pragma unroll
                for (int r = 0; r < ShF::R0; ++r) {
                    const int n = ShF::in_elem(t, b, r);
                    cx<T> val = x[b * ShF::R0 + r];
                    if (s) val = cmul(val, tw[n >= 0 ? n : 0]);  // W_P^n
                    v[b * ShF::R0 + r] = val;
                }
            fft_regs<T, Q, C::TPF, false, false, C::MAXR>(t, v, lds, tw, RS);
            if constexpr (KS == KS_COMPLEX) {
                const T sc = (T)cd.scale;
#pragma unroll
                for (int i = 0; i < NK; ++i) {
                    cx<T> k = kv[i];
                    if constexpr (RS > 1) k = s ? kvc[i] : kv[i];
                    v[i] = cscale(cmul(v[i], k), sc);
                }
            } else if constexpr (KS == KS_SEP) {
#pragma unroll
                for (int b = 0; b < ShF::NBL; ++b)
#pragma unroll
                    for (int r = 0; r < ShF::RL; ++r) {
                        const int m = ShF::out_elem(t, b, r);
                        const double kd = ka * gzs[m >= 0 ? m * RS + s : 0];
                        v[b * ShF::RL + r] = kscale(v[b * ShF::RL + r], kd);
                    }
            } else {
#pragma unroll
                for (int i = 0; i < NK; ++i) v[i] = cscale(v[i], s ? kv[i].y : kv[i].x);
            }
            fft_regs<T, Q, C::TPF, true, true, C::MAXR>(t, v, lds, tw, RS);
        } else if constexpr (MODE == COL_FWD) {
#pragma unroll
            for (int b = 0; b < ShF::NB0; ++b)
#pragma unroll
                for (int r = 0; r < ShF::R0; ++r) {
                    const int n = ShF::in_elem(t, b, r);
                    cx<T> val = x[b * ShF::R0 + r];
                    if (s) val = cmul(val, tw[n >= 0 ? n : 0]);  // W_P^n
                    v[b * ShF::R0 + r] = val;
                }
            fft_regs<T, Q, C::TPF, false, false, C::MAXR>(t, v, lds, tw, RS);
#pragma unroll
            for (int b = 0; b < ShF::NBL; ++b)
#pragma unroll
                for (int r = 0; r < ShF::RL; ++r) {
                    const int m = ShF::out_elem(t, b, r);
                    const bool ok = m >= 0 && live;
                    B::template st_cx<NTS>(v[b * ShF::RL + r], rwork, ok ? base + (O)(m * RS + s) * rstep : B::OOB);
                }
            continue;
        } else {
#pragma unroll
            for (int i = 0; i < NII; ++i) v[i] = x[s * NII + i];
            fft_regs<T, Q, C::TPF, true, false, C::MAXR>(t, v, lds, tw, RS);
        }
        if constexpr (MODE != COL_FWD) {
#pragma unroll
            for (int b = 0; b < ShI::NBL; ++b)
#pragma unroll
                for (int r = 0; r < ShI::RL; ++r) {
                    const int n = ShI::out_elem(t, b, r);
                    cx<T> val = v[b * ShI::RL + r];
                    if (s) val = cmulc(val, tw[n >= 0 ? n : 0]);  // W_P^-n
                    acc[b * ShI::RL + r] = cadd(acc[b * ShI::RL + r], val);
                    // row n + Q: W_P^-(n+Q)s = (-1)^s W_P^-ns
                    if constexpr (FOLD && MODE == COL_INV)
                        acq[b * ShI::RL + r] = s ? csub(acq[b * ShI::RL + r], val) : cadd(acq[b * ShI::RL + r], val);
                }
        }
    }
    if constexpr (MODE != COL_FWD) {
#pragma unroll
        for (int b = 0; b < ShI::NBL; ++b)
#pragma unroll
            for (int r = 0; r < ShI::RL; ++r) {
                const int n = ShI::out_elem(t, b, r);
                const bool ok = n >= 0 && n < cd.nout && live;
                B::template st_cx<NTS>(acc[b * ShI::RL + r], rwork, ok ? base + (O)n * rstep : B::OOB);
                if constexpr (FOLD && MODE == COL_INV) {
                    const bool okq = n >= 0 && n + Q < cd.nout && live;
                    B::template st_cx<NTS>(acq[b * ShI::RL + r], rwork, okq ? base + (O)(n + Q) * rstep : B::OOB);
                }
            }
    }
}

// ------------------------------------------------ z pass, direct form (FP32)
// With the separable spectrum the z factor gz is the DFT, on the Pz box, of
// the normalized window g_j (|j| <= R).  Multiplying by gz between the forward
// and the inverse z transforms is therefore the linear convolution of each
// column with that window (the box is at least nz + R long, nothing wraps),
// and this pass applies it directly: per column (kx, ky)
//   y[n] = ((gx * gy) * scale * Pz) * sum_{|j| <= R} g_|j| x[n - j],  n < nz,
// with x zero outside [0, nz).  Same operator as the FFT z pass, with
// 4 RM + 2 flops per output instead of four length-Pz/2 transforms.
//
// One lane per column, 64 consecutive columns per wave (512-byte rows).  A
// lane streams its column once: the window of 2 RM + 1 values lives in
// registers, rotated at compile time inside blocks of U = 2 RM + 1 outputs,
// and row n + RM is loaded before row n is written, so the pass is in place.
// RM is a compile-time bound on R; the table holds zero taps beyond R.  A
// launch covers every field of the batch: a work-group whose table's radius
// is not in this RM's range (zform_of) exits at once.

template <int RM>
__global__ void __launch_bounds__(64)
zdirect_kernel(const FieldBatch fb, ColDesc cd) {
    using B = BufOps<float>;
    using O = typename B::O;
    using BD = BufOps<double>;
    using OD = typename BD::O;
    constexpr int U = 2 * RM + 1;
    constexpr unsigned CB = sizeof(cx<float>);
    const FieldDesc& fd = fb.f[blockIdx.z];
    if (zform_of_table(fd.kspec, cd.kspec_bytes, cd.sep_g) != RM) return;
    const auto rwork = B::make(fd.work, cd.work_bytes);
    const long col = (long)blockIdx.x * 64 + threadIdx.x;
    const long ky = blockIdx.y;
    const bool live = col < cd.ncols;
    const O base = (O)((live ? ky * cd.pstride + col : 0) * CB);
    const O rstep = (O)cd.rstride * CB;
    const int nz = cd.nin;
    const auto tk = BD::make(fd.kspec, cd.kspec_bytes);
    float g[RM + 1];
#pragma unroll
    for (int j = 0; j <= RM; ++j) g[j] = (float)BD::ld_r(tk, (OD)(cd.sep_g + 1 + j) * 8u);
    const double sc = BD::ld_r(tk, OD(0));
    const double gyv = BD::ld_r(tk, live ? (OD)(cd.sep_gy + ky) * 8u : BD::OOB);
    const double gxv = BD::ld_r(tk, live ? (OD)(1 + col) * 8u : BD::OOB);
    // The taps as rounded to FP32 sum (exactly) to gs rather than 1: the column
    // factor takes up the difference, so the window conserves mass
    double gs = (double)g[0];
#pragma unroll
    for (int j = 1; j <= RM; ++j) gs += 2.0 * (double)g[j];
    const double kad = (((gxv * gyv) * sc) * cd.scale) / gs;  // cd.scale = Pz here
    const float ka = (float)kad, kl = (float)(kad - (double)ka);
    auto row = [&](int m) -> O { return (live && m >= 0 && m < nz) ? base + (O)m * rstep : B::OOB; };
    auto body = [&](auto pol) {
        constexpr int A = decltype(pol)::value;
        // w[(n - RM + k) mod U] holds x[n - RM + k]; x[-RM .. RM] to start
        cx<float> w[U];
#pragma unroll
        for (int k = 0; k < U; ++k) w[k] = B::template ld_cx<A>(rwork, row(k - RM));
        for (int n0 = 0; n0 < nz; n0 += U) {
#pragma unroll
            for (int i = 0; i < U; ++i) {
                // output n = n0 + i; its centre x[n] sits at slot (i + RM) mod U
                const int n = n0 + i;
                const int c = (i + RM) % U;
                cx<float> acc{g[0] * w[c].x, g[0] * w[c].y};
#pragma unroll
                for (int j = 1; j <= RM; ++j) {
                    const cx<float> a = w[(c + U - j) % U], b = w[(c + j) % U];
                    acc.x = fmaf(g[j], a.x + b.x, acc.x);
                    acc.y = fmaf(g[j], a.y + b.y, acc.y);
                }
                // x[n - RM] is no longer needed: its slot takes x[n + RM + 1]
                w[i] = B::template ld_cx<A>(rwork, row(n + RM + 1));
                B::template st_cx<A>(cx<float>{fmaf(ka, acc.x, kl * acc.x), fmaf(ka, acc.y, kl * acc.y)}, rwork,
                                     row(n));
            }
        }
    };
    // Each element is read once and written once: non-temporal both ways for
    // a wave's full 512-byte row segments.  A work-group with part of its 64
    // columns past the row's end (the last one, 1025 = 16 * 64 + 1 columns on
    // a 2048 box) goes through the L2: a narrow non-temporal store costs a
    // partial-line write to HBM (this form is 8 % faster at radius 16).
    if ((long)(blockIdx.x + 1) * 64 <= cd.ncols) body(std::integral_constant<int, kStream>{});
    else body(std::integral_constant<int, kCacheDefault>{});
}

}  // namespace vfd
