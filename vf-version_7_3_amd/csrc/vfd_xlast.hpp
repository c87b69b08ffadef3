// vfd_xlast.hpp -- the x-last pass order for boxes padded most along x.
//
//   zr   : pad + R2C along z of two real columns at a time (x, x+1 of one row)
//   col  : C2C along y, forward and inverse (col_kernel, vfd_passes.hpp)
//   xc   : x rows: forward, x K^, inverse, fused (zero pad nx -> px, crop back)
//   zri  : C2R along z of the column pairs + unpad (+ the constant-pad restore,
//          the fused chem update and the visualisation downsample)
//
// Same operator as the x-first order (vfd_passes.hpp): the 3D transform of the
// padded field is factored z, y, x instead of x, y, z.  Each pass streams its
// array once, so the HBM bytes of a pass are the sizes of what it reads and
// writes; transforming the most-padded axis last keeps its padding out of HBM:
// on the 1152x1536x256 box of a 1006x1390x110 field the arrays of the middle
// passes are 577x1536x110 instead of 1006x1536x129 complex elements.
//
// Reference semantics: fftDiffuse3D (src/Diffusion/Diffusion.cu:1622-1934):
// padDataConstantVF3D (convolutionFFT.cuh:590-638), R2C (Diffusion.cu:1814),
// modulateAndNormalize (convolutionFFT.cuh:690-716), C2R (Diffusion.cu:1842),
// unpadResult3D (convolutionFFT.cuh:152-171).  The spectrum is the reference's
// R2C layout [pz][py][px/2+1]; the x pass needs K^ at every kx of px, which the
// real kernel's Hermitian symmetry gives: K^(kz,ky,kx) = conj K^(-kz,-ky,px-kx)
// for kx > px/2 (the extension the reference's C2R applies implicitly).
#pragma once
#include "vfd_passes.hpp"

namespace vfd {

// ------------------------------------------------------------- z passes
// Two real columns a (x) and b (x+1) of data row y form one complex column
// c = a + i b.  Forward: C = FFT_pz(c), then for k = 0..pz/2
//   A_k = (C_k + conj C_{pz-k}) / 2,   B_k = -i (C_k - conj C_{pz-k}) / 2,
// the half spectra of a and b, stored side by side at (k, y, x) and (k, y, x+1).
// Inverse: C_k = A_k + i B_k with A_{pz-k} = conj A_k (k > pz/2; the imaginary
// parts of the self-conjugate bins 0 and pz/2 are dropped, as a C2R does), an
// unnormalized inverse FFT, a = Re c, b = Im c cropped to nz.
// A work-group owns TC column pairs; narrow tiles use the grouped 1-D launch
// of col_kernel (the SH tiles sharing a 128-byte input segment run on one XCD).
template <typename T, int Q, int V, bool FWD>
__global__ void __launch_bounds__((ColCfg<T, Q, V, 1>::NT))
zr_kernel(const FieldBatch fb, XlDesc d, const cx<T>* __restrict__ twg) {
    using C = ColCfg<T, Q, V, 1>;
    using Sh = FftShape<T, Q, C::TPF, false, C::MAXR>;
    using B = BufOps<T>;
    using O = typename B::O;
    constexpr unsigned CB = sizeof(cx<T>), RB = sizeof(T);
    __shared__ cx<T> smem[C::TC * Q];
    __shared__ cx<T> tw[Q];
    unsigned bx = blockIdx.x, bf = blockIdx.y;
    if constexpr (C::SH > 1) {
        const unsigned lin = blockIdx.x;
        const unsigned grp = (lin / (8u * C::SH)) * 8u + lin % 8u;
        bx = (grp % d.grid_segs) * C::SH + (lin / 8u) % C::SH;
        bf = grp / d.grid_segs;
        // whole-block exits, before any barrier: padding ids, tiles past the end
        if (bf >= d.grid_fields || (long)bx * C::TC >= d.ncolp) return;
    }
    const FieldDesc& fd = fb.f[bf];
    const int c = threadIdx.x % C::TC;
    const int t = threadIdx.x / C::TC;
    const long j = (long)bx * C::TC + c;
    const bool live = j < d.ncolp;
    const int y = live ? (int)(j / d.nxp) : 0;
    const int x0 = live ? 2 * (int)(j - (long)y * d.nxp) : 0;
    ColLds<T, C::TC> lds{smem + c};
    const auto rwork = B::make(fd.work, d.work_bytes);
    const O wcol = (O)((long)y * d.ldx + x0) * CB;  // (k, y, x0) = wcol + k * wstep
    const O wstep = (O)d.plane * CB;
    const O dcol = (O)((long)y * d.nx + x0) * RB;   // data (z, y, x0) = dcol + z * dstep
    const O dstep = (O)((long)d.ny * d.nx) * RB;
    const bool vec = d.nx % 2 == 0;                  // x0, x0+1 in one 8-byte word
    // non-temporal accesses where a tile row spans whole 128-byte lines (loads)
    // or 64-byte segments (stores); narrower rows share lines with neighbouring
    // tiles, which must find them in (or merge them through) the L2
    constexpr int NTP = C::TC * (int)RB * 2 >= 128 ? kStream : kCacheDefault;   // field rows
    constexpr int NTW = C::TC * (int)CB * 2 >= 128 ? kStream : kCacheDefault;   // work rows
    constexpr int NTSF = C::TC * (int)RB * 2 >= 128 ? kStream : kCacheDefault;   // field stores
    constexpr int NTSW = C::TC * (int)CB * 2 >= 128 ? kStream : kCacheDefault;   // work stores
    cx<T> v[Sh::E];

    if constexpr (FWD) {
        const auto rin = B::make(fd.in, d.in_bytes);
        // constant-pad mode: interior 0<z<nz-1, 0<y<ny-1, 0<x<epi keeps u-b, the rest 0
        const bool cvf = d.pad_mode == PAD_CONSTANT_VF;
        const int xlo = cvf ? 1 : 0;
        const int xhi = cvf ? (fd.epi < d.nx ? fd.epi : d.nx) : d.nx;
        const bool yin = live && (!cvf || (y > 0 && y < d.ny - 1));
        const bool in0 = yin && x0 >= xlo && x0 < xhi;
        const bool in1 = yin && x0 + 1 >= xlo && x0 + 1 < xhi;
        const T base = cvf ? (T)fd.base : T(0);
#pragma unroll
        for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
            for (int r = 0; r < Sh::R0; ++r) {
                const int n = Sh::in_elem(t, b, r);
                const bool zin = n >= 0 && n < d.nz && (!cvf || (n > 0 && n < d.nz - 1));
                const O off = dcol + (O)(zin ? n : 0) * dstep;
                cx<T> val;
                if (vec && !cvf) {
                    val = B::template ld_cx<NTP>(rin, zin && in0 ? off : B::OOB);
                } else {
                    const T a0 = B::template ld_r<NTP>(rin, zin && in0 ? off : B::OOB);
                    const T a1 = B::template ld_r<NTP>(rin, zin && in1 ? off + RB : B::OOB);
                    val = cx<T>{zin && in0 ? a0 - base : T(0), zin && in1 ? a1 - base : T(0)};
                }
                v[b * Sh::R0 + r] = val;
            }
        stage_table_unrolled<Q, C::NT>(tw, twg, Q);
        __syncthreads();
        fft_regs<T, Q, C::TPF, false, false, C::MAXR>(t, v, lds, tw, 1);
        __syncthreads();  // every LDS read of the last stage is done
#pragma unroll
        for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
            for (int r = 0; r < Sh::RL; ++r) {
                const int k = Sh::out_elem(t, b, r);
                if (k >= 0) lds.at(k) = v[b * Sh::RL + r];
            }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
            for (int r = 0; r < Sh::RL; ++r) {
                const int k = Sh::out_elem(t, b, r);
                const bool ok = k >= 0 && 2 * k <= Q && live;
                const cx<T> ck = v[b * Sh::RL + r];
                const cx<T> cc = cconj(lds.at(ok && k > 0 ? Q - k : 0));
                const cx<T> s = cscale(cadd(ck, cc), T(0.5));
                const cx<T> dd = cscale(csub(ck, cc), T(0.5));
                B::template st_cx2<NTSW>(s, cx<T>{dd.y, -dd.x}, rwork, ok ? wcol + (O)k * wstep : B::OOB);
            }
    } else {
        // The half spectrum (rows 0..Q/2) of the tile's column pairs is read
        // once, each row by one lane, and staged in the LDS strip, which holds
        // exactly Q complex values per pair: (Re A_0, Re B_0) at 0, (Re A_{Q/2},
        // Re B_{Q/2}) at 1 (even Q; the imaginary parts of these self-conjugate
        // bins are dropped, as a C2R does), A_m and B_m at 2m and 2m+1
        // (0 < m < Q/2; odd Q: at 2m-1 and 2m).  Every lane then forms its
        // C_k = A_k + i B_k from LDS, conjugating the rows above Q/2.
        constexpr int HZ = Q / 2 + 1;
        constexpr int NLD = (HZ + C::TPF - 1) / C::TPF;
        constexpr bool EVEN = Q % 2 == 0;
        auto slot = [](int m) { return EVEN ? 2 * m : 2 * m - 1; };  // A_m at slot(m), B_m at slot(m) + 1
        cx<T> ha[NLD], hb[NLD];
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int m = t + i * C::TPF;
            B::template ld_cx2<NTW>(rwork, m < HZ && live ? wcol + (O)m * wstep : B::OOB, ha[i], hb[i]);
        }
        stage_table_unrolled<Q, C::NT>(tw, twg, Q);
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int m = t + i * C::TPF;
            if (m == 0) {
                lds.at(0) = cx<T>{ha[i].x, hb[i].x};
            } else if (EVEN && m == Q / 2) {
                lds.at(1) = cx<T>{ha[i].x, hb[i].x};
            } else if (m < HZ) {
                lds.at(slot(m)) = ha[i];
                lds.at(slot(m) + 1) = hb[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < Sh::NB0; ++b)
#pragma unroll
            for (int r = 0; r < Sh::R0; ++r) {
                const int k = Sh::in_elem(t, b, r);
                cx<T> a{T(0), T(0)}, bb{T(0), T(0)};
                if (k == 0 || (EVEN && 2 * k == Q)) {
                    const cx<T> e = lds.at(k == 0 ? 0 : 1);
                    a.x = e.x;
                    bb.x = e.y;
                } else if (k > 0) {
                    const bool upper = 2 * k > Q;
                    const int m = upper ? Q - k : k;
                    a = lds.at(slot(m));
                    bb = lds.at(slot(m) + 1);
                    if (upper) {
                        a = cconj(a);
                        bb = cconj(bb);
                    }
                }
                v[b * Sh::R0 + r] = cx<T>{a.x - bb.y, a.y + bb.x};
            }
        // the first LDS store of fft_regs follows a __syncthreads, after every
        // lane has read its inputs above
        fft_regs<T, Q, C::TPF, true, false, C::MAXR>(t, v, lds, tw, 1);
        T addc = T(0);
        if (d.pad_mode == PAD_CONSTANT_VF) {
            // base * sum(k): the complex spectrum's K^(0,0,0) is sum(k) itself
            const auto rk = B::make(fd.kspec, d.kspec_bytes);
            addc = (T)fd.base * B::ld_r(rk, 0);
        }
        const auto rout = B::make(fd.out, d.out_bytes);
        const bool has1 = x0 + 1 < d.nx;
        // fused chem update (FP32): p = d + t * decay with the product rounded
        // before the add (WHChemical.cpp:425-428), and d cleared
        [[maybe_unused]] const bool upd = sizeof(T) == 4 && fd.delta != nullptr;
        [[maybe_unused]] const auto rdel = B::make(fd.delta, d.out_bytes);
        [[maybe_unused]] const T dec = (T)fd.decay;
        // fused visualisation downsample (FP32): sampleChem's every vs-th voxel,
        // min(t * scale, 1), into the dense [nz/vs][ny/vs][nx/vs] array
        const bool vis = sizeof(T) == 4 && fd.vis != nullptr && live;
        const int vs = vis ? fd.vis_stride : 1;
        const int vD = d.nz / vs, vH = d.ny / vs, vW = d.nx / vs;
        const bool vcol = vis && y % vs == 0 && y / vs < vH;
        const auto rvis = B::make(fd.vis, vis ? (unsigned long)vD * vH * vW * RB : 0ul);
        auto vis_put = [&](int z, int x, T tv) {
            if (vcol && z % vs == 0 && z / vs < vD && x % vs == 0 && x / vs < vW) {
                float s = (float)tv * fd.vis_scale;
                if (s > 1.0f) s = 1.0f;
                B::st_r((T)s, rvis, (O)(((long)(z / vs) * vH + y / vs) * vW + x / vs) * RB);
            }
        };
#pragma unroll
        for (int b = 0; b < Sh::NBL; ++b)
#pragma unroll
            for (int r = 0; r < Sh::RL; ++r) {
                const int n = Sh::out_elem(t, b, r);
                const bool ok = n >= 0 && n < d.nz && live;
                const O off = dcol + (O)(ok ? n : 0) * dstep;
                const cx<T> val = v[b * Sh::RL + r];
                T o0 = val.x + addc, o1 = val.y + addc;
                if (ok && vis) {
                    vis_put(n, x0, o0);
                    if (has1) vis_put(n, x0 + 1, o1);
                }
                if (vec) {
                    if constexpr (sizeof(T) == 4) {
                        if (upd) {
                            const cx<T> dl = B::ld_cx(rdel, ok ? off : B::OOB);
                            o0 = upd_level(dl.x, o0, dec);
                            o1 = upd_level(dl.y, o1, dec);
                            B::template st_cx<NTSF>(cx<T>{T(0), T(0)}, rdel, ok ? off : B::OOB);
                        }
                    }
                    B::template st_cx<NTSF>(cx<T>{o0, o1}, rout, ok ? off : B::OOB);
                } else {
                    const O off0 = ok ? off : B::OOB;
                    const O off1 = ok && has1 ? off + RB : B::OOB;
                    if constexpr (sizeof(T) == 4) {
                        if (upd) {
                            o0 = upd_level(B::ld_r(rdel, off0), o0, dec);
                            o1 = upd_level(B::ld_r(rdel, off1), o1, dec);
                            B::template st_r<NTSF>(T(0), rdel, off0);
                            B::template st_r<NTSF>(T(0), rdel, off1);
                        }
                    }
                    B::template st_r<NTSF>(o0, rout, off0);
                    B::template st_r<NTSF>(o1, rout, off1);
                }
            }
    }
}

// ------------------------------------------------------------- x rows
// Row (kz, ky) of the work array: nx elements, zero padded to px, forward FFT,
// x K^(kz, ky, kx) * scale, inverse FFT (reversed radix plan: it starts on the
// elements the forward left in registers), the first nx written back.  A
// work-group holds G rows of TPF threads each and loops over row groups of
// every field; the spectrum of a row is loaded before its transform starts.
template <typename T, int P, int V> struct XcCfg {
    static constexpr int MAXR = var_maxr(V);
    static constexpr int TPF = tpf_for(P, MAXR);
    static constexpr int NTT = var_ntmax(V) / 2;
    static constexpr int G = (NTT / TPF) > 0 ? (NTT / TPF) : 1;
    static constexpr int NT = TPF * G;
    static constexpr int PITCH = P + P / 32 + 1;
    static constexpr bool WS = wave_local<TPF>();
    static constexpr bool TW_LDS = (P + G * PITCH) * (int)sizeof(cx<T>) <= 64 * 1024;
    static constexpr int TW_N = TW_LDS ? P : 1;
};

template <typename T, int P, int V>
__global__ void __launch_bounds__((XcCfg<T, P, V>::NT))
xc_kernel(const FieldBatch fb, XlDesc d, const cx<T>* __restrict__ twg, int nfields) {
    using C = XcCfg<T, P, V>;
    using ShF = FftShape<T, P, C::TPF, false, C::MAXR>;
    using ShI = FftShape<T, P, C::TPF, true, C::MAXR>;
    using B = BufOps<T>;
    using O = typename B::O;
    constexpr unsigned CB = sizeof(cx<T>);
    constexpr int NK = ShF::NBL * ShF::RL;
    __shared__ cx<T> smem[C::G * C::PITCH];
    __shared__ cx<T> twl[C::TW_N];
    const cx<T>* tw = twg;
    if constexpr (C::TW_LDS) {
        stage_table(twl, twg, P);
        __syncthreads();
        tw = twl;
    }
    const int t = threadIdx.x % C::TPF;
    const int gr = threadIdx.x / C::TPF;
    const long nrows = (long)d.hz * d.py;
    const long ngroups = (nrows + C::G - 1) / C::G;
    RowLds<T> lds{smem + gr * C::PITCH};
    const T sc = (T)d.scale;
    for (long item = blockIdx.x; item < ngroups * nfields; item += gridDim.x) {
        const int f = (int)(item / ngroups);
        const long row = (item % ngroups) * C::G + gr;
        const FieldDesc& fd = fb.f[f];
        const auto rwork = B::make(fd.work, d.work_bytes);
        const auto rk = B::make(fd.kspec, d.kspec_bytes);
        const bool live = row < nrows;
        const int kz = live ? (int)(row / d.py) : 0;
        const int ky = live ? (int)(row % d.py) : 0;
        const O roff = (O)((long)kz * d.plane + (long)ky * d.ldx) * CB;
        // K^ row (kz, ky) for kx <= px/2; row (-kz, -ky) mirrored and conjugated above
        const int kzm = kz == 0 ? 0 : d.pz - kz, kym = ky == 0 ? 0 : d.py - ky;
        const O kd = (O)(((long)kz * d.py + ky) * d.ks_hx) * CB;
        const O km = (O)(((long)kzm * d.py + kym) * d.ks_hx) * CB;
        cx<T> kv[NK];
#pragma unroll
        for (int b = 0; b < ShF::NBL; ++b)
#pragma unroll
            for (int r = 0; r < ShF::RL; ++r) {
                const int k = ShF::out_elem(t, b, r);
                const bool ok = k >= 0 && live;
                const bool upper = 2 * k > P;
                const O off = !ok ? B::OOB : upper ? km + (O)(P - k) * CB : kd + (O)k * CB;
                const cx<T> kk = B::template ld_cx<kStream>(rk, off);
                kv[b * ShF::RL + r] = upper ? cconj(kk) : kk;
            }
        cx<T> v[ShF::E > ShI::E ? ShF::E : ShI::E];
#pragma unroll
        for (int b = 0; b < ShF::NB0; ++b)
#pragma unroll
            for (int r = 0; r < ShF::R0; ++r) {
                const int n = ShF::in_elem(t, b, r);
                const bool ok = n >= 0 && n < d.nx && live;
                v[b * ShF::R0 + r] = B::template ld_cx<kStream>(rwork, ok ? roff + (O)n * CB : B::OOB);
            }
        fft_regs<T, P, C::TPF, false, false, C::MAXR, C::WS>(t, v, lds, tw, 1);
#pragma unroll
        for (int i = 0; i < NK; ++i) v[i] = cscale(cmul(v[i], kv[i]), sc);
        fft_regs<T, P, C::TPF, true, true, C::MAXR, C::WS>(t, v, lds, tw, 1);
#pragma unroll
        for (int b = 0; b < ShI::NBL; ++b)
#pragma unroll
            for (int r = 0; r < ShI::RL; ++r) {
                const int n = ShI::out_elem(t, b, r);
                const bool ok = n >= 0 && n < d.nx && live;
                B::template st_cx<kStream>(v[b * ShI::RL + r], rwork, ok ? roff + (O)n * CB : B::OOB);
            }
        // the next item's first LDS store happens after a stage sync inside fft_regs
    }
}

}  // namespace vfd
