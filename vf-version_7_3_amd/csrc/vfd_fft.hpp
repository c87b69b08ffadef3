// vfd_fft.hpp -- device-side building blocks of the CDNA4 FFT passes.
//
// A batch of length-Q complex DFTs is computed per work-group by a mixed-radix
// Stockham autosort (radices 2,3,4,5,7,8,11,13).  Each thread owns a fixed set
// of butterflies per stage; data moves between stages through LDS, so only the
// first stage's inputs and the last stage's outputs touch registers that the
// surrounding pass kernel fills from / drains to HBM.  Twiddles come from a
// per-length table W_P^k (P = Q * split) computed in FP64 on the host.
//
// Convention: forward = exp(-2 pi i nk/Q), inverse = exp(+2 pi i nk/Q), both
// unnormalized (the reference's cuFFT plans, src/Diffusion/Diffusion.cu:1814,1842).
#pragma once
#include <hip/hip_runtime.h>

namespace vfd {

// ---------------------------------------------------------------- complex
template <typename T> struct cx_of;
template <> struct cx_of<float> { using type = float2; };
template <> struct cx_of<double> { using type = double2; };
template <typename T> using cx = typename cx_of<T>::type;

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename V> __device__ __forceinline__ V cmulc(V a, V b) {
    return V{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename V, typename S> __device__ __forceinline__ V cscale(V a, S s) { return V{a.x * s, a.y * s}; }
template <typename V> __device__ __forceinline__ V cconj(V a) { return V{a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename V> __device__ __forceinline__ V mul_mi(V a) {
    return INV ? V{-a.y, a.x} : V{a.y, -a.x};
}
// twiddle multiply: a * w (forward) or a * conj(w) (inverse)
template <bool INV, typename V> __device__ __forceinline__ V ctw(V a, V w) {
    return INV ? cmulc(a, w) : cmul(a, w);
}

// ------------------------------------------------- compile-time constants
namespace ce {
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double reduce(double x) {  // to [-pi, pi]
    while (x > PI) x -= 2 * PI;
    while (x < -PI) x += 2 * PI;
    return x;
}
constexpr double sin_(double x) {
    x = reduce(x);
    double term = x, s = x;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += term;
    }
    return s;
}
constexpr double cos_(double x) {
    x = reduce(x);
    double term = 1.0, s = 1.0;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += term;
    }
    return s;
}
}  // namespace ce

// ------------------------------------------------------------ butterflies
// In-place length-R DFT of v[0..R).
template <int R, bool INV, typename T> struct Bfly;

template <bool INV, typename T> struct Bfly<2, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV, typename T> struct Bfly<4, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        cx<T> t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <bool INV, typename T> struct Bfly<8, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        const T h = T(0.70710678118654752440084436210484903928);
        // b_k = a_k + a_{k+4}; c_k = (a_k - a_{k+4}) W8^k
        cx<T> b0 = cadd(v[0], v[4]), c0 = csub(v[0], v[4]);
        cx<T> b1 = cadd(v[1], v[5]), c1 = csub(v[1], v[5]);
        cx<T> b2 = cadd(v[2], v[6]), c2 = csub(v[2], v[6]);
        cx<T> b3 = cadd(v[3], v[7]), c3 = csub(v[3], v[7]);
        // W8^1 = h(1 - i) fwd / h(1 + i) inv ; W8^2 = -i / +i ; W8^3 = h(-1 - i) / h(-1 + i)
        c1 = INV ? cx<T>{h * (c1.x - c1.y), h * (c1.x + c1.y)} : cx<T>{h * (c1.x + c1.y), h * (c1.y - c1.x)};
        c2 = mul_mi<INV>(c2);
        c3 = INV ? cx<T>{-h * (c3.x + c3.y), h * (c3.x - c3.y)} : cx<T>{h * (c3.y - c3.x), -h * (c3.x + c3.y)};
        cx<T> e[4] = {b0, b1, b2, b3};
        cx<T> o[4] = {c0, c1, c2, c3};
        Bfly<4, INV, T>::run(e);
        Bfly<4, INV, T>::run(o);
        v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
        v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
    }
};

template <bool INV, typename T> struct Bfly<3, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        const T s3 = T(0.86602540378443864676372317075293618347);
        cx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        cx<T> a0 = v[0];
        v[0] = cadd(a0, s);
        cx<T> t = cx<T>{a0.x - T(0.5) * s.x, a0.y - T(0.5) * s.y};
        // fwd: y1 = t - i s3 d ; y2 = t + i s3 d
        cx<T> m = mul_mi<INV>(cscale(d, s3));
        v[1] = cadd(t, m);
        v[2] = csub(t, m);
    }
};

// Generic odd-prime DFT using the (n, R-n) symmetry; constants are compile time.
template <int R, bool INV, typename T> struct BflyOdd {
    static __device__ __forceinline__ void run(cx<T>* v) {
        constexpr int H = (R - 1) / 2;
        cx<T> a[H + 1], b[H + 1];
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            a[n] = cadd(v[n], v[R - n]);
            b[n] = csub(v[n], v[R - n]);
        }
        cx<T> x0 = v[0];
        cx<T> y0 = x0;
#pragma unroll
        for (int n = 1; n <= H; ++n) y0 = cadd(y0, a[n]);
        cx<T> out[R];
        out[0] = y0;
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            T cr = x0.x, ci = x0.y, sr = 0, si = 0;
#pragma unroll
            for (int n = 1; n <= H; ++n) {
                const T c = T(ce::cos_(2.0 * ce::PI * double((n * k) % R) / double(R)));
                const T s = T(ce::sin_(2.0 * ce::PI * double((n * k) % R) / double(R)));
                cr += a[n].x * c;
                ci += a[n].y * c;
                sr += b[n].x * s;
                si += b[n].y * s;
            }
            // fwd: y_k = sum a cos - i sum b sin ; y_{R-k} = sum a cos + i sum b sin
            // -i (sr + i si) = si - i sr
            if (!INV) {
                out[k] = cx<T>{cr + si, ci - sr};
                out[R - k] = cx<T>{cr - si, ci + sr};
            } else {
                out[k] = cx<T>{cr - si, ci + sr};
                out[R - k] = cx<T>{cr + si, ci - sr};
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = out[k];
    }
};
template <bool INV, typename T> struct Bfly<5, INV, T> : BflyOdd<5, INV, T> {};
template <bool INV, typename T> struct Bfly<7, INV, T> : BflyOdd<7, INV, T> {};
template <bool INV, typename T> struct Bfly<11, INV, T> : BflyOdd<11, INV, T> {};
template <bool INV, typename T> struct Bfly<13, INV, T> : BflyOdd<13, INV, T> {};

// ------------------------------------------------------------------ plans
// Radix of stage i for length Q: 8s first, then the 2-power remainder (4 or 2),
// then 3, 5, 7, 11, 13.
constexpr int pow2_part(int q) { int p = 1; while (q % 2 == 0) { q /= 2; p *= 2; } return p; }
constexpr int num_stages(int q) {
    int n = 0;
    int p2 = pow2_part(q);
    while (p2 >= 8) { p2 /= 8; ++n; }
    if (p2 > 1) ++n;
    q /= pow2_part(q);
    const int odd[5] = {3, 5, 7, 11, 13};
    for (int i = 0; i < 5; ++i) while (q % odd[i] == 0) { q /= odd[i]; ++n; }
    return q == 1 ? n : -1;
}
constexpr int stage_radix(int q, int s) {
    int p2 = pow2_part(q);
    int i = 0;
    while (p2 >= 8) { if (i == s) return 8; p2 /= 8; ++i; }
    if (p2 > 1) { if (i == s) return p2; ++i; }
    q /= pow2_part(q);
    const int odd[5] = {3, 5, 7, 11, 13};
    for (int k = 0; k < 5; ++k)
        while (q % odd[k] == 0) { if (i == s) return odd[k]; q /= odd[k]; ++i; }
    return 0;
}
constexpr bool supported_length(int q) { return q >= 2 && num_stages(q) > 0; }
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
// registers needed per thread: max over stages of (butterflies per thread) * radix
constexpr int max_elems(int q, int tpf) {
    int m = 0;
    for (int s = 0; s < num_stages(q); ++s) {
        int r = stage_radix(q, s);
        int e = cdiv(q / r, tpf) * r;
        if (e > m) m = e;
    }
    return m;
}

// ------------------------------------------------------------ the engine
// LDS accessor concept: struct with  __device__ cx<T>& at(int elem)  for the
// thread's own FFT.  TW: pointer to W_P^k table (k in [0,P)), TWS = P/Q.
//
// Element ownership in a stage of radix R: butterfly j = t + b*TPF (b < NB),
// inputs j + r*(Q/R), outputs ((j/NS)*NS*R + j%NS) + r*NS.
template <typename T, int Q, int TPF> struct FftShape {
    static constexpr int NST = num_stages(Q);
    static constexpr int R0 = stage_radix(Q, 0);
    static constexpr int RL = stage_radix(Q, NST - 1);
    static constexpr int E = max_elems(Q, TPF);
    static constexpr int NB0 = cdiv(Q / R0, TPF);
    static constexpr int NBL = cdiv(Q / RL, TPF);
    // first-stage input element of slot (b, r), or -1 if the slot is idle
    static __device__ __forceinline__ int in_elem(int t, int b, int r) {
        int j = t + b * TPF;
        return (j < Q / R0) ? j + r * (Q / R0) : -1;
    }
    // last-stage output element of slot (b, r), or -1 if idle
    static __device__ __forceinline__ int out_elem(int t, int b, int r) {
        int j = t + b * TPF;
        return (j < Q / RL) ? j + r * (Q / RL) : -1;
    }
};

template <typename T, int Q, int TPF, bool INV, int S, int NS> struct StageRunner {
    static constexpr int NST = num_stages(Q);
    static constexpr int R = stage_radix(Q, S);
    static constexpr int NBF = Q / R;
    static constexpr int NB = cdiv(NBF, TPF);

    // butterflies of this stage on registers already holding the stage inputs
    template <typename TW>
    static __device__ __forceinline__ void compute(int t, cx<T>* v, const TW* tw, int tws) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * TPF;
            if (NBF % TPF == 0 || j < NBF) {
                if (NS > 1) {
                    const int k = j % NS;
                    const int step = (Q / (NS * R)) * tws;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const TW w = tw[r * k * step];
                        v[b * R + r] = ctw<INV>(v[b * R + r], cx<T>{T(w.x), T(w.y)});
                    }
                }
                Bfly<R, INV, T>::run(v + b * R);
            }
        }
    }
    template <typename LDS>
    static __device__ __forceinline__ void load(int t, cx<T>* v, LDS& lds) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * TPF;
            if (NBF % TPF == 0 || j < NBF) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b * R + r] = lds.at(j + r * NBF);
            }
        }
    }
    template <typename LDS>
    static __device__ __forceinline__ void store(int t, const cx<T>* v, LDS& lds) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * TPF;
            if (NBF % TPF == 0 || j < NBF) {
                const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
                for (int r = 0; r < R; ++r) lds.at(base + r * NS) = v[b * R + r];
            }
        }
    }
    // run this stage and the remaining ones.  On entry v holds this stage's
    // inputs (stage 0) or the data is in LDS (later stages, FROM_LDS).
    template <bool FROM_LDS, typename LDS, typename TW>
    static __device__ __forceinline__ void run(int t, cx<T>* v, LDS& lds, const TW* tw, int tws) {
        if (FROM_LDS) load(t, v, lds);
        compute(t, v, tw, tws);
        if constexpr (S + 1 < NST) {
            __syncthreads();  // every thread finished reading this stage's inputs
            store(t, v, lds);
            __syncthreads();  // stage outputs visible
            StageRunner<T, Q, TPF, INV, S + 1, NS * R>::template run<true>(t, v, lds, tw, tws);
        }
    }
};

// Full length-Q transform of one FFT per thread group.  On entry v holds the
// first-stage inputs (FftShape::in_elem); on exit v holds the last-stage
// outputs (FftShape::out_elem).  Contains __syncthreads(): every thread of the
// work-group must call it.
template <typename T, int Q, int TPF, bool INV, typename LDS, typename TW>
__device__ __forceinline__ void fft_regs(int t, cx<T>* v, LDS& lds, const TW* tw, int tws) {
    StageRunner<T, Q, TPF, INV, 0, 1>::template run<false>(t, v, lds, tw, tws);
}

}  // namespace vfd
