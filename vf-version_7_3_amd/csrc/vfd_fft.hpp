// vfd_fft.hpp -- device-side building blocks of the CDNA4 FFT passes.
//
// A batch of length-Q complex DFTs is computed per work-group by a mixed-radix
// Stockham autosort (radices 2,3,4,5,7,8,11,13,16).  Each thread owns a fixed
// set of butterflies per stage; data moves between stages through LDS, so only
// the first stage's inputs and the last stage's outputs live in registers that
// the surrounding pass kernel fills from / drains to HBM.  Twiddles are read
// from a per-length table W_P^k (P = Q * split) computed in FP64 on the host
// and staged in LDS by the pass kernel.
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

// FP32 rotation constants.  A rotation w = c - i s rounded componentwise to
// FP32 has |w|^2 = c^2 + s^2 != 1: applied on every pass, its magnitude error
// is a gain of the frequencies whose path crosses it, the same every tick, so
// it accumulates linearly over ticks (with the plain roundings W16^1 has
// |w|^2 - 1 = -5.7e-8 and W8^1 -3.4e-8; a radix-16 stage then loses about
// 1.1e-8 of amplitude on average, about 9e-8 per tick over the x and y
// passes).  Phase errors cancel between a forward rotation and the inverse
// one (its conjugate); magnitude errors do not.  So each pair is picked among
// the FP32 neighbours (+-2 ulp) of the rounded (c, s) for c^2 + s^2 closest
// to a target magnitude, the smaller phase error breaking ties.
struct Rot {
    float c, s;
};
constexpr double ulp_at(float x) {  // spacing of the floats just above |x|
    double a = x < 0 ? -(double)x : (double)x;
    if (a == 0.0) return 1.4e-45;
    double p = 1.0;
    while (p > a) p *= 0.5;
    while (p * 2.0 <= a) p *= 2.0;
    return p * (1.0 / 8388608.0);  // 2^-23
}
constexpr float step_float(float x, int k) {  // k floats away from x (k in -2..2)
    float y = x;
    for (int i = 0; i < (k < 0 ? -k : k); ++i) {
        const double u = ulp_at(y);
        // below a power of two the spacing halves: try the half step first
        const float half = (float)((double)y + (k < 0 ? -0.5 : 0.5) * u);
        y = half != y ? half : (float)((double)y + (k < 0 ? -u : u));
    }
    return y;
}
constexpr double absd(double x) { return x < 0 ? -x : x; }
constexpr Rot best_rot(double c, double s, double target = 1.0) {
    const float c0 = (float)c, s0 = (float)s;
    Rot best{c0, s0};
    double be = 1e300, ba = 1e300;
    for (int i = -2; i <= 2; ++i)
        for (int j = -2; j <= 2; ++j) {
            const float cc = step_float(c0, i), ss = step_float(s0, j);
            const double e = absd((double)cc * cc + (double)ss * ss - target);
            const double a = absd((double)cc - c) + absd((double)ss - s);
            if (e < be || (e == be && a < ba)) {
                best = Rot{cc, ss};
                be = e;
                ba = a;
            }
        }
    return best;
}
// W16^1 and the radix-3 pair (-1/2, sqrt(3)/2), both directions
constexpr Rot kW16 = best_rot(0.92387953251128675613, 0.38268343236508977173);
constexpr Rot kW3 = best_rot(-0.5, 0.86602540378443864676);
// W8^1 = (1 - i)/sqrt(2): the plain rounding (near 45 degrees the FP32 lattice
// moves |w|^2 in steps of about 8e-8)
constexpr Rot kW8 = Rot{(float)0.70710678118654752440, (float)0.70710678118654752440};
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

// The rotation pairs (c, s) of the butterflies: FP32 takes the magnitude-picked
// pairs of ce, the same in both directions, FP64 the correctly rounded values.
// (Inverse pairs picked so that their magnitudes cancel the forward ones over a
// round trip gave each FP32 inverse pass a gain of its own, and the drift from
// FP64 over many ticks grew: 100-tick rel-L2 9.8e-6 against 6.9e-6.)
template <typename T, bool INV> struct Rots {
    static constexpr T w8c = T(0.70710678118654752440), w8s = T(0.70710678118654752440);
    static constexpr T w16c = T(0.92387953251128675613), w16s = T(0.38268343236508977173);
    static constexpr T w3c = T(-0.5), w3s = T(0.86602540378443864676);
};
template <bool INV> struct Rots<float, INV> {
    static constexpr ce::Rot w8 = ce::kW8, w16 = ce::kW16, w3 = ce::kW3;
    static constexpr float w8c = w8.c, w8s = w8.s, w16c = w16.c, w16s = w16.s, w3c = w3.c, w3s = w3.s;
};

// FP32: 1/sqrt(2) as hi + lo.  The FP32 value alone is the only pair near 45
// degrees and has |w|^2 - 1 = -3.4e-8, the same on every path through a W8
// rotation; (x -+ y) hi + (x -+ y) lo carries no such gain.
constexpr float kRsqrt2Hi = 0.70710678118654752440f;
constexpr float kRsqrt2Lo = (float)(0.70710678118654752440 - (double)kRsqrt2Hi);
__device__ __forceinline__ float rsqrt2_mul(float p) { return fmaf(p, kRsqrt2Hi, p * kRsqrt2Lo); }
// the same for sin(2 pi/3) of the radix-3 butterfly (its best FP32 pair is
// off by -1.1e-8 in |w|^2).  W16's pair is left single (+7e-9): carrying it
// as hi + lo too made the step 0.5 % slower for a smaller rel-L2 (100-tick
// precision run 3.4e-6 -> 2.6e-6) and a larger mass drift (2.0e-6 -> 2.5e-6).
constexpr double kS3 = 0.86602540378443864676;
constexpr float kS3Hi = (float)kS3, kS3Lo = (float)(kS3 - (double)(float)kS3);

// multiply by W8^1 = (1 - i)/sqrt(2) (fwd) / (1 + i)/sqrt(2) (inv), as the
// pair (a, b) ~ (1/sqrt(2), 1/sqrt(2)): c (a -+ i b)
template <bool INV, typename T> __device__ __forceinline__ cx<T> w8_1(cx<T> c) {
    if constexpr (sizeof(T) == 4) {
        return INV ? cx<T>{rsqrt2_mul(c.x - c.y), rsqrt2_mul(c.y + c.x)}
                   : cx<T>{rsqrt2_mul(c.x + c.y), rsqrt2_mul(c.y - c.x)};
    } else {
        constexpr T a = Rots<T, INV>::w8c, b = Rots<T, INV>::w8s;
        return INV ? cx<T>{c.x * a - c.y * b, c.y * a + c.x * b} : cx<T>{c.x * a + c.y * b, c.y * a - c.x * b};
    }
}
// multiply by W8^3 = (-1 - i)/sqrt(2) (fwd) / (-1 + i)/sqrt(2) (inv): c (-a -+ i b)
template <bool INV, typename T> __device__ __forceinline__ cx<T> w8_3(cx<T> c) {
    if constexpr (sizeof(T) == 4) {
        return INV ? cx<T>{rsqrt2_mul(-c.x - c.y), rsqrt2_mul(c.x - c.y)}
                   : cx<T>{rsqrt2_mul(c.y - c.x), rsqrt2_mul(-c.y - c.x)};
    } else {
        constexpr T a = Rots<T, INV>::w8c, b = Rots<T, INV>::w8s;
        return INV ? cx<T>{-c.x * a - c.y * b, c.x * b - c.y * a} : cx<T>{c.y * b - c.x * a, -c.y * a - c.x * b};
    }
}

template <bool INV, typename T> struct Bfly<8, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        // b_k = a_k + a_{k+4}; c_k = (a_k - a_{k+4}) W8^k; Y[2m] = DFT4(b), Y[2m+1] = DFT4(c)
        cx<T> e[4], o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            e[k] = cadd(v[k], v[k + 4]);
            o[k] = csub(v[k], v[k + 4]);
        }
        o[1] = w8_1<INV, T>(o[1]);
        o[2] = mul_mi<INV>(o[2]);
        o[3] = w8_3<INV, T>(o[3]);
        Bfly<4, INV, T>::run(e);
        Bfly<4, INV, T>::run(o);
        v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
        v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
    }
};

// 16-point DFT as 4 x 4: x[k1 + 4 k2]; DFT4 over k2, twiddle W16^{k1 m2}, DFT4 over k1.
template <bool INV, typename T> struct Bfly<16, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        constexpr T c1 = Rots<T, INV>::w16c;  // cos(pi/8)
        constexpr T s1 = Rots<T, INV>::w16s;  // sin(pi/8)
        cx<T> a[4][4];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) a[k1][k2] = v[k1 + 4 * k2];
            Bfly<4, INV, T>::run(a[k1]);  // a[k1][m2]
        }
        // twiddles W16^{k1 m2}; W16^1 = c1 - i s1 (fwd)
        const T sg = INV ? T(1) : T(-1);
        auto w1 = [&](cx<T> z) { return cx<T>{z.x * c1 - sg * z.y * s1, z.y * c1 + sg * z.x * s1}; };
        auto w3 = [&](cx<T> z) { return cx<T>{z.x * s1 - sg * z.y * c1, z.y * s1 + sg * z.x * c1}; };
        // W16^2 = W8^1, W16^4 = -i, W16^6 = W8^3, W16^9 = W16^8 * W16^1 = -W16^1
        a[1][1] = w1(a[1][1]);
        a[1][2] = w8_1<INV, T>(a[1][2]);
        a[1][3] = w3(a[1][3]);
        a[2][1] = w8_1<INV, T>(a[2][1]);
        a[2][2] = mul_mi<INV>(a[2][2]);
        a[2][3] = w8_3<INV, T>(a[2][3]);
        a[3][1] = w3(a[3][1]);
        a[3][2] = w8_3<INV, T>(a[3][2]);
        {
            cx<T> z = w1(a[3][3]);
            a[3][3] = cx<T>{-z.x, -z.y};
        }
#pragma unroll
        for (int m2 = 0; m2 < 4; ++m2) {
            cx<T> b[4] = {a[0][m2], a[1][m2], a[2][m2], a[3][m2]};
            Bfly<4, INV, T>::run(b);  // output m1: X[m2 + 4 m1]
#pragma unroll
            for (int m1 = 0; m1 < 4; ++m1) v[m2 + 4 * m1] = b[m1];
        }
    }
};

template <bool INV, typename T> struct Bfly<3, INV, T> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        constexpr T c3 = Rots<T, INV>::w3c;  // cos(2 pi/3) = -1/2
        constexpr T s3 = Rots<T, INV>::w3s;  // sin(2 pi/3)
        cx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        cx<T> a0 = v[0];
        v[0] = cadd(a0, s);
        cx<T> t, sd;
        if constexpr (sizeof(T) == 4) {
            // FP32: c3 = -1/2 exactly, s3 as hi + lo
            t = cx<T>{fmaf(-0.5f, s.x, a0.x), fmaf(-0.5f, s.y, a0.y)};
            sd = cx<T>{fmaf(d.x, kS3Hi, d.x * kS3Lo), fmaf(d.y, kS3Hi, d.y * kS3Lo)};
        } else {
            t = cx<T>{a0.x + c3 * s.x, a0.y + c3 * s.y};
            sd = cscale(d, s3);
        }
        // fwd: y1 = t - i s3 d ; y2 = t + i s3 d
        cx<T> m = mul_mi<INV>(sd);
        v[1] = cadd(t, m);
        v[2] = csub(t, m);
    }
};

// cos and sin of 2 pi q / R, q < R, as the rotation pairs of Rots (FP32:
// magnitude-picked, the same in both directions)
template <typename T, bool INV, int R> struct OddTab {
    T c[R], s[R];
    constexpr OddTab() : c(), s() {
        for (int q = 0; q < R; ++q) {
            const double a = 2.0 * ce::PI * double(q) / double(R);
            if constexpr (sizeof(T) == 4) {
                const ce::Rot r = ce::best_rot(ce::cos_(a), ce::sin_(a));
                c[q] = r.c;
                s[q] = r.s;
            } else {
                c[q] = T(ce::cos_(a));
                s[q] = T(ce::sin_(a));
            }
        }
    }
};

// Generic odd-prime DFT using the (n, R-n) symmetry; constants are compile time.
template <int R, bool INV, typename T> struct BflyOdd {
    static __device__ __forceinline__ void run(cx<T>* v) {
        constexpr OddTab<T, INV, R> tab{};
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
                const T c = tab.c[(n * k) % R];
                const T s = tab.s[(n * k) % R];
                cr += a[n].x * c;
                ci += a[n].y * c;
                sr += b[n].x * s;
                si += b[n].y * s;
            }
            // fwd: y_k = sum a cos - i sum b sin ; y_{R-k} = sum a cos + i sum b sin
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
// Radix sequence for length q: 16s first, then the 2-power remainder (8, 4 or
// 2), then 3, 5, 7, 11, 13.  REV reverses the order (used by the inverse of a
// fused forward/inverse pair so the inverse's first stage reads exactly the
// elements the forward's last stage left in each thread's registers).
struct RadixPlan {
    int n;
    int r[12];
};
constexpr int pow2_part(int q) { int p = 1; while (q % 2 == 0) { q /= 2; p *= 2; } return p; }
constexpr RadixPlan make_plan(int q, bool rev, int maxr = 16) {
    RadixPlan p{0, {0}};
    int p2 = pow2_part(q);
    while (p2 >= maxr) { p.r[p.n++] = maxr; p2 /= maxr; }
    if (p2 > 1) p.r[p.n++] = p2;
    q /= pow2_part(q);
    const int odd[5] = {3, 5, 7, 11, 13};
    for (int k = 0; k < 5; ++k)
        while (q % odd[k] == 0) { p.r[p.n++] = odd[k]; q /= odd[k]; }
    if (q != 1) p.n = -1;  // unsupported prime factor
    if (rev && p.n > 0)
        for (int i = 0; i < p.n / 2; ++i) { int t = p.r[i]; p.r[i] = p.r[p.n - 1 - i]; p.r[p.n - 1 - i] = t; }
    return p;
}
constexpr bool supported_length(int q) { return q >= 2 && make_plan(q, false).n > 0; }
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
// threads per transform: one per maxr elements
constexpr int tpf_for(int q, int maxr = 16) { return q / maxr > 0 ? q / maxr : 1; }
// registers needed per thread: max over stages of (butterflies per thread) * radix
constexpr int max_elems(int q, int tpf, int maxr = 16) {
    RadixPlan p = make_plan(q, false, maxr);
    int m = 0;
    for (int s = 0; s < p.n; ++s) {
        int e = cdiv(q / p.r[s], tpf) * p.r[s];
        if (e > m) m = e;
    }
    return m;
}

// ------------------------------------------------------------ the engine
// LDS accessor concept: struct with  __device__ cx<T>& at(int elem)  for the
// thread's own transform.  tw: LDS (or global) table W_P^k, k in [0,P); the
// length-Q stage twiddle W_Q^e is tw[e * tws] with tws = P / Q.
//
// Element ownership in a stage of radix R: butterfly j = t + b*TPF (b < NB),
// inputs j + r*(Q/R), outputs ((j/NS)*NS*R + j%NS) + r*NS.
template <typename T, int Q, int TPF, bool REV, int MAXR = 16> struct FftShape {
    static constexpr RadixPlan PL = make_plan(Q, REV, MAXR);
    static constexpr int NST = PL.n;
    static constexpr int R0 = PL.r[0];
    static constexpr int RL = PL.r[PL.n - 1];
    static constexpr int E = max_elems(Q, TPF, MAXR);
    static constexpr int NB0 = cdiv(Q / R0, TPF);
    static constexpr int NBL = cdiv(Q / RL, TPF);
    // first-stage input element of slot (b, r), or -1 if the slot is idle
    // (t < TPF always; slots are all busy when TPF divides the butterfly count)
    static __device__ __forceinline__ int in_elem(int t, int b, int r) {
        int j = t + b * TPF;
        if constexpr ((Q / R0) % TPF == 0) return j + r * (Q / R0);
        return (j < Q / R0) ? j + r * (Q / R0) : -1;
    }
    // last-stage output element of slot (b, r), or -1 if idle
    static __device__ __forceinline__ int out_elem(int t, int b, int r) {
        int j = t + b * TPF;
        if constexpr ((Q / RL) % TPF == 0) return j + r * (Q / RL);
        return (j < Q / RL) ? j + r * (Q / RL) : -1;
    }
};

// Synchronization between the stages of one transform.  When the TPF threads
// of every transform lie inside one wavefront (TPF divides 64) and each
// transform has LDS of its own, the exchange involves that wave alone: its
// LDS operations execute in issue order, so a compiler-level barrier at
// wavefront scope is enough and the other waves of the work-group never wait
// for it.  Otherwise the whole work-group synchronizes.
constexpr int kWave = 64;
template <int TPF> constexpr bool wave_local() { return TPF <= kWave && kWave % TPF == 0; }
template <bool WAVE> __device__ __forceinline__ void stage_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Mirror pairs by wavefront exchange.  In a wave-local transform whose slots
// are all busy, element k = t + TPF*b + TPF*NB*r sits at lane t of its row's
// lane group, slot (b, r) (t < TPF, b < NB, r < R).  Its mirror L - k (mod L)
// sits at lane (TPF - t) % TPF, slot (NB-1-b, R-1-r), for t > 0; lane 0 holds
// its own mirrors, at slot (0, R-r) for b = 0 < r and (NB-b, R-1-r) for b > 0
// (slot (0, 0), element 0, is its own mirror).  A lane reads the partner's
// register with ds_bpermute: one instruction per 32-bit word, no LDS storage
// and no barrier.
template <int NB, int R> struct MirrorSlots {
    static constexpr int other(int b, int r) { return (NB - 1 - b) * R + (R - 1 - r); }  // lanes t > 0
    static constexpr int lane0(int b, int r) {                                           // lane t = 0
        return b == 0 ? (r == 0 ? 0 : R - r) : (NB - b) * R + (R - 1 - r);
    }
};
// byte address for ds_bpermute of the mirror lane of lane t in its group
template <int TPF> __device__ __forceinline__ int mirror_lane_addr(int t) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    return (lane - t + (TPF - t) % TPF) << 2;
}
__device__ __forceinline__ float lane_read(float v, int addr) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float2 lane_read(float2 v, int addr) {
    return float2{lane_read(v.x, addr), lane_read(v.y, addr)};
}

template <typename T, int Q, int TPF, bool INV, bool REV, int S, int NS, int MAXR, bool WAVE = false>
struct StageRunner {
    static constexpr RadixPlan PL = make_plan(Q, REV, MAXR);
    static constexpr int NST = PL.n;
    static constexpr int R = PL.r[S];
    static constexpr int NBF = Q / R;
    static constexpr int NB = cdiv(NBF, TPF);

    // butterflies of this stage on registers already holding the stage inputs
    static __device__ __forceinline__ void compute(int t, cx<T>* v, const cx<T>* tw, int tws) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * TPF;
            if (NBF % TPF == 0 || j < NBF) {
                if (NS > 1) {
                    const int k = j % NS;
                    const int step = (Q / (NS * R)) * tws;
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b * R + r] = ctw<INV>(v[b * R + r], tw[r * k * step]);
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
    template <bool FROM_LDS, typename LDS>
    static __device__ __forceinline__ void run(int t, cx<T>* v, LDS& lds, const cx<T>* tw, int tws) {
        if (FROM_LDS) load(t, v, lds);
        compute(t, v, tw, tws);
        if constexpr (S + 1 < NST) {
            stage_sync<WAVE>();  // every thread finished reading this stage's inputs
            store(t, v, lds);
            stage_sync<WAVE>();  // stage outputs visible
            StageRunner<T, Q, TPF, INV, REV, S + 1, NS * R, MAXR, WAVE>::template run<true>(t, v, lds, tw, tws);
        }
    }
};

// Full length-Q transform of one transform per thread group.  On entry v holds
// the first-stage inputs (FftShape<..,REV>::in_elem); on exit v holds the
// last-stage outputs (FftShape<..,REV>::out_elem).  Contains __syncthreads()
// (every thread of the work-group must call it) unless WAVE, where each
// transform's threads and LDS belong to one wavefront (stage_sync).
template <typename T, int Q, int TPF, bool INV, bool REV, int MAXR = 16, bool WAVE = false, typename LDS>
__device__ __forceinline__ void fft_regs(int t, cx<T>* v, LDS& lds, const cx<T>* tw, int tws) {
    StageRunner<T, Q, TPF, INV, REV, 0, 1, MAXR, WAVE>::template run<false>(t, v, lds, tw, tws);
}

// ------------------------------------------------------- buffer memory ops
// All HBM traffic of the passes goes through raw buffer instructions: a
// per-array resource (base, size in bytes) and 32-bit byte offsets.  An offset
// at or beyond the size reads 0 and drops stores, which is how masked lanes
// are expressed (no branches).  Host code keeps every array below kBufLimit.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr unsigned kBufOOB = 0xFFFFFF00u;  // offset that is out of range of every resource

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)(unsigned)bytes, 0x00020000);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Cache-policy operand of a buffer load / store: kCacheDefault, or kStream --
// the non-temporal hint for data a pass reads exactly once or writes for a
// later pass (it keeps streamed arrays from evicting what is reused).
constexpr int kCacheDefault = 0, kStream = 2;

template <typename T> struct BufOps;
// FP32: raw buffer instructions on a per-array resource with 32-bit offsets
// (arrays below kBufLimit); a masked lane passes OOB and reads 0 / drops its store.
template <> struct BufOps<float> {
    using H = rsrc_t;
    using O = unsigned;
    static constexpr O OOB = kBufOOB;
    static __device__ __forceinline__ H make(const void* p, unsigned long bytes) { return make_rsrc(p, bytes); }
    template <int A = kCacheDefault> static __device__ __forceinline__ float2 ld_cx(H r, O off) {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, A));
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void st_cx(float2 v, H r, O off) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, A);
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ float ld_r(H r, O off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, A));
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void st_r(float v, H r, O off) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, A);
    }
    // two adjacent complex elements (16 bytes, 16-byte aligned offset)
    template <int A = kCacheDefault> static __device__ __forceinline__ void ld_cx2(H r, O off, float2& a, float2& b) {
        const float4 q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, A));
        a = float2{q.x, q.y};
        b = float2{q.z, q.w};
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void st_cx2(float2 a, float2 b, H r, O off) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, float4{a.x, a.y, b.x, b.y}), r, off, 0, A);
    }
};
// FP64: global addresses with 64-bit byte offsets, so one array may exceed
// 4 GiB (the FP64 work array of a 1536x768x192 grid is 7.4 GB).  The handle
// carries the array size: like a buffer resource, an access at or beyond it
// (a masked lane passes OOB) is skipped and a load yields 0.
struct GlobalSpan {
    char* p;
    unsigned long bytes;
};
template <> struct BufOps<double> {
    using H = GlobalSpan;
    using O = unsigned long;
    static constexpr O OOB = ~0ul;
    static __device__ __forceinline__ H make(const void* p, unsigned long bytes) {
        return GlobalSpan{(char*)const_cast<void*>(p), bytes};
    }
    static __device__ __forceinline__ bool in(const H& h, O off, unsigned long n) {
        return off < h.bytes && h.bytes - off >= n;
    }
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    template <int A = kCacheDefault> static __device__ __forceinline__ double2 ld_cx(H h, O off) {
        if (!in(h, off, 16)) return double2{0.0, 0.0};
        const f64x2* q = (const f64x2*)(h.p + off);
        if constexpr (A == kStream) return __builtin_bit_cast(double2, __builtin_nontemporal_load(q));
        return __builtin_bit_cast(double2, *q);
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void st_cx(double2 v, H h, O off) {
        if (!in(h, off, 16)) return;
        f64x2* q = (f64x2*)(h.p + off);
        if constexpr (A == kStream) __builtin_nontemporal_store(__builtin_bit_cast(f64x2, v), q);
        else *q = __builtin_bit_cast(f64x2, v);
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ double ld_r(H h, O off) {
        if (!in(h, off, 8)) return 0.0;
        const double* q = (const double*)(h.p + off);
        if constexpr (A == kStream) return __builtin_nontemporal_load(q);
        return *q;
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void st_r(double v, H h, O off) {
        if (!in(h, off, 8)) return;
        double* q = (double*)(h.p + off);
        if constexpr (A == kStream) __builtin_nontemporal_store(v, q);
        else *q = v;
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void ld_cx2(H h, O off, double2& a, double2& b) {
        a = ld_cx<A>(h, off);
        b = ld_cx<A>(h, off == OOB ? OOB : off + 16);
    }
    template <int A = kCacheDefault> static __device__ __forceinline__ void st_cx2(double2 a, double2 b, H h, O off) {
        st_cx<A>(a, h, off);
        st_cx<A>(b, h, off == OOB ? OOB : off + 16);
    }
};

// Copy a twiddle table (n entries) from global memory into LDS; caller syncs.
template <typename V>
__device__ __forceinline__ void stage_table(V* dst, const V* __restrict__ src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}
// Same with a compile-time bound (NMAX entries at most, NT threads): every load
// is issued before the first LDS store, so the table costs one memory latency
// that overlaps whatever loads the caller issued before.
template <int NMAX, int NT, typename V>
__device__ __forceinline__ void stage_table_unrolled(V* dst, const V* __restrict__ src, int n) {
    constexpr int IT = (NMAX + NT - 1) / NT;
    V tmp[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < n) tmp[i] = src[k];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < n) dst[k] = tmp[i];
    }
}

}  // namespace vfd
