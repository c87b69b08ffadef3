// vfd_legacy.hip -- the reference's diffusion C interface (include/vf_diffusion.h)
// on top of the five-pass engine.
//
// Reference: src/Diffusion/convolutionFFT_common.h:44-424, Diffusion.cu:39-1934,
// convolutionFFT.cu:22-581, convolutionFFT_gold.cpp:22-58.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/vf_diffusion.h"
#include "vfd_internal.h"

using namespace vfd;

extern "C" {
int gpu_id[N_CHEM] = {0, 0, 0, 0, 0, 0, 0, 0};
}

namespace {

constexpr int kThreads = 256;
inline unsigned nblocks(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// ---------------------------------------------------------------- 3D pads
// padDataConstantVF3D (convolutionFFT.cuh:590-638)
__global__ void pad_constant_vf3d(float* dst, const float* src, int fD, int fH, int fW, int D, int H, int W,
                                  int epi, float base) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)fD * fH * fW;
    if (i >= n) return;
    const int x = (int)(i % fW);
    const int y = (int)((i / fW) % fH);
    const int z = (int)(i / ((long)fW * fH));
    const bool in = z < D - 1 && z > 0 && y < H - 1 && y > 0 && x < epi && x > 0;
    dst[i] = in ? src[((long)z * H + y) * W + x] : base;
}

// padDataClampToBorder3D (convolutionFFT.cuh:510-585): [0,N) data, [N,N+K) last
// slice, [N+K,P) first slice on each axis.
template <typename T>
__global__ void pad_clamp3d(T* dst, const T* src, int fD, int fH, int fW, int D, int H, int W, int KZ, int KY,
                            int KX) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)fD * fH * fW;
    if (i >= n) return;
    const int x = (int)(i % fW);
    const int y = (int)((i / fW) % fH);
    const int z = (int)(i / ((long)fW * fH));
    const int dz = z < D ? z : (z < D + KZ ? D - 1 : 0);
    const int dy = y < H ? y : (y < H + KY ? H - 1 : 0);
    const int dx = x < W ? x : (x < W + KX ? W - 1 : 0);
    dst[i] = src[((long)dz * H + dy) * W + dx];
}

// padKernel3D (convolutionFFT.cuh:402-445): centre to the origin, wrap negatives.
template <typename TD, typename TS>
__global__ void pad_kernel3d(TD* dst, const TS* src, int fD, int fH, int fW, int kD, int kH, int kW, int KZ,
                             int KY, int KX) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)kD * kH * kW;
    if (i >= n) return;
    const int x = (int)(i % kW);
    const int y = (int)((i / kW) % kH);
    const int z = (int)(i / ((long)kW * kH));
    int kz = z - KZ; if (kz < 0) kz += fD;
    int ky = y - KY; if (ky < 0) ky += fH;
    int kx = x - KX; if (kx < 0) kx += fW;
    dst[((long)kz * fH + ky) * fW + kx] = (TD)src[i];
}

// unpadResult3D (convolutionFFT.cuh:152-171)
template <typename T>
__global__ void unpad3d(T* dst, const T* src, int D, int H, int W, int fH, int fW) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)D * H * W;
    if (i >= n) return;
    const int x = (int)(i % W);
    const int y = (int)((i / W) % H);
    const int z = (int)(i / ((long)W * H));
    dst[i] = src[((long)z * fH + y) * fW + x];
}

// extractCenter3D (convolutionFFT.cuh:77-112)
template <typename TD, typename TS>
__global__ void extract_center3d(TD* dst, const TS* src, int D, int H, int W, int wD, int wH, int wW) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)wD * wH * wW;
    if (i >= n) return;
    const int x = (int)(i % wW);
    const int y = (int)((i / wW) % wH);
    const int z = (int)(i / ((long)wW * wH));
    const int bD = (D - wD) / 2, bH = (H - wH) / 2, bW = (W - wW) / 2;
    dst[i] = (TD)src[((long)(z + bD) * H + (y + bH)) * W + (x + bW)];
}

// modulateAndNormalize_kernel / mulAndScale (convolutionFFT.cuh:690-716):
// dst[i] = c * (src[i] * dst[i])
__global__ void modulate_scale(float2* dst, const float2* src, long n, float c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = src[i], b = dst[i];
    dst[i] = float2{c * (a.x * b.x - a.y * b.y), c * (a.y * b.x + a.x * b.y)};
}

// complexPower_kernel / cpow (convolutionFFT.cuh:645-684): |z|^n cis(n atan(b/a)).
__device__ __forceinline__ void pow_quirk(double a, double b, int n, int quirk, double* re, double* im) {
    const double ab2 = a * a + b * b;
    if (ab2 == 0.0) {  // 0^n (the reference's atan(0/0) yields NaN here)
        *re = n == 0 ? 1.0 : 0.0;
        *im = 0.0;
        return;
    }
    const double ph = quirk ? atan(b / a) : atan2(b, a);
    const double mag = pow(ab2, 0.5 * n);
    *re = mag * cos(n * ph);
    *im = mag * sin(n * ph);
}

__global__ void complex_power(float2* d, long n, int p) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // FP32 in, FP32 out, evaluated like cpow: powf(|z|^2, n/2) cis(n atanf(b/a))
    const float a = d[i].x, b = d[i].y;
    const float ab2 = a * a + b * b;
    const float nf = (float)p;
    const float ang = nf * atanf(b / a);
    const float mag = powf(ab2, nf / 2.0f);
    d[i] = float2{mag * cosf(ang), mag * sinf(ang)};
}

// Stencil spectrum power for the kernel build: data_spec *= P(k) * scale with
// P = pow(S^(k), n) and S^ the DFT of the 7-point stencil wrapped to the origin,
// S^ = 1 - 6 lambda - gamma dt + 2 lambda (cos tz + cos ty + cos tx).
__global__ void stencil_power_multiply(double2* spec, int fD, int fH, int fW, int hx, double lambda,
                                       double c0, int n, int quirk, double scale) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)fD * fH * hx;
    if (i >= total) return;
    const int kx = (int)(i % hx);
    const int ky = (int)((i / hx) % fH);
    const int kz = (int)(i / ((long)hx * fH));
    const double tp = 6.283185307179586476925286766559005768;
    const double s = c0 + 2.0 * lambda * (cos(tp * kz / fD) + cos(tp * ky / fH) + cos(tp * kx / fW));
    double pr, pi;
    pow_quirk(s, 0.0, n, quirk, &pr, &pi);
    const double2 v = spec[i];
    spec[i] = double2{scale * (v.x * pr - v.y * pi), scale * (v.x * pi + v.y * pr)};
}

// 2D kernel build: x S^n * scale for the 5-point stencil centred at (KY, KX)
// of the padded kernel, on every kz plane of a z-constant spectrum:
// S(ty, tx) = exp(-i (ty (1 - KY) + tx (1 - KX))) (c0 + 2 lambda (cos ty + cos tx))
__global__ void stencil2d_power_multiply(double2* spec, int fD, int fH, int fW, int hx, double lambda, double c0,
                                         int n, int KY, int KX, double scale) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)fD * fH * hx;
    if (i >= total) return;
    const int kx = (int)(i % hx);
    const int ky = (int)((i / hx) % fH);
    const double tp = 6.283185307179586476925286766559005768;
    const double ty = tp * ky / fH, tx = tp * kx / fW;
    const double mag = pow(c0 + 2.0 * lambda * (cos(ty) + cos(tx)), (double)n);  // signed for odd n
    double sp, cp;
    sincos(-(double)n * (ty * (1 - KY) + tx * (1 - KX)), &sp, &cp);
    const double pr = mag * cp, pi = mag * sp;
    const double2 v = spec[i];
    spec[i] = double2{scale * (v.x * pr - v.y * pi), scale * (v.x * pi + v.y * pr)};
}

// The 2D path's kernel spectrum is the C2C transform Z of the padded kernel p
// (H x W real) read as H x W/2 complex z[h][m] = p[h][2m] + i p[h][2m+1]
// (Diffusion.cu:677-771, spProcess2D at convolutionFFT.cuh:948-1021).  With
// M = W/2, E and O the transforms of the even and odd columns and P the R2C
// spectrum of p:  Z = E + i O,  E(a,b) = (P(a,b) + P(a,b+M)) / 2,
// O(a,b) = W_W^-b (P(a,b) - P(a,b+M)) / 2,  P(a,b+M) = conj P(-a, M-b) (b > 0);
// and back, P(a,k) = E(a,k mod M) + W_W^k O(a,k mod M) with
// E = (Z(a,b) + conj Z(-a,-b)) / 2,  O = -i (Z(a,b) - conj Z(-a,-b)) / 2.
__global__ void standard_to_packed2d(float2* Z, const double2* P, int H, int W, int hx, double s) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int M = W / 2;
    if (i >= (long)H * M) return;
    const int a = (int)(i / M), b = (int)(i % M);
    const double2 pa = P[(long)a * hx + b];
    double2 pb = P[(long)a * hx + M];
    if (b > 0) {
        const double2 q = P[(long)((H - a) % H) * hx + (M - b)];
        pb = double2{q.x, -q.y};
    }
    const double tp = 6.283185307179586476925286766559005768;
    double sw, cw;
    sincos(tp * b / W, &sw, &cw);  // W_W^-b = exp(+2 pi i b / W)
    const double ex = 0.5 * (pa.x + pb.x), ey = 0.5 * (pa.y + pb.y);
    const double dx = 0.5 * (pa.x - pb.x), dy = 0.5 * (pa.y - pb.y);
    const double ox = dx * cw - dy * sw, oy = dx * sw + dy * cw;
    Z[i] = float2{(float)(s * (ex - oy)), (float)(s * (ey + ox))};  // E + i O
}

__global__ void packed_to_standard2d(float2* P, const float2* Z, int fD, int H, int W) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int M = W / 2, hx = M + 1;
    if (i >= (long)H * hx) return;
    const int a = (int)(i / hx), k = (int)(i % hx);
    const int b = k % M;
    const float2 z1 = Z[(long)a * M + b];
    const float2 z2 = Z[(long)((H - a) % H) * M + (M - b) % M];
    const double ex = 0.5 * ((double)z1.x + z2.x), ey = 0.5 * ((double)z1.y - z2.y);  // (z1 + conj z2) / 2
    const double ox = 0.5 * ((double)z1.y + z2.y), oy = -0.5 * ((double)z1.x - z2.x); // -i (z1 - conj z2) / 2
    const double tp = 6.283185307179586476925286766559005768;
    double sw, cw;
    sincos(-tp * k / W, &sw, &cw);  // W_W^k
    const float2 v{(float)(ex + ox * cw - oy * sw), (float)(ey + ox * sw + oy * cw)};
    for (int z = 0; z < fD; ++z) P[(long)z * H * hx + i] = v;
}

__global__ void to_fcomplex(float2* dst, const double2* src, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = float2{(float)src[i].x, (float)src[i].y};
}

template <typename T>
__global__ void fill(T* d, long n, T v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = v;
}

// ---------------------------------------------------------------- 2D pads
// padKernel_kernel (convolutionFFT.cuh:52-75)
__global__ void pad_kernel2d(float* dst, const float* src, int fH, int fW, int kH, int kW, int KY, int KX) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)kH * kW) return;
    const int x = (int)(i % kW), y = (int)(i / kW);
    int ky = y - KY; if (ky < 0) ky += fH;
    int kx = x - KX; if (kx < 0) kx += fW;
    dst[(long)ky * fW + kx] = src[i];
}

// padDataConstantVF_kernel (convolutionFFT.cuh:179-218); like the 3D pad, an
// epithelium index beyond the row is bounded at W (the reference reads the
// next rows there)
__global__ void pad_constant_vf2d(float* dst, const float* src, int fH, int fW, int H, int W, int epi,
                                  float base) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)fH * fW) return;
    const int x = (int)(i % fW), y = (int)(i / fW);
    const bool in = y < H - 1 && y > 0 && x < epi && x < W && x > 0;
    dst[i] = in ? src[(long)y * W + x] : base;
}

// padDataRightWall (convolutionFFT.cuh:261-336), restated from its case table.
// Each axis of the FFT box splits into the data [0, N), a band [N, N + K) of
// the kernel centre's width, and the rest.  With the mirrors mx = 2W - 1 - x,
// my = 2H - 1 - y and the far-side mirror fy = fH - 1 - y, the source is
//                    x < W      W <= x < W + KX    x >= W + KX
//   y < H            (y, x)     (y, mx)            c
//   H <= y < H + KY  c          (my, mx)           c
//   y >= H + KY      c          (fy, mx)           c
// and c = 0.0001 also where a mirror falls outside the data: below 0 for a
// band wider than the data, fy >= H for a box taller than 2H + KY (there the
// reference reads outside the field).
__global__ void pad_right_wall2d(float* dst, const float* src, int fH, int fW, int H, int W, int KY, int KX) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)fH * fW) return;
    const int x = (int)(i % fW), y = (int)(i / fW);
    const int row_region = y < H ? 0 : (y < H + KY ? 1 : 2);
    const int col_region = x < W ? 0 : (x < W + KX ? 1 : 2);
    int sy = -1, sx = -1;
    if (col_region == 0 && row_region == 0) {
        sy = y;
        sx = x;
    } else if (col_region == 1) {
        sx = 2 * W - 1 - x;
        sy = row_region == 0 ? y : (row_region == 1 ? 2 * H - 1 - y : fH - 1 - y);
    }
    dst[i] = (sx < 0 || sy < 0 || sy >= H) ? 0.0001f : src[(long)sy * W + sx];
}

// padDataMirror_kernel (convolutionFFT.cuh:315-363)
__global__ void pad_mirror2d(float* dst, const float* src, int fH, int fW, int H, int W, int KY, int KX) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)fH * fW) return;
    const int x = (int)(i % fW), y = (int)(i / fW);
    const int bH = H + KY, bW = W + KX;
    int dy = 0, dx = 0;
    if (y < H) dy = y;
    if (x < W) dx = x;
    if (y >= H && y < bH) dy = H - (y - H) - 1;
    if (x >= W && x < bW) dx = W - (x - W) - 1;
    if (y >= bH) dy = fH - y - 1;
    if (x >= bW) dx = fW - x - 1;
    // a mirror that falls outside the data (a band wider than the data, a box
    // beyond 2N + K: the reference reads outside the field there) takes the
    // nearest data index
    dy = dy < 0 ? 0 : (dy >= H ? H - 1 : dy);
    dx = dx < 0 ? 0 : (dx >= W ? W - 1 : dx);
    dst[i] = src[(long)dy * W + dx];
}

// ------------------------------------------------------------------ state
std::mutex g_mu;  // serializes fftDiffuse3D / kernel precompute (shared workspace)
// a vfdTickChemsBegin whose diffusions may still run on the shared workspace
// (guarded by g_mu); the other diffusion entry points refuse to run meanwhile
bool g_split_pending = false;
int refuse_if_split_pending(const char* who) {
    return g_split_pending ? set_error(VFD_ERR_INVALID, "%s: a vfdTickChemsBegin is pending; call vfdTickChemsEnd first",
                                       who)
                           : VFD_OK;
}
int g_quirk = 1;

struct PlanRec {
    int nz, ny, nx, type;
};
std::map<int, PlanRec> g_plans;
int g_next_plan = 1;

struct WsKey {
    int dev, D, H, W, FD, FH, FW;
    bool operator<(const WsKey& o) const {
        return std::tie(dev, D, H, W, FD, FH, FW) < std::tie(o.dev, o.D, o.H, o.W, o.FD, o.FH, o.FW);
    }
};
struct Ws {
    Plan plan;
    void* work = nullptr;
};
std::map<WsKey, Ws> g_ws;

bool fail(int rc) {
    (void)rc;
    fprintf(stderr, "vfd: %s\n", vfd_last_error());
    return false;
}

bool hip_ok(hipError_t e, const char* what) { return hip_check(e, what) == VFD_OK || fail(-1); }

// The persistent workspace of a (device, data box, FFT box): the plan and one
// work array, made on first use.  The plan's schedule is autotuned then (about
// a second, once per box and process) unless VFD_SCHEDULE fixes it or
// VFD_NO_AUTOTUNE is set (as for the engine's shapes).
int ws_for(int dev, const c_ctx& cctx, Ws** out) {
    const WsKey key{dev, cctx.DD, cctx.DH, cctx.DW, cctx.FFTD, cctx.FFTH, cctx.FFTW};
    auto it = g_ws.find(key);
    if (it == g_ws.end()) {
        vfd_plan_desc d{cctx.DD, cctx.DH, cctx.DW, cctx.FFTD, cctx.FFTH, cctx.FFTW, VFD_F32, VFD_PAD_CONSTANT_VF,
                        VFD_KSPEC_COMPLEX_STD};
        Ws ws;
        if (int rc = plan_init(&d, &ws.plan)) return rc;
        if (int rc = hip_check(hipMalloc(&ws.work, ws.plan.work_bytes), "legacy workspace")) return rc;
        if (!getenv("VFD_SCHEDULE") && !getenv("VFD_NO_AUTOTUNE")) {
            // timed on a stream of its own, away from the caller's work
            hipStream_t ts = nullptr;
            int rc = hip_check(hipStreamCreateWithFlags(&ts, hipStreamNonBlocking), "autotune stream");
            if (!rc) rc = plan_autotune(ws.plan, 1, ts, nullptr);
            if (ts) (void)hipStreamDestroy(ts);
            if (rc) {
                (void)hipFree(ws.work);
                return rc;
            }
        }
        if (getenv("VFD_VERBOSE"))
            fprintf(stderr, "vfd: workspace %dx%dx%d in %dx%dx%d, schedule %d,%d,%d,%d,%d\n", cctx.DW, cctx.DH, cctx.DD,
                    cctx.FFTW, cctx.FFTH, cctx.FFTD, ws.plan.var[0], ws.plan.var[1], ws.plan.var[2], ws.plan.var[3],
                    ws.plan.var[4]);
        it = g_ws.emplace(key, ws).first;
    }
    *out = &it->second;
    return VFD_OK;
}

}  // namespace

// ====================================================================== C-ABI
extern "C" {

int snapTransformSize(int dataSize) {
    // Diffusion.cu:39-70
    if (dataSize % 16) dataSize = dataSize - dataSize % 16 + 16;
    int hiBit;
    for (hiBit = 31; hiBit >= 0; hiBit--)
        if ((unsigned)dataSize & (1U << hiBit)) break;
    const unsigned lowPOT = 1U << hiBit;
    if (lowPOT == (unsigned)dataSize) return dataSize;
    const unsigned hiPOT = 1U << (hiBit + 1);
    if (hiPOT <= 1024) return (int)hiPOT;
    return (int)pow(2.0, ceil(log((double)dataSize) / log(2.0)));
}

void vfdSetKernelPowerQuirk(int on) { g_quirk = on ? 1 : 0; }

int vfdPlan3d(cufftHandle* plan, int nz, int ny, int nx, int type) {
    if (!plan || nz < 1 || ny < 1 || nx < 1) return set_error(VFD_ERR_INVALID, "vfdPlan3d: bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    const int id = g_next_plan++;
    g_plans[id] = PlanRec{nz, ny, nx, type};
    *plan = id;
    return VFD_OK;
}

int vfdDestroy(cufftHandle plan) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_plans.erase(plan) ? VFD_OK : set_error(VFD_ERR_NOTFOUND, "vfdDestroy: unknown plan %d", plan);
}

// memory helpers behind the diffusion_helper layer (DiffusionHelper.cpp)
int vfdDeviceCount(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int vfdSetDevice(int dev) { return hip_check(hipSetDevice(dev), "vfdSetDevice"); }

void* vfdMalloc(size_t bytes) {
    void* p = nullptr;
    return hip_check(hipMalloc(&p, bytes), "vfdMalloc") == VFD_OK ? p : nullptr;
}

void vfdFree(void* p) {
    if (p) (void)hip_check(hipFree(p), "vfdFree");
}

void* vfdHostAlloc(size_t bytes) {
    void* p = nullptr;
    return hip_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "vfdHostAlloc") == VFD_OK ? p : nullptr;
}

void vfdHostFree(void* p) {
    if (p) (void)hip_check(hipHostFree(p), "vfdHostFree");
}

int vfdMemcpy(void* dst, const void* src, size_t bytes) {
    if ((!dst || !src) && bytes) return set_error(VFD_ERR_INVALID, "vfdMemcpy: null pointer");
    return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "vfdMemcpy");
}

// H2D / D2H (Diffusion.cu:72-122; under M40 the spectrum stays on the device)
void H2D(int ic, c_ctx* cc, int np) {
    if (!cc || ic < 0 || ic >= N_CHEM || np < 0) {
        set_error(VFD_ERR_INVALID, "H2D: bad arguments");
        fail(-1);
        return;
    }
    if (!hip_ok(hipSetDevice(cc->dev_id[ic]), "H2D hipSetDevice")) return;
    hip_ok(hipMemcpy(cc->d_data[ic], cc->h_ibuffs[ic], (size_t)np * sizeof(float), hipMemcpyHostToDevice), "H2D");
}

void D2H(int ic, c_ctx* cc, int np) {
    if (!cc || ic < 0 || ic >= N_CHEM || np < 0) {
        set_error(VFD_ERR_INVALID, "D2H: bad arguments");
        fail(-1);
        return;
    }
    if (!hip_ok(hipSetDevice(cc->dev_id[ic]), "D2H hipSetDevice")) return;
    hip_ok(hipMemcpy(cc->h_obuffs[ic], cc->d_data[ic], (size_t)np * sizeof(float), hipMemcpyDeviceToHost), "D2H");
}

void reportMemUsageGPU(void) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        printf("GPU memory usage: unavailable\n");
        return;
    }
    const double gb = 1024.0 * 1024.0 * 1024.0;
    printf("GPU memory usage: used = %f GB, free = %f GB, total = %f GB\n\n", (double)(tot - fr) / gb,
           (double)fr / gb, (double)tot / gb);
}

// --------------------------------------------------------- the hot path
bool fftDiffuse3D(float* d_Data, fComplex* d_KernelSpectrum, cufftHandle fftPlanFwd, cufftHandle fftPlanInv,
                  c_ctx cctx, int epiBoundary, float baseChem, int chemIndex, int iter) {
    (void)chemIndex;
    (void)iter;
    if (!d_Data || !d_KernelSpectrum) return fail(set_error(VFD_ERR_INVALID, "fftDiffuse3D: null buffer"));
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = refuse_if_split_pending("fftDiffuse3D")) return fail(rc);
    for (cufftHandle h : {fftPlanFwd, fftPlanInv}) {
        if (h == 0) continue;  // 0: no plan supplied, the box comes from cctx
        auto it = g_plans.find(h);
        if (it == g_plans.end()) return fail(set_error(VFD_ERR_NOTFOUND, "fftDiffuse3D: unknown plan %d", h));
        const PlanRec& r = it->second;
        if (r.nz != cctx.FFTD || r.ny != cctx.FFTH || r.nx != cctx.FFTW)
            return fail(set_error(VFD_ERR_INVALID, "fftDiffuse3D: plan %d box %dx%dx%d != cctx FFT box %dx%dx%d", h,
                                  r.nz, r.ny, r.nx, cctx.FFTD, cctx.FFTH, cctx.FFTW));
    }
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    Ws* ws = nullptr;
    if (int rc = ws_for(dev, cctx, &ws)) return fail(rc);
    vfd_field f{d_Data, d_Data, ws->work, d_KernelSpectrum, (double)baseChem, epiBoundary, 0};
    if (int rc = plan_diffuse(ws->plan, &f, 1, nullptr)) return fail(rc);
    return hip_ok(hipStreamSynchronize(nullptr), "fftDiffuse3D");
}

// Batched, pipelined form of the per-chem loop of Diffuser::diffuseChemGPU
// (Diffuser.cpp:300-421, which runs H2D -> fftDiffuse3D -> D2H for one chem at
// a time with a device synchronization after every step).  Here chem ic's copy
// in runs on one stream, its diffusion (the fftDiffuse3D operator) on a second
// and its copy out on a third, ordered by events, so the copies of one chem
// overlap the diffusion of its neighbours.  A device buffer d_data[ic] shared
// by several chems is reused only after the previous chem's copy out.
namespace {
struct Pipe {
    hipStream_t in = nullptr, comp = nullptr, out = nullptr;
    hipEvent_t e_in[N_CHEM] = {}, e_comp[N_CHEM] = {}, e_out[N_CHEM] = {};
    hipEvent_t e_aux[2] = {};
};
std::map<int, Pipe> g_pipes;

int pipe_for(int dev, Pipe** out) {
    auto it = g_pipes.find(dev);
    if (it != g_pipes.end()) {
        *out = &it->second;
        return VFD_OK;
    }
    Pipe p;
    for (hipStream_t* s : {&p.in, &p.comp, &p.out})
        if (int rc = hip_check(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "chem pipeline stream")) return rc;
    for (int i = 0; i < N_CHEM; ++i)
        for (hipEvent_t* e : {&p.e_in[i], &p.e_comp[i], &p.e_out[i]})
            if (int rc = hip_check(hipEventCreateWithFlags(e, hipEventDisableTiming), "chem pipeline event")) return rc;
    for (hipEvent_t& e : p.e_aux)
        if (int rc = hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "chem pipeline event")) return rc;
    *out = &g_pipes.emplace(dev, p).first->second;
    return VFD_OK;
}
}  // namespace

extern "C" bool vfdDiffuseChems(c_ctx cctx, int nchem, int epiBoundary, const float* baseChem) {
    if (nchem < 0 || nchem > N_CHEM || (nchem > 0 && !baseChem))
        return fail(set_error(VFD_ERR_INVALID, "vfdDiffuseChems: nchem must be 0..%d with baseChem given", N_CHEM));
    if (cctx.DD < 1 || cctx.DH < 1 || cctx.DW < 1)
        return fail(set_error(VFD_ERR_INVALID, "vfdDiffuseChems: empty data box"));
    for (int ic = 0; ic < nchem; ++ic)
        if (!cctx.d_data[ic] || !cctx.d_kernelspectrum_h[ic] || !cctx.h_ibuffs[ic] || !cctx.h_obuffs[ic])
            return fail(set_error(VFD_ERR_INVALID, "vfdDiffuseChems: chem %d has a null buffer", ic));
    if (nchem == 0) return true;
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = refuse_if_split_pending("vfdDiffuseChems")) return fail(rc);
    if (!hip_ok(hipSetDevice(cctx.dev_id[0]), "hipSetDevice")) return false;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    Ws* ws = nullptr;
    if (int rc = ws_for(dev, cctx, &ws)) return fail(rc);
    Pipe* pp = nullptr;
    if (int rc = pipe_for(dev, &pp)) return fail(rc);
    const size_t bytes = (size_t)cctx.DD * cctx.DH * cctx.DW * sizeof(float);
    int rc = VFD_OK;
    for (int ic = 0; ic < nchem && !rc; ++ic) {
        float* d = cctx.d_data[ic];
        for (int jc = ic - 1; jc >= 0 && !rc; --jc)
            if (cctx.d_data[jc] == d) {  // reuse only after that chem's copy out
                rc = hip_check(hipStreamWaitEvent(pp->in, pp->e_out[jc], 0), "pipeline wait");
                break;
            }
        if (!rc) rc = hip_check(hipMemcpyAsync(d, cctx.h_ibuffs[ic], bytes, hipMemcpyHostToDevice, pp->in), "H2D");
        if (!rc) rc = hip_check(hipEventRecord(pp->e_in[ic], pp->in), "pipeline event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(pp->comp, pp->e_in[ic], 0), "pipeline wait");
        if (!rc) {
            vfd_field f{d, d, ws->work, cctx.d_kernelspectrum_h[ic], (double)baseChem[ic], epiBoundary, 0};
            rc = plan_diffuse(ws->plan, &f, 1, pp->comp);
        }
        if (!rc) rc = hip_check(hipEventRecord(pp->e_comp[ic], pp->comp), "pipeline event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(pp->out, pp->e_comp[ic], 0), "pipeline wait");
        if (!rc) rc = hip_check(hipMemcpyAsync(cctx.h_obuffs[ic], d, bytes, hipMemcpyDeviceToHost, pp->out), "D2H");
        if (!rc) rc = hip_check(hipEventRecord(pp->e_out[ic], pp->out), "pipeline event");
    }
    for (hipStream_t s : {pp->in, pp->comp, pp->out}) {
        const int src = hip_check(hipStreamSynchronize(s), "pipeline sync");
        if (!rc) rc = src;
    }
    return rc ? fail(rc) : true;
}

namespace {
// device arrays of the resident tick, per workspace: the uploaded deltas, the
// two ECM fragment fields and the three gradients
struct TickBufs {
    float* delta[N_CHEM] = {};
    float* frag[2] = {};
    float* grad[3] = {};
};
std::map<WsKey, TickBufs> g_tick;

int tick_bufs(const WsKey& key, size_t bytes, TickBufs** out) {
    auto it = g_tick.find(key);
    if (it == g_tick.end()) {
        TickBufs b;
        std::vector<float**> all;
        for (float*& p : b.delta) all.push_back(&p);
        for (float*& p : b.frag) all.push_back(&p);
        for (float*& p : b.grad) all.push_back(&p);
        int rc = VFD_OK;
        for (float** p : all)
            if (!rc) rc = hip_check(hipMalloc((void**)p, bytes), "tick buffers");
        if (rc) {
            for (float** p : all)
                if (*p) (void)hipFree(*p);
            return rc;
        }
        it = g_tick.emplace(key, b).first;
    }
    *out = &it->second;
    return VFD_OK;
}
}  // namespace

namespace {
// arguments of the resident tick: the diffusion half (base) and the update
// half (deltas, fragments, gradients); a null group is not checked
int tick_check(const char* who, const c_ctx& cctx, const float* baseChem, bool diffuse, float* const* h_dchem,
               const float* decay, const float* h_frag_col, const float* h_frag_ela, float* const* h_grads,
               bool update) {
    if ((diffuse && !baseChem) || (update && (!h_dchem || !decay || !h_frag_col || !h_frag_ela || !h_grads)))
        return set_error(VFD_ERR_INVALID, "%s: null argument", who);
    if (cctx.DD < 1 || cctx.DH < 1 || cctx.DW < 1) return set_error(VFD_ERR_INVALID, "%s: empty data box", who);
    for (int ic = 0; ic < N_CHEM; ++ic) {
        if (!cctx.d_data[ic] || (diffuse && !cctx.d_kernelspectrum_h[ic]) ||
            (update && (!cctx.h_ibuffs[ic] || !h_dchem[ic])))
            return set_error(VFD_ERR_INVALID, "%s: chem %d has a null buffer", who, ic);
        for (int jc = 0; jc < ic; ++jc)
            if (cctx.d_data[jc] == cctx.d_data[ic])
                return set_error(VFD_ERR_INVALID,
                                 "%s: chems %d and %d share d_data (levels are resident, one buffer per chem)", who, jc,
                                 ic);
    }
    if (update)
        for (int g = 0; g < 3; ++g)
            if (!h_grads[g]) return set_error(VFD_ERR_INVALID, "%s: gradient %d is null", who, g);
    return VFD_OK;
}

// the host deltas are cleared once their copy has landed, while the device works
static std::vector<std::thread> clear_host_deltas(Pipe* pp, float* const* h_dchem, size_t bytes) {
    std::vector<std::thread> clear;
    for (int ic = 0; ic < N_CHEM; ++ic)
        clear.emplace_back([pp, ic, h_dchem, bytes] {
            (void)hipEventSynchronize(pp->e_in[ic]);
            std::memset(h_dchem[ic], 0, bytes);
        });
    return clear;
}

// levels d_data -> h_ibuffs and the three gradients -> h_grads on the out
// stream, after the compute stream's work so far
int download_levels_grads(const c_ctx& cctx, Pipe* pp, TickBufs* tb, float* const* h_grads, size_t bytes,
                          bool levels) {
    int rc = hip_check(hipEventRecord(pp->e_aux[1], pp->comp), "pipeline event");
    if (!rc) rc = hip_check(hipStreamWaitEvent(pp->out, pp->e_aux[1], 0), "pipeline wait");
    for (int ic = 0; ic < N_CHEM && levels && !rc; ++ic)
        rc = hip_check(hipMemcpyAsync(cctx.h_ibuffs[ic], cctx.d_data[ic], bytes, hipMemcpyDeviceToHost, pp->out),
                       "D2H levels");
    for (int g = 0; g < 3 && !rc; ++g)
        rc = hip_check(hipMemcpyAsync(h_grads[g], tb->grad[g], bytes, hipMemcpyDeviceToHost, pp->out), "D2H gradient");
    return rc;
}

int sync_pipe(Pipe* pp, int rc) {
    for (hipStream_t s : {pp->in, pp->comp, pp->out}) {
        const int src = hip_check(hipStreamSynchronize(s), "pipeline sync");
        if (!rc) rc = src;
    }
    return rc;
}
}  // namespace

extern "C" bool vfdTickChems(c_ctx cctx, int epiBoundary, const float* baseChem, float* const* h_dchem,
                             const float* decay, const float* h_frag_col, const float* h_frag_ela,
                             float* const* h_grads, double* totals) {
    if (int rc = tick_check("vfdTickChems", cctx, baseChem, true, h_dchem, decay, h_frag_col, h_frag_ela, h_grads,
                            true))
        return fail(rc);
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = refuse_if_split_pending("vfdTickChems")) return fail(rc);
    if (!hip_ok(hipSetDevice(cctx.dev_id[0]), "hipSetDevice")) return false;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    Ws* ws = nullptr;
    if (int rc = ws_for(dev, cctx, &ws)) return fail(rc);
    Pipe* pp = nullptr;
    if (int rc = pipe_for(dev, &pp)) return fail(rc);
    const long n = (long)cctx.DD * cctx.DH * cctx.DW;
    const size_t bytes = (size_t)n * sizeof(float);
    TickBufs* tb = nullptr;
    if (int rc = tick_bufs(WsKey{dev, cctx.DD, cctx.DH, cctx.DW, cctx.FFTD, cctx.FFTH, cctx.FFTW}, bytes, &tb))
        return fail(rc);
    int rc = VFD_OK;
    for (int ic = 0; ic < N_CHEM && !rc; ++ic) {
        float* d = cctx.d_data[ic];
        rc = hip_check(hipMemcpyAsync(tb->delta[ic], h_dchem[ic], bytes, hipMemcpyHostToDevice, pp->in), "H2D delta");
        if (!rc) rc = hip_check(hipEventRecord(pp->e_in[ic], pp->in), "pipeline event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(pp->comp, pp->e_in[ic], 0), "pipeline wait");
        if (!rc) {
            vfd_field f{d, d, ws->work, cctx.d_kernelspectrum_h[ic], (double)baseChem[ic], epiBoundary, 0};
            Epilogue ep;
            ep.deltas = &tb->delta[ic];
            ep.decay = &decay[ic];
            rc = plan_diffuse_passes(ws->plan, &f, 1, 0, 5, pp->comp, &ep);
        }
        if (!rc) rc = hip_check(hipEventRecord(pp->e_comp[ic], pp->comp), "pipeline event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(pp->out, pp->e_comp[ic], 0), "pipeline wait");
        if (!rc) rc = hip_check(hipMemcpyAsync(cctx.h_ibuffs[ic], d, bytes, hipMemcpyDeviceToHost, pp->out), "D2H levels");
        if (!rc) rc = hip_check(hipEventRecord(pp->e_out[ic], pp->out), "pipeline event");
    }
    if (!rc) rc = hip_check(hipMemcpyAsync(tb->frag[0], h_frag_col, bytes, hipMemcpyHostToDevice, pp->in), "H2D frag");
    if (!rc) rc = hip_check(hipMemcpyAsync(tb->frag[1], h_frag_ela, bytes, hipMemcpyHostToDevice, pp->in), "H2D frag");
    if (!rc) rc = hip_check(hipEventRecord(pp->e_aux[0], pp->in), "pipeline event");
    if (!rc) rc = hip_check(hipStreamWaitEvent(pp->comp, pp->e_aux[0], 0), "pipeline wait");
    std::vector<std::thread> clear;
    if (!rc) clear = clear_host_deltas(pp, h_dchem, bytes);
    if (!rc) {
        float* levels[N_CHEM];
        for (int ic = 0; ic < N_CHEM; ++ic) levels[ic] = cctx.d_data[ic];
        rc = chem_grads_totals(levels, tb->frag[0], tb->frag[1], tb->grad, n, totals, pp->comp);
    }
    if (!rc) rc = download_levels_grads(cctx, pp, tb, h_grads, bytes, false);
    for (std::thread& t : clear) t.join();
    rc = sync_pipe(pp, rc);
    return rc ? fail(rc) : true;
}

// The resident tick split at the agents' step (the go() split of
// woundHealingWorld.cpp:2013-2055): Begin queues the diffusion of every chem's
// resident levels in place (t, the reference's tChem) and returns at once, so
// it runs while the agents work on the host; End, after the agents, brings the
// deltas and fragments in, forms p = d + t * decay, the gradients and totals,
// and copies levels and gradients out.  Same numbers as vfdTickChems.
extern "C" bool vfdTickChemsBegin(c_ctx cctx, int epiBoundary, const float* baseChem) {
    if (int rc = tick_check("vfdTickChemsBegin", cctx, baseChem, true, nullptr, nullptr, nullptr, nullptr, nullptr,
                            false))
        return fail(rc);
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = refuse_if_split_pending("vfdTickChemsBegin")) return fail(rc);
    if (!hip_ok(hipSetDevice(cctx.dev_id[0]), "hipSetDevice")) return false;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    Ws* ws = nullptr;
    if (int rc = ws_for(dev, cctx, &ws)) return fail(rc);
    Pipe* pp = nullptr;
    if (int rc = pipe_for(dev, &pp)) return fail(rc);
    int rc = VFD_OK;
    for (int ic = 0; ic < N_CHEM && !rc; ++ic) {
        vfd_field f{cctx.d_data[ic], cctx.d_data[ic], ws->work, cctx.d_kernelspectrum_h[ic], (double)baseChem[ic],
                    epiBoundary, 0};
        rc = plan_diffuse(ws->plan, &f, 1, pp->comp);
    }
    if (rc) {
        (void)sync_pipe(pp, rc);
        return fail(rc);
    }
    g_split_pending = true;
    return true;
}

extern "C" bool vfdTickChemsEnd(c_ctx cctx, float* const* h_dchem, const float* decay, const float* h_frag_col,
                                const float* h_frag_ela, float* const* h_grads, double* totals) {
    if (int rc = tick_check("vfdTickChemsEnd", cctx, nullptr, false, h_dchem, decay, h_frag_col, h_frag_ela, h_grads,
                            true))
        return fail(rc);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_split_pending) return fail(set_error(VFD_ERR_INVALID, "vfdTickChemsEnd: no vfdTickChemsBegin is pending"));
    // End may run on another host thread than Begin did
    if (!hip_ok(hipSetDevice(cctx.dev_id[0]), "hipSetDevice")) return false;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    Pipe* pp = nullptr;
    if (int rc = pipe_for(dev, &pp)) return fail(rc);
    const long n = (long)cctx.DD * cctx.DH * cctx.DW;
    const size_t bytes = (size_t)n * sizeof(float);
    TickBufs* tb = nullptr;
    int rc = tick_bufs(WsKey{dev, cctx.DD, cctx.DH, cctx.DW, cctx.FFTD, cctx.FFTH, cctx.FFTW}, bytes, &tb);
    // per chem, pipelined: delta in, p = d + t * decay in place over the
    // diffused levels (d cleared), levels out
    for (int ic = 0; ic < N_CHEM && !rc; ++ic) {
        rc = hip_check(hipMemcpyAsync(tb->delta[ic], h_dchem[ic], bytes, hipMemcpyHostToDevice, pp->in), "H2D delta");
        if (!rc) rc = hip_check(hipEventRecord(pp->e_in[ic], pp->in), "pipeline event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(pp->comp, pp->e_in[ic], 0), "pipeline wait");
        if (!rc) rc = chem_levels(cctx.d_data[ic], cctx.d_data[ic], tb->delta[ic], decay[ic], n, pp->comp);
        if (!rc) rc = hip_check(hipEventRecord(pp->e_comp[ic], pp->comp), "pipeline event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(pp->out, pp->e_comp[ic], 0), "pipeline wait");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(cctx.h_ibuffs[ic], cctx.d_data[ic], bytes, hipMemcpyDeviceToHost, pp->out),
                           "D2H levels");
    }
    if (!rc) rc = hip_check(hipMemcpyAsync(tb->frag[0], h_frag_col, bytes, hipMemcpyHostToDevice, pp->in), "H2D frag");
    if (!rc) rc = hip_check(hipMemcpyAsync(tb->frag[1], h_frag_ela, bytes, hipMemcpyHostToDevice, pp->in), "H2D frag");
    if (!rc) rc = hip_check(hipEventRecord(pp->e_aux[0], pp->in), "pipeline event");
    if (!rc) rc = hip_check(hipStreamWaitEvent(pp->comp, pp->e_aux[0], 0), "pipeline wait");
    std::vector<std::thread> clear;
    if (!rc) clear = clear_host_deltas(pp, h_dchem, bytes);
    if (!rc) {
        float* lv[N_CHEM];
        for (int ic = 0; ic < N_CHEM; ++ic) lv[ic] = cctx.d_data[ic];
        rc = chem_grads_totals(lv, tb->frag[0], tb->frag[1], tb->grad, n, totals, pp->comp);
    }
    if (!rc) rc = download_levels_grads(cctx, pp, tb, h_grads, bytes, false);
    for (std::thread& t : clear) t.join();
    rc = sync_pipe(pp, rc);
    g_split_pending = false;
    return rc ? fail(rc) : true;
}

// Frees what the interface keeps between calls: the per-box workspaces, the
// resident-tick arrays and the copy/compute pipeline's streams and events.
// The next call that needs them makes them again.
extern "C" int vfdReleaseWorkspaces(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = refuse_if_split_pending("vfdReleaseWorkspaces")) return rc;
    int rc = VFD_OK;
    auto keep = [&rc](int r) {
        if (!rc) rc = r;
    };
    for (auto& kv : g_pipes) {
        Pipe& p = kv.second;
        for (hipStream_t s : {p.in, p.comp, p.out}) {
            keep(hip_check(hipStreamSynchronize(s), "release: pipeline sync"));
            keep(hip_check(hipStreamDestroy(s), "release: pipeline stream"));
        }
        for (int i = 0; i < N_CHEM; ++i)
            for (hipEvent_t e : {p.e_in[i], p.e_comp[i], p.e_out[i]}) keep(hip_check(hipEventDestroy(e), "release: event"));
        for (hipEvent_t e : p.e_aux) keep(hip_check(hipEventDestroy(e), "release: event"));
    }
    g_pipes.clear();
    for (auto& kv : g_tick) {
        TickBufs& b = kv.second;
        for (float* q : b.delta) keep(hip_check(hipFree(q), "release: tick arrays"));
        for (float* q : b.frag) keep(hip_check(hipFree(q), "release: tick arrays"));
        for (float* q : b.grad) keep(hip_check(hipFree(q), "release: tick arrays"));
    }
    g_tick.clear();
    for (auto& kv : g_ws)
        if (kv.second.work) keep(hip_check(hipFree(kv.second.work), "release: workspace"));
    g_ws.clear();
    return rc;
}

// --------------------------------------------------------- kernel build
// computeKernel3D (Diffusion.cu:1078-1531), FP64 on the device: the 7-point
// stencil at the centre of the kernel domain, niter rounds of clamp-to-border
// pad -> R2C -> x S^(kernelRadius-1-[first]) / N -> C2R -> unpad, then the
// centred window.  S^ is evaluated in closed form (the exact DFT of the wrapped
// stencil) instead of an R2C of the padded stencil.
bool computeKernel3D(int kernelRadius, float lambda, float gamma, float dt, c_ctx cctx, cufftHandle fftPlanFwd,
                     cufftHandle fftPlanInv, short int ic) {
    (void)fftPlanFwd;
    (void)fftPlanInv;
    if (ic < 0 || ic >= N_CHEM) return fail(set_error(VFD_ERR_INVALID, "computeKernel3D: bad chem index"));
    const int D = cctx.DD, H = cctx.DH, W = cctx.DW;
    const int FD = cctx.FFTD, FH = cctx.FFTH, FW = cctx.FFTW;
    const int wD = cctx.windowD, wH = cctx.windowH, wW = cctx.windowW;
    const int KD = cctx.KD, KH = cctx.KH, KW = cctx.KW;
    if (D < 3 || H < 3 || W < 3 || FD < D || FH < H || FW < W || wD > D || wH > H || wW > W || cctx.niter < 1)
        return fail(set_error(VFD_ERR_INVALID, "computeKernel3D: bad kernel context dimensions"));
    if (!cctx.d_data[ic]) return fail(set_error(VFD_ERR_INVALID, "computeKernel3D: d_data[%d] is null", ic));
    std::lock_guard<std::mutex> lk(g_mu);
    vfd_plan_desc d{FD, FH, FW, FD, FH, FW, VFD_F64, VFD_PAD_ZERO, VFD_KSPEC_REAL};
    Plan plan;
    if (int rc = plan_init(&d, &plan)) return fail(rc);
    const long dn = (long)D * H * W, fn = (long)FD * FH * FW, sn = (long)FD * FH * plan.hx;
    // host: stencil in the middle block of the domain (Diffusion.cu:1187-1229)
    std::vector<double> h_data(dn, 0.0);
    const double lam = lambda, c0 = 1.0 - 6.0 * (double)lambda - (double)gamma * (double)dt;
    double hk[27] = {0};
    hk[0 * 9 + 1 * 3 + 1] = lam; hk[2 * 9 + 1 * 3 + 1] = lam;
    hk[1 * 9 + 0 * 3 + 1] = lam; hk[1 * 9 + 2 * 3 + 1] = lam;
    hk[1 * 9 + 1 * 3 + 0] = lam; hk[1 * 9 + 1 * 3 + 2] = lam;
    hk[1 * 9 + 1 * 3 + 1] = c0;
    (void)KD; (void)KH; (void)KW;
    const int sk = D / 2 - 1, si = H / 2 - 1, sj = W / 2 - 1;
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) h_data[((long)(sk + k) * H + (si + i)) * W + (sj + j)] = hk[(k * 3 + i) * 3 + j];
    double *d_data = nullptr, *d_pad = nullptr;
    double2* d_spec = nullptr;
    bool ok = hip_ok(hipMalloc(&d_data, dn * sizeof(double)), "kernel build alloc") &&
              hip_ok(hipMalloc(&d_pad, fn * sizeof(double)), "kernel build alloc") &&
              hip_ok(hipMalloc(&d_spec, sn * sizeof(double2)), "kernel build alloc") &&
              hip_ok(hipMemcpy(d_data, h_data.data(), dn * sizeof(double), hipMemcpyHostToDevice), "kernel upload");
    const double scale = 1.0 / ((double)FD * FH * FW);
    for (int it = 0; ok && it < cctx.niter; ++it) {
        hipLaunchKernelGGL(pad_clamp3d<double>, dim3(nblocks(fn)), dim3(kThreads), 0, nullptr, d_pad, d_data, FD, FH,
                           FW, D, H, W, 1, 1, 1);
        ok = plan_r2c(plan, d_pad, d_spec, nullptr) == VFD_OK || fail(-1);
        if (!ok) break;
        const int n = kernelRadius - (it == 0 ? 1 : 0) - 1;
        hipLaunchKernelGGL(stencil_power_multiply, dim3(nblocks(sn)), dim3(kThreads), 0, nullptr, d_spec, FD, FH, FW,
                           plan.hx, lam, c0, n, g_quirk, scale);
        ok = plan_c2r(plan, d_spec, d_pad, nullptr) == VFD_OK || fail(-1);
        if (!ok) break;
        hipLaunchKernelGGL(unpad3d<double>, dim3(nblocks(dn)), dim3(kThreads), 0, nullptr, d_data, d_pad, D, H, W,
                           FH, FW);
        ok = hip_ok(hipGetLastError(), "kernel build launch");
    }
    const long wn = (long)wD * wH * wW;
    if (ok) {
        hipLaunchKernelGGL((extract_center3d<float, double>), dim3(nblocks(wn)), dim3(kThreads), 0, nullptr,
                           cctx.d_data[ic], d_data, D, H, W, wD, wH, wW);
        ok = hip_ok(hipGetLastError(), "extract window") && hip_ok(hipDeviceSynchronize(), "kernel build");
    }
    if (ok && cctx.h_data[ic])
        ok = hip_ok(hipMemcpy(cctx.h_data[ic], cctx.d_data[ic], wn * sizeof(float), hipMemcpyDeviceToHost),
                    "window download");
    (void)hipFree(d_data);
    (void)hipFree(d_pad);
    (void)hipFree(d_spec);
    return ok;
}

// computeKernelSpectrum3D (Diffusion.cu:1533-1619): window (kernel_cctx.d_data[ic],
// chem KD x KH x KW, centre KZ,KY,KX) padded into the chem FFT box, R2C in FP64,
// stored as fComplex in chem_cctx.d_kernelspectrum_h[ic] (and h_kernelspectrum).
bool computeKernelSpectrum3D(c_ctx kernel_cctx, c_ctx chem_cctx, short int ic) {
    if (ic < 0 || ic >= N_CHEM) return fail(set_error(VFD_ERR_INVALID, "computeKernelSpectrum3D: bad chem index"));
    const int FD = chem_cctx.FFTD, FH = chem_cctx.FFTH, FW = chem_cctx.FFTW;
    const int KD = chem_cctx.KD, KH = chem_cctx.KH, KW = chem_cctx.KW;
    const float* d_win = kernel_cctx.d_data[ic];
    fComplex* d_out = chem_cctx.d_kernelspectrum_h[ic];
    if (!d_win || !d_out) return fail(set_error(VFD_ERR_INVALID, "computeKernelSpectrum3D: null buffer"));
    if (KD > FD || KH > FH || KW > FW)
        return fail(set_error(VFD_ERR_INVALID, "computeKernelSpectrum3D: window larger than the FFT box"));
    std::lock_guard<std::mutex> lk(g_mu);
    vfd_plan_desc d{FD, FH, FW, FD, FH, FW, VFD_F64, VFD_PAD_ZERO, VFD_KSPEC_REAL};
    Plan plan;
    if (int rc = plan_init(&d, &plan)) return fail(rc);
    const long fn = (long)FD * FH * FW, sn = (long)FD * FH * plan.hx, kn = (long)KD * KH * KW;
    double* d_pad = nullptr;
    double2* d_spec = nullptr;
    bool ok = hip_ok(hipMalloc(&d_pad, fn * sizeof(double)), "spectrum alloc") &&
              hip_ok(hipMalloc(&d_spec, sn * sizeof(double2)), "spectrum alloc");
    if (ok) {
        hipLaunchKernelGGL(fill<double>, dim3(nblocks(fn)), dim3(kThreads), 0, nullptr, d_pad, fn, 0.0);
        hipLaunchKernelGGL((pad_kernel3d<double, float>), dim3(nblocks(kn)), dim3(kThreads), 0, nullptr, d_pad,
                           d_win, FD, FH, FW, KD, KH, KW, chem_cctx.KZ, chem_cctx.KY, chem_cctx.KX);
        ok = hip_ok(hipGetLastError(), "pad kernel") && (plan_r2c(plan, d_pad, d_spec, nullptr) == VFD_OK || fail(-1));
    }
    if (ok) {
        hipLaunchKernelGGL(to_fcomplex, dim3(nblocks(sn)), dim3(kThreads), 0, nullptr, (float2*)d_out, d_spec, sn);
        ok = hip_ok(hipGetLastError(), "spectrum convert") && hip_ok(hipDeviceSynchronize(), "spectrum");
    }
    if (ok && chem_cctx.h_kernelspectrum[ic])
        ok = hip_ok(hipMemcpy(chem_cctx.h_kernelspectrum[ic], d_out, sn * sizeof(fComplex), hipMemcpyDeviceToHost),
                    "spectrum download");
    (void)hipFree(d_pad);
    (void)hipFree(d_spec);
    return ok;
}

// computeKernel3DBatch (Diffusion.cu:988-1076)
bool computeKernel3DBatch(int kernelRadius, float lambda[N_CHEM], float gamma[N_CHEM], float dt, c_ctx kern_cctx,
                          c_ctx chem_cctx) {
    bool ok = true;
    for (int ic = 0; ic < N_CHEM && ok; ++ic) {
        if (!hip_ok(hipSetDevice(chem_cctx.dev_id[ic]), "hipSetDevice")) return false;
        ok = computeKernel3D(kernelRadius, lambda[ic], gamma[ic], dt, kern_cctx, 0, 0, (short)ic) &&
             computeKernelSpectrum3D(kern_cctx, chem_cctx, (short)ic);
    }
    return ok;
}

// ------------------------------------------------------------ 2D path
// The 2D entry points (Diffusion.cu:303-907, the !MODEL_3D build) run on the
// 3D passes: an H x W array on an (fftH, fftW) box is plane 0 of a box two
// planes deep, and a 2D kernel k is k (x) delta(z), whose spectrum is the 2D
// one on both kz planes.  The kernel build works on a z-constant array (both
// planes equal), whose spectrum is twice the 2D one on plane kz = 0.

// computeKernel (Diffusion.cu:303-675): the 5-point stencil (lambda at the four
// neighbours, 1 - 4 lambda - gamma dt at the centre of its KH x KW array) is
// placed at the centre of the H x W domain, padded once to the FFT box by
// clamp-to-border, convolved kernelRadius times with the stencil on the box
// (the reference repeats R2C, x S^ / (fftH fftW), C2R on the padded array),
// cropped to the domain, and the centred window is written to d_Window.  Here
// the kernelRadius convolutions are one multiplication by S^kernelRadius, in
// FP64.  The reference copies 2 (KH/2) + 1 stencil rows and columns (one past
// the array for even KH or KW); here only the KH x KW array is copied.
bool computeKernel(float* d_Window, int kernelRadius, float lambda, float gamma, float dt, c_ctx cctx) {
    const int H = cctx.DH, W = cctx.DW, fH = cctx.FFTH, fW = cctx.FFTW;
    const int wH = cctx.windowH, wW = cctx.windowW, KH = cctx.KH, KW = cctx.KW;
    if (!d_Window) return fail(set_error(VFD_ERR_INVALID, "computeKernel: null window"));
    if (H < 3 || W < 3 || KH < 3 || KW < 3 || fH < H || fW < W || wH < 1 || wW < 1 || wH > H || wW > W ||
        kernelRadius < 0)
        return fail(set_error(VFD_ERR_INVALID, "computeKernel: bad kernel context dimensions"));
    std::lock_guard<std::mutex> lk(g_mu);
    vfd_plan_desc d{2, fH, fW, 2, fH, fW, VFD_F64, VFD_PAD_ZERO, VFD_KSPEC_REAL};
    Plan plan;
    if (int rc = plan_init(&d, &plan)) return fail(rc);
    const long dn = (long)H * W, pn = (long)fH * fW, sn = 2L * fH * plan.hx;
    const double lam = lambda, c0 = 1.0 - 4.0 * (double)lambda - (double)gamma * (double)dt;
    std::vector<double> hk((size_t)KH * KW, 0.0);
    hk[0 * KW + 1] = lam;
    hk[1 * KW + 0] = lam;
    hk[1 * KW + 2] = lam;
    hk[2 * KW + 1] = lam;
    hk[1 * KW + 1] = c0;
    std::vector<double> h_data(dn, 0.0);
    const int i0 = H / 2 - KH / 2, j0 = W / 2 - KW / 2;
    for (int ki = 0; ki < KH && i0 + ki < H; ++ki)
        for (int kj = 0; kj < KW && j0 + kj < W; ++kj)
            if (i0 + ki >= 0 && j0 + kj >= 0) h_data[(long)(i0 + ki) * W + (j0 + kj)] = hk[ki * KW + kj];
    double *d_data = nullptr, *d_pad = nullptr;
    double2* d_spec = nullptr;
    bool ok = hip_ok(hipMalloc(&d_data, dn * sizeof(double)), "kernel build alloc") &&
              hip_ok(hipMalloc(&d_pad, 2 * pn * sizeof(double)), "kernel build alloc") &&
              hip_ok(hipMalloc(&d_spec, sn * sizeof(double2)), "kernel build alloc") &&
              hip_ok(hipMemcpy(d_data, h_data.data(), dn * sizeof(double), hipMemcpyHostToDevice), "kernel upload");
    if (ok) {
        hipLaunchKernelGGL(pad_clamp3d<double>, dim3(nblocks(pn)), dim3(kThreads), 0, nullptr, d_pad, d_data, 1, fH,
                           fW, 1, H, W, 0, cctx.KY, cctx.KX);
        ok = hip_ok(hipGetLastError(), "kernel build pad") &&
             hip_ok(hipMemcpy(d_pad + pn, d_pad, pn * sizeof(double), hipMemcpyDeviceToDevice), "kernel build copy") &&
             (plan_r2c(plan, d_pad, d_spec, nullptr) == VFD_OK || fail(-1));
    }
    if (ok) {
        hipLaunchKernelGGL(stencil2d_power_multiply, dim3(nblocks(sn)), dim3(kThreads), 0, nullptr, d_spec, 2, fH, fW,
                           plan.hx, lam, c0, kernelRadius, cctx.KY, cctx.KX, 1.0 / (2.0 * fH * fW));
        ok = hip_ok(hipGetLastError(), "kernel build power") &&
             (plan_c2r(plan, d_spec, d_pad, nullptr) == VFD_OK || fail(-1));
    }
    if (ok) {
        hipLaunchKernelGGL(unpad3d<double>, dim3(nblocks(dn)), dim3(kThreads), 0, nullptr, d_data, d_pad, 1, H, W, fH,
                           fW);
        hipLaunchKernelGGL((extract_center3d<float, double>), dim3(nblocks((long)wH * wW)), dim3(kThreads), 0, nullptr,
                           d_Window, d_data, 1, H, W, 1, wH, wW);
        ok = hip_ok(hipGetLastError(), "extract window") && hip_ok(hipDeviceSynchronize(), "kernel build");
    }
    (void)hipFree(d_data);
    (void)hipFree(d_pad);
    (void)hipFree(d_spec);
    return ok;
}

// computeKernelSpectrum (Diffusion.cu:677-771): the KH x KW window d_Kernel
// (chem_cctx KH, KW, centre KY, KX) padded into the chem FFT box and stored as
// the packed spectrum fftH x fftW/2 fComplex that fftDiffuse2D reads.
bool computeKernelSpectrum(fComplex* d_KernelSpectrum, float* d_Kernel, c_ctx kernel_cctx, c_ctx chem_cctx) {
    (void)kernel_cctx;
    const int fH = chem_cctx.FFTH, fW = chem_cctx.FFTW, KH = chem_cctx.KH, KW = chem_cctx.KW;
    if (!d_KernelSpectrum || !d_Kernel) return fail(set_error(VFD_ERR_INVALID, "computeKernelSpectrum: null buffer"));
    if (KH < 1 || KW < 1 || KH > fH || KW > fW || fW % 2)
        return fail(set_error(VFD_ERR_INVALID, "computeKernelSpectrum: bad kernel or FFT box"));
    std::lock_guard<std::mutex> lk(g_mu);
    vfd_plan_desc d{2, fH, fW, 2, fH, fW, VFD_F64, VFD_PAD_ZERO, VFD_KSPEC_REAL};
    Plan plan;
    if (int rc = plan_init(&d, &plan)) return fail(rc);
    const long pn = (long)fH * fW, sn = 2L * fH * plan.hx;
    double* d_pad = nullptr;
    double2* d_spec = nullptr;
    bool ok = hip_ok(hipMalloc(&d_pad, 2 * pn * sizeof(double)), "spectrum alloc") &&
              hip_ok(hipMalloc(&d_spec, sn * sizeof(double2)), "spectrum alloc");
    if (ok) {
        hipLaunchKernelGGL(fill<double>, dim3(nblocks(pn)), dim3(kThreads), 0, nullptr, d_pad, pn, 0.0);
        hipLaunchKernelGGL((pad_kernel3d<double, float>), dim3(nblocks((long)KH * KW)), dim3(kThreads), 0, nullptr,
                           d_pad, d_Kernel, 1, fH, fW, 1, KH, KW, 0, chem_cctx.KY, chem_cctx.KX);
        ok = hip_ok(hipGetLastError(), "pad kernel") &&
             hip_ok(hipMemcpy(d_pad + pn, d_pad, pn * sizeof(double), hipMemcpyDeviceToDevice), "spectrum copy") &&
             (plan_r2c(plan, d_pad, d_spec, nullptr) == VFD_OK || fail(-1));
    }
    if (ok) {
        // plane kz = 0 of the z-constant spectrum is twice the 2D one
        hipLaunchKernelGGL(standard_to_packed2d, dim3(nblocks((long)fH * (fW / 2))), dim3(kThreads), 0, nullptr,
                           (float2*)d_KernelSpectrum, d_spec, fH, fW, plan.hx, 0.5);
        ok = hip_ok(hipGetLastError(), "spectrum convert") && hip_ok(hipDeviceSynchronize(), "spectrum");
    }
    (void)hipFree(d_pad);
    (void)hipFree(d_spec);
    return ok;
}

// fftDiffuse2D (Diffusion.cu:773-907): constant-VF pad of the H x W field
// (interior 0 < y < H-1, 0 < x < min(epiBoundary, W), baseChem elsewhere),
// circular convolution on the FFT box with the kernel of the packed spectrum,
// normalized, cropped back to H x W in place.  Buffers are allocated per call,
// as the reference does.
bool fftDiffuse2D(float* d_Data, fComplex* d_KernelSpectrum0, c_ctx cctx, int epiBoundary, float baseChem) {
    const int H = cctx.DH, W = cctx.DW, fH = cctx.FFTH, fW = cctx.FFTW;
    if (!d_Data || !d_KernelSpectrum0) return fail(set_error(VFD_ERR_INVALID, "fftDiffuse2D: null buffer"));
    if (H < 1 || W < 1 || fH < H || fW < W || fW % 2)
        return fail(set_error(VFD_ERR_INVALID, "fftDiffuse2D: bad data or FFT box"));
    std::lock_guard<std::mutex> lk(g_mu);
    vfd_plan_desc d{1, fH, fW, 2, fH, fW, VFD_F32, VFD_PAD_ZERO, VFD_KSPEC_COMPLEX_STD};
    Plan plan;
    if (int rc = plan_init(&d, &plan)) return fail(rc);
    const long pn = (long)fH * fW, sn = 2L * fH * (fW / 2 + 1);
    float* d_pad = nullptr;
    float2* d_spec = nullptr;
    void* d_work = nullptr;
    bool ok = hip_ok(hipMalloc(&d_pad, pn * sizeof(float)), "fftDiffuse2D alloc") &&
              hip_ok(hipMalloc(&d_spec, sn * sizeof(float2)), "fftDiffuse2D alloc") &&
              hip_ok(hipMalloc(&d_work, plan.work_bytes), "fftDiffuse2D alloc");
    if (ok) {
        hipLaunchKernelGGL(pad_constant_vf2d, dim3(nblocks(pn)), dim3(kThreads), 0, nullptr, d_pad, d_Data, fH, fW, H,
                           W, epiBoundary, baseChem);
        hipLaunchKernelGGL(packed_to_standard2d, dim3(nblocks((long)fH * (fW / 2 + 1))), dim3(kThreads), 0, nullptr,
                           d_spec, (const float2*)d_KernelSpectrum0, 2, fH, fW);
        ok = hip_ok(hipGetLastError(), "fftDiffuse2D pad");
    }
    if (ok) {
        vfd_field f{d_pad, d_pad, d_work, d_spec, 0.0, 0, 0};
        ok = plan_diffuse(plan, &f, 1, nullptr) == VFD_OK || fail(-1);
    }
    if (ok) {
        hipLaunchKernelGGL(unpad3d<float>, dim3(nblocks((long)H * W)), dim3(kThreads), 0, nullptr, d_Data, d_pad, 1, H,
                           W, fH, fW);
        ok = hip_ok(hipGetLastError(), "fftDiffuse2D unpad") && hip_ok(hipDeviceSynchronize(), "fftDiffuse2D");
    }
    (void)hipFree(d_pad);
    (void)hipFree(d_spec);
    (void)hipFree(d_work);
    return ok;
}

// ------------------------------------------------------------ 3D helpers
void extractCenter3D(float* d_Dst, float* d_Src, int dataD, int dataH, int dataW, int windowD, int windowH,
                     int windowW) {
    const long n = (long)windowD * windowH * windowW;
    hipLaunchKernelGGL((extract_center3d<float, float>), dim3(nblocks(n)), dim3(kThreads), 0, nullptr, d_Dst, d_Src,
                       dataD, dataH, dataW, windowD, windowH, windowW);
    hip_ok(hipGetLastError(), "extractCenter3D");
}

void unpadResult3D(float* d_Dst, float* d_Src, int dataD, int dataH, int dataW, int fftD, int fftH, int fftW) {
    (void)fftD;
    const long n = (long)dataD * dataH * dataW;
    hipLaunchKernelGGL(unpad3d<float>, dim3(nblocks(n)), dim3(kThreads), 0, nullptr, d_Dst, d_Src, dataD, dataH,
                       dataW, fftH, fftW);
    hip_ok(hipGetLastError(), "unpadResult3D");
}

void padKernel3D(float* d_Dst, float* d_Src, int fftD, int fftH, int fftW, int kernelD, int kernelH, int kernelW,
                 int kernelZ, int kernelY, int kernelX) {
    const long n = (long)kernelD * kernelH * kernelW;
    hipLaunchKernelGGL((pad_kernel3d<float, float>), dim3(nblocks(n)), dim3(kThreads), 0, nullptr, d_Dst, d_Src,
                       fftD, fftH, fftW, kernelD, kernelH, kernelW, kernelZ, kernelY, kernelX);
    hip_ok(hipGetLastError(), "padKernel3D");
}

void padDataClampToBorder3D(float* d_PaddedData, float* d_Data, int fftD, int fftH, int fftW, int dataD, int dataH,
                            int dataW, int kernelD, int kernelH, int kernelW, int kernelZ, int kernelY, int kernelX) {
    (void)kernelD; (void)kernelH; (void)kernelW;
    const long n = (long)fftD * fftH * fftW;
    hipLaunchKernelGGL(pad_clamp3d<float>, dim3(nblocks(n)), dim3(kThreads), 0, nullptr, d_PaddedData, d_Data, fftD,
                       fftH, fftW, dataD, dataH, dataW, kernelZ, kernelY, kernelX);
    hip_ok(hipGetLastError(), "padDataClampToBorder3D");
}

void padDataConstantVF3D(float* d_Dst, float* d_Src, int fftD, int fftH, int fftW, int dataD, int dataH, int dataW,
                         int kernelD, int kernelH, int kernelW, int kernelZ, int kernelY, int kernelX,
                         int epiBoundary, float baseChem) {
    (void)kernelD; (void)kernelH; (void)kernelW; (void)kernelZ; (void)kernelY; (void)kernelX;
    const long n = (long)fftD * fftH * fftW;
    hipLaunchKernelGGL(pad_constant_vf3d, dim3(nblocks(n)), dim3(kThreads), 0, nullptr, d_Dst, d_Src, fftD, fftH,
                       fftW, dataD, dataH, dataW, epiBoundary, baseChem);
    hip_ok(hipGetLastError(), "padDataConstantVF3D");
}

void modulateAndNormalize3D(fComplex* d_Dst, fComplex* d_Src, int fftD, int fftH, int fftW, int padding) {
    const long n = (long)fftD * fftH * (fftW / 2 + padding);
    hipLaunchKernelGGL(modulate_scale, dim3(nblocks(n)), dim3(kThreads), 0, nullptr, (float2*)d_Dst,
                       (const float2*)d_Src, n, 1.0f / (float)(fftW * fftH * fftD));
    hip_ok(hipGetLastError(), "modulateAndNormalize3D");
}

void modulate3D(fComplex* d_Dst, fComplex* d_Src, int fftD, int fftH, int fftW, int padding) {
    const long n = (long)fftD * fftH * (fftW / 2 + padding);
    hipLaunchKernelGGL(modulate_scale, dim3(nblocks(n)), dim3(kThreads), 0, nullptr, (float2*)d_Dst,
                       (const float2*)d_Src, n, 1.0f);
    hip_ok(hipGetLastError(), "modulate3D");
}

void complexPower(fComplex* d_Dst, int fftD, int fftH, int fftW, int padding, int n) {
    const long cnt = (long)fftD * fftH * (fftW / 2 + padding);
    hipLaunchKernelGGL(complex_power, dim3(nblocks(cnt)), dim3(kThreads), 0, nullptr, (float2*)d_Dst, cnt, n);
    hip_ok(hipGetLastError(), "complexPower");
}

// ------------------------------------------------------------ 2D helpers
void extractCenter(float* d_Dst, float* d_Src, int dataH, int dataW, int windowH, int windowW) {
    extractCenter3D(d_Dst, d_Src, 1, dataH, dataW, 1, windowH, windowW);
}

void unpadResult(float* d_Dst, float* d_Src, int dataH, int dataW, int fftH, int fftW) {
    unpadResult3D(d_Dst, d_Src, 1, dataH, dataW, 1, fftH, fftW);
}

void padKernel(float* d_PaddedKernel, float* d_Kernel, int fftH, int fftW, int kernelH, int kernelW, int kernelY,
               int kernelX) {
    hipLaunchKernelGGL(pad_kernel2d, dim3(nblocks((long)kernelH * kernelW)), dim3(kThreads), 0, nullptr,
                       d_PaddedKernel, d_Kernel, fftH, fftW, kernelH, kernelW, kernelY, kernelX);
    hip_ok(hipGetLastError(), "padKernel");
}

void padDataConstantVF(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW, int kernelH,
                       int kernelW, int kernelY, int kernelX, int epiBoundary, float baseChem) {
    (void)kernelH; (void)kernelW; (void)kernelY; (void)kernelX;
    hipLaunchKernelGGL(pad_constant_vf2d, dim3(nblocks((long)fftH * fftW)), dim3(kThreads), 0, nullptr, d_PaddedData,
                       d_Data, fftH, fftW, dataH, dataW, epiBoundary, baseChem);
    hip_ok(hipGetLastError(), "padDataConstantVF");
}

void padDataRightWall(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW, int kernelH,
                      int kernelW, int kernelY, int kernelX) {
    (void)kernelH; (void)kernelW;
    hipLaunchKernelGGL(pad_right_wall2d, dim3(nblocks((long)fftH * fftW)), dim3(kThreads), 0, nullptr, d_PaddedData,
                       d_Data, fftH, fftW, dataH, dataW, kernelY, kernelX);
    hip_ok(hipGetLastError(), "padDataRightWall");
}

void padDataMirror(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW, int kernelH,
                   int kernelW, int kernelY, int kernelX) {
    (void)kernelH; (void)kernelW;
    hipLaunchKernelGGL(pad_mirror2d, dim3(nblocks((long)fftH * fftW)), dim3(kThreads), 0, nullptr, d_PaddedData,
                       d_Data, fftH, fftW, dataH, dataW, kernelY, kernelX);
    hip_ok(hipGetLastError(), "padDataMirror");
}

void padDataClampToBorder(float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW, int kernelH,
                          int kernelW, int kernelY, int kernelX) {
    (void)kernelH; (void)kernelW;
    hipLaunchKernelGGL(pad_clamp3d<float>, dim3(nblocks((long)fftH * fftW)), dim3(kThreads), 0, nullptr,
                       d_PaddedData, d_Data, 1, fftH, fftW, 1, dataH, dataW, 0, kernelY, kernelX);
    hip_ok(hipGetLastError(), "padDataClampToBorder");
}

void padData2D(pad_t pt, float* d_PaddedData, float* d_Data, int fftH, int fftW, int dataH, int dataW, int kernelH,
               int kernelW, int kernelY, int kernelX, int epiBoundary, float baseChem) {
    switch (pt) {
        case pClampToBorder:
            padDataClampToBorder(d_PaddedData, d_Data, fftH, fftW, dataH, dataW, kernelH, kernelW, kernelY, kernelX);
            break;
        case pRightWall:
            padDataRightWall(d_PaddedData, d_Data, fftH, fftW, dataH, dataW, kernelH, kernelW, kernelY, kernelX);
            break;
        case pMirror:
            padDataMirror(d_PaddedData, d_Data, fftH, fftW, dataH, dataW, kernelH, kernelW, kernelY, kernelX);
            break;
        case pConstantVF:
            padDataConstantVF(d_PaddedData, d_Data, fftH, fftW, dataH, dataW, kernelH, kernelW, kernelY, kernelX,
                              epiBoundary, baseChem);
            break;
    }
}

void modulateAndNormalize(fComplex* d_Dst, fComplex* d_Src, int fftH, int fftW, int padding) {
    const long n = (long)fftH * (fftW / 2 + padding);
    hipLaunchKernelGGL(modulate_scale, dim3(nblocks(n)), dim3(kThreads), 0, nullptr, (float2*)d_Dst,
                       (const float2*)d_Src, n, 1.0f / (float)(fftW * fftH));
    hip_ok(hipGetLastError(), "modulateAndNormalize");
}

// convolutionClampToBorderCPU (convolutionFFT_gold.cpp:22-58), restated from
// its definition.  For kernel taps k[r][c] (r < kH, c < kW) centred at
// (KY, KX) and clampN(i) = min(max(i, 0), N - 1):
//   out(y, x) = sum over r = kH-1 down to 0, then c = kW-1 down to 0, of
//               fp32( u(clampH(y + KY - r), clampW(x + KX - c)) * k[r][c] ),
// each product rounded to FP32 and summed in FP64 in that order, the sum
// rounded to FP32.  (The order and the product's rounding decide the last bits.)
void convolutionClampToBorderCPU(float* h_Result, float* h_Data, float* h_Kernel, int dataH, int dataW, int kernelH,
                                 int kernelW, int kernelY, int kernelX) {
    auto clampi = [](int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); };
    const long n = (long)dataH * dataW;
    for (long o = 0; o < n; ++o) {
        const int y = (int)(o / dataW), x = (int)(o % dataW);
        double acc = 0.0;
        for (int r = kernelH - 1; r >= 0; --r) {
            const float* urow = h_Data + (long)clampi(y + kernelY - r, dataH) * dataW;
            const float* krow = h_Kernel + (long)r * kernelW;
            for (int c = kernelW - 1; c >= 0; --c) {
                const float prod = urow[clampi(x + kernelX - c, dataW)] * krow[c];
                acc += (double)prod;
            }
        }
        h_Result[o] = (float)acc;
    }
}

}  // extern "C"
