// vfd_chem.hip -- the post-diffusion chem update on the device.
//
// Restates WHChemical::update with gpuUpdate (src/FieldVariable/Usr_FieldVariables/
// WHChemical.cpp:395-450), which runs on the host after every diffusion tick:
// per voxel and chem, p = d + t * decay[ic] (the diffused field t scaled by the
// chem's decay, plus the agents' deltas d), d = 0, per-chem totals, then the
// three chemotaxis gradient fields (NEU, MAC, FIB; enums.h:134-136) from the
// new levels and the ECM fragment fields.  With the fields resident on the
// device, this removes the host pass over 8 + 8 + 8 + 3 arrays per tick.
//
// Arithmetic follows the reference in FP32 with the multiply and the add
// rounded separately (no fused multiply-add), and the gradient sums in the
// reference's left-to-right order, so levels and gradients are bit-identical
// to a host evaluation of the same expressions.  Totals are compensated
// (Neumaier) FP64 sums with a fixed reduction order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/vfd.h"
#include "vfd_internal.h"

using namespace vfd;

namespace {

constexpr int kChems = 8;
constexpr int kThreads = 256;
constexpr int kBlocks = 1024;
// chem indices (enums.h:124-133)
constexpr int TNF = 0, TGF = 1, FGF = 2, IL1beta = 4, IL6 = 5, IL8 = 6;

struct ChemPtrs {
    const float* t[kChems];
    float* d[kChems];
    float* p[kChems];
    float decay[kChems];
};

__device__ __forceinline__ void nadd(double& s, double& c, double x) {
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

// Each block owns one contiguous chunk of voxels; lanes stride through it and
// a fixed-order tree combines the per-lane (sum, compensation) pairs.
// LEVELS: the levels are already in cp.p (the x-inverse pass wrote them);
// only the gradients and totals are formed.
template <bool LEVELS>
__global__ void __launch_bounds__(kThreads)
chem_update_kernel(ChemPtrs cp, const float* __restrict__ frag_col, const float* __restrict__ frag_ela,
                   float* __restrict__ neu, float* __restrict__ mac, float* __restrict__ fib, long n, long chunk,
                   double2* __restrict__ partial) {
#pragma clang fp contract(off)  // the product is rounded before the add, as on the host
    const long lo = (long)blockIdx.x * chunk;
    const long hi = lo + chunk < n ? lo + chunk : n;
    double s[kChems], c[kChems];
#pragma unroll
    for (int ic = 0; ic < kChems; ++ic) s[ic] = c[ic] = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kThreads) {
        float lv[kChems];
#pragma unroll
        for (int ic = 0; ic < kChems; ++ic) {
            if constexpr (LEVELS) {
                lv[ic] = cp.p[ic][i];
            } else {
                const float scaled = cp.t[ic][i] * cp.decay[ic];
                lv[ic] = cp.d[ic][i] + scaled;
            }
        }
#pragma unroll
        for (int ic = 0; ic < kChems; ++ic) {
            if constexpr (!LEVELS) {
                cp.d[ic][i] = 0.0f;
                cp.p[ic][i] = lv[ic];
            }
            nadd(s[ic], c[ic], (double)lv[ic]);
        }
        // gradients (WHChemical.cpp:437-445), left-to-right FP32 sums
        const float grad = lv[IL1beta] + lv[TNF] + lv[TGF] + lv[FGF] + frag_col[i];
        fib[i] = lv[TGF];
        neu[i] = grad + lv[IL6] + lv[IL8];
        mac[i] = grad + frag_ela[i];
    }
    __shared__ double ss[kChems][kThreads], sc[kChems][kThreads];
#pragma unroll
    for (int ic = 0; ic < kChems; ++ic) {
        ss[ic][threadIdx.x] = s[ic];
        sc[ic][threadIdx.x] = c[ic];
    }
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int ic = 0; ic < kChems; ++ic) {
                double a = ss[ic][threadIdx.x], ca = sc[ic][threadIdx.x];
                const double b = ss[ic][threadIdx.x + w], cb = sc[ic][threadIdx.x + w];
                nadd(a, ca, b);
                ss[ic][threadIdx.x] = a;
                sc[ic][threadIdx.x] = ca + cb;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < kChems)
        partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = double2{ss[threadIdx.x][0], sc[threadIdx.x][0]};
}

// one block per chem: the block partials in a fixed order
__global__ void __launch_bounds__(kThreads) chem_total_kernel(const double2* __restrict__ partial, int nparts,
                                                              double* __restrict__ totals) {
    const double2* pp = partial + (long)blockIdx.x * nparts;
    double s = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        nadd(s, c, pp[i].x);
        c += pp[i].y;
    }
    __shared__ double ss[kThreads], sc[kThreads];
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = ss[threadIdx.x], ca = sc[threadIdx.x];
            nadd(a, ca, ss[threadIdx.x + w]);
            ss[threadIdx.x] = a;
            sc[threadIdx.x] = ca + sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = ss[0] + sc[0];
}

// sampleChem_kernel (src/Visualization/3D/volumeRender_kernel.cu:215-300):
// every stride-th voxel along each axis, times the chem's display scale,
// clamped above at 1, into a dense [D/stride][H/stride][W/stride] array.

__global__ void sample_chem_kernel(const float* __restrict__ src, int H, int W, int sD, int sH, int sW, int stride,
                                   float scale, float* __restrict__ dst) {
    const long total = (long)sD * sH * sW;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long dx = i % sW, r = i / sW;
        const long dy = r % sH, dz = r / sH;
        float v = src[(dz * stride * H + dy * stride) * (long)W + dx * stride] * scale;
        if (v > 1.0f) v = 1.0f;
        dst[i] = v;
    }
}

// one chem's levels: p = d + t * decay (product rounded before the add), d = 0;
// p may alias t
__global__ void __launch_bounds__(kThreads)
chem_levels_kernel(float* p, const float* t, float* __restrict__ d, float decay, long n) {
#pragma clang fp contract(off)
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float scaled = t[i] * decay;
        const float lv = d[i] + scaled;
        d[i] = 0.0f;
        p[i] = lv;
    }
}

}  // namespace

extern "C" int vfd_sample_chem(const float* d_src, int D, int H, int W, int chem_index, int stride, float* d_dst,
                               void* stream) {
    if (!d_src || !d_dst) return set_error(VFD_ERR_INVALID, "sample_chem: null buffer");
    if (D < 1 || H < 1 || W < 1 || stride < 1) return set_error(VFD_ERR_INVALID, "sample_chem: bad dimensions");
    if (chem_index < 0 || chem_index >= kChems) return set_error(VFD_ERR_INVALID, "sample_chem: chem index %d", chem_index);
    const int sD = D / stride, sH = H / stride, sW = W / stride;
    const long total = (long)sD * sH * sW;
    if (total == 0) return VFD_OK;
    const unsigned blocks = (unsigned)std::min<long>(4096, (total + 255) / 256);
    hipLaunchKernelGGL(sample_chem_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, d_src, H, W, sD, sH, sW,
                       stride, vfd::chem_display_scale(chem_index), d_dst);
    return hip_check(hipGetLastError(), "sample_chem launch");
}

// Device bytes chem_pass needs for its partial sums: one (sum, compensation)
// pair per chem and work-group (at most kBlocks work-groups, whatever n is),
// then the 8 totals.
constexpr size_t kPassScratchBytes = (size_t)kChems * kBlocks * sizeof(double2) + kChems * sizeof(double);

// levels (unless already formed), gradients and totals for 8 chems of n voxels.
// buf: kPassScratchBytes of device memory for the partial sums, or null to
// allocate them on the stream; d_totals: a device destination of the 8 totals
// instead of the host array `totals` (both may be null).
static int chem_pass(const ChemPtrs& cp, bool levels_in_p, const float* d_frag_col, const float* d_frag_ela,
                     float* const* d_grads, long n, double* totals, hipStream_t st, void* buf = nullptr,
                     double* d_totals = nullptr) {
    const int nb = (int)std::min<long>(kBlocks, std::max<long>(1, (n + kThreads - 1) / kThreads));
    const long chunk = (n + nb - 1) / nb;
    double2* partial = nullptr;
    double* d_tot = nullptr;
    const size_t pbytes = (size_t)kChems * nb * sizeof(double2);
    if (buf) {
        if (pbytes + kChems * sizeof(double) > kPassScratchBytes)
            return set_error(VFD_ERR_INVALID, "chem update: %d partial sums exceed the scratch array", kChems * nb);
        partial = (double2*)buf;
    } else if (int rc = hip_check(hipMallocAsync((void**)&partial, pbytes + kChems * sizeof(double), st),
                                  "hipMallocAsync(chem update)")) {
        return rc;
    }
    d_tot = d_totals ? d_totals : (double*)((char*)partial + pbytes);
    if (levels_in_p)
        hipLaunchKernelGGL(chem_update_kernel<true>, dim3(nb), dim3(kThreads), 0, st, cp, d_frag_col, d_frag_ela,
                           d_grads[0], d_grads[1], d_grads[2], n, chunk, partial);
    else
        hipLaunchKernelGGL(chem_update_kernel<false>, dim3(nb), dim3(kThreads), 0, st, cp, d_frag_col, d_frag_ela,
                           d_grads[0], d_grads[1], d_grads[2], n, chunk, partial);
    int rc = hip_check(hipGetLastError(), "chem update launch");
    if (!rc) {
        hipLaunchKernelGGL(chem_total_kernel, dim3(kChems), dim3(kThreads), 0, st, partial, nb, d_tot);
        rc = hip_check(hipGetLastError(), "chem totals launch");
    }
    if (!rc && totals) {
        rc = hip_check(hipMemcpyAsync(totals, d_tot, kChems * sizeof(double), hipMemcpyDeviceToHost, st),
                       "chem totals download");
        if (!rc) rc = hip_check(hipStreamSynchronize(st), "chem totals sync");
    }
    if (!buf) (void)hipFreeAsync(partial, st);
    return rc;
}

// ------------------------------------------------ allocation-free variant
// The partial sums of vfd_diffuse_update_noalloc live in this file-scope
// array, which each GPU the library runs on holds until the process ends.  In
// host code `scratch` names a host-side stand-in: kernels get the device
// address of each GPU's copy, looked up once per device (scratch_addr).
static __device__ __align__(256) unsigned char scratch[kPassScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices];

// device address of the current device's `scratch`, looked up on the first
// call on that device with the device current
static int scratch_addr(void** out) {
    int dev = 0;
    if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices) return set_error(VFD_ERR_UNSUPPORTED, "scratch: device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return rc;
        if (!p) return set_error(VFD_ERR_DEVICE, "scratch: null device address");
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return VFD_OK;
}

namespace vfd {
// one chem's level update (the first half of vfd_chem_update for that chem)
int chem_levels(float* d_p, const float* d_t, float* d_d, float decay, long n, hipStream_t st) {
    if (n <= 0) return VFD_OK;
    const unsigned blocks = (unsigned)std::min<long>(kBlocks * 4, (n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(chem_levels_kernel, dim3(blocks), dim3(kThreads), 0, st, d_p, d_t, d_d, decay, n);
    return hip_check(hipGetLastError(), "chem levels launch");
}

// display scales per chem (volumeRender_kernel.cu:258-294)
float chem_display_scale(int chem) {
    static const float scale[kChems] = {50000.0f, 10000.0f, 1000000.0f, 100.0f, 10000.0f, 100000.0f, 100000.0f, 10000.0f};
    return scale[chem];
}

// gradients and totals of 8 levels the x-inverse pass has already written
int chem_grads_totals(float* const* d_p, const float* d_frag_col, const float* d_frag_ela, float* const* d_grads,
                      long n, double* totals, hipStream_t st) {
    if (!d_p || !d_frag_col || !d_frag_ela || !d_grads) return set_error(VFD_ERR_INVALID, "chem update: null argument");
    ChemPtrs cp{};
    for (int ic = 0; ic < kChems; ++ic) {
        if (!d_p[ic]) return set_error(VFD_ERR_INVALID, "chem update: chem %d has a null field", ic);
        cp.p[ic] = d_p[ic];
    }
    for (int g = 0; g < 3; ++g)
        if (!d_grads[g]) return set_error(VFD_ERR_INVALID, "chem update: gradient %d is null", g);
    return chem_pass(cp, true, d_frag_col, d_frag_ela, d_grads, n, totals, st);
}
}  // namespace vfd

extern "C" int vfd_chem_update(const float* const* d_t, float* const* d_d, float* const* d_p, const float* decay,
                               const float* d_frag_col, const float* d_frag_ela, float* const* d_grads, long n,
                               double* totals, void* stream) {
    if (!d_t || !d_d || !d_p || !decay || !d_frag_col || !d_frag_ela || !d_grads)
        return set_error(VFD_ERR_INVALID, "chem_update: null argument");
    if (n < 0) return set_error(VFD_ERR_INVALID, "chem_update: negative voxel count");
    ChemPtrs cp;
    for (int ic = 0; ic < kChems; ++ic) {
        if (!d_t[ic] || !d_d[ic] || !d_p[ic])
            return set_error(VFD_ERR_INVALID, "chem_update: chem %d has a null field", ic);
        cp.t[ic] = d_t[ic];
        cp.d[ic] = d_d[ic];
        cp.p[ic] = d_p[ic];
        cp.decay[ic] = decay[ic];
    }
    for (int g = 0; g < 3; ++g)
        if (!d_grads[g]) return set_error(VFD_ERR_INVALID, "chem_update: gradient %d is null", g);
    return chem_pass(cp, false, d_frag_col, d_frag_ela, d_grads, n, totals, (hipStream_t)stream);
}

// vfd_diffuse_update without any device allocation: the same kernels (and so
// the same bits), with the partial sums of the totals in `scratch` and the
// totals written to device memory.  Calls that overlap on one GPU share its
// `scratch` and race; nothing in it needs clearing (every partial sum a call
// reads, that call has written).
extern "C" int vfd_diffuse_update_noalloc(const vfd_plan* plan, const vfd_field* fields, float* const* d_delta,
                                          const float* decay, const float* d_frag_col, const float* d_frag_ela,
                                          float* const* d_grads, double* d_totals, void* stream) {
    if (!plan || !fields || !d_delta || !decay || !d_frag_col || !d_frag_ela || !d_grads)
        return set_error(VFD_ERR_INVALID, "diffuse_update_noalloc: null argument");
    const Plan& p = plan->p;
    if (p.prec != VFD_F32) return set_error(VFD_ERR_UNSUPPORTED, "diffuse_update_noalloc: FP32 plans only");
    ChemPtrs cp{};
    for (int ic = 0; ic < kChems; ++ic) {
        if (!d_delta[ic]) return set_error(VFD_ERR_INVALID, "diffuse_update_noalloc: chem %d has no delta field", ic);
        if (!fields[ic].out) return set_error(VFD_ERR_INVALID, "diffuse_update_noalloc: chem %d has no output", ic);
        cp.p[ic] = (float*)fields[ic].out;
    }
    for (int g = 0; g < 3; ++g)
        if (!d_grads[g]) return set_error(VFD_ERR_INVALID, "diffuse_update_noalloc: gradient %d is null", g);
    void* buf = nullptr;
    if (int rc = scratch_addr(&buf)) return rc;
    hipStream_t st = (hipStream_t)stream;
    Epilogue ep;
    ep.deltas = d_delta;
    ep.decay = decay;
    if (int rc = plan_diffuse_passes(p, fields, kChems, 0, 5, st, &ep)) return rc;
    const long n = (long)p.nx * p.ny * p.nz;
    return chem_pass(cp, true, d_frag_col, d_frag_ela, d_grads, n, nullptr, st, buf, d_totals);
}
