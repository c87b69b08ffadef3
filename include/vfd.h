/*
 * vfd.h -- native C-ABI of the MI355X cytokine-diffusion engine.
 *
 * One diffusion tick of a 3D field u (row-major [nz][ny][nx], x fastest) is the
 * operator of the reference's fftDiffuse3D (src/Diffusion/Diffusion.cu:1622-1934):
 *
 *     v = unpad( C2R( R2C(pad(u)) * K^ ) / (px*py*pz) )
 *
 * computed as five HBM passes (pad+R2C-x, C2C-y, C2C-z*K^*C2C-z^-1, C2C-y^-1,
 * C2R-x+unpad).  Everything here is additive to the reference's interface:
 * the reference-compatible entry points (c_ctx, fftDiffuse3D, ...) are in
 * vf_diffusion.h and are implemented on top of these.
 *
 * Conventions: all device pointers are HIP device memory on the current
 * device; `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Functions return VFD_OK (0) or a negative VFD_ERR_*; vfd_last_error()
 * returns the message of the calling thread's last failure.  Nothing in the
 * library calls exit() or abort().
 */
#ifndef VFD_H
#define VFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VFD_OK = 0,
    VFD_ERR_INVALID = -1,     /* bad argument / dimensions */
    VFD_ERR_UNSUPPORTED = -2, /* transform length not compiled in */
    VFD_ERR_DEVICE = -3,      /* HIP runtime error */
    VFD_ERR_NOMEM = -4,       /* device allocation failed */
    VFD_ERR_BUDGET = -5,      /* HBM budget of an engine exceeded */
    VFD_ERR_NOTFOUND = -6,    /* unknown run / ticket */
    VFD_ERR_STATE = -7        /* call not valid in the current state */
};

enum { VFD_F32 = 4, VFD_F64 = 8 };
/* pad modes: VFD_PAD_CONSTANT_VF is the reference's pConstantVF
 * (src/enums.h:39, convolutionFFT.cuh:590-638); VFD_PAD_ZERO is the zero pad of
 * a linear convolution. */
enum { VFD_PAD_CONSTANT_VF = 3, VFD_PAD_ZERO = 100 };
/* kernel-spectrum formats:
 *  VFD_KSPEC_REAL        real K^ of an even kernel, [pz][py][ld] (vfd_plan_info ld)
 *  VFD_KSPEC_COMPLEX_STD complex K^ in the R2C layout [pz][py][px/2+1] of the
 *                        reference (d_kernelspectrum_h, Diffusion.cu:1605)
 *  VFD_KSPEC_SEPARABLE   a separable even kernel as FP64 tables
 *                        [scale | gx[px/2+1] | gy[py] | gz[pz] | R | g_0..g_16],
 *                        gz being the DFT of the normalized z window g of radius
 *                        R.  FP64 plans (and FP32 windows with R > 16): the z
 *                        pass forms K^(kz,ky,kx) = ((gx*gy)*scale)*gz in FP64
 *                        and rounds it to the plan's precision, the same numbers
 *                        VFD_KSPEC_REAL stores for the Gaussian.  FP32 plans with
 *                        R <= 16: the z pass convolves each z column with g
 *                        directly (the same operator; agrees with the FFT form to
 *                        FP32 rounding).  The form is chosen on the device from
 *                        the R stored in the table, so a table copied or moved
 *                        by any means runs the form its own R selects.
 *                        VFD_PAD_ZERO plans only. */
enum { VFD_KSPEC_REAL = 0, VFD_KSPEC_COMPLEX_STD = 1, VFD_KSPEC_SEPARABLE = 2 };

typedef struct vfd_plan vfd_plan;

typedef struct {
    int nz, ny, nx;  /* data box */
    int pz, py, px;  /* FFT box; 0 picks the smallest compiled length >= 2*n (a
                        zero-padded linear convolution for kernel radius <= P - n) */
    int precision;   /* VFD_F32 or VFD_F64 */
    int pad_mode;    /* VFD_PAD_ZERO or VFD_PAD_CONSTANT_VF */
    int kspec_format;/* VFD_KSPEC_REAL, VFD_KSPEC_COMPLEX_STD or VFD_KSPEC_SEPARABLE */
} vfd_plan_desc;

typedef struct {
    int nz, ny, nx, pz, py, px;
    int L;              /* px/2: packed row length of the x passes */
    int hx;             /* px/2+1: half-spectrum row length */
    long ld;            /* work row stride in complex elements */
    int rs_y, rs_z;     /* even/odd split used on y and z (1 or 2) */
    size_t work_bytes;  /* per-field work buffer */
    size_t kspec_bytes; /* per-field kernel spectrum in the plan's format */
    double bytes_per_field_tick; /* algorithmic HBM bytes of one tick (pruned model) */
    int order;          /* pass order, VFD_ORDER_XFIRST or VFD_ORDER_XLAST */
    int y_inv_fold;     /* 1: the inverse y pass splits its py-long columns even/odd
                           with the data rows past py/2 folded in (FP32, ny > py/2) */
} vfd_plan_info;

/* Pass orders.  XFIRST: x R2C, y, z (fused forward * K^ * inverse), y inverse,
 * x C2R.  XLAST: z R2C (two real columns per complex transform), y, x fused,
 * y inverse, z C2R -- the same operator factored the other way, which keeps
 * the x padding out of HBM; FP32 plans with VFD_KSPEC_COMPLEX_STD take it when
 * it moves at most 0.9x the bytes of XFIRST (a box padded most along x, as the
 * reference's 1152x1536x256 box of a 1006x1390x110 field) and both its
 * lengths are compiled.  The environment variable VFD_ORDER=0 (1) at plan
 * creation forces XFIRST (XLAST where possible). */
enum { VFD_ORDER_XFIRST = 0, VFD_ORDER_XLAST = 1 };

typedef struct {
    const void* in;    /* real field [nz][ny][nx] */
    void* out;         /* real field [nz][ny][nx]; may equal in */
    void* work;        /* >= work_bytes, private to this field during the call */
    const void* kspec; /* kernel spectrum in the plan's format */
    double base;       /* constant-pad value (VFD_PAD_CONSTANT_VF) */
    int epi;           /* epithelium index: interior x < epi (VFD_PAD_CONSTANT_VF) */
    int reserved;
} vfd_field;

const char* vfd_last_error(void);
int vfd_version(void);
/* 1 when a transform of length q is compiled in. */
int vfd_length_supported(int q);

int vfd_plan_create(const vfd_plan_desc* desc, vfd_plan** plan);
void vfd_plan_destroy(vfd_plan* plan);
int vfd_plan_get_info(const vfd_plan* plan, vfd_plan_info* info);

/* One tick for nfields fields of this plan, asynchronous on stream.  Fields are
 * batched into launches of up to 32. */
int vfd_diffuse(const vfd_plan* plan, const vfd_field* fields, int nfields, void* stream);

/* The five passes of vfd_diffuse, one at a time (same arguments), in order
 * VFD_PASS_XFWD, VFD_PASS_YFWD, VFD_PASS_ZCONV, VFD_PASS_YINV, VFD_PASS_XINV;
 * for per-pass timing.  In the XLAST order the same five indices run z R2C,
 * y forward, the fused x pass, y inverse and z C2R.  vfd_plan_pass_bytes gives the algorithmic HBM bytes
 * of one pass for one field (sum over passes = bytes_per_field_tick). */
enum { VFD_PASS_XFWD = 0, VFD_PASS_YFWD = 1, VFD_PASS_ZCONV = 2, VFD_PASS_YINV = 3, VFD_PASS_XINV = 4 };
/* The single launches of the z pass, for per-kernel timing: the direct form
 * of radius bound 8, 12 and 16 (FP32 separable plans; no launch otherwise) and
 * the FFT form.  Each covers every field and acts on the fields of its own
 * form, so issuing all four equals VFD_PASS_ZCONV.  vfd_plan_pass_bytes gives
 * the bytes of one field of that form. */
enum { VFD_PASS_Z_DIRECT8 = 5, VFD_PASS_Z_DIRECT12 = 6, VFD_PASS_Z_DIRECT16 = 7, VFD_PASS_Z_FFT = 8 };
int vfd_diffuse_pass(const vfd_plan* plan, const vfd_field* fields, int nfields, int pass, void* stream);
double vfd_plan_pass_bytes(const vfd_plan* plan, int pass);

/* Device-to-device copy of `bytes` (a multiple of 16, below 4 GiB) with one
 * 16-byte non-temporal load and store per lane: the streaming-bandwidth
 * reference the passes are measured against.  Asynchronous. */
int vfd_stream_copy(void* dst, const void* src, size_t bytes, void* stream);
/* The same streaming form for other read:write mixes, the references for
 * passes that read one byte per two written (dst1 given: dst0 = dst1 = src0)
 * or two per one (src1 given: dst0 = src0 + src1 as 32-bit integer lanes).
 * With neither it is vfd_stream_copy.  Buffers of `bytes` each.  Asynchronous. */
int vfd_stream_mix(void* dst0, void* dst1, const void* src0, const void* src1, size_t bytes, void* stream);

/* Schedule autotuner: every compiled schedule variant (radix 16/8, work-group
 * size) of every pass is timed on scratch buffers for nfields fields; the
 * fastest per pass is kept in the plan and remembered for plans of the same
 * shape created later in the process.  best_ms (optional, 5 floats) receives
 * the winning per-pass times.  Allocates scratch; synchronizes stream. */
int vfd_plan_autotune(vfd_plan* plan, int nfields, void* stream, float* best_ms);
int vfd_plan_get_schedule(const vfd_plan* plan, int* variants);  /* 5 ints */
int vfd_plan_set_schedule(vfd_plan* plan, const int* variants);  /* 5 ints, 0..3 */

/* Kernel spectrum of the normalized sampled Gaussian g(j) ~ exp(-j^2/(2 sigma2)),
 * |j| <= ceil(6 sqrt(sigma2)) per axis, times decay, with 1/(px py pz) folded
 * in: the 3D array in the plan's precision (VFD_KSPEC_REAL) or the FP64 tables
 * (VFD_KSPEC_SEPARABLE), computed in FP64.  Both formats give bit-identical
 * ticks in FP64 and agree to FP32 rounding in FP32 (direct z pass). */
int vfd_gauss_spectrum(const vfd_plan* plan, double sigma2, double decay, void* d_kspec, void* stream);

/* Compensated (Neumaier, FP64) sum of each of nfields real fields of the plan's
 * data box; fixed reduction order (bitwise reproducible).  Synchronizes stream. */
int vfd_field_mass(const vfd_plan* plan, const void* const* d_fields, int nfields, double* mass,
                   void* stream);

/* Standalone 3D transforms on a full box (px,py,pz of the plan, data box = box):
 * R2C real [pz][py][px] -> complex R2C layout [pz][py][px/2+1]; C2R the inverse
 * (unnormalized).  d_out / d_in of the complex side must hold
 * pz*py*(px/2+1) complex elements.  Plans for these must have n == p. */
int vfd_r2c_3d(const vfd_plan* plan, const void* d_real, void* d_complex, void* stream);
int vfd_c2r_3d(const vfd_plan* plan, void* d_complex, void* d_real, void* stream);

/* Post-diffusion chem update: WHChemical::update with gpuUpdate
 * (src/FieldVariable/Usr_FieldVariables/WHChemical.cpp:395-450) on the device.
 * For the 8 chems (enums.h:124-133: TNF, TGF, FGF, MMP8, IL1beta, IL6, IL8,
 * IL10) and n voxels: p = d + t * decay[ic] (FP32, multiply and add rounded
 * separately), d = 0; then the gradients (d_grads in the order NEU, MAC, FIB,
 * enums.h:134-136):
 *   g   = p[IL1beta] + p[TNF] + p[TGF] + p[FGF] + frag_col   (left to right)
 *   NEU = g + p[IL6] + p[IL8];  MAC = g + frag_ela;  FIB = p[TGF].
 * d_t: the diffused fields (tChem); d_d: the agents' deltas (dChem, cleared);
 * d_p: the new levels (pChem; may alias d_t).  totals (8 doubles, optional)
 * receives compensated FP64 sums of each p; a non-null totals synchronizes
 * stream. */
int vfd_chem_update(const float* const* d_t, float* const* d_d, float* const* d_p, const float* decay,
                    const float* d_frag_col, const float* d_frag_ela, float* const* d_grads, long n,
                    double* totals, void* stream);

/* One tick of the 8 chem fields with the chem update fused into it: the
 * diffusion of vfd_diffuse, whose x-inverse pass writes fields[ic].out =
 * d_delta[ic] + t * decay[ic] (t the diffused field, FP32 product rounded
 * before the add) and clears d_delta[ic]; then the gradients and compensated
 * totals of vfd_chem_update from those levels.  Bit-identical to vfd_diffuse
 * followed by vfd_chem_update, with one pass over the fields fewer.  FP32
 * plans; fields and deltas in the chem order of vfd_chem_update. */
int vfd_diffuse_update(const vfd_plan* plan, const vfd_field* fields, float* const* d_delta, const float* decay,
                       const float* d_frag_col, const float* d_frag_ela, float* const* d_grads, double* totals,
                       void* stream);

/* vfd_diffuse_update without any device allocation, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops).  No counterpart
 * in the reference.  Same kernels and bits as vfd_diffuse_update; the partial
 * sums of the totals go to a fixed library-owned device array (128 KiB + 64 B
 * per GPU, held until the process ends; any voxel count fits), and the 8
 * totals to d_totals, device memory (optional).  Asynchronous: never
 * synchronizes.  Calls that overlap on one GPU share that array and race. */
int vfd_diffuse_update_noalloc(const vfd_plan* plan, const vfd_field* fields, float* const* d_delta,
                               const float* decay, const float* d_frag_col, const float* d_frag_ela,
                               float* const* d_grads, double* d_totals, void* stream);

/* Visualisation downsample: sampleChem (src/Visualization/3D/volumeRender_kernel.cu:
 * 215-322) into a dense array instead of a surface.  d_dst[dz][dy][dx] =
 * min(d_src[stride*dz][stride*dy][stride*dx] * scale[chem_index], 1) for
 * dz < D/stride, dy < H/stride, dx < W/stride, with the reference's per-chem
 * display scales {5e4, 1e4, 1e6, 100, 1e4, 1e5, 1e5, 1e4}.  Asynchronous. */
int vfd_sample_chem(const float* d_src, int D, int H, int W, int chem_index, int stride, float* d_dst,
                    void* stream);

/* vfd_diffuse with the visualisation downsample of vfd_sample_chem fused into
 * the x-inverse pass (the unpad epilogue, where the reference samples the
 * diffused field: Diffusion.cu:1878-1903): d_vis[i] receives the downsample of
 * fields[i]'s diffused output with chem chem_index[i]'s display scale.
 * Bit-identical to vfd_diffuse followed by vfd_sample_chem on each output.
 * FP32 plans; stride >= 1.  Asynchronous. */
int vfd_diffuse_sample(const vfd_plan* plan, const vfd_field* fields, int nfields, const int* chem_index, int stride,
                       float* const* d_vis, void* stream);

/* ======================================================================
 * Engine: resident runs, batched ticks, serving.
 *
 * A run is one ABM instance: nfields fields of one grid (zero-padded linear
 * convolution on the 2n box) with their Gaussian kernels (separable FP64
 * spectrum tables, VFD_KSPEC_SEPARABLE), resident in
 * HBM for the run's lifetime.  Runs of every grid shape share per-shape plans
 * and work-buffer pools.  Everything the engine allocates is charged against
 * its HBM budget; a run that does not fit is refused with VFD_ERR_BUDGET.
 * ====================================================================== */
typedef struct vfd_engine vfd_engine;

typedef struct {
    int nz, ny, nx;        /* grid */
    int nfields;           /* 1..32 */
    int precision;         /* VFD_F32 or VFD_F64 */
    const double* sigma2;  /* nfields Gaussian variances, sigma^2 = 2 D dt */
    const double* decay;   /* nfields decay factors exp(-lambda dt) */
} vfd_run_desc;

typedef struct {
    size_t budget_bytes, used_bytes;
    int runs, shapes;
    uint64_t batches, ticks, requests_done, requests_pending;
    size_t peak_used_bytes;            /* high-water mark of used_bytes */
    uint64_t runs_admitted_after_wait; /* vfd_run_create_wait calls that waited for memory */
} vfd_engine_stats;

int vfd_engine_create(size_t hbm_budget_bytes, vfd_engine** engine);
void vfd_engine_destroy(vfd_engine* engine);
/* Device bytes a run of this description holds while resident (fields and
 * spectra; work buffers are pooled per shape and charged when grown). */
size_t vfd_run_bytes(const vfd_run_desc* desc);
int vfd_run_create(vfd_engine* engine, const vfd_run_desc* desc, int* run_id);
/* vfd_run_create that waits for memory: when the run does not fit the budget,
 * the call blocks until runs destroyed on other threads free enough of it,
 * for at most timeout_s seconds (then VFD_ERR_BUDGET).  A run larger than
 * the whole budget is refused at once. */
int vfd_run_create_wait(vfd_engine* engine, const vfd_run_desc* desc, double timeout_s, int* run_id);
int vfd_run_destroy(vfd_engine* engine, int run_id);
/* device pointer of a resident field (row-major [nz][ny][nx], precision of the run) */
int vfd_run_field_ptr(vfd_engine* engine, int run_id, int field, void** d_field);
/* Copy one field in / out (host or device memory, nz*ny*nx values of the run's
 * precision).  Asynchronous on stream; host memory must stay valid until the
 * stream reaches the copy. */
int vfd_run_write_field(vfd_engine* engine, int run_id, int field, const void* src, void* stream);
int vfd_run_read_field(vfd_engine* engine, int run_id, int field, void* dst, void* stream);

/* Advance runs by `ticks` ticks: every field of every listed run, batched by
 * grid shape (one launch per pass and shape, shapes on parallel branches).
 * use_graph = 1 captures the tick into a hipGraph (cached per run set) and
 * replays it.  Asynchronous on stream. */
int vfd_step(vfd_engine* engine, const int* run_ids, int nruns, int ticks, int use_graph, void* stream);
/* The same work one field at a time with per-pass launches and no graph
 * (the plain per-field form, kept for comparison). */
int vfd_step_plain(vfd_engine* engine, const int* run_ids, int nruns, int ticks, void* stream);

/* Serving: a background loop takes one-tick requests for resident runs,
 * admits them into the next batch as soon as an execution slot is free
 * (continuous batching, `slots` batches in flight on separate streams), and
 * records per-request latency from submission to result ready.  A run is in
 * at most one in-flight batch; its requests run in submission order. */
int vfd_serve_start(vfd_engine* engine, int slots, int max_batch_fields);
int vfd_serve_stop(vfd_engine* engine);
int vfd_submit(vfd_engine* engine, int run_id, uint64_t* ticket);
/* done = 1 once the tick is complete; latency_ms from submission to completion */
int vfd_poll(vfd_engine* engine, uint64_t ticket, int* done, double* latency_ms);
int vfd_wait(vfd_engine* engine, uint64_t ticket, double* latency_ms);
int vfd_engine_get_stats(vfd_engine* engine, vfd_engine_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* VFD_H */
