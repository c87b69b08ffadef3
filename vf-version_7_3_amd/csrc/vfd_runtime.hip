// vfd_runtime.hip -- plans, twiddle tables, kernel spectra, mass sums and the
// native C-ABI (include/vfd.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vfd.h"
#include "vfd_fft.hpp"
#include "vfd_internal.h"
#include "vfd_launch.h"

namespace vfd {

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return VFD_OK;
    return set_error(e == hipErrorOutOfMemory ? VFD_ERR_NOMEM : VFD_ERR_DEVICE, "%s: %s", what,
                     hipGetErrorString(e));
}

// --------------------------------------------------------- twiddle tables
// W_P^k = exp(-2 pi i k / P), k in [0, P), computed in long double on the host,
// one table per (device, P, precision), kept for the life of the process.
struct TwKey {
    int dev, p, prec;
    bool operator<(const TwKey& o) const {
        return dev != o.dev ? dev < o.dev : (p != o.p ? p < o.p : prec < o.prec);
    }
};
static std::mutex g_tw_mu;
static std::map<TwKey, void*> g_tw;

int twiddles(int p, int prec, const void** out) {
    int dev = 0;
    if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice")) return rc;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_tw.find({dev, p, prec});
    if (it != g_tw.end()) {
        *out = it->second;
        return VFD_OK;
    }
    const size_t bytes = (size_t)p * 2 * prec;
    std::vector<unsigned char> host(bytes);
    const long double tp = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < p; ++k) {
        long double a = tp * (long double)k / (long double)p;
        double c = (double)cosl(a), s = (double)-sinl(a);
        if (prec == VFD_F32) {
            // correctly rounded components (the nearest complex value): table
            // twiddles differ between a frequency's forward and inverse paths
            // in the x and y passes, so their phase errors matter as much as
            // their magnitude errors (picking pairs for |w| = 1 made the
            // 100-tick precision run's rel-L2 worse, 9.6e-6 -> 1.1e-5); their
            // magnitude errors have no common sign (mean |w|^2 - 1 ~ -3e-9)
            float* f = (float*)host.data();
            f[2 * k] = (float)c;
            f[2 * k + 1] = (float)s;
        } else {
            double* d = (double*)host.data();
            d[2 * k] = c;
            d[2 * k + 1] = s;
        }
    }
    void* d = nullptr;
    if (int rc = hip_check(hipMalloc(&d, bytes), "hipMalloc(twiddles)")) return rc;
    if (int rc = hip_check(hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice), "twiddle upload")) {
        (void)hipFree(d);
        return rc;
    }
    g_tw[{dev, p, prec}] = d;
    *out = d;
    return VFD_OK;
}

// ------------------------------------------------------- schedule cache
// Tuned per-pass schedule variants, keyed by device and plan shape.
struct ShapeKey {
    int dev, prec, pad, kfmt, nz, ny, nx, pz, py, px, order;
    bool operator<(const ShapeKey& o) const {
        const int a[11] = {dev, prec, pad, kfmt, nz, ny, nx, pz, py, px, order};
        const int b[11] = {o.dev, o.prec, o.pad, o.kfmt, o.nz, o.ny, o.nx, o.pz, o.py, o.px, o.order};
        for (int i = 0; i < 11; ++i)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};
static std::mutex g_sched_mu;
static std::map<ShapeKey, std::array<int, 5>> g_sched;

static ShapeKey shape_key(const Plan& q) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return ShapeKey{dev, q.prec, q.pad_mode, q.kspec_format, q.nz, q.ny, q.nx, q.pz, q.py, q.px, q.order};
}

void tuned_lookup(Plan& q) {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    auto it = g_sched.find(shape_key(q));
    if (it != g_sched.end())
        for (int i = 0; i < 5; ++i) q.var[i] = it->second[i];
}

bool tuned_known(const Plan& q) {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    return g_sched.count(shape_key(q)) != 0;
}

// ------------------------------------------------------------------ plans
static long round_up(long a, long b) { return (a + b - 1) / b * b; }

// Smallest FFT length >= 2n that a pass can run: even with a compiled half
// length for x (R2C via the half-length complex FFT); for y and z a compiled
// half length (even/odd split) or a compiled full length.
static int default_box(int n, bool x_axis) {
    if (n < 1) return 2 * n;
    for (int p = 2 * n; p <= 2 * n + 4096; ++p) {
        if (x_axis) {
            if (p % 2 == 0 && length_supported(p / 2)) return p;
        } else if ((p % 2 == 0 && length_supported(p / 2)) || length_supported(p)) {
            return p;
        }
    }
    return 2 * n;
}

int plan_init(const vfd_plan_desc* d, Plan* p) {
    if (!d) return set_error(VFD_ERR_INVALID, "plan: null descriptor");
    Plan q{};
    q.nz = d->nz; q.ny = d->ny; q.nx = d->nx;
    q.pz = d->pz ? d->pz : default_box(d->nz, false);
    q.py = d->py ? d->py : default_box(d->ny, false);
    q.px = d->px ? d->px : default_box(d->nx, true);
    q.prec = d->precision;
    q.pad_mode = d->pad_mode;
    q.kspec_format = d->kspec_format;
    if (q.nx < 1 || q.ny < 1 || q.nz < 1)
        return set_error(VFD_ERR_INVALID, "plan: data box %dx%dx%d must be positive", q.nx, q.ny, q.nz);
    if (q.px < q.nx || q.py < q.ny || q.pz < q.nz)
        return set_error(VFD_ERR_INVALID, "plan: FFT box %dx%dx%d smaller than data box %dx%dx%d", q.px, q.py,
                         q.pz, q.nx, q.ny, q.nz);
    if (q.px % 2)
        return set_error(VFD_ERR_UNSUPPORTED, "plan: px=%d must be even (R2C via half-length complex FFT)", q.px);
    if (q.prec != VFD_F32 && q.prec != VFD_F64)
        return set_error(VFD_ERR_INVALID, "plan: precision must be VFD_F32 or VFD_F64");
    if (q.pad_mode != VFD_PAD_ZERO && q.pad_mode != VFD_PAD_CONSTANT_VF)
        return set_error(VFD_ERR_INVALID, "plan: unknown pad mode %d", q.pad_mode);
    if (q.kspec_format != VFD_KSPEC_REAL && q.kspec_format != VFD_KSPEC_COMPLEX_STD &&
        q.kspec_format != VFD_KSPEC_SEPARABLE)
        return set_error(VFD_ERR_INVALID, "plan: unknown kernel-spectrum format %d", q.kspec_format);
    if (q.kspec_format == VFD_KSPEC_SEPARABLE && q.pad_mode != VFD_PAD_ZERO)
        return set_error(VFD_ERR_INVALID, "plan: the separable spectrum format needs VFD_PAD_ZERO");
    q.L = q.px / 2;
    q.hx = q.L + 1;
    if (!length_supported(q.L))
        return set_error(VFD_ERR_UNSUPPORTED, "plan: x transform length %d (px/2) not compiled in", q.L);
    auto pick = [](int n, int p, int* rs, int* qq) -> bool {
        if (p % 2 == 0 && n <= p / 2 && length_supported(p / 2)) {
            *rs = 2;
            *qq = p / 2;
            return true;
        }
        *rs = 1;
        *qq = p;
        return length_supported(p);
    };
    if (!pick(q.ny, q.py, &q.rs_y, &q.qy))
        return set_error(VFD_ERR_UNSUPPORTED, "plan: y transform length %d not compiled in", q.py);
    // FP32 y columns longer than half their box (the ABM's 1390 rows on 1536):
    // the inverse y pass takes the folded even/odd split where it is compiled,
    // half the LDS per column and twice the columns per tile (production box:
    // 0.38 -> 0.35 ms per field).  The forward pass keeps the unsplit
    // transform: folded, its two half-passes write alternate rows of 64-byte
    // segments and it ran slower (0.40 -> 0.53 ms).
    q.fold_y = q.prec == VFD_F32 && q.rs_y == 1 && q.py % 2 == 0 && q.ny > q.py / 2 && fold_supported(q.py / 2);
    if (!pick(q.nz, q.pz, &q.rs_z, &q.qz))
        return set_error(VFD_ERR_UNSUPPORTED, "plan: z transform length %d not compiled in", q.pz);
    q.full_box = (q.nx == q.px && q.ny == q.py && q.nz == q.pz);
    // rows start on 128-byte lines for both the complex work rows and the real
    // kernel-spectrum rows (32 elements)
    q.ld = q.full_box ? q.hx : round_up(q.hx, 32);
    // (padding the plane stride by 16 to 2064 elements to spread the z-column
    // rows over HBM channels was measured: no gain)
    q.plane = (long)q.py * q.ld;
    const size_t cxb = 2 * (size_t)q.prec;
    q.work_bytes = (size_t)q.nz * q.plane * cxb;
    if (q.kspec_format == VFD_KSPEC_REAL) {
        // element (kz,ky,kx) at ((kz/rs)*py + ky)*ld*rs + kx*rs + kz%rs: the rs
        // frequencies one z sub-transform pair needs sit side by side
        q.ks_row = (long)q.py * q.ld * q.rs_z;  // per kz/rs
        q.ks_plane = q.ld * q.rs_z;             // per ky
        q.kspec_bytes = (size_t)q.pz * q.py * q.ld * q.prec;
    } else if (q.kspec_format == VFD_KSPEC_SEPARABLE) {
        // FP64 tables [scale | gx[hx] | gy[py] | gz[pz] | R | g_0 .. g_kZDirectMaxR]
        // (the last part: the z window's radius and taps for the direct z pass)
        q.ks_row = 0;
        q.ks_plane = 0;
        q.kspec_bytes = round_up((long)(1 + q.hx + q.py + q.pz + 2 + kZDirectMaxR), 32) * sizeof(double);
    } else {
        q.ks_row = (long)q.py * q.hx;
        q.ks_plane = q.hx;
        q.kspec_bytes = (size_t)q.pz * q.py * q.hx * cxb;
    }
    q.field_bytes = (size_t)q.nx * q.ny * q.nz * q.prec;
    // FP32 passes address each array with 32-bit buffer offsets; FP64 passes
    // use 64-bit addresses
    if (q.prec == VFD_F32 &&
        (q.work_bytes >= kBufLimit || q.kspec_bytes >= kBufLimit || q.field_bytes >= kBufLimit))
        return set_error(VFD_ERR_UNSUPPORTED,
                         "plan: a per-field FP32 array exceeds 4 GiB (work %zu, spectrum %zu, field %zu bytes)",
                         q.work_bytes, q.kspec_bytes, q.field_bytes);
    // pruned 5-pass algorithmic bytes: 8N + 32 A1 + 32 A2 + 4 A3 (FP32); the
    // complex terms scale with the precision and K^ with its element size.
    const double N = (double)q.nx * q.ny * q.nz;
    const double A1 = (double)q.hx * q.ny * q.nz, A2 = (double)q.hx * q.py * q.nz,
                 A3 = (double)q.hx * q.py * q.pz;
    const double r = (double)q.prec, c = 2.0 * q.prec;
    const double kbytes = q.kspec_format == VFD_KSPEC_REAL          ? r * A3
                          : q.kspec_format == VFD_KSPEC_COMPLEX_STD ? c * A3
                                                                     : (double)q.kspec_bytes;
    q.bytes_per_field_tick = 2 * r * N + 4 * c * A1 + 4 * c * A2 + kbytes;
    // x-last order: z R2C over column pairs (hz planes of ny rows of 2 nxp
    // columns), y to py rows, x fused with K^ read at every kx (px per row)
    q.order = VFD_ORDER_XFIRST;
    const char* so = getenv("VFD_ORDER");
    const int want = so && *so ? atoi(so) : -1;
    if (want != 0 && q.prec == VFD_F32 && q.kspec_format == VFD_KSPEC_COMPLEX_STD && !q.full_box &&
        xlast_supported(q.pz, q.px)) {
        const int hz = q.pz / 2 + 1;
        const long ldx = round_up(q.nx, 16);  // even, > nx when nx is odd; rows on 128-byte lines
        const double nxp2 = 2.0 * ((q.nx + 1) / 2);
        const double B1 = (double)q.nx * q.ny * hz, B1p = nxp2 * q.ny * hz, B2 = (double)q.nx * q.py * hz;
        const double xl = 2 * r * N + 2 * c * B1p + 2 * c * B1 + 4 * c * B2 + c * (double)hz * q.py * q.px;
        const size_t wb = (size_t)hz * q.py * ldx * cxb;
        if ((want == 1 || xl <= 0.9 * q.bytes_per_field_tick) && wb < kBufLimit) {
            q.order = VFD_ORDER_XLAST;
            q.hz = hz;
            q.ldx = ldx;
            q.xplane = (long)q.py * ldx;
            q.work_bytes = wb;
            q.bytes_per_field_tick = xl;
        }
    }
    // schedule variants per pass: tuned choice for this shape if known, else
    // VFD_SCHEDULE="x,y,z,yi,xi" or schedule 0
    for (int i = 0; i < 5; ++i) q.var[i] = 0;
    if (const char* sv = getenv("VFD_SCHEDULE")) {
        int v[5];
        if (sscanf(sv, "%d,%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3], &v[4]) == 5)
            for (int i = 0; i < 5; ++i) q.var[i] = (v[i] >= 0 && v[i] < kScheduleVariants) ? v[i] : 0;
    } else {
        tuned_lookup(q);
    }
    if (int rc = twiddles(q.px, q.prec, &q.tw_x)) return rc;
    if (int rc = twiddles(q.py, q.prec, &q.tw_y)) return rc;
    if (int rc = twiddles(q.pz, q.prec, &q.tw_z)) return rc;
    *p = q;
    return VFD_OK;
}

// z pass.  FP32 fields with a separable table take the direct form when the
// table's window radius R is at most kZDirectMaxR, else the FFT form.  The
// form is decided on the device from the R the table stores: the direct
// launches (radius bounds 8, 12, 16) and the FFT launch each cover every field
// of the batch, and a work-group whose field is not of its form exits at once
// (zform_of in vfd_passes.hpp), so no host-side record of the tables is needed.
// part selects one launch (0, 1, 2: the direct form of radius bound 8, 12,
// 16; 3: the FFT form) for per-kernel timing, -1 all of them.
template <typename T>
static hipError_t launch_z(const Plan& p, const FieldBatch& fb, int nf, const ColDesc& z, hipStream_t st,
                           int part = -1) {
    if constexpr (sizeof(T) == 4) {
        if (p.kspec_format == VFD_KSPEC_SEPARABLE) {
            ColDesc zd = z;
            zd.scale = (double)p.pz;
            const int rms[3] = {8, 12, 16};
            for (int i = 0; i < 3; ++i)
                if (part < 0 || part == i)
                    if (hipError_t e = launch_zdirect(rms[i], fb, nf, zd, p.py, st)) return e;
        }
    }
    if (part >= 0 && part != 3) return hipSuccess;
    return launch_col<T>(p.qz, fb, nf, z, p.py, p.tw_z, st, p.var[VFD_PASS_ZCONV]);
}

// ------------------------------------------------------------- execution
// The x-last order's passes (FP32) over one launch's fields; the y passes are
// the x-first order's column pass on the x-last work layout.
static int run_xlast(const Plan& p, const FieldBatch& fb, int nf, int pass_lo, int pass_hi, hipStream_t st) {
    XlDesc d{};
    d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
    d.px = p.px; d.py = p.py; d.pz = p.pz;
    d.hz = p.hz;
    d.nxp = (p.nx + 1) / 2;
    d.ncolp = (long)p.ny * d.nxp;
    d.ldx = p.ldx;
    d.plane = p.xplane;
    d.ks_hx = p.hx;
    d.pad_mode = p.pad_mode;
    d.scale = 1.0 / ((double)p.px * p.py * p.pz);
    d.in_bytes = d.out_bytes = p.field_bytes;
    d.work_bytes = p.work_bytes;
    d.kspec_bytes = p.kspec_bytes;
    ColDesc y{};
    y.mode = COL_FWD; y.rs = p.rs_y; y.nin = p.ny; y.nout = p.py; y.ncols = p.nx;
    y.rstride = p.ldx; y.pstride = p.xplane; y.scale = 1.0;
    y.work_bytes = p.work_bytes; y.kspec_bytes = p.kspec_bytes;
    ColDesc yi = y;
    yi.mode = COL_INV; yi.nin = p.py; yi.nout = p.ny;
    if (p.fold_y) { yi.rs = 2; yi.fold = 1; }
    const int qyi = p.fold_y ? p.py / 2 : p.qy;
    for (int pass = pass_lo; pass < pass_hi; ++pass) {
        hipError_t e = hipSuccess;
        switch (pass) {
            case VFD_PASS_XFWD: e = launch_zr(true, p.pz, fb, nf, d, p.tw_z, st, p.var[pass]); break;
            case VFD_PASS_YFWD: e = launch_col<float>(p.qy, fb, nf, y, p.hz, p.tw_y, st, p.var[pass]); break;
            case VFD_PASS_ZCONV: e = launch_xc(p.px, fb, nf, d, p.tw_x, st, p.var[pass]); break;
            case VFD_PASS_YINV: e = launch_col<float>(qyi, fb, nf, yi, p.hz, p.tw_y, st, p.var[pass]); break;
            case VFD_PASS_XINV: e = launch_zr(false, p.pz, fb, nf, d, p.tw_z, st, p.var[pass]); break;
        }
        if (e != hipSuccess) return hip_check(e, "x-last pass launch");
    }
    return VFD_OK;
}

template <typename T>
static int run_tick(const Plan& p, const vfd_field* fields, int nfields, int pass_lo, int pass_hi,
                    hipStream_t st, const Epilogue* epi, int zpart = -1) {
    for (int f0 = 0; f0 < nfields; f0 += kMaxFieldsPerLaunch) {
        const int nf = std::min(kMaxFieldsPerLaunch, nfields - f0);
        FieldBatch fb;
        memset(&fb, 0, sizeof fb);
        for (int i = 0; i < nf; ++i) {
            const vfd_field& f = fields[f0 + i];
            fb.f[i].in = f.in;
            fb.f[i].out = f.out;
            fb.f[i].work = f.work;
            fb.f[i].kspec = f.kspec;
            fb.f[i].base = f.base;
            fb.f[i].epi = f.epi;
            if (epi && epi->deltas) {
                fb.f[i].delta = epi->deltas[f0 + i];
                fb.f[i].decay = epi->decay[f0 + i];
            }
            if (epi && epi->vis) {
                fb.f[i].vis = epi->vis[f0 + i];
                fb.f[i].vis_scale = epi->vis_scale[f0 + i];
                fb.f[i].vis_stride = epi->vis_stride;
            }
        }
        GridDesc g{p.nx, p.ny, p.nz, p.L, p.ld, p.plane, p.pad_mode, p.kspec_format == VFD_KSPEC_COMPLEX_STD,
                   p.field_bytes, p.work_bytes, p.kspec_bytes, p.field_bytes,
                   p.kspec_format == VFD_KSPEC_COMPLEX_STD ? 1.0 : (double)p.px * p.py * p.pz};
        ColDesc y{};
        y.mode = COL_FWD; y.rs = p.rs_y; y.nin = p.ny; y.nout = p.py; y.ncols = p.hx;
        y.rstride = p.ld; y.pstride = p.plane; y.scale = 1.0;
        y.work_bytes = p.work_bytes; y.kspec_bytes = p.kspec_bytes;
        ColDesc z{};
        z.mode = COL_FUSED; z.rs = p.rs_z; z.nin = p.nz; z.nout = p.nz; z.ncols = p.hx;
        z.rstride = p.plane; z.pstride = p.ld; z.ks_row = p.ks_row; z.ks_plane = p.ks_plane;
        z.kspec_complex = p.kspec_format == VFD_KSPEC_COMPLEX_STD;
        z.kspec_sep = p.kspec_format == VFD_KSPEC_SEPARABLE;
        z.sep_gy = 1 + p.hx;
        z.sep_gz = 1 + p.hx + p.py;
        z.sep_g = 1 + p.hx + p.py + p.pz;
        // the real native spectrum carries the 1/(px py pz) normalization
        z.scale = z.kspec_complex ? 1.0 / ((double)p.px * p.py * p.pz) : 1.0;
        z.work_bytes = p.work_bytes; z.kspec_bytes = p.kspec_bytes;
        ColDesc yi = y;
        yi.mode = COL_INV; yi.nin = p.py; yi.nout = p.ny;
        if (p.fold_y) { yi.rs = 2; yi.fold = 1; }
        const int qyi = p.fold_y ? p.py / 2 : p.qy;
        if constexpr (sizeof(T) == 4) {
            if (p.order == VFD_ORDER_XLAST) {
                if (int rc = run_xlast(p, fb, nf, pass_lo, pass_hi, st)) return rc;
                continue;
            }
        }
        for (int pass = pass_lo; pass < pass_hi; ++pass) {
            hipError_t e = hipSuccess;
            switch (pass) {
                case VFD_PASS_XFWD: e = launch_xfwd<T>(p.L, fb, nf, g, p.tw_x, st, p.var[pass]); break;
                case VFD_PASS_YFWD: e = launch_col<T>(p.qy, fb, nf, y, p.nz, p.tw_y, st, p.var[pass]); break;
                case VFD_PASS_ZCONV: e = launch_z<T>(p, fb, nf, z, st, zpart); break;
                case VFD_PASS_YINV: e = launch_col<T>(qyi, fb, nf, yi, p.nz, p.tw_y, st, p.var[pass]); break;
                case VFD_PASS_XINV:
                    e = launch_xinv<T>(p.L, fb, nf, g, p.tw_x, st, p.var[pass], epi && epi->vis);
                    break;
            }
            if (e != hipSuccess) return hip_check(e, "pass launch");
        }
    }
    return VFD_OK;
}

static int check_fields(const vfd_field* fields, int nfields) {
    if (nfields < 0 || (nfields > 0 && !fields)) return set_error(VFD_ERR_INVALID, "diffuse: bad field list");
    for (int i = 0; i < nfields; ++i)
        if (!fields[i].in || !fields[i].out || !fields[i].work || !fields[i].kspec)
            return set_error(VFD_ERR_INVALID, "diffuse: field %d has a null pointer", i);
    return VFD_OK;
}

int plan_diffuse_passes(const Plan& p, const vfd_field* fields, int nfields, int lo, int hi, hipStream_t st,
                        const Epilogue* epi) {
    if (int rc = check_fields(fields, nfields)) return rc;
    if (lo < 0 || hi > 5 || lo >= hi) return set_error(VFD_ERR_INVALID, "diffuse: bad pass range [%d,%d)", lo, hi);
    if (nfields == 0) return VFD_OK;
    return p.prec == VFD_F32 ? run_tick<float>(p, fields, nfields, lo, hi, st, epi)
                             : run_tick<double>(p, fields, nfields, lo, hi, st, epi);
}

// one launch of the z pass (part 0..3, see launch_z)
static int plan_diffuse_zpart(const Plan& p, const vfd_field* fields, int nfields, int part, hipStream_t st) {
    if (int rc = check_fields(fields, nfields)) return rc;
    if (nfields == 0) return VFD_OK;
    return p.prec == VFD_F32 ? run_tick<float>(p, fields, nfields, 2, 3, st, nullptr, part)
                             : run_tick<double>(p, fields, nfields, 2, 3, st, nullptr, part);
}

double plan_pass_bytes(const Plan& p, int pass) {
    const double N = (double)p.nx * p.ny * p.nz;
    if (p.order == VFD_ORDER_XLAST) {
        // z R2C writes both columns of every pair (2 nxp per row)
        const double c = 2.0 * p.prec, r = (double)p.prec;
        const double B1 = (double)p.nx * p.ny * p.hz, B1p = 2.0 * ((p.nx + 1) / 2) * p.ny * p.hz,
                     B2 = (double)p.nx * p.py * p.hz, K = (double)p.hz * p.py * p.px;
        switch (pass) {
            case VFD_PASS_XFWD: return r * N + c * B1p;
            case VFD_PASS_YFWD: return c * B1 + c * B2;
            case VFD_PASS_ZCONV:
            case VFD_PASS_Z_FFT: return 2 * c * B2 + c * K;
            case VFD_PASS_YINV: return c * B2 + c * B1;
            case VFD_PASS_XINV: return c * B1p + r * N;
            default: return 0.0;
        }
    }
    const double A1 = (double)p.hx * p.ny * p.nz, A2 = (double)p.hx * p.py * p.nz,
                 A3 = (double)p.hx * p.py * p.pz;
    const double r = (double)p.prec, c = 2.0 * p.prec;
    const double kbytes = p.kspec_format == VFD_KSPEC_REAL          ? r * A3
                          : p.kspec_format == VFD_KSPEC_COMPLEX_STD ? c * A3
                                                                     : (double)p.kspec_bytes;
    switch (pass) {
        case VFD_PASS_XFWD: return r * N + c * A1;
        case VFD_PASS_YFWD: return c * A1 + c * A2;
        case VFD_PASS_ZCONV:
        case VFD_PASS_Z_DIRECT8:
        case VFD_PASS_Z_DIRECT12:
        case VFD_PASS_Z_DIRECT16:
        case VFD_PASS_Z_FFT: return 2 * c * A2 + kbytes;  // per field of the launch's form
        case VFD_PASS_YINV: return c * A2 + c * A1;
        case VFD_PASS_XINV: return c * A1 + r * N;
        default: return 0.0;
    }
}

int plan_diffuse(const Plan& p, const vfd_field* fields, int nfields, hipStream_t st) {
    return plan_diffuse_passes(p, fields, nfields, 0, 5, st);
}

// Fills scratch with hashed values in [0, 1) (as FP32 words): timings on
// all-zero buffers are not representative of real fields.
__global__ void scratch_fill_kernel(unsigned* __restrict__ p, size_t nwords, unsigned seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)i * 0x9E3779B1u ^ seed;
        h ^= h >> 15;
        h *= 0x85EBCA77u;
        h ^= h >> 13;
        p[i] = 0x3F800000u | (h >> 9);  // 1 + mantissa: [1, 2)
    }
}

// Autotuner: time every compiled schedule variant of every pass on scratch
// buffers of nfields fields filled with non-zero data, keep the fastest per
// pass (median of 5 timed launches within steps), and remember it for this
// shape.  Every
// scratch field has its own field, spectrum and work arrays, as in a real
// step: fields sharing one input would be read from the last-level cache and
// time the row passes at a bandwidth no real step sees.
int plan_autotune(Plan& p, int nfields, hipStream_t st, float* best_ms) {
    if (nfields < 1) nfields = 1;
    if (nfields > kMaxFieldsPerLaunch) nfields = kMaxFieldsPerLaunch;
    const size_t fbytes = (size_t)p.nx * p.ny * p.nz * p.prec;
    const size_t per_field = fbytes + p.kspec_bytes + p.work_bytes;  // each a multiple of 8
    const size_t total = (size_t)nfields * per_field;
    char* buf = nullptr;
    if (int rc = hip_check(hipMalloc((void**)&buf, total), "autotune scratch")) return rc;
    hipLaunchKernelGGL(scratch_fill_kernel, dim3(2048), dim3(256), 0, st, (unsigned*)buf, total / 4, 0x5EEDu);
    int rc = hip_check(hipGetLastError(), "autotune fill");
    std::vector<vfd_field> fl(nfields);
    for (int i = 0; i < nfields; ++i) {
        char* f = buf + (size_t)i * per_field;
        fl[i] = vfd_field{f, f, f + fbytes + p.kspec_bytes, f + fbytes, 0.0, 0, 0};
    }
    if (!rc && p.kspec_format == VFD_KSPEC_SEPARABLE) {
        // real separable tables, so that the z pass runs the form their radius
        // selects: the schedule variants only matter for the FFT form, which
        // FP32 tables take beyond radius kZDirectMaxR, so the window is made
        // wider than that where the zero padding allows
        const int rmax = std::min({p.px - p.nx, p.py - p.ny, p.pz - p.nz});
        const int R = std::min(rmax, kZDirectMaxR + 3);
        for (int i = 0; i < nfields && !rc && R >= 1; ++i)
            rc = plan_gauss_spectrum(p, ((R - 0.5) / 6.0) * ((R - 0.5) / 6.0), 1.0, (void*)fl[i].kspec, st);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = hip_check(hipEventCreate(&e0), "event");
    if (!rc) rc = hip_check(hipEventCreate(&e1), "event");
    const int saved[5] = {p.var[0], p.var[1], p.var[2], p.var[3], p.var[4]};
    // a variant is timed inside whole steps (the pass between events, the
    // other passes around it): a pass repeated on its own meets the caches in
    // a state no step leaves them in (x forward: 1.54 ms for the variant
    // chosen that way against 1.31 ms for the best one within steps)
    auto step_timed = [&](int pass, float* ms) {
        int rc2 = VFD_OK;
        for (int q = 0; q < 5 && !rc2; ++q) {
            if (q == pass) rc2 = hip_check(hipEventRecord(e0, st), "event");
            if (!rc2) rc2 = plan_diffuse_passes(p, fl.data(), nfields, q, q + 1, st);
            if (!rc2 && q == pass) rc2 = hip_check(hipEventRecord(e1, st), "event");
        }
        if (!rc2 && ms) rc2 = hip_check(hipEventSynchronize(e1), "event sync");
        if (!rc2 && ms) rc2 = hip_check(hipEventElapsedTime(ms, e0, e1), "event time");
        return rc2;
    };
    for (int pass = 0; pass < 5 && !rc; ++pass) {
        float best = 1e30f;
        int bestv = saved[pass];
        for (int v = 0; v < kScheduleVariants && !rc; ++v) {
            p.var[pass] = v;
            rc = step_timed(pass, nullptr);  // warm-up
            constexpr int kReps = 5;
            float t[kReps];
            for (int k = 0; k < kReps && !rc; ++k) rc = step_timed(pass, &t[k]);
            if (rc) break;
            std::sort(t, t + kReps);
            if (t[kReps / 2] < best) {
                best = t[kReps / 2];
                bestv = v;
            }
        }
        p.var[pass] = bestv;
        if (best_ms) best_ms[pass] = best;
    }
    if (rc)
        for (int i = 0; i < 5; ++i) p.var[i] = saved[i];
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(buf);
    if (!rc) {
        std::lock_guard<std::mutex> lk(g_sched_mu);
        g_sched[shape_key(p)] = {p.var[0], p.var[1], p.var[2], p.var[3], p.var[4]};
    }
    return rc;
}

// Standalone full-box transforms (used by the kernel-spectrum precompute).
template <typename T>
static int run_r2c(const Plan& p, const void* in, void* out, hipStream_t st) {
    FieldBatch fb;
    memset(&fb, 0, sizeof fb);
    fb.f[0].in = in;
    fb.f[0].work = out;
    const size_t cbytes = (size_t)p.pz * p.py * p.hx * 2 * p.prec;
    GridDesc g{p.nx, p.ny, p.nz, p.L, p.ld, p.plane, VFD_PAD_ZERO, 0, p.field_bytes, cbytes, 0, 0, 1.0};
    hipError_t e = launch_xfwd<T>(p.L, fb, 1, g, p.tw_x, st);
    if (e != hipSuccess) return hip_check(e, "r2c x launch");
    ColDesc y{};
    y.mode = COL_FWD; y.rs = 1; y.nin = p.ny; y.nout = p.py; y.ncols = p.hx;
    y.rstride = p.ld; y.pstride = p.plane; y.scale = 1.0;
    y.work_bytes = cbytes;
    e = launch_col<T>(p.py, fb, 1, y, p.nz, p.tw_y, st);
    if (e != hipSuccess) return hip_check(e, "r2c y launch");
    ColDesc z = y;
    z.nin = p.nz; z.nout = p.pz; z.rstride = p.plane; z.pstride = p.ld;
    e = launch_col<T>(p.pz, fb, 1, z, p.py, p.tw_z, st);
    return e == hipSuccess ? VFD_OK : hip_check(e, "r2c z launch");
}

template <typename T>
static int run_c2r(const Plan& p, void* in, void* out, hipStream_t st) {
    FieldBatch fb;
    memset(&fb, 0, sizeof fb);
    fb.f[0].out = out;
    fb.f[0].work = in;
    const size_t cbytes = (size_t)p.pz * p.py * p.hx * 2 * p.prec;
    ColDesc z{};
    z.mode = COL_INV; z.rs = 1; z.nin = p.pz; z.nout = p.pz; z.ncols = p.hx;
    z.rstride = p.plane; z.pstride = p.ld; z.scale = 1.0;
    z.work_bytes = cbytes;
    hipError_t e = launch_col<T>(p.pz, fb, 1, z, p.py, p.tw_z, st);
    if (e != hipSuccess) return hip_check(e, "c2r z launch");
    ColDesc y = z;
    y.nin = p.py; y.nout = p.py; y.rstride = p.ld; y.pstride = p.plane;
    e = launch_col<T>(p.py, fb, 1, y, p.nz, p.tw_y, st);
    if (e != hipSuccess) return hip_check(e, "c2r y launch");
    GridDesc g{p.nx, p.ny, p.nz, p.L, p.ld, p.plane, VFD_PAD_ZERO, 0, 0, cbytes, 0, p.field_bytes, 1.0};
    e = launch_xinv<T>(p.L, fb, 1, g, p.tw_x, st);
    return e == hipSuccess ? VFD_OK : hip_check(e, "c2r x launch");
}

// --------------------------------------------------- Gaussian spectrum
// K^(kz,ky,kx) = decay * Gz(kz) Gy(ky) Gx(kx), G(k) = sum_{|j|<=R} g_j cos(2 pi j k / P):
// the DFT of the centred, normalized, sampled Gaussian wrapped to the origin.
template <typename T>
__global__ void gauss_product_kernel(T* __restrict__ out, const double* __restrict__ gx,
                                     const double* __restrict__ gy, const double* __restrict__ gz,
                                     int hx, long ld, int py, int pz, int rs, double scale) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)pz * py * ld;
    if (i >= total) return;
    // storage order ((kz/rs)*py + ky)*ld*rs + kx*rs + kz%rs
    const int s = (int)(i % rs);
    const long rest = i / rs;
    const long kx = rest % ld;
    const long r2 = rest / ld;
    const int ky = (int)(r2 % py);
    const int kz = (int)(r2 / py) * rs + s;
    // ((gx*gy)*scale)*gz: the order the separable format evaluates in the z pass
    out[i] = kx < hx ? (T)(((gx[kx] * gy[ky]) * scale) * gz[kz]) : T(0);
}

// the normalized window g_j, |j| <= R = ceil(6 sigma), as g[j + R]
static void gauss_window(double sigma2, std::vector<double>& g) {
    const int R = (int)std::ceil(6.0 * std::sqrt(sigma2));
    g.assign(2 * R + 1, 0.0);
    double s = 0.0;
    for (int j = -R; j <= R; ++j) s += (g[j + R] = std::exp(-(double)j * j / (2.0 * sigma2)));
    for (auto& v : g) v /= s;
}

static void gauss_dft(double sigma2, int p, int nk, std::vector<double>& out) {
    std::vector<double> g;
    gauss_window(sigma2, g);
    const int R = ((int)g.size() - 1) / 2;
    out.assign(nk, 0.0);
    const long double tp = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < nk; ++k) {
        long double acc = 0.0L;
        for (int j = -R; j <= R; ++j) {
            long e = (((long)j * k) % p + p) % p;
            acc += (long double)g[j + R] * cosl(tp * (long double)e / (long double)p);
        }
        out[k] = (double)acc;
    }
}

int plan_gauss_spectrum(const Plan& p, double sigma2, double decay, void* d_kspec, hipStream_t st) {
    if (p.kspec_format != VFD_KSPEC_REAL && p.kspec_format != VFD_KSPEC_SEPARABLE)
        return set_error(VFD_ERR_INVALID,
                         "gauss spectrum: plan kernel-spectrum format must be VFD_KSPEC_REAL or VFD_KSPEC_SEPARABLE");
    if (!(sigma2 > 0.0) || !d_kspec) return set_error(VFD_ERR_INVALID, "gauss spectrum: sigma2 must be > 0");
    const int R = (int)std::ceil(6.0 * std::sqrt(sigma2));
    if (R > p.px - p.nx || R > p.py - p.ny || R > p.pz - p.nz)
        return set_error(VFD_ERR_INVALID,
                         "gauss spectrum: kernel radius %d exceeds the zero padding (FFT box %dx%dx%d, data %dx%dx%d): "
                         "the product would wrap around instead of being a linear convolution",
                         R, p.px, p.py, p.pz, p.nx, p.ny, p.nz);
    std::vector<double> gx, gy, gz;
    gauss_dft(sigma2, p.px, p.hx, gx);
    gauss_dft(sigma2, p.py, p.py, gy);
    gauss_dft(sigma2, p.pz, p.pz, gz);
    if (p.kspec_format == VFD_KSPEC_SEPARABLE) {
        // the tables themselves: [scale | gx | gy | gz], zero-padded to kspec_bytes
        std::vector<double> tab(p.kspec_bytes / sizeof(double), 0.0);
        tab[0] = decay / ((double)p.px * p.py * p.pz);
        std::copy(gx.begin(), gx.end(), tab.begin() + 1);
        std::copy(gy.begin(), gy.end(), tab.begin() + 1 + p.hx);
        std::copy(gz.begin(), gz.end(), tab.begin() + 1 + p.hx + p.py);
        // the z window for the direct z pass: R, then g_0 .. g_min(R, kZDirectMaxR)
        std::vector<double> win;
        gauss_window(sigma2, win);
        const size_t g0 = 1 + p.hx + p.py + p.pz;
        tab[g0] = (double)R;
        for (int j = 0; j <= std::min(R, kZDirectMaxR); ++j) tab[g0 + 1 + j] = win[R + j];
        if (R <= kZDirectMaxR) {
            // the direct pass multiplies by the taps rounded to FP32; g_0 takes
            // up their rounding so that g_0 + 2 sum_j g_j stays 1 to FP32
            // precision (otherwise the taps' rounding is a per-tick gain)
            double tail = 0.0;
            for (int j = 1; j <= R; ++j) tail += (double)(float)win[R + j];
            tab[g0 + 1] = 1.0 - 2.0 * tail;
        }
        int rc = hip_check(hipMemcpyAsync(d_kspec, tab.data(), p.kspec_bytes, hipMemcpyHostToDevice, st),
                           "gauss tables upload");
        // the host vector must outlive the async copy
        if (!rc) rc = hip_check(hipStreamSynchronize(st), "gauss tables sync");
        return rc;
    }
    double* d = nullptr;
    const size_t n = gx.size() + gy.size() + gz.size();
    if (int rc = hip_check(hipMallocAsync((void**)&d, n * sizeof(double), st), "hipMallocAsync(gauss)")) return rc;
    std::vector<double> all;
    all.insert(all.end(), gx.begin(), gx.end());
    all.insert(all.end(), gy.begin(), gy.end());
    all.insert(all.end(), gz.begin(), gz.end());
    int rc = hip_check(hipMemcpyAsync(d, all.data(), n * sizeof(double), hipMemcpyHostToDevice, st), "gauss upload");
    if (!rc) {
        const long total = (long)p.pz * p.py * p.ld;
        const unsigned blocks = (unsigned)((total + 255) / 256);
        const double scale = decay / ((double)p.px * p.py * p.pz);
        if (p.prec == VFD_F32)
            hipLaunchKernelGGL(gauss_product_kernel<float>, dim3(blocks), dim3(256), 0, st, (float*)d_kspec, d,
                               d + gx.size(), d + gx.size() + gy.size(), p.hx, p.ld, p.py, p.pz, p.rs_z, scale);
        else
            hipLaunchKernelGGL(gauss_product_kernel<double>, dim3(blocks), dim3(256), 0, st, (double*)d_kspec, d,
                               d + gx.size(), d + gx.size() + gy.size(), p.hx, p.ld, p.py, p.pz, p.rs_z, scale);
        rc = hip_check(hipGetLastError(), "gauss spectrum launch");
        // the host vector must outlive the async copy
        if (!rc) rc = hip_check(hipStreamSynchronize(st), "gauss spectrum sync");
    }
    (void)hipFreeAsync(d, st);
    return rc;
}

// ------------------------------------------------------ compensated mass
template <typename T>
__device__ __forceinline__ void neumaier_add(double& s, double& c, T xv) {
    const double x = (double)xv;
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

// Each block sums a fixed contiguous chunk of one field; lanes accumulate
// strided elements, then a fixed-order tree combines (s, c) pairs.
template <typename T>
__global__ void mass_partial_kernel(const T* const* __restrict__ fields, long n, long chunk,
                                    double2* __restrict__ partial) {
    const T* f = fields[blockIdx.y];
    const long lo = (long)blockIdx.x * chunk;
    const long hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0, c = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) neumaier_add(s, c, f[i]);
    __shared__ double ss[256], sc[256];
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = ss[threadIdx.x], ca = sc[threadIdx.x];
            double b = ss[threadIdx.x + w], cb = sc[threadIdx.x + w];
            double t = a + b;
            double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
            ss[threadIdx.x] = t;
            sc[threadIdx.x] = ca + cb + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = double2{ss[0], sc[0]};
}

__global__ void mass_final_kernel(const double2* __restrict__ partial, int nblocks, double* __restrict__ out) {
    // one thread per field, fixed order
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (int)gridDim.x * (int)blockDim.x) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        const double2 p = partial[(long)f * nblocks + b];
        neumaier_add(s, c, p.x);
        c += p.y;
    }
    out[f] = s + c;
}

int plan_mass(const Plan& p, const void* const* d_fields, int nfields, double* mass, hipStream_t st) {
    if (nfields <= 0 || !d_fields || !mass) return set_error(VFD_ERR_INVALID, "mass: bad arguments");
    const long n = (long)p.nx * p.ny * p.nz;
    const int nblocks = (int)std::min<long>(1024, (n + 4095) / 4096);
    const long chunk = (n + nblocks - 1) / nblocks;
    const size_t ptr_bytes = sizeof(void*) * nfields;
    const size_t part_bytes = sizeof(double2) * (size_t)nblocks * nfields;
    const size_t out_bytes = sizeof(double) * nfields;
    char* d = nullptr;
    if (int rc = hip_check(hipMallocAsync((void**)&d, ptr_bytes + part_bytes + out_bytes + 64, st), "hipMallocAsync(mass)"))
        return rc;
    void** dptr = (void**)d;
    double2* dpart = (double2*)(d + ((ptr_bytes + 15) / 16) * 16);
    double* dout = (double*)((char*)dpart + part_bytes);
    int rc = hip_check(hipMemcpyAsync(dptr, d_fields, ptr_bytes, hipMemcpyHostToDevice, st), "mass upload");
    if (!rc) {
        if (p.prec == VFD_F32)
            hipLaunchKernelGGL(mass_partial_kernel<float>, dim3(nblocks, nfields), dim3(256), 0, st,
                               (const float* const*)dptr, n, chunk, dpart);
        else
            hipLaunchKernelGGL(mass_partial_kernel<double>, dim3(nblocks, nfields), dim3(256), 0, st,
                               (const double* const*)dptr, n, chunk, dpart);
        hipLaunchKernelGGL(mass_final_kernel, dim3(nfields), dim3(1), 0, st, dpart, nblocks, dout);
        rc = hip_check(hipGetLastError(), "mass launch");
    }
    if (!rc) rc = hip_check(hipMemcpyAsync(mass, dout, out_bytes, hipMemcpyDeviceToHost, st), "mass download");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "mass sync");
    (void)hipFreeAsync(d, st);
    return rc;
}

int plan_r2c(const Plan& p, const void* in, void* out, hipStream_t st) {
    if (!p.full_box) return set_error(VFD_ERR_INVALID, "r2c: plan data box must equal its FFT box");
    return p.prec == VFD_F32 ? run_r2c<float>(p, in, out, st) : run_r2c<double>(p, in, out, st);
}
int plan_c2r(const Plan& p, void* in, void* out, hipStream_t st) {
    if (!p.full_box) return set_error(VFD_ERR_INVALID, "c2r: plan data box must equal its FFT box");
    return p.prec == VFD_F32 ? run_c2r<float>(p, in, out, st) : run_c2r<double>(p, in, out, st);
}

// The HBM streaming reference: one 16-byte element per lane over a full grid,
// non-temporal loads and stores (measured on MI355X: this form streams at
// about 6.4 TB/s where grid-stride loops and 8-byte lanes stay near 5.5).
__global__ void __launch_bounds__(256) stream_copy_kernel(const void* src, void* dst, unsigned nvec) {
    const rsrc_t rs = make_rsrc(src, (unsigned long)nvec * 16u);
    const rsrc_t rd = make_rsrc(dst, (unsigned long)nvec * 16u);
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    const unsigned off = i < nvec ? i * 16u : kBufOOB;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, kStream), rd, off, 0,
                                           kStream);
}

// The same form with a second source (added) or a second destination: the
// streaming references for passes that read one byte per two written (x and
// y forward) or two per one (their inverses).
template <bool R2, bool W2>
__global__ void __launch_bounds__(256) stream_mix_kernel(const void* src0, const void* src1, void* dst0, void* dst1,
                                                         unsigned nvec) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    const unsigned off = i < nvec ? i * 16u : kBufOOB;
    u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(make_rsrc(src0, (unsigned long)nvec * 16u), off, 0, kStream);
    if constexpr (R2) v += __builtin_amdgcn_raw_buffer_load_b128(make_rsrc(src1, (unsigned long)nvec * 16u), off, 0,
                                                                 kStream);
    __builtin_amdgcn_raw_buffer_store_b128(v, make_rsrc(dst0, (unsigned long)nvec * 16u), off, 0, kStream);
    if constexpr (W2) __builtin_amdgcn_raw_buffer_store_b128(v, make_rsrc(dst1, (unsigned long)nvec * 16u), off, 0,
                                                             kStream);
}

bool length_supported(int q) {
    switch (q) {
#define VFD_CASE(Q) case Q:
        VFD_CASE(2) VFD_CASE(4) VFD_CASE(8) VFD_CASE(16) VFD_CASE(32) VFD_CASE(64) VFD_CASE(128)
        VFD_CASE(256) VFD_CASE(512) VFD_CASE(1024) VFD_CASE(2048) VFD_CASE(3) VFD_CASE(6) VFD_CASE(12)
        VFD_CASE(24) VFD_CASE(48) VFD_CASE(96) VFD_CASE(192) VFD_CASE(384) VFD_CASE(768) VFD_CASE(1536)
        VFD_CASE(9) VFD_CASE(18) VFD_CASE(36) VFD_CASE(72) VFD_CASE(144) VFD_CASE(288) VFD_CASE(576)
        VFD_CASE(1152) VFD_CASE(2304) VFD_CASE(616) VFD_CASE(308) VFD_CASE(5) VFD_CASE(10) VFD_CASE(20)
        VFD_CASE(40) VFD_CASE(80) VFD_CASE(160) VFD_CASE(320) VFD_CASE(640) VFD_CASE(1280)
#undef VFD_CASE
        return true;
        default:
            return false;
    }
}

}  // namespace vfd

// ====================================================================== C-ABI
using namespace vfd;


extern "C" {

const char* vfd_last_error(void) { return g_last_error.c_str(); }
int vfd_version(void) { return 10000; }
int vfd_length_supported(int q) { return length_supported(q) ? 1 : 0; }

int vfd_plan_create(const vfd_plan_desc* desc, vfd_plan** plan) {
    if (!plan) return set_error(VFD_ERR_INVALID, "plan_create: null output");
    *plan = nullptr;
    Plan p;
    if (int rc = plan_init(desc, &p)) return rc;
    *plan = new vfd_plan{p};
    return VFD_OK;
}

void vfd_plan_destroy(vfd_plan* plan) { delete plan; }

int vfd_plan_get_info(const vfd_plan* plan, vfd_plan_info* info) {
    if (!plan || !info) return set_error(VFD_ERR_INVALID, "plan_get_info: null argument");
    const Plan& p = plan->p;
    info->nz = p.nz; info->ny = p.ny; info->nx = p.nx;
    info->pz = p.pz; info->py = p.py; info->px = p.px;
    info->L = p.L; info->hx = p.hx; info->ld = p.ld;
    info->rs_y = p.rs_y; info->rs_z = p.rs_z;
    info->y_inv_fold = p.fold_y;
    info->work_bytes = p.work_bytes;
    info->kspec_bytes = p.kspec_bytes;
    info->bytes_per_field_tick = p.bytes_per_field_tick;
    info->order = p.order;
    return VFD_OK;
}

int vfd_diffuse(const vfd_plan* plan, const vfd_field* fields, int nfields, void* stream) {
    if (!plan) return set_error(VFD_ERR_INVALID, "diffuse: null plan");
    return plan_diffuse(plan->p, fields, nfields, (hipStream_t)stream);
}

int vfd_diffuse_update(const vfd_plan* plan, const vfd_field* fields, float* const* d_delta, const float* decay,
                       const float* d_frag_col, const float* d_frag_ela, float* const* d_grads, double* totals,
                       void* stream) {
    if (!plan || !fields || !d_delta || !decay) return set_error(VFD_ERR_INVALID, "diffuse_update: null argument");
    if (plan->p.prec != VFD_F32) return set_error(VFD_ERR_UNSUPPORTED, "diffuse_update: FP32 plans only");
    for (int ic = 0; ic < 8; ++ic)
        if (!d_delta[ic]) return set_error(VFD_ERR_INVALID, "diffuse_update: chem %d has no delta field", ic);
    hipStream_t st = (hipStream_t)stream;
    Epilogue ep;
    ep.deltas = d_delta;
    ep.decay = decay;
    if (int rc = plan_diffuse_passes(plan->p, fields, 8, 0, 5, st, &ep)) return rc;
    float* levels[8];
    for (int ic = 0; ic < 8; ++ic) levels[ic] = (float*)fields[ic].out;
    const long n = (long)plan->p.nx * plan->p.ny * plan->p.nz;
    return chem_grads_totals(levels, d_frag_col, d_frag_ela, d_grads, n, totals, st);
}

int vfd_diffuse_sample(const vfd_plan* plan, const vfd_field* fields, int nfields, const int* chem_index, int stride,
                       float* const* d_vis, void* stream) {
    if (!plan || !fields || !chem_index || !d_vis) return set_error(VFD_ERR_INVALID, "diffuse_sample: null argument");
    if (plan->p.prec != VFD_F32) return set_error(VFD_ERR_UNSUPPORTED, "diffuse_sample: FP32 plans only");
    if (stride < 1) return set_error(VFD_ERR_INVALID, "diffuse_sample: stride %d", stride);
    if (nfields < 0) return set_error(VFD_ERR_INVALID, "diffuse_sample: negative field count");
    std::vector<float> scale(nfields);
    for (int i = 0; i < nfields; ++i) {
        if (chem_index[i] < 0 || chem_index[i] >= 8)
            return set_error(VFD_ERR_INVALID, "diffuse_sample: chem index %d", chem_index[i]);
        if (!d_vis[i]) return set_error(VFD_ERR_INVALID, "diffuse_sample: field %d has no output array", i);
        scale[i] = chem_display_scale(chem_index[i]);
    }
    Epilogue ep;
    ep.vis = d_vis;
    ep.vis_scale = scale.data();
    ep.vis_stride = stride;
    return plan_diffuse_passes(plan->p, fields, nfields, 0, 5, (hipStream_t)stream, &ep);
}

int vfd_diffuse_pass(const vfd_plan* plan, const vfd_field* fields, int nfields, int pass, void* stream) {
    if (!plan) return set_error(VFD_ERR_INVALID, "diffuse_pass: null plan");
    if (pass >= VFD_PASS_Z_DIRECT8 && pass <= VFD_PASS_Z_FFT)
        return plan_diffuse_zpart(plan->p, fields, nfields, pass - VFD_PASS_Z_DIRECT8, (hipStream_t)stream);
    return plan_diffuse_passes(plan->p, fields, nfields, pass, pass + 1, (hipStream_t)stream);
}

int vfd_stream_mix(void* dst0, void* dst1, const void* src0, const void* src1, size_t bytes, void* stream) {
    if (!dst0 || !src0) return set_error(VFD_ERR_INVALID, "stream_mix: null buffer");
    if (bytes % 16 || bytes >= kBufLimit)
        return set_error(VFD_ERR_INVALID, "stream_mix: %zu bytes (a multiple of 16 below 4 GiB)", bytes);
    if (bytes == 0) return VFD_OK;
    const unsigned nvec = (unsigned)(bytes / 16);
    const dim3 grid((nvec + 255) / 256);
    hipStream_t st = (hipStream_t)stream;
    if (src1 && dst1) hipLaunchKernelGGL((stream_mix_kernel<true, true>), grid, dim3(256), 0, st, src0, src1, dst0, dst1, nvec);
    else if (src1) hipLaunchKernelGGL((stream_mix_kernel<true, false>), grid, dim3(256), 0, st, src0, src1, dst0, dst0, nvec);
    else if (dst1) hipLaunchKernelGGL((stream_mix_kernel<false, true>), grid, dim3(256), 0, st, src0, src0, dst0, dst1, nvec);
    else hipLaunchKernelGGL((stream_mix_kernel<false, false>), grid, dim3(256), 0, st, src0, src0, dst0, dst0, nvec);
    return hip_check(hipGetLastError(), "stream_mix launch");
}

int vfd_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return set_error(VFD_ERR_INVALID, "stream_copy: null buffer");
    if (bytes % 16 || bytes >= kBufLimit)
        return set_error(VFD_ERR_INVALID, "stream_copy: %zu bytes (a multiple of 16 below 4 GiB)", bytes);
    if (bytes == 0) return VFD_OK;
    const unsigned nvec = (unsigned)(bytes / 16);
    hipLaunchKernelGGL(stream_copy_kernel, dim3((nvec + 255) / 256), dim3(256), 0, (hipStream_t)stream, src, dst,
                       nvec);
    return hip_check(hipGetLastError(), "stream_copy launch");
}

double vfd_plan_pass_bytes(const vfd_plan* plan, int pass) { return plan ? plan_pass_bytes(plan->p, pass) : 0.0; }

int vfd_plan_autotune(vfd_plan* plan, int nfields, void* stream, float* best_ms) {
    if (!plan) return set_error(VFD_ERR_INVALID, "plan_autotune: null plan");
    return plan_autotune(plan->p, nfields, (hipStream_t)stream, best_ms);
}

int vfd_plan_get_schedule(const vfd_plan* plan, int* variants) {
    if (!plan || !variants) return set_error(VFD_ERR_INVALID, "plan_get_schedule: null argument");
    for (int i = 0; i < 5; ++i) variants[i] = plan->p.var[i];
    return VFD_OK;
}

int vfd_plan_set_schedule(vfd_plan* plan, const int* variants) {
    if (!plan || !variants) return set_error(VFD_ERR_INVALID, "plan_set_schedule: null argument");
    for (int i = 0; i < 5; ++i)
        if (variants[i] < 0 || variants[i] >= kScheduleVariants)
            return set_error(VFD_ERR_INVALID, "plan_set_schedule: variant %d out of range", variants[i]);
    for (int i = 0; i < 5; ++i) plan->p.var[i] = variants[i];
    return VFD_OK;
}

int vfd_gauss_spectrum(const vfd_plan* plan, double sigma2, double decay, void* d_kspec, void* stream) {
    if (!plan) return set_error(VFD_ERR_INVALID, "gauss_spectrum: null plan");
    return plan_gauss_spectrum(plan->p, sigma2, decay, d_kspec, (hipStream_t)stream);
}

int vfd_field_mass(const vfd_plan* plan, const void* const* d_fields, int nfields, double* mass, void* stream) {
    if (!plan) return set_error(VFD_ERR_INVALID, "field_mass: null plan");
    return plan_mass(plan->p, d_fields, nfields, mass, (hipStream_t)stream);
}

int vfd_r2c_3d(const vfd_plan* plan, const void* d_real, void* d_complex, void* stream) {
    if (!plan || !d_real || !d_complex) return set_error(VFD_ERR_INVALID, "r2c_3d: null argument");
    return plan_r2c(plan->p, d_real, d_complex, (hipStream_t)stream);
}

int vfd_c2r_3d(const vfd_plan* plan, void* d_complex, void* d_real, void* stream) {
    if (!plan || !d_real || !d_complex) return set_error(VFD_ERR_INVALID, "c2r_3d: null argument");
    return plan_c2r(plan->p, d_complex, d_real, (hipStream_t)stream);
}

}  // extern "C"
