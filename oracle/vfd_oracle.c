/*
 * vfd_oracle.c -- CPU restatement of the reference's cytokine-diffusion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls this
 * file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker (or, for cpu_fft_diffuse3d_f32, as the timed CPU
 * baseline).
 *
 * Conventions (all follow the reference, src/Diffusion/):
 *   - fields are row-major [D][H][W] (W fastest), index z*H*W + y*W + x
 *     (WHChemical.cpp:345, World.cpp:83);
 *   - forward transforms are unnormalized e^{-2 pi i nk/N}; inverse transforms are
 *     unnormalized e^{+2 pi i nk/N} (cuFFT semantics used at Diffusion.cu:1814,1842);
 *   - R2C half layout is [FD][FH][FW/2+1] complex (interleaved re,im);
 *   - the diffusion operator is pad -> R2C -> (x K^ / (FD*FH*FW)) -> C2R -> unpad
 *     (Diffusion.cu:1778-1868, convolutionFFT.cu:578).
 *
 * Pinning (how each function is checked; see tests/test_oracle.py):
 *   orc_fft_c2c / orc_fft_c2c_f32 : O(n^2) DFT on every n <= 64, numpy golden vectors,
 *                                   impulse/cosine/Parseval/round-trip identities.
 *   orc_rfft3 / orc_irfft3        : numpy rfftn/irfftn golden fixtures (tests/golden),
 *                                   Hermitian round trip.
 *   orc_pad_* / orc_unpad3d       : exhaustive index rules on tiny boxes.
 *   orc_fft_diffuse3d             : brute-force direct 3D convolution of the padded box
 *                                   (orc_direct_conv3d) on tiny boxes, all pad modes;
 *                                   delta-field / delta-kernel identities; mass invariant.
 *   orc_direct_separable3d        : closed-form Gaussian facts (symmetry, mass on an
 *                                   unbounded domain), brute-force 3D direct conv.
 *   orc_compute_kernel3d          : closed-form stencil spectrum S(theta)^M on a box where
 *                                   clamp padding cannot reach the kernel, and the
 *                                   TEST_KERNEL mass check (Diffusion.cu:1439-1454).
 *   cpu_fft_diffuse3d_f32         : the FP64 orc_fft_diffuse3d (rel-L2 <= 1e-5).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338327950288

/* pad modes: the reference's pad_t (src/enums.h:35-40) plus ZERO used by the
 * tissue-scale configuration (zero-padded linear convolution). */
enum { ORC_PAD_CLAMP = 0, ORC_PAD_RIGHTWALL = 1, ORC_PAD_MIRROR = 2, ORC_PAD_CONSTANT_VF = 3,
       ORC_PAD_ZERO = 100 };

/* ------------------------------------------------------------------------- */
/* snapTransformSize (Diffusion.cu:39-70): align to 16, next power of two.     */
/* ------------------------------------------------------------------------- */
int orc_snap_transform_size(int dataSize)
{
    if (dataSize % 16) dataSize = dataSize - dataSize % 16 + 16;
    int hiBit;
    for (hiBit = 31; hiBit >= 0; hiBit--)
        if ((unsigned)dataSize & (1U << hiBit)) break;
    unsigned lowPOT = 1U << hiBit;
    if (lowPOT == (unsigned)dataSize) return dataSize;
    unsigned hiPOT = 1U << (hiBit + 1);
    if (hiPOT <= 1024) return (int)hiPOT;
    return (int)pow(2.0, ceil(log((double)dataSize) / log(2.0)));
}

/* ------------------------------------------------------------------------- */
/* 1D complex FFT, any length: recursive mixed radix (decimation in time).     */
/* ------------------------------------------------------------------------- */
static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

/* out[k] = sum_n in[n*s] w^{nk}, n < len; w = exp(sign*2*pi*i/len) */
static void fft_rec_d(int len, const double *in, int s, double *out, int sign, double *tmp)
{
    if (len == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = smallest_factor(len);
    int m = len / p;
    /* sub-transforms of the p decimated sequences into out (contiguous blocks) */
    for (int q = 0; q < p; q++)
        fft_rec_d(m, in + 2 * q * s, s * p, out + 2 * q * m, sign, tmp);
    /* combine: X[k + m r] = sum_q W_len^{q(k + m r)} F_q[k] */
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            double acc_re = 0.0, acc_im = 0.0;
            int kk = k + m * r;
            for (int q = 0; q < p; q++) {
                long e = ((long)q * kk) % len;
                double ang = sign * 2.0 * ORC_PI * (double)e / (double)len;
                double c = cos(ang), sn = sin(ang);
                double fr = out[2 * (q * m + k)], fi = out[2 * (q * m + k) + 1];
                acc_re += fr * c - fi * sn;
                acc_im += fr * sn + fi * c;
            }
            tmp[2 * kk] = acc_re;
            tmp[2 * kk + 1] = acc_im;
        }
    }
    memcpy(out, tmp, sizeof(double) * 2 * len);
}

/* in-place unnormalized complex FFT of n interleaved doubles */
int orc_fft_c2c(int n, double *x, int inverse)
{
    if (n <= 0) return -1;
    double *buf = (double *)malloc(sizeof(double) * 4 * n);
    if (!buf) return -2;
    double *out = buf, *tmp = buf + 2 * n;
    fft_rec_d(n, x, 1, out, inverse ? +1 : -1, tmp);
    memcpy(x, out, sizeof(double) * 2 * n);
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* 3D transforms on [nz][ny][nx] boxes via 1D lines.                           */
/* ------------------------------------------------------------------------- */
static void fft3_c2c_d(int nz, int ny, int nx, double *a, int inverse)
{
    int nmax = nx > ny ? nx : ny; nmax = nmax > nz ? nmax : nz;
    double *line = (double *)malloc(sizeof(double) * 2 * nmax);
    /* x */
    for (long r = 0; r < (long)nz * ny; r++) orc_fft_c2c(nx, a + 2 * r * nx, inverse);
    /* y */
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            for (int y = 0; y < ny; y++) {
                long i = ((long)z * ny + y) * nx + x;
                line[2 * y] = a[2 * i]; line[2 * y + 1] = a[2 * i + 1];
            }
            orc_fft_c2c(ny, line, inverse);
            for (int y = 0; y < ny; y++) {
                long i = ((long)z * ny + y) * nx + x;
                a[2 * i] = line[2 * y]; a[2 * i + 1] = line[2 * y + 1];
            }
        }
    /* z */
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            for (int z = 0; z < nz; z++) {
                long i = ((long)z * ny + y) * nx + x;
                line[2 * z] = a[2 * i]; line[2 * z + 1] = a[2 * i + 1];
            }
            orc_fft_c2c(nz, line, inverse);
            for (int z = 0; z < nz; z++) {
                long i = ((long)z * ny + y) * nx + x;
                a[2 * i] = line[2 * z]; a[2 * i + 1] = line[2 * z + 1];
            }
        }
    free(line);
}

/* R2C 3D: real [nz][ny][nx] -> half spectrum [nz][ny][nx/2+1] (cuFFT layout) */
int orc_rfft3(int nz, int ny, int nx, const double *in, double *out_half)
{
    long n = (long)nz * ny * nx;
    int hx = nx / 2 + 1;
    double *a = (double *)malloc(sizeof(double) * 2 * n);
    if (!a) return -2;
    for (long i = 0; i < n; i++) { a[2 * i] = in[i]; a[2 * i + 1] = 0.0; }
    fft3_c2c_d(nz, ny, nx, a, 0);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int k = 0; k < hx; k++) {
                long s = ((long)z * ny + y) * nx + k;
                long d = ((long)z * ny + y) * hx + k;
                out_half[2 * d] = a[2 * s]; out_half[2 * d + 1] = a[2 * s + 1];
            }
    free(a);
    return 0;
}

/* C2R 3D (unnormalized): half spectrum -> real [nz][ny][nx].  The missing half is
 * rebuilt from Hermitian symmetry X[-k] = conj X[k]; the imaginary parts of the
 * self-conjugate bins are ignored, as a C2R transform does. */
int orc_irfft3(int nz, int ny, int nx, const double *in_half, double *out)
{
    long n = (long)nz * ny * nx;
    int hx = nx / 2 + 1;
    double *a = (double *)malloc(sizeof(double) * 2 * n);
    if (!a) return -2;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int k = 0; k < nx; k++) {
                long d = ((long)z * ny + y) * nx + k;
                if (k < hx) {
                    long s = ((long)z * ny + y) * hx + k;
                    a[2 * d] = in_half[2 * s]; a[2 * d + 1] = in_half[2 * s + 1];
                } else {
                    int zz = (nz - z) % nz, yy = (ny - y) % ny, kk = nx - k;
                    long s = ((long)zz * ny + yy) * hx + kk;
                    a[2 * d] = in_half[2 * s]; a[2 * d + 1] = -in_half[2 * s + 1];
                }
            }
    /* Enforce Hermitian symmetry on the self-conjugate planes kx = 0 and kx = nx/2
     * (even nx) the way a C2R transform does: the result is real, so take the
     * Hermitian part.  x -> x then the inverse of the full box, keep Re. */
    fft3_c2c_d(nz, ny, nx, a, 1);
    for (long i = 0; i < n; i++) out[i] = a[2 * i];
    free(a);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Pad / unpad rules (restated index maps).                                   */
/* ------------------------------------------------------------------------- */

/* padDataConstantVF3D (convolutionFFT.cuh:590-638): the interior
 * 0<z<D-1, 0<y<H-1, 0<x<epi copies the field, every other voxel of the whole FFT
 * box is baseChem.  The reference does not bound x by the row length W (an
 * epi above W reads the next row); here x >= W is baseChem, as outside the data. */
void orc_pad_constant_vf3d(double *dst, const float *src, int FD, int FH, int FW,
                           int D, int H, int W, int epi, double base)
{
    if (epi > W) epi = W;
    for (int z = 0; z < FD; z++)
        for (int y = 0; y < FH; y++)
            for (int x = 0; x < FW; x++) {
                long d = ((long)z * FH + y) * FW + x;
                if (z < D - 1 && z > 0 && y < H - 1 && y > 0 && x < epi && x > 0)
                    dst[d] = src[((long)z * H + y) * W + x];
                else
                    dst[d] = base;
            }
}

/* Zero pad: field on the data box, zero elsewhere (the linear-convolution box). */
void orc_pad_zero3d(double *dst, const float *src, int FD, int FH, int FW, int D, int H, int W)
{
    for (int z = 0; z < FD; z++)
        for (int y = 0; y < FH; y++)
            for (int x = 0; x < FW; x++) {
                long d = ((long)z * FH + y) * FW + x;
                dst[d] = (z < D && y < H && x < W) ? (double)src[((long)z * H + y) * W + x] : 0.0;
            }
}

/* padDataClampToBorder3D (convolutionFFT.cuh:510-585): [0,N) data, [N,N+K*) the
 * last slice, [N+K*,P) the first slice, per axis. */
void orc_pad_clamp3d(double *dst, const double *src, int FD, int FH, int FW, int D, int H, int W,
                     int KZ, int KY, int KX)
{
    for (int z = 0; z < FD; z++)
        for (int y = 0; y < FH; y++)
            for (int x = 0; x < FW; x++) {
                int dz = z < D ? z : (z < D + KZ ? D - 1 : 0);
                int dy = y < H ? y : (y < H + KY ? H - 1 : 0);
                int dx = x < W ? x : (x < W + KX ? W - 1 : 0);
                dst[((long)z * FH + y) * FW + x] = src[((long)dz * H + dy) * W + dx];
            }
}

/* padKernel3D (convolutionFFT.cuh:402-445): kernel centre (KZ,KY,KX) to the origin,
 * negative offsets wrapped.  dst must be zeroed by the caller. */
void orc_pad_kernel3d(double *dst, const double *k, int FD, int FH, int FW, int KD, int KH, int KW,
                      int KZ, int KY, int KX)
{
    for (int z = 0; z < KD; z++)
        for (int y = 0; y < KH; y++)
            for (int x = 0; x < KW; x++) {
                int kz = z - KZ; if (kz < 0) kz += FD;
                int ky = y - KY; if (ky < 0) ky += FH;
                int kx = x - KX; if (kx < 0) kx += FW;
                dst[((long)kz * FH + ky) * FW + kx] = k[((long)z * KH + y) * KW + x];
            }
}

/* unpadResult3D (convolutionFFT.cuh:152-171): origin-aligned crop. */
void orc_unpad3d(double *dst, const double *src, int D, int H, int W, int FD, int FH, int FW)
{
    (void)FD;
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                dst[((long)z * H + y) * W + x] = src[((long)z * FH + y) * FW + x];
}

/* extractCenter3D (convolutionFFT.cuh:77-112): centred window of the data box. */
void orc_extract_center3d(double *dst, const double *src, int D, int H, int W, int wD, int wH, int wW)
{
    int bD = (D - wD) / 2, bH = (H - wH) / 2, bW = (W - wW) / 2;
    for (int z = 0; z < wD; z++)
        for (int y = 0; y < wH; y++)
            for (int x = 0; x < wW; x++)
                dst[((long)z * wH + y) * wW + x] =
                    src[((long)(z + bD) * H + (y + bH)) * W + (x + bW)];
}

/* ------------------------------------------------------------------------- */
/* Kernel spectrum (computeKernelSpectrum3D, Diffusion.cu:1533-1619):         */
/* padKernel3D into the chem FFT box, then R2C.                              */
/* ------------------------------------------------------------------------- */
int orc_kernel_spectrum3d(const double *window, int KD, int KH, int KW, int KZ, int KY, int KX,
                          int FD, int FH, int FW, double *spec_half)
{
    long n = (long)FD * FH * FW;
    double *pk = (double *)calloc(n, sizeof(double));
    if (!pk) return -2;
    orc_pad_kernel3d(pk, window, FD, FH, FW, KD, KH, KW, KZ, KY, KX);
    int rc = orc_rfft3(FD, FH, FW, pk, spec_half);
    free(pk);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* The diffusion operator, fftDiffuse3D (Diffusion.cu:1622-1934), in FP64.    */
/* ------------------------------------------------------------------------- */
int orc_fft_diffuse3d(const float *u, int D, int H, int W, const double *spec_half,
                      int FD, int FH, int FW, int pad_mode, int epi, double base, double *out)
{
    if (FD < D || FH < H || FW < W) return -1;
    long n = (long)FD * FH * FW;
    int hx = FW / 2 + 1;
    long ns = (long)FD * FH * hx;
    double *pad = (double *)malloc(sizeof(double) * n);
    double *spec = (double *)malloc(sizeof(double) * 2 * ns);
    if (!pad || !spec) { free(pad); free(spec); return -2; }
    if (pad_mode == ORC_PAD_CONSTANT_VF)
        orc_pad_constant_vf3d(pad, u, FD, FH, FW, D, H, W, epi, base);
    else if (pad_mode == ORC_PAD_ZERO)
        orc_pad_zero3d(pad, u, FD, FH, FW, D, H, W);
    else { free(pad); free(spec); return -3; }
    orc_rfft3(FD, FH, FW, pad, spec);
    /* modulateAndNormalize (convolutionFFT.cuh:690-716): spec <- c (K^ spec) */
    double c = 1.0 / ((double)FD * FH * FW);
    for (long i = 0; i < ns; i++) {
        double ar = spec_half[2 * i], ai = spec_half[2 * i + 1];
        double br = spec[2 * i], bi = spec[2 * i + 1];
        spec[2 * i] = c * (ar * br - ai * bi);
        spec[2 * i + 1] = c * (ai * br + ar * bi);
    }
    orc_irfft3(FD, FH, FW, spec, pad);
    orc_unpad3d(out, pad, D, H, W, FD, FH, FW);
    free(pad); free(spec);
    return 0;
}

/* Brute-force circular convolution on the padded box, evaluated on the data box:
 * out[p] = sum_q kwin[q] * pad[(p - (q - Kc)) mod F].  Independent of every FFT;
 * used to pin orc_fft_diffuse3d on tiny boxes. */
int orc_direct_conv3d(const float *u, int D, int H, int W, const double *kwin, int KD, int KH, int KW,
                      int KZ, int KY, int KX, int FD, int FH, int FW, int pad_mode, int epi,
                      double base, double *out)
{
    long n = (long)FD * FH * FW;
    double *pad = (double *)malloc(sizeof(double) * n);
    if (!pad) return -2;
    if (pad_mode == ORC_PAD_CONSTANT_VF)
        orc_pad_constant_vf3d(pad, u, FD, FH, FW, D, H, W, epi, base);
    else if (pad_mode == ORC_PAD_ZERO)
        orc_pad_zero3d(pad, u, FD, FH, FW, D, H, W);
    else { free(pad); return -3; }
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double acc = 0.0;
                for (int a = 0; a < KD; a++)
                    for (int b = 0; b < KH; b++)
                        for (int c = 0; c < KW; c++) {
                            int sz = ((z - (a - KZ)) % FD + FD) % FD;
                            int sy = ((y - (b - KY)) % FH + FH) % FH;
                            int sx = ((x - (c - KX)) % FW + FW) % FW;
                            acc += kwin[((long)a * KH + b) * KW + c] * pad[((long)sz * FH + sy) * FW + sx];
                        }
                out[((long)z * H + y) * W + x] = acc;
            }
    free(pad);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Gaussian Green's-function kernel of the tissue-scale configuration.        */
/* g[j] = exp(-j^2/(2 sigma^2)), |j| <= R, normalized to sum 1.                */
/* ------------------------------------------------------------------------- */
void orc_gauss_window1d(double sigma2, int R, double *g)
{
    double s = 0.0;
    for (int j = -R; j <= R; j++) { g[j + R] = exp(-(double)j * j / (2.0 * sigma2)); s += g[j + R]; }
    for (int j = 0; j < 2 * R + 1; j++) g[j] /= s;
}

int orc_gauss_radius(double sigma2) { return (int)ceil(6.0 * sqrt(sigma2)); }

/* Separable direct zero-padded linear convolution with the Gaussian window and
 * the decay factor, evaluated on the data box.  Independent of every FFT. */
static int separable_core(const float *uf, const double *ud, int D, int H, int W, double sigma2,
                          double decay, double *out);

int orc_direct_separable3d(const float *u, int D, int H, int W, double sigma2, double decay,
                           double *out)
{
    return separable_core(u, NULL, D, H, W, sigma2, decay, out);
}

/* The same with an FP64 input field: chains ticks of the FP64 path without
 * rounding the intermediate field to FP32. */
int orc_direct_separable3d_f64(const double *u, int D, int H, int W, double sigma2, double decay,
                               double *out)
{
    return separable_core(NULL, u, D, H, W, sigma2, decay, out);
}

static int separable_core(const float *uf, const double *ud, int D, int H, int W, double sigma2,
                          double decay, double *out)
{
    int R = orc_gauss_radius(sigma2);
    double *g = (double *)malloc(sizeof(double) * (2 * R + 1));
    long n = (long)D * H * W;
    double *a = (double *)malloc(sizeof(double) * n);
    double *b = (double *)malloc(sizeof(double) * n);
    if (!g || !a || !b) { free(g); free(a); free(b); return -2; }
    orc_gauss_window1d(sigma2, R, g);
    for (long i = 0; i < n; i++) a[i] = uf ? (double)uf[i] : ud[i];
    /* x */
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)D * H; r++)
        for (int x = 0; x < W; x++) {
            double acc = 0.0;
            for (int j = -R; j <= R; j++) {
                int s = x - j;
                if (s >= 0 && s < W) acc += g[j + R] * a[r * W + s];
            }
            b[r * W + x] = acc;
        }
    /* y */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double acc = 0.0;
                for (int j = -R; j <= R; j++) {
                    int s = y - j;
                    if (s >= 0 && s < H) acc += g[j + R] * b[((long)z * H + s) * W + x];
                }
                a[((long)z * H + y) * W + x] = acc;
            }
    /* z */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double acc = 0.0;
                for (int j = -R; j <= R; j++) {
                    int s = z - j;
                    if (s >= 0 && s < D) acc += g[j + R] * a[((long)s * H + y) * W + x];
                }
                out[((long)z * H + y) * W + x] = decay * acc;
            }
    free(g); free(a); free(b);
    return 0;
}

/* Full 3D window of the Gaussian kernel (KD=KH=KW=2R+1, centre R) times decay. */
int orc_gauss_window3d(double sigma2, double decay, double *win, int R)
{
    int K = 2 * R + 1;
    double *g = (double *)malloc(sizeof(double) * K);
    if (!g) return -2;
    orc_gauss_window1d(sigma2, R, g);
    for (int z = 0; z < K; z++)
        for (int y = 0; y < K; y++)
            for (int x = 0; x < K; x++)
                win[((long)z * K + y) * K + x] = decay * g[z] * g[y] * g[x];
    free(g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* computeKernel3D (Diffusion.cu:1078-1531), restated in FP64.                */
/* ------------------------------------------------------------------------- */

/* complexPower's cpow (convolutionFFT.cuh:645-666): |z|^n cis(n atan(b/a)).
 * quirk=1 keeps the reference's atan(b/a) (phase off by pi when Re<0);
 * quirk=0 uses the principal-value power via atan2. */
static void orc_cpow(double *re, double *im, int n, int quirk)
{
    double a = *re, b = *im;
    double ab2 = a * a + b * b;
    double ph = quirk ? atan(b / a) : atan2(b, a);
    double mag = pow(ab2, (double)n / 2.0);
    *re = mag * cos(n * ph);
    *im = mag * sin(n * ph);
}

/* kernel_cctx fields used: dt, kernelRadius, niter, KD..KX (3,3,3 / 1,1,1),
 * DD..DW (kernel domain), FFTD..FFTW (kernel FFT box), window*.  Produces the
 * extracted window (windowD x windowH x windowW). */
int orc_compute_kernel3d(int kernelRadius, double lambda, double gamma, double dt, int niter,
                         int DD, int DH, int DW, int FD, int FH, int FW, int wD, int wH, int wW,
                         int quirk, double *window_out)
{
    const int KD = 3, KH = 3, KW = 3, KZ = 1, KY = 1, KX = 1;
    double h_kernel[27];
    memset(h_kernel, 0, sizeof h_kernel);
    /* 7-point stencil (Diffusion.cu:1187-1195) */
    int hs = KW, ds = KH * KW;
    h_kernel[0 * ds + 1 * hs + 1] = lambda;
    h_kernel[2 * ds + 1 * hs + 1] = lambda;
    h_kernel[1 * ds + 0 * hs + 1] = lambda;
    h_kernel[1 * ds + 2 * hs + 1] = lambda;
    h_kernel[1 * ds + 1 * hs + 0] = lambda;
    h_kernel[1 * ds + 1 * hs + 2] = lambda;
    h_kernel[1 * ds + 1 * hs + 1] = 1.0 - 6.0 * lambda - gamma * dt;

    long dn = (long)DD * DH * DW, fn = (long)FD * FH * FW;
    int hx = FW / 2 + 1;
    long sn = (long)FD * FH * hx;
    double *data = (double *)calloc(dn, sizeof(double));
    double *pdata = (double *)malloc(sizeof(double) * fn);
    double *pker = (double *)calloc(fn, sizeof(double));
    double *kspec = (double *)malloc(sizeof(double) * 2 * sn);
    double *dspec = (double *)malloc(sizeof(double) * 2 * sn);
    if (!data || !pdata || !pker || !kspec || !dspec) {
        free(data); free(pdata); free(pker); free(kspec); free(dspec); return -2;
    }
    /* stencil copied into the middle block of the domain (Diffusion.cu:1205-1229) */
    int sk = DD / 2 - KD / 2, si = DH / 2 - KH / 2, sj = DW / 2 - KW / 2;
    for (int k = 0; k < KD; k++)
        for (int i = 0; i < KH; i++)
            for (int j = 0; j < KW; j++)
                data[((long)(sk + k) * DH + (si + i)) * DW + (sj + j)] = h_kernel[(k * KH + i) * KW + j];

    double c = 1.0 / ((double)FD * FH * FW);
    for (int it = 0; it < niter; it++) {
        memset(pker, 0, sizeof(double) * fn);
        orc_pad_kernel3d(pker, h_kernel, FD, FH, FW, KD, KH, KW, KZ, KY, KX);
        orc_pad_clamp3d(pdata, data, FD, FH, FW, DD, DH, DW, KZ, KY, KX);
        orc_rfft3(FD, FH, FW, pker, kspec);
        orc_rfft3(FD, FH, FW, pdata, dspec);
        int n = kernelRadius - (it == 0 ? 1 : 0) - 1;
        for (long i = 0; i < sn; i++) orc_cpow(&kspec[2 * i], &kspec[2 * i + 1], n, quirk);
        for (long i = 0; i < sn; i++) {
            double ar = kspec[2 * i], ai = kspec[2 * i + 1];
            double br = dspec[2 * i], bi = dspec[2 * i + 1];
            dspec[2 * i] = c * (ar * br - ai * bi);
            dspec[2 * i + 1] = c * (ai * br + ar * bi);
        }
        orc_irfft3(FD, FH, FW, dspec, pdata);
        orc_unpad3d(data, pdata, DD, DH, DW, FD, FH, FW);
    }
    orc_extract_center3d(window_out, data, DD, DH, DW, wD, wH, wW);
    free(data); free(pdata); free(pker); free(kspec); free(dspec);
    return 0;
}

/* Closed-form stencil heat kernel: inverse DFT of S(theta)^M on an F^3 box,
 * S = 1 - 6 lambda - gamma dt + 2 lambda (cos tx + cos ty + cos tz), evaluated on
 * the centred window (wD x wH x wW).  Independent of the iterative restatement. */
int orc_stencil_heat_window(double lambda, double gamma, double dt, int M, int F,
                            int wD, int wH, int wW, double *window_out)
{
    long fn = (long)F * F * F;
    double *a = (double *)malloc(sizeof(double) * 2 * fn);
    if (!a) return -2;
    double c0 = 1.0 - 6.0 * lambda - gamma * dt;
    for (int z = 0; z < F; z++)
        for (int y = 0; y < F; y++)
            for (int x = 0; x < F; x++) {
                double s = c0 + 2.0 * lambda * (cos(2 * ORC_PI * z / F) + cos(2 * ORC_PI * y / F) +
                                                 cos(2 * ORC_PI * x / F));
                long i = ((long)z * F + y) * F + x;
                a[2 * i] = pow(s, M);
                a[2 * i + 1] = 0.0;
            }
    fft3_c2c_d(F, F, F, a, 1);
    for (int z = 0; z < wD; z++)
        for (int y = 0; y < wH; y++)
            for (int x = 0; x < wW; x++) {
                int sz = ((z - wD / 2) % F + F) % F, sy = ((y - wH / 2) % F + F) % F,
                    sx = ((x - wW / 2) % F + F) % F;
                window_out[((long)z * wH + y) * wW + x] = a[2 * (((long)sz * F + sy) * F + sx)] / (double)fn;
            }
    free(a);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Compensated (Neumaier) FP64 sums and error metrics.                        */
/* ------------------------------------------------------------------------- */
double orc_neumaier_sum_f32(const float *v, long n)
{
    double s = 0.0, c = 0.0;
    for (long i = 0; i < n; i++) {
        double x = v[i], t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
    return s + c;
}

double orc_neumaier_sum_f64(const double *v, long n)
{
    double s = 0.0, c = 0.0;
    for (long i = 0; i < n; i++) {
        double x = v[i], t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
    return s + c;
}

/* ------------------------------------------------------------------------- */
/* Synthetic inputs (counter-based, SplitMix64): identical on every host.     */
/* ------------------------------------------------------------------------- */
static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static double u01(uint64_t key, uint64_t ctr)
{
    uint64_t r = splitmix64(key ^ splitmix64(ctr));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
/* background: lognormal(mu, sigma) (Box-Muller on two counters) or, when
 * loguniform_lo > 0, log-uniform in [lo, hi]; plus nsrc point sources with
 * amplitude U[amp_lo, amp_hi] at uniform voxels. */
void orc_synth_field(float *f, int D, int H, int W, uint64_t key, double mu, double sigma,
                     double lo, double hi, int nsrc, double amp_lo, double amp_hi)
{
    long n = (long)D * H * W;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double v;
        if (lo > 0.0) {
            double t = u01(key, 2 * (uint64_t)i);
            v = exp(log(lo) + t * (log(hi) - log(lo)));
        } else {
            double a = u01(key, 2 * (uint64_t)i), b = u01(key, 2 * (uint64_t)i + 1);
            double g = sqrt(-2.0 * log(a)) * cos(2.0 * ORC_PI * b);
            v = exp(mu + sigma * g);
        }
        f[i] = (float)v;
    }
    uint64_t skey = splitmix64(key ^ 0x5EED5EED5EEDull);
    for (int s = 0; s < nsrc; s++) {
        long pos = (long)(u01(skey, 2 * (uint64_t)s) * (double)n);
        if (pos >= n) pos = n - 1;
        double amp = amp_lo + (amp_hi - amp_lo) * u01(skey, 2 * (uint64_t)s + 1);
        f[pos] += (float)amp;
    }
}

/* Per-field diffusion parameters: D ~ U[dlo,dhi], decay rate ~ U[llo,lhi]. */
void orc_synth_params(uint64_t key, int field, double dlo, double dhi, double llo, double lhi,
                      double *Dout, double *lamout)
{
    uint64_t k = splitmix64(key ^ (0xD1FFu + (uint64_t)field * 0x1000193ull));
    *Dout = dlo + (dhi - dlo) * u01(k, 1);
    *lamout = llo + (lhi - llo) * u01(k, 2);
}

/* ------------------------------------------------------------------------- */
/* FP32 CPU port (timed CPU baseline): the same pruned pipeline as the GPU      */
/* path, with its own FP32 mixed-radix FFT and OpenMP over lines.               */
/*   pad/R2C-x on the Ny*Nz data rows, C2C-y on Nz planes, C2C-z . K^ . C2C-z^-1,  */
/*   C2C-y^-1, C2R-x + unpad.  ZERO pad, P = 2N per axis, real spectrum K^.    */
/* ------------------------------------------------------------------------- */
typedef struct { float re, im; } cf;

static void fft_rec_f(int len, const cf *in, int s, cf *out, int sign, cf *tmp, const double *twr,
                      const double *twi, int Nfull, int tstride)
{
    if (len == 1) { out[0] = in[0]; return; }
    int p = (len % 4 == 0) ? 4 : smallest_factor(len);
    int m = len / p;
    for (int q = 0; q < p; q++)
        fft_rec_f(m, in + q * s, s * p, out + q * m, sign, tmp, twr, twi, Nfull, tstride * p);
    for (int k = 0; k < m; k++)
        for (int r = 0; r < p; r++) {
            int kk = k + m * r;
            float ar = 0.f, ai = 0.f;
            for (int q = 0; q < p; q++) {
                long e = ((long)q * kk) % len;
                /* W_len^e = W_Nfull^(e * Nfull/len) */
                long ti = e * (long)tstride;
                float c = (float)twr[ti], sn = (float)(sign * twi[ti]);
                cf f = out[q * m + k];
                ar += f.re * c - f.im * sn;
                ai += f.re * sn + f.im * c;
            }
            tmp[kk].re = ar; tmp[kk].im = ai;
        }
    memcpy(out, tmp, sizeof(cf) * len);
}

typedef struct { int n; double *twr, *twi; } ftab;
static void ftab_init(ftab *t, int n)
{
    t->n = n;
    t->twr = (double *)malloc(sizeof(double) * n);
    t->twi = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) {
        t->twr[k] = cos(2.0 * ORC_PI * k / n);
        t->twi[k] = sin(2.0 * ORC_PI * k / n);
    }
}
static void ftab_free(ftab *t) { free(t->twr); free(t->twi); }
/* unnormalized in-place FFT of x[0..n) using scratch of 2n */
static void fft_f32(const ftab *t, cf *x, int inverse, cf *scratch)
{
    fft_rec_f(t->n, x, 1, scratch, inverse ? +1 : -1, scratch + t->n, t->twr, t->twi, t->n, 1);
    memcpy(x, scratch, sizeof(cf) * t->n);
}

int cpu_fft_diffuse3d_f32(const float *u, float *out, int Nz, int Ny, int Nx, const float *kreal,
                          int nthreads)
{
    /* kreal: real kernel spectrum [Pz][Py][Px/2+1], already scaled by 1/(Px Py Pz). */
    int Px = 2 * Nx, Py = 2 * Ny, Pz = 2 * Nz, L = Px / 2, Hx = L + 1;
    long A = (long)Nz * Py * Hx; /* working spectrum [Nz][Py][Hx] */
    cf *w = (cf *)malloc(sizeof(cf) * A);
    if (!w) return -2;
    ftab tL, tY, tZ;
    ftab_init(&tL, L); ftab_init(&tY, Py); ftab_init(&tZ, Pz);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    /* pass 1: R2C-x of each data row (pack pairs, length-L FFT, post-process) */
#pragma omp parallel
    {
        int mx = Py > Pz ? Py : Pz; if (L + 1 > mx) mx = L + 1;
        cf *line = (cf *)malloc(sizeof(cf) * mx);
        cf *v = (cf *)malloc(sizeof(cf) * mx);
        cf *scr = (cf *)malloc(sizeof(cf) * 2 * mx);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)Nz * Ny; r++) {
            int z = (int)(r / Ny), y = (int)(r % Ny);
            const float *row = u + r * Nx;
            for (int n = 0; n < L; n++) {
                int a = 2 * n, b = 2 * n + 1;
                line[n].re = a < Nx ? row[a] : 0.f;
                line[n].im = b < Nx ? row[b] : 0.f;
            }
            fft_f32(&tL, line, 0, scr);
            cf *dst = w + ((long)z * Py + y) * Hx;
            for (int k = 0; k <= L; k++) {
                cf zk = line[k % L], zc = line[(L - k) % L];
                zc.im = -zc.im;
                double ang = -2.0 * ORC_PI * k / Px;
                float c = (float)cos(ang), s = (float)sin(ang);
                float er = 0.5f * (zk.re + zc.re), ei = 0.5f * (zk.im + zc.im);
                float dr = 0.5f * (zk.re - zc.re), di = 0.5f * (zk.im - zc.im);
                float wdr = c * dr - s * di, wdi = c * di + s * dr; /* W d */
                dst[k].re = er + wdi;                                /* e - i W d */
                dst[k].im = ei - wdr;
            }
        }
        /* pass 2: forward C2C-y on each (z, kx) column, Ny inputs -> Py outputs */
#pragma omp for schedule(static)
        for (long c = 0; c < (long)Nz * Hx; c++) {
            int z = (int)(c / Hx), kx = (int)(c % Hx);
            for (int y = 0; y < Py; y++) {
                if (y < Ny) v[y] = w[((long)z * Py + y) * Hx + kx];
                else { v[y].re = 0.f; v[y].im = 0.f; }
            }
            fft_f32(&tY, v, 0, scr);
            for (int y = 0; y < Py; y++) w[((long)z * Py + y) * Hx + kx] = v[y];
        }
        /* pass 3: C2C-z . K^ . C2C-z^-1 on each (ky, kx) column, Nz in, Nz out */
#pragma omp for schedule(static)
        for (long c = 0; c < (long)Py * Hx; c++) {
            int ky = (int)(c / Hx), kx = (int)(c % Hx);
            for (int z = 0; z < Pz; z++) {
                if (z < Nz) v[z] = w[((long)z * Py + ky) * Hx + kx];
                else { v[z].re = 0.f; v[z].im = 0.f; }
            }
            fft_f32(&tZ, v, 0, scr);
            for (int z = 0; z < Pz; z++) {
                float kk = kreal[((long)z * Py + ky) * Hx + kx];
                v[z].re *= kk; v[z].im *= kk;
            }
            fft_f32(&tZ, v, 1, scr);
            for (int z = 0; z < Nz; z++) w[((long)z * Py + ky) * Hx + kx] = v[z];
        }
        /* pass 4: inverse C2C-y, Py in, Ny out */
#pragma omp for schedule(static)
        for (long c = 0; c < (long)Nz * Hx; c++) {
            int z = (int)(c / Hx), kx = (int)(c % Hx);
            for (int y = 0; y < Py; y++) v[y] = w[((long)z * Py + y) * Hx + kx];
            fft_f32(&tY, v, 1, scr);
            for (int y = 0; y < Ny; y++) w[((long)z * Py + y) * Hx + kx] = v[y];
        }
        /* pass 5: C2R-x + unpad */
#pragma omp for schedule(static)
        for (long r = 0; r < (long)Nz * Ny; r++) {
            const cf *src = w + ((long)(r / Ny) * Py + (r % Ny)) * Hx;
            for (int k = 0; k < L; k++) {
                cf ak = src[k], ac = src[L - k];
                ac.im = -ac.im;
                double ang = 2.0 * ORC_PI * k / Px;
                float c = (float)cos(ang), s = (float)sin(ang);
                float er = ak.re + ac.re, ei = ak.im + ac.im;
                float dr = ak.re - ac.re, di = ak.im - ac.im;
                float odr = c * dr - s * di, odi = c * di + s * dr; /* (A - A*) W^-k */
                line[k].re = er - odi;                               /* Xe + i Xo */
                line[k].im = ei + odr;
            }
            fft_f32(&tL, line, 1, scr);
            float *orow = out + r * Nx;
            for (int n = 0; n < L; n++) {
                int a = 2 * n, b = 2 * n + 1;
                if (a < Nx) orow[a] = line[n].re;
                if (b < Nx) orow[b] = line[n].im;
            }
        }
        free(line); free(v); free(scr);
    }
    ftab_free(&tL); ftab_free(&tY); ftab_free(&tZ);
    free(w);
    return 0;
}

/* Real kernel spectrum of the separable Gaussian on the padded box, FP64:
 * K^(kz,ky,kx) = scale * decay * Gz(kz) Gy(ky) Gx(kx), G(k) = sum_j g[j] cos(2 pi j k/P)
 * over |j| <= R, layout [Pz][Py][Px/2+1]. */
static double gauss_dft_1d(const double *g, int R, int P, int k)
{
    double s = 0.0;
    for (int j = -R; j <= R; j++) {
        long e = (((long)j * k) % P + P) % P;
        s += g[j + R] * cos(2.0 * ORC_PI * (double)e / (double)P);
    }
    return s;
}

int orc_gauss_spectrum_real(double sigma2, double decay, int Pz, int Py, int Px, double scale,
                            float *out)
{
    int R = orc_gauss_radius(sigma2), Hx = Px / 2 + 1;
    double *g = (double *)malloc(sizeof(double) * (2 * R + 1));
    double *Gx = (double *)malloc(sizeof(double) * Hx);
    double *Gy = (double *)malloc(sizeof(double) * Py);
    double *Gz = (double *)malloc(sizeof(double) * Pz);
    if (!g || !Gx || !Gy || !Gz) { free(g); free(Gx); free(Gy); free(Gz); return -2; }
    orc_gauss_window1d(sigma2, R, g);
    for (int k = 0; k < Hx; k++) Gx[k] = gauss_dft_1d(g, R, Px, k);
    for (int k = 0; k < Py; k++) Gy[k] = gauss_dft_1d(g, R, Py, k);
    for (int k = 0; k < Pz; k++) Gz[k] = gauss_dft_1d(g, R, Pz, k);
    for (int z = 0; z < Pz; z++)
        for (int y = 0; y < Py; y++)
            for (int x = 0; x < Hx; x++)
                out[((long)z * Py + y) * Hx + x] = (float)(scale * decay * Gz[z] * Gy[y] * Gx[x]);
    free(g); free(Gx); free(Gy); free(Gz);
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}


/* ------------------------------------------------------------------------- */
/* WHChemical::update with gpuUpdate (WHChemical.cpp:395-450): per voxel, for  */
/* the 8 chems (enums.h:124-133) level = d + t * decay (FP32, the product      */
/* rounded before the add), d = 0, p = level; gradients (WHChemical.cpp:       */
/* 437-445) in the reference's left-to-right FP32 order; totals as            */
/* compensated FP64 sums of p.  grads = NEU, MAC, FIB (enums.h:134-136).       */
/* ------------------------------------------------------------------------- */
int orc_chem_update(const float *const *t, float *const *d, float *const *p, const float *decay,
                    const float *frag_col, const float *frag_ela, float *neu, float *mac, float *fib,
                    long n, double *totals)
{
    enum { TNF = 0, TGF = 1, FGF = 2, IL1beta = 4, IL6 = 5, IL8 = 6 };
    double s[8] = {0}, c[8] = {0};
    for (long i = 0; i < n; i++) {
        float lv[8];
        for (int ic = 0; ic < 8; ic++) {
            volatile float scaled = t[ic][i] * decay[ic];
            lv[ic] = d[ic][i] + scaled;
            d[ic][i] = 0.0f;
            p[ic][i] = lv[ic];
            double x = lv[ic], tt = s[ic] + x;
            c[ic] += fabs(s[ic]) >= fabs(x) ? (s[ic] - tt) + x : (x - tt) + s[ic];
            s[ic] = tt;
        }
        float grad = lv[IL1beta] + lv[TNF] + lv[TGF] + lv[FGF] + frag_col[i];
        fib[i] = lv[TGF];
        neu[i] = grad + lv[IL6] + lv[IL8];
        mac[i] = grad + frag_ela[i];
    }
    if (totals)
        for (int ic = 0; ic < 8; ic++) totals[ic] = s[ic] + c[ic];
    return 0;
}


/* sampleChem_kernel (Visualization/3D/volumeRender_kernel.cu:215-300): every   */
/* stride-th voxel times the chem's display scale, clamped above at 1.          */
int orc_sample_chem(const float *src, int D, int H, int W, int chem, int stride, float *dst)
{
    static const float scale[8] = {50000.0f, 10000.0f, 1000000.0f, 100.0f,
                                   10000.0f, 100000.0f, 100000.0f, 10000.0f};
    if (chem < 0 || chem > 7 || stride < 1) return -1;
    int sD = D / stride, sH = H / stride, sW = W / stride;
    for (int z = 0; z < sD; z++)
        for (int y = 0; y < sH; y++)
            for (int x = 0; x < sW; x++) {
                float v = src[((long)z * stride * H + (long)y * stride) * W + (long)x * stride] * scale[chem];
                if (v > 1.0f) v = 1.0f;
                dst[((long)z * sH + y) * sW + x] = v;
            }
    return 0;
}
