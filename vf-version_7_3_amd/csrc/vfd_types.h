// vfd_types.h -- descriptors shared by the host runtime and the device passes.
#pragma once

namespace vfd {

// pad modes: the reference's pad_t (src/enums.h:35-40) value for the constant
// pad, plus the zero pad of the linear-convolution configuration.
enum PadMode { PAD_CONSTANT_VF = 3, PAD_ZERO = 100 };

enum ColMode { COL_FWD = 0, COL_INV = 1, COL_FUSED = 2 };

// Per-field pointers of one launch.
struct FieldDesc {
    const void* in;     // real field [Nz][Ny][Nx] (x-forward input)
    void* out;          // real field [Nz][Ny][Nx] (x-inverse output; may alias in)
    void* work;         // complex work [planes][Py][ld]
    const void* kspec;  // kernel spectrum (real or complex), see ColDesc strides
    double base;        // constant pad value (PAD_CONSTANT_VF), else 0
    int epi;            // epithelium index: interior x < epi (PAD_CONSTANT_VF)
    int pad_;
    float* delta;       // fused chem update: out = delta + t * decay, delta = 0 (null: plain output)
    float decay;
    int vis_stride;     // fused visualisation downsample of t (vis non-null): every vis_stride-th voxel
    float* vis;         //   min(t * vis_scale, 1) into a dense [Nz/s][Ny/s][Nx/s] array
    float vis_scale;
    int pad2_;
};

// Fields of one launch, passed by value as a kernel argument (graph-capture
// friendly, no device-side descriptor copies).  blockIdx.z selects the field.
constexpr int kMaxFieldsPerLaunch = 32;
struct FieldBatch {
    FieldDesc f[kMaxFieldsPerLaunch];
};

// Shape of the x passes (identical for every field of a launch).
struct GridDesc {
    int Nx, Ny, Nz;     // data box
    int L;              // Px / 2 (packed complex row length)
    long ld;            // work row stride (complex elements), >= L + 1
    long plane;         // work plane stride = Py * ld
    int pad_mode;       // PadMode
    int kspec_complex;  // 1: complex spectrum, 0: real
    unsigned long in_bytes, work_bytes, kspec_bytes, out_bytes;  // per-field array sizes
    // constant-pad restore term base * K^(0) * kdc_scale: the complex spectrum
    // is stored unnormalized (1), the real one with 1/(Px Py Pz) folded in
    // (Px Py Pz)
    double kdc_scale;
};

// One column pass (y or z).
struct ColDesc {
    int mode;           // ColMode
    int rs;             // 1 or 2 (even/odd split of the padded length)
    int nin, nout;      // valid input rows / rows written back
    long ncols;         // columns per plane (= L + 1)
    long rstride;       // element stride between rows of a column
    long pstride;       // element stride between planes (blockIdx.y)
    long ks_row;        // kernel-spectrum stride along this axis' frequency
    long ks_plane;      // kernel-spectrum stride along the plane index
    int kspec_complex;
    int kspec_sep;      // K^ from FP64 tables [scale | gx | gy | gz]
    int sep_gy, sep_gz; // table offsets (elements) of gy and gz
    double scale;       // applied to K^ products (1 / (Px Py Pz))
    unsigned long work_bytes, kspec_bytes;  // per-field array sizes
    // grouped 1-D launch of narrow tiles (set by the launcher): planes and
    // fields of the launch, 128-byte row segments per plane row
    unsigned grid_planes, grid_fields, grid_segs;
    int sep_g;          // separable: table offset (elements) of [R | g_0 .. g_kZDirectMaxR]
    int fold;           // rs = 2 with output rows past P/2 (FP32 inverse-only, fold_supported)
};

// The x-last pass order (z R2C, y C2C, x fused, y inverse, z C2R) for boxes
// padded most along x.  Work of one field: complex [hz][py][ldx], hz = pz/2+1
// z half-spectrum planes of py rows; element (kz, y, x) at kz*plane + y*ldx + x.
struct XlDesc {
    int nx, ny, nz;      // data box
    int px, py, pz;      // FFT box
    int hz;              // pz / 2 + 1
    int nxp;             // real column pairs per data row, (nx + 1) / 2
    long ncolp;          // column pairs per field, ny * nxp
    long ldx;            // work row stride (complex elements), even, > nx when nx is odd
    long plane;          // work plane stride py * ldx
    long ks_hx;          // spectrum row length px/2 + 1 (the R2C layout [pz][py][px/2+1])
    int pad_mode;        // PadMode
    double scale;        // applied to K^ products (1 / (px py pz))
    unsigned long in_bytes, work_bytes, kspec_bytes, out_bytes;  // per-field array sizes
    // grouped 1-D launch of narrow tiles (set by the launcher)
    unsigned grid_fields, grid_segs;
};

// Largest window radius the direct z pass takes (separable spectrum, FP32);
// the separable table carries the taps g_0 .. g_kZDirectMaxR.  Wider windows
// keep the FFT form: with a loop-rolled window, radius 20 and 24 compiled to
// scratch-memory window accesses; fully unrolled (every window index a
// constant), radius 20 held its window in 194 VGPRs without scratch but, at two
// waves per SIMD, ran slower than the FFT form (0.60 against 0.53 ms per field
// on 1024x512x128), and capped at three or four waves it spilled.
constexpr int kZDirectMaxR = 16;

// Every per-field array is addressed with 32-bit byte offsets.
constexpr unsigned long kBufLimit = 0xFFFFFF00ul;

}  // namespace vfd
