// vfd_internal.h -- host-side plan and helpers shared by the runtime files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/vfd.h"
#include "vfd_types.h"

namespace vfd {

struct Plan {
    int nz, ny, nx, pz, py, px;
    int prec, pad_mode, kspec_format;
    int L, hx;
    int rs_y, rs_z, qy, qz;
    int fold_y;  // y inverse pass split and folded (ny > py/2, col_kernel FOLD with rs = 2)
    bool full_box;
    long ld, plane;
    long ks_row, ks_plane;
    size_t work_bytes, kspec_bytes, field_bytes;
    double bytes_per_field_tick;
    const void *tw_x, *tw_y, *tw_z;
    int var[5];  // schedule variant per pass
    // pass order: VFD_ORDER_XFIRST (x R2C, y, z fused, y, x C2R) or
    // VFD_ORDER_XLAST (z R2C, y, x fused, y, z C2R; FP32 complex spectrum)
    int order;
    int hz;        // x-last: pz/2 + 1 z half-spectrum planes
    long ldx;      // x-last: work row stride (complex elements)
    long xplane;   // x-last: work plane stride py * ldx
};

int set_error(int code, const char* fmt, ...);
int hip_check(hipError_t e, const char* what);
int twiddles(int p, int prec, const void** out);

int plan_init(const vfd_plan_desc* d, Plan* p);
void tuned_lookup(Plan& q);
bool tuned_known(const Plan& q);  // a tuned schedule of this shape is cached
int plan_autotune(Plan& p, int nfields, hipStream_t st, float* best_ms);
int plan_diffuse(const Plan& p, const vfd_field* fields, int nfields, hipStream_t st);
float chem_display_scale(int chem);  // sampleChem's per-chem display scale, chem < 8
int chem_levels(float* d_p, const float* d_t, float* d_d, float decay, long n, hipStream_t st);
int chem_grads_totals(float* const* d_p, const float* d_frag_col, const float* d_frag_ela, float* const* d_grads,
                      long n, double* totals, hipStream_t st);
// Optional work of the x-inverse epilogue (FP32 plans), per field f:
//  deltas/decay: out = deltas[f] + t * decay[f], deltas[f] cleared (the chem update);
//  vis/vis_scale/vis_stride: the visualisation downsample min(t * vis_scale[f], 1)
//  of every vis_stride-th voxel into vis[f].
struct Epilogue {
    float* const* deltas = nullptr;
    const float* decay = nullptr;
    float* const* vis = nullptr;
    const float* vis_scale = nullptr;
    int vis_stride = 0;
};
int plan_diffuse_passes(const Plan& p, const vfd_field* fields, int nfields, int lo, int hi, hipStream_t st,
                        const Epilogue* epi = nullptr);
double plan_pass_bytes(const Plan& p, int pass);
int plan_gauss_spectrum(const Plan& p, double sigma2, double decay, void* d_kspec, hipStream_t st);
int plan_mass(const Plan& p, const void* const* d_fields, int nfields, double* mass, hipStream_t st);
int plan_r2c(const Plan& p, const void* in, void* out, hipStream_t st);
int plan_c2r(const Plan& p, void* in, void* out, hipStream_t st);

}  // namespace vfd

// the opaque plan handle of the C-ABI (include/vfd.h)
struct vfd_plan {
    vfd::Plan p;
};
