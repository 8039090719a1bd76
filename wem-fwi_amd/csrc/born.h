// Descriptors and launchers of the Born / imaging kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace wemfwi {

// One (shot, frequency) row of a LinDown / LinUp application.
struct LinRow {
  float2* bg;            // [nz][nx] background wavefield
  const float2* m;       // forward: perturbation [nz][nx]; adjoint: wavefield [nz][nx]
  float2* out;           // forward: wavefield (+=); adjoint: model contribution (+=)
  const float2* ptab;    // phase table of the row's frequency
  float w;
  int model;
  int zbase;             // depth of the first row of bg / m / out (0: whole arrays)
  const float2* post;    // adjoint only, or NULL: after use, bg[d] <- post[d] * bg[d] (post: [nz][nx]
                         // by absolute depth, e.g. the model's reflection coefficient)
  int assign;            // adjoint only: 1 stores the contribution instead of adding it
};

// Injection of many devices as a gather per touched cell (CSR by cell).
struct InjPlan {
  int ncell;
  const long long* cell;  // [ncell] offset z*nx + x
  const int* ptr;         // [ncell+1]
  const int* dev;         // [nc] device of each contribution
  const float* wgt;       // [nc] cos^2 weight
};

// Recording of many devices (CSR by device).
struct RecPlan {
  int ndev;
  const int* ptr;         // [ndev+1]
  const long long* cell;  // [nc] offset z*nx + x of each contribution
  const float* wgt;       // [nc]
};

// depths d of [dlo, dlo + dcount) (dlo < 0: every depth of the direction)
hipError_t launch_lin(int N, const LinRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz,
                      int dir, bool adj, bool exact, const float2* tw, hipStream_t s, int dlo = -1, int dcount = -1);
hipError_t launch_scatter(int N, const StepRow* rows, int nrows, const DepthTables* models, StepGeom g, bool adj,
                          bool exact, const float2* tw, hipStream_t s);
// One row of a batched product out (=|+=) op(a) * b over n samples.
struct MulRow {
  float2* out;
  const float2* a;
  const float2* b;
};

hipError_t launch_mul_rows(const MulRow* rows, int nrows, long long n, bool conjA, bool acc, hipStream_t s);
hipError_t launch_mul(float2* out, long long so, const float2* A, long long sa, const float2* B, long long sb,
                      long long n, int nrows, bool conjA, bool acc, hipStream_t s);
// Rows of a reduction are summed in runs of kRunRows rows of their global
// order (runs start at multiples of kRunRows): each run in row order, then
// the run sums in run order.  Chunks of rows that start at multiples of
// kRunRows therefore give the same sums however the rows are chunked.
constexpr int kRunRows = 64;
// out[i] (+)= sum_r in[r * stride + i] for i < n over nrows rows, the first of
// which sits at position phase of its run
// in2 non-NULL: the summand is conj(in) * in2
hipError_t launch_reduce_rows(double2* out, const float2* in, long long stride, long long n, int nrows, int phase,
                              bool acc, hipStream_t s, const float2* in2 = nullptr);
hipError_t launch_add(float2* out, const float2* in, long long n, hipStream_t s);
// zero rows x width 8-byte words at base + row * pitch (a kernel, capturable)
hipError_t launch_zero2d(void* base, long long pitch_words, long long width_words, long long rows, hipStream_t s);
// res -= obs over n samples and *half_sum = 0.5 sum |res|^2 in FP64, in a
// fixed order (kMisfitParts partial sums)
constexpr int kMisfitParts = 256;
hipError_t launch_residual_sumsq(float2* res, const float2* obs, long long n, double* part, double* half_sum,
                                 hipStream_t s);
// Phshift rows (w, sref per row; device arrays)
hipError_t launch_phshift(float2* out, const float2* in, int nk, double dk, double dz, double eps, const double* w,
                          const double2* sref, int nrows, bool adjoint, bool add, hipStream_t s);
// out[r][loc[j]] = in[r][loc[j]] * mult[r][j]
hipError_t launch_gather_mul(float2* out, const float2* in, const float2* mult, const int* loc, int nloc, int nx,
                             int nrows, hipStream_t s);
hipError_t launch_gather_sum(float2* out, long long n, const long long* ptr, const long long* src, const float2* in,
                             bool add, hipStream_t s);
hipError_t launch_d2f(float2* out, const double2* in, long long n, bool acc, hipStream_t s);
// out[z][a] = in[z][x0 + a] (0 outside [0, nxg)), z < nz, a < nxw; and its
// adjoint acc[z][x0 + a] += in[z][a] (FP64 accumulator [nz][nxg])
hipError_t launch_window(float2* out, const float2* in, int nz, int nxw, int nxg, int x0, hipStream_t s);
hipError_t launch_window_add(double2* acc, const float2* in, int nz, int nxw, int nxg, int x0, hipStream_t s);
// W holds rows zbase.. of each [nz][nx] field; sources outside [zlo, zhi]
// (cells outside [clo, chi), with cbase = zbase*nx) are skipped
hipError_t launch_inject_point(float2* W, long long wstride, int nx, const int* iz, const int* ix,
                               const float2* val, const float* gw, int ng, int tap, int nrows, hipStream_t s,
                               int zbase = 0, int zlo = 0, int zhi = 0x7fffffff);
hipError_t launch_inject_csr(float2* W, long long wstride, InjPlan p, const float2* val, const long long* vbase,
                             long long vdev, int nrows, hipStream_t s, long long cbase = 0, long long clo = 0,
                             long long chi = 0x7fffffffffffffffLL);
hipError_t launch_record(float2* val, const long long* vbase, long long vdev, const float2* W, long long wstride,
                         RecPlan p, int nrows, hipStream_t s);
hipError_t launch_dreflect_fwd(float2* dr, const float2* coef, const float2* dS, const float2* dR, int nz, int nx,
                               bool acc, hipStream_t s);
hipError_t launch_dreflect_adj(float2* gS, float2* gR, const float2* coef, const float2* dr, int nz, int nx,
                               hipStream_t s);
hipError_t launch_offset_ic_adj(double2* img, const float2* const* bg, const float2* const* sc, int nrows, int nh,
                                int nz, int nx, hipStream_t s);
hipError_t launch_offset_ic_fwd(float2* src, const float2* bg, const float2* img, int nh, int nz, int nx,
                                hipStream_t s);

hipError_t launch_trace_twiddles(double2* tw, int n, hipStream_t s);
hipError_t launch_trace_to_freq(const float* x, int ntr, int nt, int n, const int* bins, int nw, const double2* tw,
                                float2* f, int nbins, hipStream_t s);
hipError_t launch_freq_to_trace(const float2* f, int ntr, int nbins, const int* bins, int nw, int nt, int n,
                                const double2* tw, float* x, int add, hipStream_t s);

}  // namespace wemfwi
