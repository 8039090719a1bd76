// HIP kernels of the linearised (Born) operators, imaging conditions,
// injection/recording and the deterministic frequency/shot reduction.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "born.h"
#include "sweep.hpp"

namespace wemfwi {

constexpr int kLinWaves = 4;

// row loads/accumulates of the general variant (padding allowed)
template <class V>
__device__ __forceinline__ void lrow(float2 (&v)[V::E], const float2* __restrict__ p, const StepGeom& g, int t) {
#pragma unroll
  for (int e = 0; e < V::E; ++e) v[e] = row_sample<V>(p, g, t, e);
}

// p[x] += v for the model samples of the row (through the row's buffer
// descriptor: padding samples read 0 and their stores are dropped)
template <class V>
__device__ __forceinline__ void acc_row(float2* __restrict__ p, const float2 (&v)[V::E], const StepGeom& g,
                                        int t) {
#pragma unroll
  for (int e = 0; e < V::E; ++e) row_store<V>(p, cadd(row_sample<V>(p, g, t, e), v[e]), g, t, e);
}

// LinDown / LinUp (LinOneWay.cpp:33-142) as independent (row, depth) items.
// dir = +1 (LinDown): depths d = 0..nz-2, neighbour d+1; dir = -1 (LinUp):
// d = 1..nz-1, neighbour d-1.
//   forward: out[d+dir] += P_d Sc_d (bg[d] * m[d])
//   adjoint: out[d]     += conj(bg[d]) * Sc^H_d P^H_d (in[d+dir])
// FX: the compile-time pass chain of GEN rows (GenFixed id, 0: runtime)
template <int N, bool ADJ, bool EXACT, bool GEN, int FX = 0>
__global__ void __launch_bounds__(64 * kLinWaves)
    __attribute__((amdgpu_waves_per_eu(N <= 256 ? 4 : 2, 8)))
    lin_kernel(const LinRow* __restrict__ rows, int nrows, const DepthTables* __restrict__ models, StepGeom g,
               int dlo, int nd, int dir, const float2* __restrict__ twg) {
  using V = Var<N, false, true, GEN, FX>;
  constexpr int E = N / 64;
  __shared__ float2 tw[TwV<V>::value];
  __shared__ float2 bufs[kLinWaves][LdsV<V>::value];
  const int t = threadIdx.x & 63;
  const int wv = wave_id();
  for (int i = threadIdx.x; i < TwV<V>::value; i += blockDim.x) tw[i] = ldg(twg, (size_t)i);
  __syncthreads();
  const int item = blockIdx.x * kLinWaves + wv;
  if (item >= nrows * nd) return;
  const int row = item / nd;
  const int d = dlo + (item - row * nd);
  const LinRow R = ldu_struct(rows, row);
  const DepthTables dt = ldu_struct(models, R.model);
  // the row's width, padding and taper come from its model (a launch may
  // hold rows of different apertures; it fixes the FFT length)
  g.nx = dt.nx;
  g.tap = dt.tap;
  g.taper = dt.taper;
  float2* buf = bufs[wv];
  const size_t nx = g.nx;
  const int r0 = d - R.zbase;  // buffer row of depth d
  float2 st[E];
  if constexpr (!ADJ) {
    float2 b[E];
    lrow<V>(st, R.m + r0 * nx, g, t);
    lrow<V>(b, R.bg + r0 * nx, g, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmul(b[e], st[e]);
    scatter_row<V, false, false>(st, d, R.w, dt, g, buf, tw, t);
    step_forward<V>(st, d, R.w, dt, g, R.ptab, buf, tw, t);
    acc_row<V>(R.out + (r0 + dir) * nx, st, g, t);
  } else {
    lrow<V>(st, R.m + (r0 + dir) * nx, g, t);
    step_adjoint<V>(st, d, R.w, dt, g, R.ptab, buf, tw, t);
    scatter_row<V, true, EXACT>(st, d, R.w, dt, g, buf, tw, t);
    float2 b[E];
    lrow<V>(b, R.bg + r0 * nx, g, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmulc(st[e], b[e]);
    if (R.assign) {
#pragma unroll
      for (int e = 0; e < E; ++e) row_store<V>(R.out + r0 * nx, st[e], g, t, e);
    } else {
      acc_row<V>(R.out + r0 * nx, st, g, t);
    }
    if (R.post) {
      // bg[d] <- post[d] * bg[d]: this item is the only reader of bg row d
      float2 q[E];
      lrow<V>(q, R.post + (size_t)d * nx, g, t);
#pragma unroll
      for (int e = 0; e < E; ++e) row_store<V>(R.bg + r0 * nx, cmul(q[e], b[e]), g, t, e);
    }
  }
}

// Scatter as a unit operator on independent rows: out (+)= Sc_d(in).
template <int N, bool ADJ, bool EXACT, bool GEN, int FX = 0>
__global__ void __launch_bounds__(64 * kLinWaves)
    scatter_kernel(const StepRow* __restrict__ rows, int nrows, const DepthTables* __restrict__ models,
                   StepGeom g, const float2* __restrict__ twg) {
  using V = Var<N, false, true, GEN, FX>;
  constexpr int E = N / 64;
  __shared__ float2 tw[TwV<V>::value];
  __shared__ float2 bufs[kLinWaves][LdsV<V>::value];
  const int t = threadIdx.x & 63;
  const int wv = wave_id();
  for (int i = threadIdx.x; i < TwV<V>::value; i += blockDim.x) tw[i] = ldg(twg, (size_t)i);
  __syncthreads();
  const int row = blockIdx.x * kLinWaves + wv;
  if (row >= nrows) return;
  const StepRow R = ldu_struct(rows, row);
  const DepthTables dt = ldu_struct(models, R.model);
  float2 st[E];
  lrow<V>(st, R.in, g, t);
  scatter_row<V, ADJ, EXACT>(st, R.d, R.w, dt, g, bufs[wv], tw, t);
  if (R.add) {
    acc_row<V>(R.out, st, g, t);
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) row_store<V>(R.out, st[e], g, t, e);
  }
}

// out[r so + i] (= or +=) op(A[r sa + i]) * B[r sb + i]; op = identity or conj
__global__ void mul_kernel(float2* __restrict__ out, long long so, const float2* __restrict__ A, long long sa,
                           const float2* __restrict__ B, long long sb, long long n, int nrows, int conjA,
                           int acc) {
  const long long total = n * nrows;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / n, k = i - r * n;
    const float2 a = ldg(A, (size_t)(r * sa + k));
    const float2 b = ldg(B, (size_t)(r * sb + k));
    float2 v = conjA ? cmulc(b, a) : cmul(a, b);
    float2* o = out + r * so + k;
    if (acc) v = cadd(*o, v);
    *o = v;
  }
}

// Batched rows of out = op(a) * b (or +=): x over the samples of a row, y over
// rows (strided when there are more rows than the grid's y extent).
__global__ void mul_rows_kernel(const MulRow* __restrict__ rows, int nrows, long long n, int conjA, int acc) {
  for (int r = blockIdx.y; r < nrows; r += gridDim.y) {
    const MulRow R = rows[r];
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
      const float2 a = ldg(R.a, (size_t)k);
      const float2 b = ldg(R.b, (size_t)k);
      float2 v = conjA ? cmulc(b, a) : cmul(a, b);
      if (acc) v = cadd(R.out[k], v);
      R.out[k] = v;
    }
  }
}

// out[i] (+)= sum_r in[r stride + i] in double precision, rows in order
// out[i] (+)= sum over rows of in[r * stride + i], summed in runs: the rows
// are cut at multiples of kRunRows of their global order (phase = position
// of row 0 within its run), each run is summed in double in row order and
// the run sums are added to out in run order.  One wave per run and 64
// consecutive samples (one per lane, coalesced 512-byte row loads); the
// waves of a workgroup take kRedWaves runs at a time and wave 0 adds their
// sums in order from LDS.  PROD: the summand is conj(in) * in2, formed in
// FP32 as a product row would be (the product is never stored).
constexpr int kRedWaves = 8;
template <bool PROD>
__global__ void __launch_bounds__(64 * kRedWaves)
    reduce_runs_kernel(double2* __restrict__ out, const float2* __restrict__ in, const float2* __restrict__ in2,
                       long long stride, long long n, int nrows, int phase, int acc) {
  __shared__ double2 part[kRedWaves][64];
  const int t = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long i = (long long)blockIdx.x * 64 + t;
  const bool live = i < n;
  const int nruns = (phase + nrows + kRunRows - 1) / kRunRows;
  double tx = 0.0, ty = 0.0;
  if (acc && live && wv == 0) {
    const double2 o = out[i];
    tx = o.x;
    ty = o.y;
  }
  for (int g = 0; g < nruns; g += kRedWaves) {
    const int m = g + wv;
    if (m < nruns) {
      const int r0 = max(0, m * kRunRows - phase), r1 = min(nrows, (m + 1) * kRunRows - phase);
      double sx = 0.0, sy = 0.0;
      if (live) {
#pragma unroll 4
        for (int r = r0; r < r1; ++r) {
          float2 v = ldg(in, (size_t)((long long)r * stride + i));
          if constexpr (PROD) v = cmulc(ldg(in2, (size_t)((long long)r * stride + i)), v);
          sx += v.x;
          sy += v.y;
        }
      }
      part[wv][t] = make_double2(sx, sy);
    }
    __syncthreads();
    if (wv == 0) {
      const int mend = min(kRedWaves, nruns - g);
      for (int w = 0; w < mend; ++w) {
        tx += part[w][t].x;
        ty += part[w][t].y;
      }
    }
    __syncthreads();
  }
  if (wv == 0 && live) out[i] = make_double2(tx, ty);
}

// Phshift::forward/adjoint (phase_shift.ispc:2-32) on independent k rows:
// out[k] (+)= in[k] P(k) (forward) or in[k] conj(P(k)) (adjoint) with
// P = exp(-i kz dz), kz = re + i im of phase_shift.ispc:8-12 (im <= 0), on
// the k grid of Phshift_cmplx.cpp:12-24: k[i] = i dk for i < nk/2,
// k[nk-i] = -k[i] for 1 <= i < nk/2, other entries 0.  Evaluated in FP64 as
// the phase tables are (see phase_table_kernel).
__global__ void phshift_rows_kernel(float2* __restrict__ out, const float2* __restrict__ in, int nk, double dk,
                                    double dz, double eps, const double* __restrict__ w,
                                    const double2* __restrict__ sref, int nrows, int adjoint, int add) {
  const long long n = (long long)nrows * nk;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(i / nk), ik = (int)(i - (long long)row * nk);
    const int h = nk / 2;
    double kk = 0.0;
    if (ik < h) kk = ik * dk;
    else if (nk - ik >= 1 && nk - ik < h) kk = -(nk - ik) * dk;
    const double ww = w[row];
    const double2 sr = sref[row];
    const double a = ww * ww * sr.x - kk * kk;
    const double b = ww * ww * (sr.y - eps * sr.x);
    const double c = hypot(a, b);
    const double big = sqrt(0.5 * (c + fabs(a)));
    const double small = big > 0.0 ? fabs(b) / (2.0 * big) : 0.0;
    const double re = a >= 0.0 ? big : small;
    const double im = a >= 0.0 ? -small : -big;
    const double amp = exp(im * dz);
    double sn, cs;
    sincos(re * dz, &sn, &cs);
    const float2 p = make_float2((float)(amp * cs), (float)(adjoint ? amp * sn : -amp * sn));
    const float2 v = cmul(in[i], p);
    out[i] = add ? cadd(out[i], v) : v;
  }
}

// out[r][loc[j]] = in[r][loc[j]] * mult[r][j] for nloc positions per row
// (rows of width nx); other samples of out untouched
__global__ void gather_mul_kernel(float2* __restrict__ out, const float2* __restrict__ in,
                                  const float2* __restrict__ mult, const int* __restrict__ loc, int nloc, int nx,
                                  int nrows) {
  const long long n = (long long)nrows * nloc;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / nloc), j = (int)(i - (long long)r * nloc);
    const long long x = (long long)r * nx + loc[j];
    out[x] = cmul(in[x], mult[i]);
  }
}

// rows x width 8-byte words at base + row * pitch set to zero (a kernel,
// not a memset: the clears of an operator call may be captured in a graph)
__global__ void zero2d_kernel(unsigned long long* __restrict__ base, long long pitch, long long width,
                              long long rows) {
  const long long n = width * rows;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / width;
    base[r * pitch + (i - r * width)] = 0ull;
  }
}

// res[i] -= obs[i] and the FP64 partial sums of |res[i]|^2: part[b] = the
// sum over the elements i = b * blockDim + t + k * gridDim * blockDim of
// block b, in a fixed order (a strided loop per thread, then a tree over
// the block)
constexpr int kSumThreads = 256;
__global__ void __launch_bounds__(kSumThreads)
    residual_sumsq_kernel(float2* __restrict__ res, const float2* __restrict__ obs, long long n,
                          double* __restrict__ part) {
  __shared__ double red[kSumThreads];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float2 r = csub(res[i], obs[i]);
    res[i] = r;
    acc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[0] = 0.5 * sum of the nparts partial sums, in order
__global__ void half_sum_kernel(double* __restrict__ out, const double* __restrict__ part, int nparts) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nparts; ++b) s += part[b];
  out[0] = 0.5 * s;
}

// out += in
__global__ void add_kernel(float2* __restrict__ out, const float2* __restrict__ in, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = cadd(out[i], in[i]);
}

// out[j] (+)= sum of in[src[c]] over c in [ptr[j], ptr[j+1]), in list order,
// in double precision; an empty list gives zero (or leaves out[j] with add):
// the gathers that carry a model between grids, and their sums over shots.
__global__ void gather_sum_kernel(float2* __restrict__ out, long long n, const long long* __restrict__ ptr,
                                  const long long* __restrict__ src, const float2* __restrict__ in, int add) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (long long)gridDim.x * blockDim.x) {
    double sx = 0.0, sy = 0.0;
    for (long long c = ptr[j]; c < ptr[j + 1]; ++c) {
      const float2 v = in[src[c]];
      sx += v.x;
      sy += v.y;
    }
    if (add) {
      const float2 o = out[j];
      sx += o.x;
      sy += o.y;
    }
    out[j] = make_float2((float)sx, (float)sy);
  }
}

// out (= or +=) (float2) in
__global__ void d2f_kernel(float2* __restrict__ out, const double2* __restrict__ in, long long n, int acc) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double x = in[i].x, y = in[i].y;
    if (acc) {
      x += out[i].x;
      y += out[i].y;
    }
    out[i] = make_float2((float)x, (float)y);
  }
}

// Aperture window of a perturbation grid: out[z][a] = in[z][x0 + a] for
// 0 <= x0 + a < nxg, 0 outside (the zero padding of born_prop.cpp:52-58).
__global__ void window_kernel(float2* __restrict__ out, const float2* __restrict__ in, int nz, int nxw, int nxg,
                              int x0) {
  const long long n = (long long)nz * nxw;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i / nxw;
    const int x = x0 + (int)(i - z * nxw);
    out[i] = (x >= 0 && x < nxg) ? in[z * nxg + x] : make_float2(0.f, 0.f);
  }
}

// Its adjoint into an FP64 accumulator: acc[z][x0 + a] += in[z][a] for the
// columns inside [0, nxg).
__global__ void window_add_kernel(double2* __restrict__ acc, const float2* __restrict__ in, int nz, int nxw,
                                  int nxg, int x0) {
  const long long n = (long long)nz * nxw;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long z = i / nxw;
    const int x = x0 + (int)(i - z * nxw);
    if (x < 0 || x >= nxg) continue;
    const float2 v = in[i];
    double2& o = acc[z * nxg + x];
    o.x += (double)v.x;
    o.y += (double)v.y;
  }
}

// Point source of each row (Injection.cpp:8-19): W[row][iz][ix+ig+tap] +=
// g[ig] * val[row], ig = -ng..ng, skipping cells outside [0,nx).
__global__ void inject_point_kernel(float2* __restrict__ W, long long wstride, int nx, const int* __restrict__ iz,
                                    const int* __restrict__ ix, const float2* __restrict__ val,
                                    const float* __restrict__ gw, int ng, int tap, int nrows, int zbase, int zlo,
                                    int zhi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int span = 2 * ng + 1;
  if (i >= nrows * span) return;
  const int row = i / span, ig = i - row * span - ng;
  const int off = ix[row] + ig + tap;
  if (off < 0 || off >= nx) return;
  const int z = iz[row];
  if (z < zlo || z > zhi) return;  // outside the rows W holds
  float2* p = W + row * wstride + (long long)(z - zbase) * nx + off;
  *p = cadd(*p, cscale(val[row], gw[ig + ng]));
}

// Receiver injection (Injection.cpp:56-67) as a gather over cells:
// W[row][cell] += sum_c w_c * val[vbase[row] + dev_c * vdev].
__global__ void inject_csr_kernel(float2* __restrict__ W, long long wstride, InjPlan p,
                                  const float2* __restrict__ val, const long long* __restrict__ vbase,
                                  long long vdev, int nrows, long long cbase, long long clo, long long chi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nrows * p.ncell) return;
  const int row = (int)(i / p.ncell), cell = (int)(i - (long long)row * p.ncell);
  if (p.cell[cell] < clo || p.cell[cell] >= chi) return;  // outside the rows W holds
  float sx = 0.f, sy = 0.f;
  for (int c = p.ptr[cell]; c < p.ptr[cell + 1]; ++c) {
    const float2 v = val[vbase[row] + (long long)p.dev[c] * vdev];
    sx += p.wgt[c] * v.x;
    sy += p.wgt[c] * v.y;
  }
  float2* q = W + row * wstride + (p.cell[cell] - cbase);
  *q = make_float2(q->x + sx, q->y + sy);
}

// Receiver recording (Injection.cpp:69-80): val[vbase[row] + i*vdev] +=
// sum_ig g[ig] W[row][iz_i][ix_i+ig+tap].
__global__ void record_kernel(float2* __restrict__ val, const long long* __restrict__ vbase, long long vdev,
                              const float2* __restrict__ W, long long wstride, RecPlan p, int nrows) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nrows * p.ndev) return;
  const int row = (int)(i / p.ndev), dev = (int)(i - (long long)row * p.ndev);
  float sx = 0.f, sy = 0.f;
  for (int c = p.ptr[dev]; c < p.ptr[dev + 1]; ++c) {
    const float2 v = ldg(W, (size_t)(row * wstride + p.cell[c]));
    sx += p.wgt[c] * v.x;
    sy += p.wgt[c] * v.y;
  }
  float2* q = val + vbase[row] + dev * vdev;
  *q = make_float2(q->x + sx, q->y + sy);
}

// dReflect::forward (dReflect.h:29-56) with precomputed coefficients
// a1 = c s1s2sq, b1 = s1s2sq / c, a2 = rho[iz-1] r1r2sq, b2 = rho[iz] r1r2sq:
// dr[iz] = a1 dS[iz-1] - b1 dS[iz] + a2 dR[iz] - b2 dR[iz-1], iz = 1..nz-2.
__global__ void dreflect_fwd_kernel(float2* __restrict__ dr, const float2* __restrict__ coef,
                                    const float2* __restrict__ dS, const float2* __restrict__ dR, int nz, int nx,
                                    int acc) {
  const long long n = (long long)nz * nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(i / nx);
    float2 v = make_float2(0.f, 0.f);
    if (iz >= 1 && iz <= nz - 2) {
      const float2 a1 = coef[i], b1 = coef[n + i], a2 = coef[2 * n + i], b2 = coef[3 * n + i];
      v = csub(cmul(a1, dS[i - nx]), cmul(b1, dS[i]));
      v = cadd(v, csub(cmul(a2, dR[i]), cmul(b2, dR[i - nx])));
    }
    if (acc) v = cadd(dr[i], v);
    dr[i] = v;
  }
}

// dReflect::adjoint (dReflect.h:58-84) as a gather: gS[z] += conj(a1[z+1])
// dr[z+1] - conj(b1[z]) dr[z]; gR[z] += conj(a2[z]) dr[z] - conj(b2[z+1]) dr[z+1]
// (terms only where the row index lies in 1..nz-2).
__global__ void dreflect_adj_kernel(float2* __restrict__ gS, float2* __restrict__ gR,
                                    const float2* __restrict__ coef, const float2* __restrict__ dr, int nz,
                                    int nx) {
  const long long n = (long long)nz * nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(i / nx);
    float2 s = gS[i], r = gR[i];
    if (z + 1 >= 1 && z + 1 <= nz - 2) {
      const long long j = i + nx;
      s = cadd(s, cmulc(dr[j], coef[j]));
      r = csub(r, cmulc(dr[j], coef[3 * n + j]));
    }
    if (z >= 1 && z <= nz - 2) {
      s = csub(s, cmulc(dr[i], coef[n + i]));
      r = cadd(r, cmulc(dr[i], coef[2 * n + i]));
    }
    gS[i] = s;
    gR[i] = r;
  }
}

// Subsurface-offset imaging condition (no reference counterpart; h = 0 is
// IC::adjoint, ic.cpp:17-28):
//   adjoint: I[h][z][x] += sum_rows conj(bg[z][x-h]) sc[z][x+h]
//   forward: src[z][x+h] += bg[z][x-h] I[h][z][x]
// h = -(nh/2)..nh/2; terms with x-h or x+h outside [0,nx) vanish.
// The adjoint sums rows in order per (h, z, x) in double precision.
__global__ void offset_ic_adj_kernel(double2* __restrict__ img, const float2* const* __restrict__ bg,
                                     const float2* const* __restrict__ sc, int nrows, int nh, int nz, int nx) {
  const long long n = (long long)nh * nz * nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i % nx);
    const int z = (int)((i / nx) % nz);
    const int h = (int)(i / ((long long)nz * nx)) - nh / 2;
    if (x - h < 0 || x - h >= nx || x + h < 0 || x + h >= nx) continue;
    double sx = img[i].x, sy = img[i].y;
    for (int r = 0; r < nrows; ++r) {
      const float2 b = bg[r][(long long)z * nx + x - h];
      const float2 s = sc[r][(long long)z * nx + x + h];
      const float2 v = cmulc(s, b);
      sx += v.x;
      sy += v.y;
    }
    img[i] = make_double2(sx, sy);
  }
}

__global__ void offset_ic_fwd_kernel(float2* __restrict__ src, const float2* __restrict__ bg,
                                     const float2* __restrict__ img, int nh, int nz, int nx) {
  const long long n = (long long)nz * nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int xo = (int)(i % nx);   // output position x + h
    const long long zrow = i - xo;
    float2 s = src[i];
    for (int j = 0; j < nh; ++j) {
      const int h = j - nh / 2;
      const int x = xo - h;
      if (x - h < 0 || x - h >= nx || x < 0 || x >= nx) continue;
      s = cadd(s, cmul(bg[zrow + x - h], img[(long long)j * n + zrow + x]));
    }
    src[i] = s;
  }
}

// ------------------------------------------------------------------ launchers

static dim3 grid1d(long long n, int block = 256) {
  long long b = (n + block - 1) / block;
  if (b > 1048576) b = 1048576;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

template <int N, bool GEN, int FX = 0>
static hipError_t launch_lin_v(const LinRow* rows, int nrows, const DepthTables* models, StepGeom g, int dlo,
                               int nd, int dir, bool adj, bool exact, const float2* tw, hipStream_t s) {
  const long long items = (long long)nrows * nd;
  dim3 grid((unsigned)((items + kLinWaves - 1) / kLinWaves)), block(64 * kLinWaves);
  if (!adj)
    hipLaunchKernelGGL((lin_kernel<N, false, false, GEN, FX>), grid, block, 0, s, rows, nrows, models, g, dlo, nd,
                       dir, tw);
  else if (exact)
    hipLaunchKernelGGL((lin_kernel<N, true, true, GEN, FX>), grid, block, 0, s, rows, nrows, models, g, dlo, nd,
                       dir, tw);
  else
    hipLaunchKernelGGL((lin_kernel<N, true, false, GEN, FX>), grid, block, 0, s, rows, nrows, models, g, dlo, nd,
                       dir, tw);
  return hipGetLastError();
}

template <int N>
static hipError_t launch_lin_n(const LinRow* rows, int nrows, const DepthTables* models, StepGeom g, int dlo,
                               int nd, int dir, bool adj, bool exact, const float2* tw, hipStream_t s) {
  if (g.fg.n != N)
    return gen_fixed_dispatch<N>(
        g.fg.n,
        [&](auto id) {
          return launch_lin_v<N, true, decltype(id)::value>(rows, nrows, models, g, dlo, nd, dir, adj, exact, tw, s);
        },
        [&] { return launch_lin_v<N, true>(rows, nrows, models, g, dlo, nd, dir, adj, exact, tw, s); });
  return launch_lin_v<N, false>(rows, nrows, models, g, dlo, nd, dir, adj, exact, tw, s);
}

hipError_t launch_lin(int N, const LinRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz,
                      int dir, bool adj, bool exact, const float2* tw, hipStream_t s, int dlo, int dcount) {
  if (dlo < 0) {  // LinDown: d = 0..nz-2 (neighbour d+1); LinUp: d = 1..nz-1 (neighbour d-1)
    dlo = dir > 0 ? 0 : 1;
    dcount = nz - 1;
  }
  if (dlo < (dir > 0 ? 0 : 1) || dlo + dcount > (dir > 0 ? nz - 1 : nz)) return hipErrorInvalidValue;
  if (nrows <= 0 || dcount <= 0) return hipSuccess;
  switch (N) {
    case 128: return launch_lin_n<128>(rows, nrows, models, g, dlo, dcount, dir, adj, exact, tw, s);
    case 256: return launch_lin_n<256>(rows, nrows, models, g, dlo, dcount, dir, adj, exact, tw, s);
    case 512: return launch_lin_n<512>(rows, nrows, models, g, dlo, dcount, dir, adj, exact, tw, s);
    case 1024: return launch_lin_n<1024>(rows, nrows, models, g, dlo, dcount, dir, adj, exact, tw, s);
    case 2048: return launch_lin_n<2048>(rows, nrows, models, g, dlo, dcount, dir, adj, exact, tw, s);
    default: return hipErrorInvalidValue;
  }
}

template <int N, bool GEN, int FX = 0>
static hipError_t launch_scatter_v(const StepRow* rows, int nrows, const DepthTables* models, StepGeom g,
                                   bool adj, bool exact, const float2* tw, hipStream_t s) {
  dim3 grid((nrows + kLinWaves - 1) / kLinWaves), block(64 * kLinWaves);
  if (!adj)
    hipLaunchKernelGGL((scatter_kernel<N, false, false, GEN, FX>), grid, block, 0, s, rows, nrows, models, g, tw);
  else if (exact)
    hipLaunchKernelGGL((scatter_kernel<N, true, true, GEN, FX>), grid, block, 0, s, rows, nrows, models, g, tw);
  else
    hipLaunchKernelGGL((scatter_kernel<N, true, false, GEN, FX>), grid, block, 0, s, rows, nrows, models, g, tw);
  return hipGetLastError();
}

template <int N>
static hipError_t launch_scatter_n(const StepRow* rows, int nrows, const DepthTables* models, StepGeom g,
                                   bool adj, bool exact, const float2* tw, hipStream_t s) {
  if (g.fg.n != N)
    return gen_fixed_dispatch<N>(
        g.fg.n,
        [&](auto id) { return launch_scatter_v<N, true, decltype(id)::value>(rows, nrows, models, g, adj, exact, tw, s); },
        [&] { return launch_scatter_v<N, true>(rows, nrows, models, g, adj, exact, tw, s); });
  return launch_scatter_v<N, false>(rows, nrows, models, g, adj, exact, tw, s);
}

hipError_t launch_scatter(int N, const StepRow* rows, int nrows, const DepthTables* models, StepGeom g, bool adj,
                          bool exact, const float2* tw, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  switch (N) {
    case 128: return launch_scatter_n<128>(rows, nrows, models, g, adj, exact, tw, s);
    case 256: return launch_scatter_n<256>(rows, nrows, models, g, adj, exact, tw, s);
    case 512: return launch_scatter_n<512>(rows, nrows, models, g, adj, exact, tw, s);
    case 1024: return launch_scatter_n<1024>(rows, nrows, models, g, adj, exact, tw, s);
    case 2048: return launch_scatter_n<2048>(rows, nrows, models, g, adj, exact, tw, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mul_rows(const MulRow* rows, int nrows, long long n, bool conjA, bool acc, hipStream_t s) {
  if (n <= 0 || nrows <= 0) return hipSuccess;
  long long bx = (n + 255) / 256;
  if (bx > 64) bx = 64;
  const int by = nrows < 65535 ? nrows : 65535;
  hipLaunchKernelGGL(mul_rows_kernel, dim3((unsigned)bx, (unsigned)by), dim3(256), 0, s, rows, nrows, n,
                     conjA ? 1 : 0, acc ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_mul(float2* out, long long so, const float2* A, long long sa, const float2* B, long long sb,
                      long long n, int nrows, bool conjA, bool acc, hipStream_t s) {
  if (n <= 0 || nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(mul_kernel, grid1d(n * nrows), dim3(256), 0, s, out, so, A, sa, B, sb, n, nrows,
                     conjA ? 1 : 0, acc ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_reduce_rows(double2* out, const float2* in, long long stride, long long n, int nrows, int phase,
                              bool acc, hipStream_t s, const float2* in2) {
  if (n <= 0 || nrows <= 0) return hipSuccess;
  if (phase < 0 || phase >= kRunRows) return hipErrorInvalidValue;
  const long long tiles = (n + 63) / 64;
  if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
  if (in2)
    hipLaunchKernelGGL(reduce_runs_kernel<true>, dim3((unsigned)tiles), dim3(64 * kRedWaves), 0, s, out, in, in2,
                       stride, n, nrows, phase, acc ? 1 : 0);
  else
    hipLaunchKernelGGL(reduce_runs_kernel<false>, dim3((unsigned)tiles), dim3(64 * kRedWaves), 0, s, out, in,
                       in, stride, n, nrows, phase, acc ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_gather_sum(float2* out, long long n, const long long* ptr, const long long* src, const float2* in,
                             bool add, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_sum_kernel, grid1d(n), dim3(256), 0, s, out, n, ptr, src, in, add ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_phshift(float2* out, const float2* in, int nk, double dk, double dz, double eps, const double* w,
                          const double2* sref, int nrows, bool adjoint, bool add, hipStream_t s) {
  const long long n = (long long)nrows * nk;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(phshift_rows_kernel, grid1d(n), dim3(256), 0, s, out, in, nk, dk, dz, eps, w, sref, nrows,
                     adjoint ? 1 : 0, add ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_gather_mul(float2* out, const float2* in, const float2* mult, const int* loc, int nloc, int nx,
                             int nrows, hipStream_t s) {
  const long long n = (long long)nrows * nloc;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_mul_kernel, grid1d(n), dim3(256), 0, s, out, in, mult, loc, nloc, nx, nrows);
  return hipGetLastError();
}

hipError_t launch_zero2d(void* base, long long pitch_words, long long width_words, long long rows, hipStream_t s) {
  const long long n = width_words * rows;
  if (n <= 0) return hipSuccess;
  const int grid = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero2d_kernel, dim3(grid), dim3(256), 0, s, (unsigned long long*)base, pitch_words, width_words,
                     rows);
  return hipGetLastError();
}

hipError_t launch_residual_sumsq(float2* res, const float2* obs, long long n, double* part, double* half_sum,
                                 hipStream_t s) {
  hipLaunchKernelGGL(residual_sumsq_kernel, dim3(kMisfitParts), dim3(kSumThreads), 0, s, res, obs, n, part);
  hipLaunchKernelGGL(half_sum_kernel, dim3(1), dim3(64), 0, s, half_sum, part, kMisfitParts);
  return hipGetLastError();
}

hipError_t launch_add(float2* out, const float2* in, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_kernel, grid1d(n), dim3(256), 0, s, out, in, n);
  return hipGetLastError();
}

hipError_t launch_d2f(float2* out, const double2* in, long long n, bool acc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(d2f_kernel, grid1d(n), dim3(256), 0, s, out, in, n, acc ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_window(float2* out, const float2* in, int nz, int nxw, int nxg, int x0, hipStream_t s) {
  const long long n = (long long)nz * nxw;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(window_kernel, grid1d(n), dim3(256), 0, s, out, in, nz, nxw, nxg, x0);
  return hipGetLastError();
}

hipError_t launch_window_add(double2* acc, const float2* in, int nz, int nxw, int nxg, int x0, hipStream_t s) {
  const long long n = (long long)nz * nxw;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(window_add_kernel, grid1d(n), dim3(256), 0, s, acc, in, nz, nxw, nxg, x0);
  return hipGetLastError();
}

hipError_t launch_inject_point(float2* W, long long wstride, int nx, const int* iz, const int* ix,
                               const float2* val, const float* gw, int ng, int tap, int nrows, hipStream_t s,
                               int zbase, int zlo, int zhi) {
  const long long n = (long long)nrows * (2 * ng + 1);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(inject_point_kernel, grid1d(n), dim3(256), 0, s, W, wstride, nx, iz, ix, val, gw, ng, tap,
                     nrows, zbase, zlo, zhi);
  return hipGetLastError();
}

hipError_t launch_inject_csr(float2* W, long long wstride, InjPlan p, const float2* val, const long long* vbase,
                             long long vdev, int nrows, hipStream_t s, long long cbase, long long clo,
                             long long chi) {
  const long long n = (long long)nrows * p.ncell;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(inject_csr_kernel, grid1d(n), dim3(256), 0, s, W, wstride, p, val, vbase, vdev, nrows, cbase,
                     clo, chi);
  return hipGetLastError();
}

hipError_t launch_record(float2* val, const long long* vbase, long long vdev, const float2* W, long long wstride,
                         RecPlan p, int nrows, hipStream_t s) {
  const long long n = (long long)nrows * p.ndev;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(record_kernel, grid1d(n), dim3(256), 0, s, val, vbase, vdev, W, wstride, p, nrows);
  return hipGetLastError();
}

hipError_t launch_dreflect_fwd(float2* dr, const float2* coef, const float2* dS, const float2* dR, int nz, int nx,
                               bool acc, hipStream_t s) {
  hipLaunchKernelGGL(dreflect_fwd_kernel, grid1d((long long)nz * nx), dim3(256), 0, s, dr, coef, dS, dR, nz, nx,
                     acc ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_dreflect_adj(float2* gS, float2* gR, const float2* coef, const float2* dr, int nz, int nx,
                               hipStream_t s) {
  hipLaunchKernelGGL(dreflect_adj_kernel, grid1d((long long)nz * nx), dim3(256), 0, s, gS, gR, coef, dr, nz, nx);
  return hipGetLastError();
}

hipError_t launch_offset_ic_adj(double2* img, const float2* const* bg, const float2* const* sc, int nrows, int nh,
                                int nz, int nx, hipStream_t s) {
  hipLaunchKernelGGL(offset_ic_adj_kernel, grid1d((long long)nh * nz * nx), dim3(256), 0, s, img, bg, sc, nrows,
                     nh, nz, nx);
  return hipGetLastError();
}

hipError_t launch_offset_ic_fwd(float2* src, const float2* bg, const float2* img, int nh, int nz, int nx,
                                hipStream_t s) {
  hipLaunchKernelGGL(offset_ic_fwd_kernel, grid1d((long long)nz * nx), dim3(256), 0, s, src, bg, img, nh, nz, nx);
  return hipGetLastError();
}

}  // namespace wemfwi
