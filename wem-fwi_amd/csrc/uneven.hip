// Subsurface-offset migration of shots with uneven apertures and bands: the
// adjoint of BornRefl's modelling (ext_born_refl.cpp:60-110: D = Down(inj_src),
// R = Up^H(inj_rec^H d), image += IC^H) with the subsurface-offset imaging
// condition I(h, z, x) = sum conj(D(z, x-h)) R(z, x+h) in place of the zero-lag
// one, for shots that each carry their own background on their own aperture
// (a model nx_s wide, placed at column x0_s of the image) and their own list
// of frequencies.  All (shot, frequency) rows of a group of shots are swept by
// two batched launches (rows of different widths share a launch per FFT
// length); the image is reduced in FP64, shot after shot and row after row.
//
// The same driver gives the extended least-squares objective of such shots,
// J(I) = 1/2 sum_s ||L_s I - d_s||^2, and its gradient sum_s L_s^H (L_s I - d_s):
// L_s models each row as BornRefl's forward (ext_born_refl.cpp:8-56) with the
// offset scatter U(z, y) = sum_h I(h, z, y - h) D(z, y - 2h) in place of IC, so
// a group runs D = Down(sources) once and shares it between L and L^H.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "born.h"
#include "host.h"
#include "sweep.hpp"

namespace wemfwi {
namespace {

using cf = std::complex<float>;

// device buffer released at scope exit
struct ScopedWs : Workspace {
  ~ScopedWs() { release(); }
};

// Source row of the batch: W[iz][ix + ig] += g[ig] * val for ix + ig in
// [0, nx) (Injection.cpp:8-32 with offset 0, as BornRefl injects).
struct SrcRow {
  float2* W;
  int nx, iz, ix;
  float2 val;
};

// Receiver row: W[iz_r][ix_r] += data[r * stride + col] for every receiver
// r in order (Injection.cpp:56-67 without spread); one thread per row keeps
// the sums of receivers sharing a cell in receiver order.
struct RecRow {
  float2* W;
  int nx, nr;
  const int* iz;
  const int* ix;
  const float2* data;
  int stride, col;
};

__global__ void src_rows_kernel(const SrcRow* __restrict__ rows, int nrows, const float* __restrict__ gw,
                                int ng) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int span = 2 * ng + 1;
  if (i >= nrows * span) return;
  const int row = i / span, ig = i - row * span - ng;
  const SrcRow R = rows[row];
  const int x = R.ix + ig;
  if (x < 0 || x >= R.nx) return;
  float2* p = R.W + (size_t)R.iz * R.nx + x;
  *p = cadd(*p, cscale(R.val, gw[ig + ng]));
}

__global__ void rec_rows_kernel(const RecRow* __restrict__ rows, int nrows) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const RecRow R = rows[row];
  for (int r = 0; r < R.nr; ++r) {
    float2* p = R.W + (size_t)R.iz[r] * R.nx + R.ix[r];
    *p = cadd(*p, R.data[(size_t)r * R.stride + R.col]);
  }
}

// One shot of a group on the device: aperture columns [x0, x0 + nx) of the
// image, nw rows of nz x nx samples from row offset roff of the group's D
// and R buffers.
struct UShot {
  int x0, nx, nw;
  long long roff;
};

// img[h][z][x] += for each shot covering image column x (in shot order), the
// sum over its rows (in row order) of conj(D[z][a-h]) R[z][a+h], a = x - x0
// the aperture column, h = j - nh/2, terms with a-h or a+h outside [0, nx)
// absent.  A workgroup owns one depth z and kTX consecutive image columns
// (one per thread, the shot's nh lags in FP64 registers, added to the FP64
// image after each shot, so the sums run in the order of one launch per
// shot).  kRows rows at a time have their D and R segments (nh/2 halos)
// loaded together and staged in LDS, so every sample is read from HBM once
// per row with kRows rows of loads in flight.  One launch covers every shot
// of the group; block b takes tile order[b / nz] at depth b % nz, the tiles
// ordered by decreasing work (rows covering them) so the longest ones start
// first.
constexpr int kTX = 256;
constexpr int kHalo = 20;  // nh <= 41
constexpr int kRows = 4;
constexpr int kSeg = kTX + 2 * kHalo;
template <int NHM>
__global__ void __launch_bounds__(kTX)
    offset_image_shots_kernel(double2* __restrict__ img, long long hstride, int nxi, const UShot* __restrict__ shots,
                              int nshots, const int* __restrict__ order, const float2* __restrict__ D,
                              const float2* __restrict__ R, int nz, int nh) {
  __shared__ float2 sd[kRows * kSeg], sr[kRows * kSeg];
  const int z = blockIdx.x % nz;
  const int X = order[blockIdx.x / nz] * kTX;
  const int t = threadIdx.x;
  const int hh = nh / 2;
  const int seg = kTX + 2 * hh;
  for (int sh = 0; sh < nshots; ++sh) {
    const UShot S = shots[sh];
    if (S.nw == 0 || S.x0 >= X + kTX || S.x0 + S.nx <= X) continue;  // the same for the whole workgroup
    const int a0 = X - S.x0;  // aperture column of thread 0
    const int x = a0 + t;
    const int nx = S.nx;
    const long long rstride = (long long)nz * nx;
    double ax[NHM], ay[NHM];
#pragma unroll
    for (int j = 0; j < NHM; ++j) ax[j] = ay[j] = 0.0;
    for (int r0 = 0; r0 < S.nw; r0 += kRows) {
      const int kr = min(kRows, S.nw - r0);
      const float2* d = D + S.roff + r0 * rstride + (long long)z * nx;
      const float2* q = R + S.roff + r0 * rstride + (long long)z * nx;
      for (int i = t; i < seg; i += kTX) {
        // column xx of depth z of row r0 + k, through that row's buffer
        // descriptor (nx samples): columns outside [0, nx) read 0
        const unsigned off = (unsigned)(a0 - hh + i) * 8u;
        float2 vd[kRows], vr[kRows];
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
          // rows past the shot's last (k >= kr) are never addressed: their
          // loads read row kr - 1 and are replaced by 0
          const int kk = k < kr ? k : kr - 1;
          vd[k] = bld2(make_rsrc(d + kk * rstride, (unsigned)nx * 8u), off, 0u);
          vr[k] = bld2(make_rsrc(q + kk * rstride, (unsigned)nx * 8u), off, 0u);
          if (k >= kr) vd[k] = vr[k] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
          sd[k * kSeg + i] = vd[k];
          sr[k * kSeg + i] = vr[k];
        }
      }
      __syncthreads();
      if (NHM > 21 && x >= 0 && x < nx) {
        // 41 lags: one at a time (their FP64 sums already fill the registers)
        for (int k = 0; k < kr; ++k) {
#pragma unroll
          for (int j = 0; j < NHM; ++j) {
            const int h = j - hh;
            if (j < nh && x - h >= 0 && x - h < nx && x + h >= 0 && x + h < nx) {
              const float2 v = cmulc(sr[k * kSeg + t + hh + h], sd[k * kSeg + t + hh - h]);
              ax[j] += v.x;
              ay[j] += v.y;
            }
          }
        }
      } else if (x >= 0 && x < nx) {
        for (int k = 0; k < kr; ++k) {
          // every lag without a branch: the staged segments are 0 outside
          // [0, nx), so an absent term is a zero product, which leaves the
          // (never negative-zero) accumulator as it is; lags past nh read a
          // valid pair and are never written out
          // (chunks of up to 11 lags in flight at a time)
          constexpr int CH = NHM < 11 ? NHM : 11;
#pragma unroll
          for (int j0 = 0; j0 < NHM; j0 += CH) {
            float2 pr[CH], pd[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              const int h = min(j0 + c - hh, hh);
              pr[c] = sr[k * kSeg + t + hh + h];
              pd[c] = sd[k * kSeg + t + hh - h];
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              if (j0 + c < NHM) {
                const float2 v = cmulc(pr[c], pd[c]);
                ax[j0 + c] += v.x;
                ay[j0 + c] += v.y;
              }
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      __syncthreads();
    }
    if (x >= 0 && x < nx) {
#pragma unroll
      for (int j = 0; j < NHM; ++j) {
        if (j >= nh) break;
        double2* o = img + j * hstride + (long long)z * nxi + X + t;
        const double2 v = *o;
        *o = make_double2(v.x + ax[j], v.y + ay[j]);
      }
    }
  }
}

template <int NHM>
hipError_t launch_ois(double2* img, long long hstride, int nxi, const UShot* shots, int nshots, const int* order,
                      int ntiles, const float2* D, const float2* R, int nz, int nh, hipStream_t s) {
  hipLaunchKernelGGL((offset_image_shots_kernel<NHM>), dim3((unsigned)(ntiles * nz)), dim3(kTX), 0, s, img, hstride,
                     nxi, shots, nshots, order, D, R, nz, nh);
  return hipGetLastError();
}

hipError_t launch_offset_image_shots(double2* img, long long hstride, int nxi, const UShot* shots, int nshots,
                                     const int* order, int ntiles, const float2* D, const float2* R, int nz, int nh,
                                     hipStream_t s) {
  if (nshots <= 0 || nz <= 0 || ntiles <= 0) return hipSuccess;
  if (nh > 2 * kHalo + 1 || (long long)ntiles * nz > 0x7fffffffLL) return hipErrorInvalidValue;
  if (nh <= 1) return launch_ois<1>(img, hstride, nxi, shots, nshots, order, ntiles, D, R, nz, nh, s);
  if (nh <= 3) return launch_ois<3>(img, hstride, nxi, shots, nshots, order, ntiles, D, R, nz, nh, s);
  if (nh <= 11) return launch_ois<11>(img, hstride, nxi, shots, nshots, order, ntiles, D, R, nz, nh, s);
  if (nh <= 21) return launch_ois<21>(img, hstride, nxi, shots, nshots, order, ntiles, D, R, nz, nh, s);
  return launch_ois<41>(img, hstride, nxi, shots, nshots, order, ntiles, D, R, nz, nh, s);
}

// U_r[z][y] = sum_h I[h][z][x0 + y - h] D_r[z][y - 2h] for every row r of
// every shot, y an aperture column, terms with y - h or y - 2h outside
// [0, nx) absent: the adjoint of offset_image_shots_kernel with respect to
// R.  A workgroup owns one depth and kTX consecutive image columns; the
// image tile of all nh lags (halo nh/2) is staged in LDS once for every shot
// that covers the tile, kRows rows' D segments (halo 2 (nh/2)) at a time.
// Every (z, y < nx) of every row is written (each aperture column lies in
// exactly one tile).  One launch covers every shot of the group, tiles in
// the order of offset_image_shots_kernel.
template <int NHM>
__global__ void __launch_bounds__(kTX)
    offset_scatter_shots_kernel(float2* __restrict__ U, const float2* __restrict__ img, long long hstride, int nxi,
                                const UShot* __restrict__ shots, int nshots, const int* __restrict__ order,
                                const float2* __restrict__ D, int nz, int nh) {
  constexpr int HM = NHM / 2;
  constexpr int WD = kTX + 4 * HM;
  __shared__ float2 si[NHM * (kTX + 2 * HM)];
  __shared__ float2 sd[kRows * WD];
  const int hh = nh / 2;
  const int wi = kTX + 2 * hh;
  const int wd = kTX + 4 * hh;
  const int z = blockIdx.x % nz;
  const int X = order[blockIdx.x / nz] * kTX;
  const int t = threadIdx.x;
  for (int i = t; i < nh * wi; i += kTX) {
    // image column X - hh + c of lag j through the image row's buffer
    // descriptor: columns outside [0, nxi) read 0
    const int j = i / wi, c = i - j * wi;
    si[i] = bld2(make_rsrc(img + j * hstride + (long long)z * nxi, (unsigned)nxi * 8u), (unsigned)(X - hh + c) * 8u,
                 0u);
  }
  for (int sh = 0; sh < nshots; ++sh) {
    const UShot S = shots[sh];
    if (S.nw == 0 || S.x0 >= X + kTX || S.x0 + S.nx <= X) continue;  // the same for the whole workgroup
    const int a0 = X - S.x0;
    const int y = a0 + t;
    const int nx = S.nx;
    const long long rstride = (long long)nz * nx;
    for (int r0 = 0; r0 < S.nw; r0 += kRows) {
      const int kr = min(kRows, S.nw - r0);
      const float2* d = D + S.roff + r0 * rstride + (long long)z * nx;
      __syncthreads();
      for (int i = t; i < wd; i += kTX) {
        // column xx of depth z of row r0 + k through the row's buffer
        // descriptor: columns outside [0, nx) read 0
        const unsigned off = (unsigned)(a0 - 2 * hh + i) * 8u;
        float2 v[kRows];
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
          const int kk = k < kr ? k : kr - 1;  // rows past the shot's last: row kr - 1, value 0
          v[k] = bld2(make_rsrc(d + kk * rstride, (unsigned)nx * 8u), off, 0u);
          if (k >= kr) v[k] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < kRows; ++k) sd[k * WD + i] = v[k];
      }
      __syncthreads();
      if (y >= 0 && y < nx) {
        for (int k = 0; k < kr; ++k) {
          // every lag without a branch: y and y - 2h inside [0, nx) put y - h
          // inside too, and the staged D segment is 0 outside [0, nx), so an
          // absent term is a zero product (the accumulator, never negative
          // zero, keeps its value); lags past nh add exact zeros
          // (chunks of up to 11 lags in flight at a time)
          constexpr int CH = NHM < 11 ? NHM : 11;
          float ax = 0.f, ay = 0.f;
#pragma unroll
          for (int j0 = 0; j0 < NHM; j0 += CH) {
            float2 pi[CH], pb[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              const int j = j0 + c;
              const int h = min(j - hh, hh);
              pi[c] = (j < nh && j < NHM) ? si[j * wi + t - h + hh] : make_float2(0.f, 0.f);
              pb[c] = sd[k * WD + t - 2 * h + 2 * hh];
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              if (j0 + c < NHM) {
                ax = fmaf(pi[c].x, pb[c].x, fmaf(-pi[c].y, pb[c].y, ax));
                ay = fmaf(pi[c].x, pb[c].y, fmaf(pi[c].y, pb[c].x, ay));
              }
            }
            __builtin_amdgcn_sched_barrier(0);
          }
          U[S.roff + (r0 + k) * rstride + (long long)z * nx + y] = make_float2(ax, ay);
        }
      }
    }
  }
}

template <int NHM>
hipError_t launch_oss(float2* U, const float2* img, long long hstride, int nxi, const UShot* shots, int nshots,
                      const int* order, int ntiles, const float2* D, int nz, int nh, hipStream_t s) {
  hipLaunchKernelGGL((offset_scatter_shots_kernel<NHM>), dim3((unsigned)(ntiles * nz)), dim3(kTX), 0, s, U, img,
                     hstride, nxi, shots, nshots, order, D, nz, nh);
  return hipGetLastError();
}

hipError_t launch_offset_scatter_shots(float2* U, const float2* img, long long hstride, int nxi, const UShot* shots,
                                       int nshots, const int* order, int ntiles, const float2* D, int nz, int nh,
                                       hipStream_t s) {
  if (nshots <= 0 || nz <= 0 || ntiles <= 0) return hipSuccess;
  if (nh > 2 * kHalo + 1 || (long long)ntiles * nz > 0x7fffffffLL) return hipErrorInvalidValue;
  if (nh <= 1) return launch_oss<1>(U, img, hstride, nxi, shots, nshots, order, ntiles, D, nz, nh, s);
  if (nh <= 3) return launch_oss<3>(U, img, hstride, nxi, shots, nshots, order, ntiles, D, nz, nh, s);
  if (nh <= 11) return launch_oss<11>(U, img, hstride, nxi, shots, nshots, order, ntiles, D, nz, nh, s);
  if (nh <= 21) return launch_oss<21>(U, img, hstride, nxi, shots, nshots, order, ntiles, D, nz, nh, s);
  return launch_oss<41>(U, img, hstride, nxi, shots, nshots, order, ntiles, D, nz, nh, s);
}

// Residual of one row at its receivers: res[r * stride + col] =
// W[iz_r][ix_r] - dobs[r * stride + col] (Injection::adjoint without spread,
// Injection.cpp:56-67, minus the observed sample).  Grid: (receiver blocks,
// rows, strided past 65535).
struct ResRow {
  const float2* W;
  int nx, nr;
  const int* iz;
  const int* ix;
  const float2* dobs;
  float2* res;
  int stride, col;
};

__global__ void residual_rows_kernel(const ResRow* __restrict__ rows, int nrows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const ResRow R = rows[row];
    if (r >= R.nr) continue;
    const size_t k = (size_t)r * R.stride + R.col;
    R.res[k] = csub(R.W[(size_t)R.iz[r] * R.nx + R.ix[r]], R.dobs[k]);
  }
}

// Sum of |a|^2 in FP64 in a fixed order: kSumBlocks blocks reduce fixed
// strided slices into part[], one block adds part[] in order to *out.
constexpr int kSumBlocks = 256;
constexpr int kSumThreads = 256;

__global__ void __launch_bounds__(kSumThreads)
    sumsq_partial_kernel(const float2* __restrict__ a, long long n, double* __restrict__ part) {
  __shared__ double red[kSumThreads];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * kSumThreads + threadIdx.x; i < n;
       i += (long long)kSumBlocks * kSumThreads) {
    const float2 v = a[i];
    acc += (double)v.x * v.x + (double)v.y * v.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kSumThreads) sumsq_final_kernel(const double* __restrict__ part, double* out) {
  __shared__ double red[kSumThreads];
  red[threadIdx.x] = (int)threadIdx.x < kSumBlocks ? part[threadIdx.x] : 0.0;
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out += red[0];
}

template <typename T>
int upload_vec(Workspace& w, const std::vector<T>& h, hipStream_t s, const char* what) {
  if (h.empty()) return WEMFWI_OK;
  if (int rc = w.reserve(h.size() * sizeof(T), what)) return rc;
  hipError_t e = hipMemcpyAsync(w.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, what);
}

#define UCHK(expr)                     \
  do {                                 \
    int _rc = (expr);                  \
    if (_rc) return _rc;               \
  } while (0)
#define UHIP(expr, what)                               \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, (what)); \
  } while (0)

// One group of shots [s0, s1): inject, sweep both wavefields of all rows,
// reduce the image shot by shot.  With a model image (fwi), the receiver
// wavefield is the residual's: the shots' data are the observed data, the
// rows are first modelled from the image (scatter, Up, record), and
// 1/2 sum |residual|^2 is added to *jacc (device).
int image_group(const wemfwi_image_shot* shots, int s0, int s1, const std::vector<float>& gw, int ngs, int nh,
                int nxi, double2* acc, hipStream_t s, const float2* fwi = nullptr, double* jacc = nullptr,
                double* jpart = nullptr) {
  std::vector<Model*> rm;
  std::vector<float> rf;
  std::vector<size_t> roff;
  size_t total = 0;
  for (int sh = s0; sh < s1; ++sh) {
    Model* m = &shots[sh].model->m;
    for (int i = 0; i < shots[sh].nw; ++i) {
      rm.push_back(m);
      rf.push_back(shots[sh].freq_hz[i]);
      roff.push_back(total);
      total += (size_t)m->d.nz * m->d.nx;
    }
  }
  const int nrows = (int)rm.size();
  if (nrows == 0) return WEMFWI_OK;
  const int nz = shots[s0].model->m.d.nz;
  ScopedWs Dw, Rw, srcw, recw, gww, devw;
  UCHK(Dw.reserve(total * sizeof(float2), "source wavefields"));
  UCHK(Rw.reserve(total * sizeof(float2), "receiver wavefields"));
  float2* D = (float2*)Dw.ptr;
  float2* R = (float2*)Rw.ptr;
  // The sweeps read only the source depth (D) and the receiver depths (R)
  // as sources, so only those rows of each wavefield are cleared.
  std::vector<int> dlo(nrows), dhi(nrows), rlo(nrows), rhi(nrows);
  auto clear_rows = [&](float2* W, const std::vector<int>& lo, const std::vector<int>& hi) -> int {
    size_t k0 = 0;
    for (int sh = s0; sh < s1; ++sh) {
      const wemfwi_image_shot& S = shots[sh];
      const int nx = S.model->m.d.nx;
      if (S.nw > 0 && lo[k0] <= hi[k0])
        UHIP(hipMemset2DAsync(W + roff[k0] + (size_t)lo[k0] * nx, (size_t)nz * nx * sizeof(float2), 0,
                              (size_t)(hi[k0] - lo[k0] + 1) * nx * sizeof(float2), (size_t)S.nw, s),
             "memset");
      k0 += S.nw;
    }
    return WEMFWI_OK;
  };
  {
    size_t k0 = 0;
    for (int sh = s0; sh < s1; ++sh) {
      const wemfwi_image_shot& S = shots[sh];
      int lo = 1 << 30, hi = -1;
      for (int r = 0; r < S.nr; ++r) {
        lo = std::min(lo, S.rec_iz[r]);
        hi = std::max(hi, S.rec_iz[r]);
      }
      for (int i = 0; i < S.nw; ++i, ++k0) {
        dlo[k0] = dhi[k0] = S.src_iz;
        rlo[k0] = lo;
        rhi[k0] = hi;
      }
    }
  }
  UCHK(clear_rows(D, dlo, dhi));
  if (!fwi) UCHK(clear_rows(R, rlo, rhi));   // with a model image, the scatter writes every row of R
  // receiver coordinates of the group, one block per shot
  std::vector<int> hdev;
  std::vector<size_t> dev_off(s1 - s0);
  for (int sh = s0; sh < s1; ++sh) {
    dev_off[sh - s0] = hdev.size();
    hdev.insert(hdev.end(), shots[sh].rec_iz, shots[sh].rec_iz + shots[sh].nr);
    hdev.insert(hdev.end(), shots[sh].rec_ix, shots[sh].rec_ix + shots[sh].nr);
  }
  UCHK(upload_vec(devw, hdev, s, "receiver positions"));
  const int* ddev = (const int*)devw.ptr;
  // the group's shots for the scatter and imaging launches
  ScopedWs ushw;
  {
    std::vector<UShot> hu;
    size_t k0 = 0;
    for (int sh = s0; sh < s1; ++sh) {
      const Model& m = shots[sh].model->m;
      hu.push_back(UShot{shots[sh].x0, m.d.nx, shots[sh].nw, (long long)(shots[sh].nw > 0 ? roff[k0] : 0)});
      k0 += shots[sh].nw;
    }
    UCHK(upload_vec(ushw, hu, s, "shot table"));
  }
  // image tiles by decreasing work (rows of the shots covering them), ties
  // by position
  const int ntiles = (nxi + kTX - 1) / kTX;
  ScopedWs ordw;
  {
    std::vector<long long> work(ntiles, 0);
    for (int sh = s0; sh < s1; ++sh) {
      if (shots[sh].nw == 0) continue;
      const int a = shots[sh].x0 / kTX;
      const int b = (shots[sh].x0 + shots[sh].model->m.d.nx - 1) / kTX;
      for (int tl = a; tl <= b; ++tl) work[tl] += shots[sh].nw;
    }
    std::vector<int> ord(ntiles);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) { return work[p] > work[q]; });
    UCHK(upload_vec(ordw, ord, s, "tile order"));
  }
  std::vector<SrcRow> hs(nrows);
  std::vector<RecRow> hr(nrows);
  std::vector<const cf*> pin(nrows);
  std::vector<cf*> pout(nrows);
  int k = 0;
  for (int sh = s0; sh < s1; ++sh) {
    const wemfwi_image_shot& S = shots[sh];
    const int nx = S.model->m.d.nx;
    const int* riz = ddev + dev_off[sh - s0];
    for (int i = 0; i < S.nw; ++i, ++k) {
      hs[k] = SrcRow{D + roff[k], nx, S.src_iz, S.src_ix, make_float2(S.wavelet[i].re, S.wavelet[i].im)};
      hr[k] = RecRow{R + roff[k], nx, S.nr, riz, riz + S.nr, (const float2*)S.data, S.nw, i};
    }
  }
  // fwi: the residual of every row, shot by shot as the data ([nr][nw])
  ScopedWs resw, resrw;
  float2* res = nullptr;
  int nr_max = 0;
  size_t nres = 0;
  if (fwi) {
    std::vector<size_t> res_off(s1 - s0);
    for (int sh = s0; sh < s1; ++sh) {
      res_off[sh - s0] = nres;
      nres += (size_t)shots[sh].nr * shots[sh].nw;
      nr_max = std::max(nr_max, shots[sh].nr);
    }
    if (nres > 0) {
      UCHK(resw.reserve(nres * sizeof(float2), "residuals"));
      res = (float2*)resw.ptr;
    }
    std::vector<ResRow> hq(nrows);
    k = 0;
    for (int sh = s0; sh < s1; ++sh) {
      const wemfwi_image_shot& S = shots[sh];
      const int* riz = ddev + dev_off[sh - s0];
      for (int i = 0; i < S.nw; ++i, ++k) {
        float2* rs = res + res_off[sh - s0];
        hq[k] = ResRow{R + roff[k], S.model->m.d.nx, S.nr, riz, riz + S.nr, (const float2*)S.data, rs, S.nw, i};
        hr[k].data = rs;   // the receiver wavefield injects the residual
      }
    }
    UCHK(upload_vec(resrw, hq, s, "residual rows"));
  }
  UCHK(upload_vec(srcw, hs, s, "source rows"));
  UCHK(upload_vec(recw, hr, s, "receiver rows"));
  UCHK(upload_vec(gww, gw, s, "spread weights"));
  const int span = 2 * ngs + 1;
  hipLaunchKernelGGL(src_rows_kernel, dim3((nrows * span + 255) / 256), dim3(256), 0, s, (const SrcRow*)srcw.ptr,
                     nrows, (const float*)gww.ptr, ngs);
  UHIP(hipGetLastError(), "inject sources");
  if (!fwi) {
    hipLaunchKernelGGL(rec_rows_kernel, dim3((nrows + 63) / 64), dim3(64), 0, s, (const RecRow*)recw.ptr, nrows);
    UHIP(hipGetLastError(), "inject receivers");
  }
  UHIP(hipStreamSynchronize(s), "inject");
  // D = Down(sources), R = Up^H(receivers), in place, every row in one plan
  for (int i = 0; i < nrows; ++i) pin[i] = pout[i] = (cf*)(D + roff[i]);
  {
    BatchPlan p;
    UCHK(p.build(nrows, rm.data(), rf.data(), pin.data(), pout.data(), s, dlo.data(), dhi.data()));
    UCHK(p.run(0, 0, s));
  }
  for (int i = 0; i < nrows; ++i) pin[i] = pout[i] = (cf*)(R + roff[i]);
  {
    BatchPlan p;   // Up^H of the receiver (or residual) rows
    UCHK(p.build(nrows, rm.data(), rf.data(), pin.data(), pout.data(), s, rlo.data(), rhi.data()));
    if (fwi) {
      BatchPlan pu;   // Up of the scattered field: every row is a source
      UCHK(pu.build(nrows, rm.data(), rf.data(), pin.data(), pout.data(), s));
      // L I: offset scatter of the image on D into every row, Up, record
      UHIP(launch_offset_scatter_shots(R, fwi, (long long)nz * nxi, nxi, (const UShot*)ushw.ptr, s1 - s0,
                                       (const int*)ordw.ptr, ntiles, D, nz, nh, s),
           "offset scatter");
      UCHK(pu.run(1, 0, s));
      if (nr_max > 0) {
        hipLaunchKernelGGL(residual_rows_kernel, dim3((unsigned)((nr_max + 255) / 256), (unsigned)std::min(nrows, 65535)),
                           dim3(256), 0, s, (const ResRow*)resrw.ptr, nrows);
        UHIP(hipGetLastError(), "residual");
        hipLaunchKernelGGL(sumsq_partial_kernel, dim3(kSumBlocks), dim3(kSumThreads), 0, s, (const float2*)res,
                           (long long)nres, jpart);
        UHIP(hipGetLastError(), "misfit");
        hipLaunchKernelGGL(sumsq_final_kernel, dim3(1), dim3(kSumThreads), 0, s, (const double*)jpart, jacc);
        UHIP(hipGetLastError(), "misfit");
      }
      // R^H residual, then Up^H as in the migration
      UCHK(clear_rows(R, rlo, rhi));
      hipLaunchKernelGGL(rec_rows_kernel, dim3((nrows + 63) / 64), dim3(64), 0, s, (const RecRow*)recw.ptr, nrows);
      UHIP(hipGetLastError(), "inject residuals");
    }
    UCHK(p.run(1, 1, s));
  }
  // image, every shot in one launch, sums in shot order
  UHIP(launch_offset_image_shots(acc, (long long)nz * nxi, nxi, (const UShot*)ushw.ptr, s1 - s0,
                                 (const int*)ordw.ptr, ntiles, D, R, nz, nh, s),
       "offset image");
  UHIP(hipStreamSynchronize(s), "offset image");
  return WEMFWI_OK;
}

}  // namespace
}  // namespace wemfwi

using namespace wemfwi;

namespace {

// Validation, grouping and the FP64 image of both entry points; fwi (device
// image [nh][nz][nx_image]) selects the objective and gradient, whose misfit
// goes to *misfit (host).
int uneven_run(const char* name, const wemfwi_image_shot* shots, int ns, int ngs, int nh, int nz, int nx_image,
               const float2* fwi, wemfwi_complex* image, double* misfit, int add, hipStream_t s) {
  const std::string who = name;
  if (!shots || ns < 0 || ngs < 0 || nh < 1 || nh > 41 || nh % 2 == 0 || nz < 2 || nx_image < 1 || !image)
    return fail(WEMFWI_ERR_ARG, who + ": invalid argument (nh odd, 1..41)");
  size_t budget = 0;
  {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) fr = (size_t)8 << 30;
    budget = fr / 2;
  }
  // validate every shot before touching the image
  for (int sh = 0; sh < ns; ++sh) {
    const wemfwi_image_shot& S = shots[sh];
    const std::string at = who + ": shot " + std::to_string(sh);
    if (!S.model || S.nw < 0 || (S.nw > 0 && (!S.freq_hz || !S.wavelet || !S.data)) || S.nr < 0 ||
        (S.nr > 0 && (!S.rec_iz || !S.rec_ix)))
      return fail(WEMFWI_ERR_ARG, at + ": null argument");
    const Model& m = S.model->m;
    if (m.d.nz != nz) return fail(WEMFWI_ERR_ARG, at + ": model nz differs");
    if (S.x0 < 0 || S.x0 + m.d.nx > nx_image) return fail(WEMFWI_ERR_ARG, at + ": aperture outside the image");
    if (S.src_iz < 0 || S.src_iz >= nz) return fail(WEMFWI_ERR_ARG, at + ": source depth");
    for (int r = 0; r < S.nr; ++r)
      if (S.rec_iz[r] < 0 || S.rec_iz[r] >= nz || S.rec_ix[r] < 0 || S.rec_ix[r] >= m.d.nx)
        return fail(WEMFWI_ERR_ARG, at + ": receiver outside the aperture");
    if (S.nw > 0) {
      if (int rc = m.require_fft()) return rc;
    }
  }
  const size_t npix = (size_t)nh * nz * nx_image;
  ScopedWs accw, jw;
  UCHK(accw.reserve(npix * sizeof(double2), "image accumulator"));
  double2* acc = (double2*)accw.ptr;
  UHIP(hipMemsetAsync(acc, 0, npix * sizeof(double2), s), "memset");
  double* jacc = nullptr;
  if (fwi) {
    UCHK(jw.reserve((1 + kSumBlocks) * sizeof(double), "misfit"));
    jacc = (double*)jw.ptr;
    UHIP(hipMemsetAsync(jacc, 0, sizeof(double), s), "memset");
  }
  const std::vector<float> gw = cos2_window(ngs);
  // groups of consecutive shots whose two wavefields fit half the free memory
  int s0 = 0;
  while (s0 < ns) {
    int s1 = s0;
    size_t bytes = 0;
    while (s1 < ns) {
      const Model& m = shots[s1].model->m;
      const size_t b = 2 * (size_t)shots[s1].nw * m.d.nz * m.d.nx * sizeof(float2);
      if (s1 > s0 && bytes + b > budget) break;
      bytes += b;
      ++s1;
    }
    UCHK(image_group(shots, s0, s1, gw, ngs, nh, nx_image, acc, s, fwi, jacc, jacc ? jacc + 1 : nullptr));
    s0 = s1;
  }
  UHIP(launch_d2f((float2*)image, acc, (long long)npix, add != 0, s), "image");
  if (fwi) {
    double j = 0.0;
    UHIP(hipMemcpyAsync(&j, jacc, sizeof(double), hipMemcpyDeviceToHost, s), "misfit");
    UHIP(hipStreamSynchronize(s), "misfit");
    *misfit = 0.5 * j;
  }
  UHIP(hipStreamSynchronize(s), "image");
  return WEMFWI_OK;
}

}  // namespace

extern "C" int wemfwi_uneven_offset_image_adjoint(const wemfwi_image_shot* shots, int ns, int ngs, int nh, int nz,
                                                  int nx_image, wemfwi_complex* image, int add, void* stream) {
  return uneven_run("uneven_offset_image_adjoint", shots, ns, ngs, nh, nz, nx_image, nullptr, image, nullptr, add,
                    (hipStream_t)stream);
}

extern "C" int wemfwi_uneven_offset_gradient(const wemfwi_image_shot* shots, int ns, int ngs, int nh, int nz,
                                             int nx_image, const wemfwi_complex* model, wemfwi_complex* grad,
                                             double* misfit, int add, void* stream) {
  if (!model || !misfit) return fail(WEMFWI_ERR_ARG, "uneven_offset_gradient: null model or misfit");
  if (model == grad) return fail(WEMFWI_ERR_ARG, "uneven_offset_gradient: model and gradient must differ");
  return uneven_run("uneven_offset_gradient", shots, ns, ngs, nh, nz, nx_image, (const float2*)model, grad, misfit,
                    add, (hipStream_t)stream);
}
