// Subsurface-offset migration of shots with uneven apertures and bands: the
// adjoint of BornRefl's modelling (ext_born_refl.cpp:60-110: D = Down(inj_src),
// R = Up^H(inj_rec^H d), image += IC^H) with the subsurface-offset imaging
// condition I(h, z, x) = sum conj(D(z, x-h)) R(z, x+h) in place of the zero-lag
// one, for shots that each carry their own background on their own aperture
// (a model nx_s wide, placed at column x0_s of the image) and their own list
// of frequencies.  All (shot, frequency) rows of a group of shots are swept by
// two batched launches (rows of different widths share a launch per FFT
// length); the image is reduced in FP64, shot after shot and row after row.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// img[h][z][x0 + x] += sum over the nrows rows of one shot (in order) of
// conj(D[z][x-h]) R[z][x+h], h = j - nh/2, terms with x-h or x+h outside
// [0, nx) absent.  A workgroup owns one depth z and kTX consecutive x (one
// per thread, nh lags in FP64 registers); per row the D and R segments with
// their nh/2 halos are staged in LDS once, so every sample is read from HBM
// once per row instead of once per lag.
constexpr int kTX = 256;
constexpr int kHalo = 20;  // nh <= 41
template <int NHM>
__global__ void __launch_bounds__(kTX)
    offset_image_rows_kernel(double2* __restrict__ img, long long hstride, int nxi, int x0,
                             const float2* __restrict__ D, const float2* __restrict__ R, long long rstride,
                             int nrows, int nz, int nx, int nh) {
  __shared__ float2 sd[kTX + 2 * kHalo], sr[kTX + 2 * kHalo];
  const int z = blockIdx.y;
  const int xb = blockIdx.x * kTX;
  const int t = threadIdx.x;
  const int x = xb + t;
  const int hh = nh / 2;
  const int seg = kTX + 2 * hh;
  double ax[NHM], ay[NHM];
#pragma unroll
  for (int j = 0; j < NHM; ++j) ax[j] = ay[j] = 0.0;
  for (int r = 0; r < nrows; ++r) {
    const float2* d = D + r * rstride + (long long)z * nx;
    const float2* s = R + r * rstride + (long long)z * nx;
    for (int i = t; i < seg; i += kTX) {
      const int xx = xb - hh + i;
      const bool in = xx >= 0 && xx < nx;
      sd[i] = in ? ldg(d, (size_t)xx) : make_float2(0.f, 0.f);
      sr[i] = in ? ldg(s, (size_t)xx) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (x < nx) {
#pragma unroll
      for (int j = 0; j < NHM; ++j) {
        const int h = j - hh;
        if (j < nh && x - h >= 0 && x - h < nx && x + h >= 0 && x + h < nx) {
          const float2 v = cmulc(sr[t + hh + h], sd[t + hh - h]);
          ax[j] += v.x;
          ay[j] += v.y;
        }
      }
    }
    __syncthreads();
  }
  if (x < nx) {
#pragma unroll
    for (int j = 0; j < NHM; ++j) {
      if (j >= nh) break;
      double2* o = img + j * hstride + (long long)z * nxi + x0 + x;
      const double2 a = *o;
      *o = make_double2(a.x + ax[j], a.y + ay[j]);
    }
  }
}

template <int NHM>
hipError_t launch_oir(double2* img, long long hstride, int nxi, int x0, const float2* D, const float2* R,
                      long long rstride, int nrows, int nz, int nx, int nh, hipStream_t s) {
  dim3 grid((unsigned)((nx + kTX - 1) / kTX), (unsigned)nz);
  hipLaunchKernelGGL((offset_image_rows_kernel<NHM>), grid, dim3(kTX), 0, s, img, hstride, nxi, x0, D, R, rstride,
                     nrows, nz, nx, nh);
  return hipGetLastError();
}

hipError_t launch_offset_image_rows(double2* img, long long hstride, int nxi, int x0, const float2* D,
                                   const float2* R, long long rstride, int nrows, int nz, int nx, int nh,
                                   hipStream_t s) {
  if (nrows <= 0 || nz <= 0 || nx <= 0) return hipSuccess;
  if (nh > 2 * kHalo + 1 || nz > 65535) return hipErrorInvalidValue;
  if (nh <= 1) return launch_oir<1>(img, hstride, nxi, x0, D, R, rstride, nrows, nz, nx, nh, s);
  if (nh <= 3) return launch_oir<3>(img, hstride, nxi, x0, D, R, rstride, nrows, nz, nx, nh, s);
  if (nh <= 11) return launch_oir<11>(img, hstride, nxi, x0, D, R, rstride, nrows, nz, nx, nh, s);
  if (nh <= 21) return launch_oir<21>(img, hstride, nxi, x0, D, R, rstride, nrows, nz, nx, nh, s);
  return launch_oir<41>(img, hstride, nxi, x0, D, R, rstride, nrows, nz, nx, nh, s);
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
// reduce the image shot by shot.
int image_group(const wemfwi_image_shot* shots, int s0, int s1, const std::vector<float>& gw, int ngs, int nh,
                int nxi, double2* acc, hipStream_t s) {
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
  ScopedWs Dw, Rw, srcw, recw, gww, devw;
  UCHK(Dw.reserve(total * sizeof(float2), "source wavefields"));
  UCHK(Rw.reserve(total * sizeof(float2), "receiver wavefields"));
  float2* D = (float2*)Dw.ptr;
  float2* R = (float2*)Rw.ptr;
  UHIP(hipMemsetAsync(D, 0, total * sizeof(float2), s), "memset");
  UHIP(hipMemsetAsync(R, 0, total * sizeof(float2), s), "memset");
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
  UCHK(upload_vec(srcw, hs, s, "source rows"));
  UCHK(upload_vec(recw, hr, s, "receiver rows"));
  UCHK(upload_vec(gww, gw, s, "spread weights"));
  const int span = 2 * ngs + 1;
  hipLaunchKernelGGL(src_rows_kernel, dim3((nrows * span + 255) / 256), dim3(256), 0, s, (const SrcRow*)srcw.ptr,
                     nrows, (const float*)gww.ptr, ngs);
  UHIP(hipGetLastError(), "inject sources");
  hipLaunchKernelGGL(rec_rows_kernel, dim3((nrows + 63) / 64), dim3(64), 0, s, (const RecRow*)recw.ptr, nrows);
  UHIP(hipGetLastError(), "inject receivers");
  UHIP(hipStreamSynchronize(s), "inject");
  // D = Down(sources), R = Up^H(receivers), in place, every row in one plan
  for (int i = 0; i < nrows; ++i) pin[i] = pout[i] = (cf*)(D + roff[i]);
  {
    BatchPlan p;
    UCHK(p.build(nrows, rm.data(), rf.data(), pin.data(), pout.data(), s));
    UCHK(p.run(0, 0, s));
  }
  for (int i = 0; i < nrows; ++i) pin[i] = pout[i] = (cf*)(R + roff[i]);
  {
    BatchPlan p;
    UCHK(p.build(nrows, rm.data(), rf.data(), pin.data(), pout.data(), s));
    UCHK(p.run(1, 1, s));
  }
  // image, shot after shot
  k = 0;
  const long long hstride = (long long)shots[s0].model->m.d.nz * nxi;
  for (int sh = s0; sh < s1; ++sh) {
    const Model& m = shots[sh].model->m;
    UHIP(launch_offset_image_rows(acc, hstride, nxi, shots[sh].x0, D + roff[k], R + roff[k],
                                  (long long)m.d.nz * m.d.nx, shots[sh].nw, m.d.nz, m.d.nx, nh, s),
         "offset image");
    k += shots[sh].nw;
  }
  UHIP(hipStreamSynchronize(s), "offset image");
  return WEMFWI_OK;
}

}  // namespace
}  // namespace wemfwi

using namespace wemfwi;

extern "C" int wemfwi_uneven_offset_image_adjoint(const wemfwi_image_shot* shots, int ns, int ngs, int nh, int nz,
                                                  int nx_image, wemfwi_complex* image, int add, void* stream) {
  if (!shots || ns < 0 || ngs < 0 || nh < 1 || nh > 41 || nh % 2 == 0 || nz < 2 || nx_image < 1 || !image)
    return fail(WEMFWI_ERR_ARG, "uneven_offset_image_adjoint: invalid argument (nh odd, 1..41)");
  hipStream_t s = (hipStream_t)stream;
  size_t budget = 0;
  {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) fr = (size_t)8 << 30;
    budget = fr / 2;
  }
  // validate every shot before touching the image
  for (int sh = 0; sh < ns; ++sh) {
    const wemfwi_image_shot& S = shots[sh];
    if (!S.model || S.nw < 0 || (S.nw > 0 && (!S.freq_hz || !S.wavelet || !S.data)) || S.nr < 0 ||
        (S.nr > 0 && (!S.rec_iz || !S.rec_ix)))
      return fail(WEMFWI_ERR_ARG, "uneven_offset_image_adjoint: shot " + std::to_string(sh) + ": null argument");
    const Model& m = S.model->m;
    if (m.d.nz != nz)
      return fail(WEMFWI_ERR_ARG, "uneven_offset_image_adjoint: shot " + std::to_string(sh) + ": model nz differs");
    if (S.x0 < 0 || S.x0 + m.d.nx > nx_image)
      return fail(WEMFWI_ERR_ARG, "uneven_offset_image_adjoint: shot " + std::to_string(sh) +
                                      ": aperture outside the image");
    if (S.src_iz < 0 || S.src_iz >= nz)
      return fail(WEMFWI_ERR_ARG, "uneven_offset_image_adjoint: shot " + std::to_string(sh) + ": source depth");
    for (int r = 0; r < S.nr; ++r)
      if (S.rec_iz[r] < 0 || S.rec_iz[r] >= nz || S.rec_ix[r] < 0 || S.rec_ix[r] >= m.d.nx)
        return fail(WEMFWI_ERR_ARG, "uneven_offset_image_adjoint: shot " + std::to_string(sh) +
                                        ": receiver outside the aperture");
    if (S.nw > 0) {
      if (int rc = m.require_fft()) return rc;
    }
  }
  const size_t npix = (size_t)nh * nz * nx_image;
  ScopedWs accw;
  UCHK(accw.reserve(npix * sizeof(double2), "image accumulator"));
  double2* acc = (double2*)accw.ptr;
  UHIP(hipMemsetAsync(acc, 0, npix * sizeof(double2), s), "memset");
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
    UCHK(image_group(shots, s0, s1, gw, ngs, nh, nx_image, acc, s));
    s0 = s1;
  }
  UHIP(launch_d2f((float2*)image, acc, (long long)npix, add != 0, s), "image");
  UHIP(hipStreamSynchronize(s), "image");
  return WEMFWI_OK;
}
