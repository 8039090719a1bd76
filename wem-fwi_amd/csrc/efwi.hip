// Operators of the frequency-extended FWI objective (scripts/efwi.py:160-190,
// operator/src/python/Operator.py:228-292,619-775): the soft clips of the
// slowness parameterisation and the filter along the frequency axis of the
// DSO regularisation, on device-resident [n_w][n_z][n_x] complex models.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "host.h"

namespace wemfwi {
namespace {

// h(u) = l u - sqrt(l^2 u^2 + tau^2) and h'(u) (Operator.py:619-623)
__device__ __forceinline__ float hpen(float u, float l, float tau) { return l * u - sqrtf(l * l * u * u + tau * tau); }
__device__ __forceinline__ float dhpen(float u, float l, float tau) {
  return l - l * l * u / sqrtf(l * l * u * u + tau * tau);
}

// One component (0 real, 1 imaginary) of each sample through
//   value:    y = s h(s (x - x0)) + x0          (s = +1: soft maximum, -1: soft minimum)
//   jacobian: y = x h'(s (b - x0))  at the background b (self-adjoint)
// the other component passes through.
__global__ void soft_clip_kernel(float2* __restrict__ out, const float2* __restrict__ in,
                                 const float2* __restrict__ bg, long long n, int comp, float s, float x0, float l,
                                 float tau, int jacobian, int add) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float2 v = in[i];
    float x = comp ? v.y : v.x;
    if (jacobian) {
      const float2 b = bg[i];
      x *= dhpen(s * ((comp ? b.y : b.x) - x0), l, tau);
    } else {
      x = s * hpen(s * (x - x0), l, tau) + x0;
    }
    if (comp)
      v.y = x;
    else
      v.x = x;
    if (add) {
      const float2 o = out[i];
      v = make_float2(o.x + v.x, o.y + v.y);
    }
    out[i] = v;
  }
}

// out[i][p] (+)= c * sum_j M[i][j] in[j][p] + (0, alpha * mean_j Im in[j][p])
// for a real nw x nw matrix M; each thread owns one column p and kTile
// outputs i (blockIdx.y picks the tile), accumulating in double.
constexpr int kTile = 16;
__global__ void freq_filter_kernel(float2* __restrict__ out, const float2* __restrict__ in,
                                   const float* __restrict__ M, int nw, long long ncol, float cr, float ci,
                                   float alpha, int add) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncol) return;
  const int i0 = blockIdx.y * kTile;
  double ar[kTile], ai[kTile];
#pragma unroll
  for (int o = 0; o < kTile; ++o) ar[o] = ai[o] = 0.0;
  double sim = 0.0;
  for (int j = 0; j < nw; ++j) {
    const float2 x = in[(long long)j * ncol + p];
    sim += x.y;
#pragma unroll
    for (int o = 0; o < kTile; ++o) {
      const int i = i0 + o;
      const double m = i < nw ? (double)M[(long long)i * nw + j] : 0.0;
      ar[o] += m * x.x;
      ai[o] += m * x.y;
    }
  }
  const double mean = sim / nw;
#pragma unroll
  for (int o = 0; o < kTile; ++o) {
    const int i = i0 + o;
    if (i >= nw) break;
    const double yr = cr * ar[o] - ci * ai[o];
    const double yi = cr * ai[o] + ci * ar[o] + alpha * mean;
    float2* q = out + (long long)i * ncol + p;
    if (add) {
      const float2 v = *q;
      *q = make_float2((float)(v.x + yr), (float)(v.y + yi));
    } else {
      *q = make_float2((float)yr, (float)yi);
    }
  }
}

// Index of sample k of a line of n samples extended by a boundary mode, or -1
// (scipy.ndimage extension modes, repeated for extensions longer than n):
// 0 reflect (d c b a | a b c d | d c b a), 1 constant 0, 2 nearest,
// 3 mirror (d c b | a b c d | c b a), 4 wrap.
long long extend(long long k, long long n, int mode) {
  if (k >= 0 && k < n) return k;
  switch (mode) {
    case 0: {
      const long long per = 2 * n;
      long long m = ((k % per) + per) % per;
      return m < n ? m : per - 1 - m;
    }
    case 1:
      return -1;
    case 2:
      return k < 0 ? 0 : n - 1;
    case 3: {
      if (n == 1) return 0;
      const long long per = 2 * n - 2;
      long long m = ((k % per) + per) % per;
      return m < n ? m : per - m;
    }
    default:
      return ((k % n) + n) % n;
  }
}

}  // namespace
}  // namespace wemfwi

using namespace wemfwi;

extern "C" {

int wemfwi_filter_matrix(int n, int L, const double* w, int mode, int convolve, double* M) {
  if (n < 1 || L < 1 || !w || !M || mode < 0 || mode > 4) return fail(WEMFWI_ERR_ARG, "filter_matrix: invalid argument");
  // correlate1d, origin 0: out[i] = sum_j w[j] x[i + j - L/2]; convolve1d is
  // correlate1d with the taps reversed and, for even L, the centre one sample
  // to the right (scipy.ndimage: origin = -1 for even lengths)
  std::vector<double> taps(w, w + L);
  long long centre = L / 2;
  if (convolve) {
    for (int j = 0; j < L; ++j) taps[j] = w[L - 1 - j];
    if (L % 2 == 0) centre = L / 2 - 1;
  }
  for (long long i = 0; i < (long long)n * n; ++i) M[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < L; ++j) {
      const long long k = extend((long long)i + j - centre, n, mode);
      if (k >= 0) M[(long long)i * n + k] += taps[j];
    }
  return WEMFWI_OK;
}

int wemfwi_freq_filter(int nw, long long ncol, const float* M, float c_re, float c_im, float alpha,
                       const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (nw < 1 || ncol < 0 || !M || !in || !out) return fail(WEMFWI_ERR_ARG, "freq_filter: invalid argument");
  if (ncol == 0) return WEMFWI_OK;
  if (in == out) return fail(WEMFWI_ERR_ARG, "freq_filter: in and out must differ");
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((unsigned)((ncol + 255) / 256), (unsigned)((nw + kTile - 1) / kTile));
  hipLaunchKernelGGL(freq_filter_kernel, grid, dim3(256), 0, s, (float2*)out, (const float2*)in, M, nw, ncol, c_re,
                     c_im, alpha, add);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "freq_filter");
}

int wemfwi_soft_clip(long long n, int comp, int sign, float x0, float l, float tau, int jacobian,
                     const wemfwi_complex* bg, const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (n < 0 || (comp != 0 && comp != 1) || (sign != 1 && sign != -1) || !in || !out || (jacobian && !bg))
    return fail(WEMFWI_ERR_ARG, "soft_clip: invalid argument");
  if (n == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  long long b = (n + 255) / 256;
  if (b > 1048576) b = 1048576;
  hipLaunchKernelGGL(soft_clip_kernel, dim3((unsigned)b), dim3(256), 0, s, (float2*)out, (const float2*)in,
                     (const float2*)bg, n, comp, (float)sign, x0, l, tau, jacobian, add);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "soft_clip");
}

}  // extern "C"
