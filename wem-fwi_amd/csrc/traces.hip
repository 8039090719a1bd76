// Time <-> frequency transforms of data traces (the float-data boundary of
// WEM::forward, BornFull::forward/adjoint and BornRefl, WEMnl.cpp:39-65,
// BornFull.cpp:39-62,80-102, BornRefl.cpp:13-49).
//
// The reference transforms a whole padded trace with FFTW (FFT1_r.cpp:7-96)
// and the propagation reads/writes only the survey's frequency bins.  Here
// only those bins are computed, directly:
//   forward:  F[b]  = n^-1/2 sum_{t<nt} x[t] exp(-2 pi i b t / n)
//   inverse:  x[t] (+)= Re( n^-1/2 sum_{b in bins} F[b] exp(+2 pi i b t / n) ),  t < nt
// with n = nt + tpad (the zero padding of Pad(..., tpad), PropParam.h:49) and
// the inverse keeping the real part as FFT1::adjoint does (FFT1_r.cpp:86-96).
// Twiddle angles use the exact integer phase (b t mod n); sums are in double.
#include <hip/hip_runtime.h>

#include "born.h"

namespace wemfwi {

// exp(-2 pi i j / n) for j < n, double precision
__global__ void trace_twiddle_kernel(double2* __restrict__ tw, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s, c;
  sincospi(-2.0 * (double)j / (double)n, &s, &c);
  tw[j] = make_double2(c, s);
}

// One workgroup per trace; the trace is staged in LDS, each thread owns bins.
__global__ void __launch_bounds__(256)
    trace_to_freq_kernel(const float* __restrict__ x, int ntr, int nt, int n, const int* __restrict__ bins, int nw,
                         const double2* __restrict__ tw, float2* __restrict__ f, int nbins) {
  extern __shared__ float xs[];
  const int tr = blockIdx.x;
  if (tr >= ntr) return;
  const float* xt = x + (size_t)tr * nt;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) xs[t] = xt[t];
  __syncthreads();
  const double scale = 1.0 / sqrt((double)n);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) {
    const int b = bins[i];
    double re = 0.0, im = 0.0;
    int idx = 0;
    for (int t = 0; t < nt; ++t) {
      const double2 w = tw[idx];
      const double v = (double)xs[t];
      re += v * w.x;
      im += v * w.y;
      idx += b;
      if (idx >= n) idx -= n;
    }
    f[(size_t)tr * nbins + b] = make_float2((float)(re * scale), (float)(im * scale));
  }
}

// One workgroup per trace; the survey bins of the trace are staged in LDS.
__global__ void __launch_bounds__(256)
    freq_to_trace_kernel(const float2* __restrict__ f, int ntr, int nbins, const int* __restrict__ bins, int nw,
                         int nt, int n, const double2* __restrict__ tw, float* __restrict__ x, int add) {
  extern __shared__ float2 fs[];
  const int tr = blockIdx.x;
  if (tr >= ntr) return;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) fs[i] = f[(size_t)tr * nbins + bins[i]];
  __syncthreads();
  const double scale = 1.0 / sqrt((double)n);
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    double re = 0.0;
    for (int i = 0; i < nw; ++i) {
      // exp(+2 pi i b t / n) = conj(tw[(b t) mod n])
      const int idx = (int)(((long long)bins[i] * t) % n);
      const double2 w = tw[idx];
      const float2 v = fs[i];
      re += (double)v.x * w.x + (double)v.y * w.y;
    }
    float* out = x + (size_t)tr * nt + t;
    const float val = (float)(re * scale);
    *out = add ? *out + val : val;
  }
}

hipError_t launch_trace_twiddles(double2* tw, int n, hipStream_t s) {
  hipLaunchKernelGGL(trace_twiddle_kernel, dim3((n + 255) / 256), dim3(256), 0, s, tw, n);
  return hipGetLastError();
}

hipError_t launch_trace_to_freq(const float* x, int ntr, int nt, int n, const int* bins, int nw, const double2* tw,
                                float2* f, int nbins, hipStream_t s) {
  if (ntr <= 0) return hipSuccess;
  hipLaunchKernelGGL(trace_to_freq_kernel, dim3(ntr), dim3(256), (size_t)nt * sizeof(float), s, x, ntr, nt, n,
                     bins, nw, tw, f, nbins);
  return hipGetLastError();
}

hipError_t launch_freq_to_trace(const float2* f, int ntr, int nbins, const int* bins, int nw, int nt, int n,
                                const double2* tw, float* x, int add, hipStream_t s) {
  if (ntr <= 0) return hipSuccess;
  hipLaunchKernelGGL(freq_to_trace_kernel, dim3(ntr), dim3(256), (size_t)nw * sizeof(float2), s, f, ntr, nbins,
                     bins, nw, nt, n, tw, x, add);
  return hipGetLastError();
}

}  // namespace wemfwi
