// Host-side acquisition parameters: the frequency grid and wavelet spectrum of
// PropParam::calcPropParam and the regular device geometry of Device.  Float32
// arithmetic follows the reference expression by expression (the rounding
// decides which FFT bins and grid cells are used).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "born.h"
#include "host.h"
#include "../../include/wemfwi.h"

namespace wemfwi {

// Unitary forward DFT, X[k] = n^-1/2 sum_t x[t] exp(-2 pi i k t / n), in double
// precision with exact twiddle indices (k t mod n).  This is the FFTW_FORWARD
// transform scaled by 1/sqrt(n) of FFT1::forward (operator/src/FFT1_r.cpp:74-83);
// the wavelet is transformed once per operator, so O(n^2) is immaterial.
static void unitary_dft(const std::vector<double>& x, wemfwi_complex* out) {
  const size_t n = x.size();
  std::vector<double> c(n), s(n);
  for (size_t j = 0; j < n; ++j) {
    const double a = -2.0 * M_PI * (double)j / (double)n;
    c[j] = std::cos(a);
    s[j] = std::sin(a);
  }
  const double scale = 1.0 / std::sqrt((double)n);
  for (size_t k = 0; k < n; ++k) {
    double re = 0.0, im = 0.0;
    size_t idx = 0;
    for (size_t t = 0; t < n; ++t) {
      re += x[t] * c[idx];
      im += x[t] * s[idx];
      idx += k;
      if (idx >= n) idx -= n;
    }
    out[k].re = (float)(re * scale);
    out[k].im = (float)(im * scale);
  }
}

// Twiddle table exp(-2 pi i j/n), j < n, per (device, n), kept for the process.
static int trace_twiddles(int n, hipStream_t s, const double2** out) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, double2*> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "trace twiddles");
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({dev, n});
  if (it != cache.end()) {
    *out = it->second;
    return WEMFWI_OK;
  }
  double2* tw = nullptr;
  int rc = dev_alloc(&tw, (size_t)n, "trace twiddles");
  if (rc) return rc;
  e = launch_trace_twiddles(tw, n, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(tw);
    return hip_fail(e, "trace twiddles");
  }
  cache[{dev, n}] = tw;
  *out = tw;
  return WEMFWI_OK;
}

static int check_bins(int nw, const int* bins, int limit) {
  if (nw < 1 || !bins) return fail(WEMFWI_ERR_ARG, "traces: no frequency bins");
  for (int i = 0; i < nw; ++i)
    if (bins[i] < 0 || bins[i] >= limit) return fail(WEMFWI_ERR_ARG, "traces: frequency bin out of range");
  return WEMFWI_OK;
}

}  // namespace wemfwi

using namespace wemfwi;

extern "C" {

int wemfwi_prop_param(int nt, float ot, float dt, const float* wave, int tpad, float fmin, float fmax,
                      int nfreq, int cap, int* nw, float* freq, int* index, wemfwi_complex* wave_f) {
  if (nt < 1 || tpad < 0 || !nw || !(dt > 0.f)) return fail(WEMFWI_ERR_ARG, "prop_param: invalid trace axis");
  const int n = nt + tpad;
  // PropParam.h:58: tmax of the padded trace axis
  const float tmax = ot + dt * (float)(n - 1);
  int imin, imax;
  if (fmax == 0.f) {
    imin = 1;
    imax = nfreq;
    if (imax == 0) imax = n / 2 + 1;
  } else {
    imin = (int)(fmin * tmax);
    imax = (int)(fmax * tmax + .5f);
  }
  const int count = imax - imin + 1;
  if (count < 1) return fail(WEMFWI_ERR_ARG, "prop_param: empty frequency range");
  *nw = count;
  if (!freq && !index && !wave_f) return WEMFWI_OK;
  if (cap < count) return fail(WEMFWI_ERR_ARG, "prop_param: output capacity below the number of frequencies");
  for (int i = 0; i < count; ++i) {
    if (freq) freq[i] = fmin + (float)i / tmax;
    if (index) index[i] = i + imin;
  }
  if (wave_f) {
    if (!wave) return fail(WEMFWI_ERR_ARG, "prop_param: wavelet is NULL");
    std::vector<double> x((size_t)n, 0.0);
    for (int t = 0; t < nt; ++t) x[t] = (double)wave[t];
    unitary_dft(x, wave_f);
  }
  return WEMFWI_OK;
}

int wemfwi_trace_to_freq(int ntr, int nt, int tpad, const float* data_t, int nw, const int* bins, int nbins,
                         wemfwi_complex* data_f, int add, void* stream) {
  if (ntr < 0 || nt < 1 || tpad < 0 || nbins < 1 || !data_t || !data_f)
    return fail(WEMFWI_ERR_ARG, "trace_to_freq: invalid shape or NULL pointer");
  const int n = nt + tpad;
  int rc = check_bins(nw, bins, std::min(nbins, n));
  if (rc) return rc;
  if (ntr == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  const double2* tw = nullptr;
  if ((rc = trace_twiddles(n, s, &tw))) return rc;
  int* d_bins = nullptr;
  if ((rc = dev_upload(&d_bins, bins, (size_t)nw, "bins"))) return rc;
  hipError_t e = hipSuccess;
  if (!add) e = hipMemsetAsync(data_f, 0, (size_t)ntr * nbins * sizeof(wemfwi_complex), s);
  if (e == hipSuccess)
    e = launch_trace_to_freq(data_t, ntr, nt, n, d_bins, nw, tw, (float2*)data_f, nbins, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_bins);
  if (e != hipSuccess) return hip_fail(e, "trace_to_freq");
  return WEMFWI_OK;
}

int wemfwi_freq_to_trace(int ntr, int nt, int tpad, const wemfwi_complex* data_f, int nbins, int nw, const int* bins,
                         float* data_t, int add, void* stream) {
  if (ntr < 0 || nt < 1 || tpad < 0 || nbins < 1 || !data_t || !data_f)
    return fail(WEMFWI_ERR_ARG, "freq_to_trace: invalid shape or NULL pointer");
  const int n = nt + tpad;
  int rc = check_bins(nw, bins, std::min(nbins, n));
  if (rc) return rc;
  if (ntr == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  const double2* tw = nullptr;
  if ((rc = trace_twiddles(n, s, &tw))) return rc;
  int* d_bins = nullptr;
  if ((rc = dev_upload(&d_bins, bins, (size_t)nw, "bins"))) return rc;
  hipError_t e = launch_freq_to_trace((const float2*)data_f, ntr, nbins, d_bins, nw, nt, n, tw, data_t, add, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_bins);
  if (e != hipSuccess) return hip_fail(e, "freq_to_trace");
  return WEMFWI_OK;
}

int wemfwi_device_indices(int n, float o_x, float o_z, float d_x, float d_z, float ox, float oz, float dxv,
                          float dzv, int* iz, int* ix) {
  if (n <= 0) return fail(WEMFWI_ERR_ARG, "device: number of devices is zero");
  if (!iz || !ix) return fail(WEMFWI_ERR_ARG, "device: NULL output");
  if (o_x < ox || o_z < oz) return fail(WEMFWI_ERR_ARG, "device: devices start outside of the model");
  if (d_x == 0.f && d_z == 0.f) return fail(WEMFWI_ERR_ARG, "device: devices have spacing equal to zero");
  if (!(dxv > 0.f) || !(dzv > 0.f)) return fail(WEMFWI_ERR_ARG, "device: invalid model spacing");
  for (int i = 0; i < n; ++i) {
    const float x = o_x + (float)i * d_x;
    const float z = o_z + (float)i * d_z;
    const int jx = (int)((x - ox) / dxv + .5f);
    const int jz = (int)((z - oz) / dzv + .5f);
    if (jx < 0 || jz < 0)
      return fail(WEMFWI_ERR_ARG, "device {" + std::to_string(z) + "," + std::to_string(x) + "} is out of bounds");
    ix[i] = jx;
    iz[i] = jz;
  }
  return WEMFWI_OK;
}

int wemfwi_device_coords(int n, const float* x, const float* z, float ox, float oz, float dxv, float dzv, int* iz,
                         int* ix) {
  if (n < 0 || (n > 0 && (!x || !z || !iz || !ix))) return fail(WEMFWI_ERR_ARG, "device: invalid argument");
  if (!(dxv > 0.f) || !(dzv > 0.f)) return fail(WEMFWI_ERR_ARG, "device: invalid model spacing");
  for (int i = 0; i < n; ++i) {
    const int jx = (int)((x[i] - ox) / dxv + .5f);
    const int jz = (int)((z[i] - oz) / dzv + .5f);
    if (jx < 0 || jz < 0)
      return fail(WEMFWI_ERR_ARG,
                  "device {" + std::to_string(z[i]) + "," + std::to_string(x[i]) + "} is out of bounds");
    ix[i] = jx;
    iz[i] = jz;
  }
  return WEMFWI_OK;
}

}  // extern "C"
