// Model interpolation of the FWI preconditioning (scripts/fwi.py:46-68,
// efwi.py:48-80): separable Lanczos and cubic-spline filters from a coarse
// model grid to the propagation grid (operator/src/LanczosInterpolation2D.h,
// LanczosInterpolation3D.h, Spline2D.h, Spline3D.h, Interpolation3D.h), built
// on the host in the reference's float / double expression order and applied
// one axis at a time on the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "host.h"

namespace wemfwi {
namespace {

// Nonzero band of each output row of the filter: band[i] = [first, last + 1)
// of the j with F[i][j] != 0 (F[j][i] for the transpose; [0, 0) for an
// all-zero row).  One thread per output row, run once per apply: the rows of
// the interpolation filters hold a few taps near the diagonal and, from the
// even extension, a few more near the ends.
__global__ void axis_band_kernel(int2* __restrict__ band, const float* __restrict__ F, int n_out, int n_in,
                                 int transpose) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  int lo = n_in, hi = 0;
  for (int j = 0; j < n_in; ++j) {
    const float f = transpose ? F[(long long)j * n_out + i] : F[(long long)i * n_in + j];
    if (f != 0.f) {
      lo = j < lo ? j : lo;
      hi = j + 1;
    }
  }
  band[i] = lo < hi ? make_int2(lo, hi) : make_int2(0, 0);
}

// out[(a n_out + i) inner + b] (+)= sum_j F[i][j] in[(a n_in + j) inner + b]
// (transpose: F[j][i] with F stored [n_in][n_out]) over j in row i's band;
// one thread per output sample, double accumulation in increasing j (zero
// entries inside the band add nothing, as outside it).
__global__ void axis_matrix_kernel(float2* __restrict__ out, const float2* __restrict__ in,
                                   const float* __restrict__ F, const int2* __restrict__ band, int n_out, int n_in,
                                   long long outer, long long inner, int transpose, int add) {
  const long long total = outer * n_out * inner;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const long long b = q % inner;
    const long long r = q / inner;
    const int i = (int)(r % n_out);
    const long long a = r / n_out;
    const float2* x = in + a * n_in * inner + b;
    const int2 bd = band[i];
    double sr = 0.0, si = 0.0;
    for (int j = bd.x; j < bd.y; ++j) {
      const float f = transpose ? F[(long long)j * n_out + i] : F[(long long)i * n_in + j];
      if (f != 0.f) {
        const float2 v = x[(long long)j * inner];
        sr += (double)f * v.x;
        si += (double)f * v.y;
      }
    }
    float2* y = out + q;
    if (add) {
      const float2 o = *y;
      *y = make_float2((float)(o.x + sr), (float)(o.y + si));
    } else {
      *y = make_float2((float)sr, (float)si);
    }
  }
}

const double kPiD = 4. * std::atan(1);

// Lanczos window of LanczosInterpolation3D.h:29-55 (the 2-D form is the same
// with taper 0): a = a_cells * d_m.
void lanczos_filter(int nd, float dd, int nm, float dm, float a, float tap, float* F) {
  const int nf = 2 * a / dd + 1;
  const int na = (nf - 1) / 2;
  const float half = na * dd;
  const float new_d = (1.f + tap) * dm;
  for (long long k = 0; k < (long long)nd * nm; ++k) F[k] = 0.f;
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nm; ++j) {
      const float x = i * dd - j * dm;
      if (std::abs(x) <= half) {
        if (x == 0)
          F[(long long)i * nm + j] = 1;
        else
          F[(long long)i * nm + j] =
              new_d * a * std::sin(kPiD * x / new_d) * std::sin(kPiD * x / a) / (kPiD * kPiD * x * x);
      }
    }
  // taps beyond the first sample folded back (even extension at both ends)
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nm; ++j) {
      const float x = i * dd - j * dm;
      const float xx = i * dd + j * dm;
      if (std::abs(xx) < half && x != xx) {
        F[(long long)i * nm + j] += new_d * a * std::sin(kPiD * xx / new_d) * std::sin(kPiD * xx / a) / (kPiD * kPiD * xx * xx);
        F[(long long)(nd - i - 1) * nm + (nm - j - 1)] +=
            new_d * a * std::sin(kPiD * xx / new_d) * std::sin(kPiD * xx / a) / (kPiD * kPiD * xx * xx);
      }
    }
}

float cubic(float y, float a, float b) {
  if (y < 1) return ((-6 * a - 9 * b + 12) * y * y * y + (6 * a + 12 * b - 18) * y * y - 2 * b + 6) / 6;
  if (1 <= y && y < 2)
    return ((-6 * a - b) * y * y * y + (30 * a + 6 * b) * y * y + (-48 * a - 12 * b) * y + 24 * a + 8 * b) / 6;
  return 0.f;
}

// cubic (a, b) spline of Spline3D.h:22-52 (Spline2D.h:97-127 with taper 0)
void spline_filter(int nd, float dd, int nm, float dm, float a, float b, float tap, float* F) {
  const float new_d = (1.f + tap) * dm;
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nm; ++j) {
      const float x = std::abs(i * dd - j * dm);
      F[(long long)i * nm + j] = cubic(x / new_d, a, b);
    }
  for (int i = 0; i < nd; ++i)
    for (int j = 1; j < nm - 1; ++j) {
      const float xx = i * dd + j * dm;
      const float f = cubic(std::abs(xx) / new_d, a, b);
      F[(long long)i * nm + j] += f;
      F[(long long)(nd - i - 1) * nm + (nm - j - 1)] += f;
    }
}

}  // namespace
}  // namespace wemfwi

using namespace wemfwi;

extern "C" {

int wemfwi_interp_filter(int kind, int nd, float dd, int nm, float dm, float a, float b, float taper, float* F) {
  if (nd < 1 || nm < 1 || !F || !(dd > 0.f) || !(dm > 0.f)) return fail(WEMFWI_ERR_ARG, "interp_filter: invalid axes");
  if (kind == 0) {
    const float aw = a * dm;
    if (!(aw > 0.f)) return fail(WEMFWI_ERR_ARG, "interp_filter: Lanczos a must be positive");
    lanczos_filter(nd, dd, nm, dm, aw, taper, F);
  } else if (kind == 1) {
    spline_filter(nd, dd, nm, dm, a, b, taper, F);
  } else {
    return fail(WEMFWI_ERR_ARG, "interp_filter: kind must be 0 (Lanczos) or 1 (spline)");
  }
  return WEMFWI_OK;
}

int wemfwi_axis_apply(const float* F, int n_out, int n_in, long long outer, long long inner, int transpose,
                      const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (!F || !in || !out || n_out < 1 || n_in < 1 || outer < 0 || inner < 1)
    return fail(WEMFWI_ERR_ARG, "axis_apply: invalid argument");
  if ((const void*)in == (const void*)out) return fail(WEMFWI_ERR_ARG, "axis_apply: in and out must differ");
  const long long total = outer * n_out * inner;
  if (total == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  // the bands live in stream-ordered memory of this call (concurrent calls
  // on other streams keep their own)
  int2* band = nullptr;
  hipError_t e = hipMallocAsync((void**)&band, (size_t)n_out * sizeof(int2), s);
  if (e != hipSuccess) return hip_fail(e, "axis_apply band");
  hipLaunchKernelGGL(axis_band_kernel, dim3((n_out + 255) / 256), dim3(256), 0, s, band, F, n_out, n_in, transpose);
  e = hipGetLastError();
  long long blocks = (total + 255) / 256;
  if (blocks > 1048576) blocks = 1048576;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(axis_matrix_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (float2*)out,
                       (const float2*)in, F, band, n_out, n_in, outer, inner, transpose, add);
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(band, s);
  if (e == hipSuccess) e = ef;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "axis_apply");
}

}  // extern "C"
