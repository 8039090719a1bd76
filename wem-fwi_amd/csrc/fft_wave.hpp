// Wavefront-resident Stockham FFT for CDNA4 (gfx950).
//
// One 64-lane wavefront owns one row of N = 64*E complex64 samples.  Lane t
// keeps samples t + 64*e (e = 0..E-1) in registers ("strided natural
// layout").  A transform is a chain of radix-R Stockham passes
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC 2008): pass p with sub-transform length NS reads
// butterfly j's inputs at j + r*N/R and writes its outputs at
// (j/NS)*NS*R + j%NS + r*NS.  When R divides E the first pass reads and the
// last pass writes the strided natural layout directly, so only the interior
// passes exchange data through LDS.  LDS addresses are padded by one complex
// every 16 to keep the scattered first-pass writes bank-conflict free.
//
// Twiddles come from a per-workgroup LDS table tw[m] = exp(-2*pi*i*m/N)
// computed in double precision on the host.
#pragma once

#include <hip/hip_runtime.h>

namespace wemfwi {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

template <bool INV>
__device__ __forceinline__ float2 cmul_tw(float2 a, float2 w) {
  if constexpr (INV) return cmulc(a, w);
  else return cmul(a, w);
}

// forward: multiply by -i ; inverse: multiply by +i
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  if constexpr (INV) return make_float2(-a.y, a.x);
  else return make_float2(a.y, -a.x);
}

// padded LDS index (one pad slot every 16 complex)
__device__ __forceinline__ int lidx(int i) { return i + (i >> 4); }

template <int N>
struct LdsSize {
  static constexpr int value = N + N / 16;
};

// ---------------------------------------------------------------- small DFTs

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

constexpr float kSqrtHalf = 0.70710678118654752440f;
constexpr float kCos8 = 0.92387953251128675613f;   // cos(pi/8)
constexpr float kSin8 = 0.38268343236508977173f;   // sin(pi/8)

// multiply by W16^m = exp(-2*pi*i*m/16) (forward) or its conjugate (inverse)
template <bool INV, int M>
__device__ __forceinline__ float2 tw16(float2 a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) return a;
  else if constexpr (m == 4) return mul_mi<INV>(a);
  else if constexpr (m == 8) return make_float2(-a.x, -a.y);
  else if constexpr (m == 12) return mul_mi<!INV>(a);
  else {
    // cos/sin of 2*pi*m/16 from the first-octant constants
    constexpr float c[16] = {1.f, kCos8, kSqrtHalf, kSin8, 0.f, -kSin8, -kSqrtHalf, -kCos8,
                             -1.f, -kCos8, -kSqrtHalf, -kSin8, 0.f, kSin8, kSqrtHalf, kCos8};
    constexpr float s[16] = {0.f, kSin8, kSqrtHalf, kCos8, 1.f, kCos8, kSqrtHalf, kSin8,
                             0.f, -kSin8, -kSqrtHalf, -kCos8, -1.f, -kCos8, -kSqrtHalf, -kSin8};
    // forward: exp(-i*theta) = (c, -s); inverse: (c, s)
    constexpr float wr = c[m];
    constexpr float wi = INV ? s[m] : -s[m];
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

template <bool INV>
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
  dft4<INV>(v[0], v[2], v[4], v[6]);
  dft4<INV>(v[1], v[3], v[5], v[7]);
  float2 o1 = tw16<INV, 2>(v[3]);
  float2 o2 = tw16<INV, 4>(v[5]);
  float2 o3 = tw16<INV, 6>(v[7]);
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <bool INV>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
  // y[q1 + 4 q2] = sum_n2 W16^(n2 q1) W4^(n2 q2) sum_n1 v[4 n1 + n2] W4^(n1 q1)
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
  // now v[n2 + 4 q1] = B[n2][q1]
  v[5] = tw16<INV, 1>(v[5]);
  v[6] = tw16<INV, 2>(v[6]);
  v[7] = tw16<INV, 3>(v[7]);
  v[9] = tw16<INV, 2>(v[9]);
  v[10] = tw16<INV, 4>(v[10]);
  v[11] = tw16<INV, 6>(v[11]);
  v[13] = tw16<INV, 3>(v[13]);
  v[14] = tw16<INV, 6>(v[14]);
  v[15] = tw16<INV, 9>(v[15]);
#pragma unroll
  for (int q1 = 0; q1 < 4; ++q1) dft4<INV>(v[4 * q1 + 0], v[4 * q1 + 1], v[4 * q1 + 2], v[4 * q1 + 3]);
  // v[4 q1 + q2] = y[q1 + 4 q2]: transpose
  float2 t[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) t[i] = v[i];
#pragma unroll
  for (int q1 = 0; q1 < 4; ++q1)
#pragma unroll
    for (int q2 = 0; q2 < 4; ++q2) v[q1 + 4 * q2] = t[4 * q1 + q2];
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&v)[R]) {
  if constexpr (R == 2) dft2<INV>(v[0], v[1]);
  else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8<INV>(v);
  else if constexpr (R == 16) dft16<INV>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------- passes

__device__ __forceinline__ void wave_lds_order() {
  // LDS instructions of one wavefront execute in order; this keeps the
  // compiler from moving LDS accesses across a pass boundary.
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int N, int E, int R, int NS, bool INV, bool FROM_LDS, bool TO_LDS>
__device__ __forceinline__ void stockham_pass(float2 (&x)[E], float2* __restrict__ buf,
                                              const float2* __restrict__ tw, int t) {
  constexpr int NB = N / R;
  constexpr int BPL = E / R;
  static_assert(BPL * R == E, "radix must divide the per-lane count");
  if constexpr (FROM_LDS) {
#pragma unroll
    for (int b = 0; b < BPL; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) x[b + r * BPL] = buf[lidx(t + 64 * b + r * NB)];
    wave_lds_order();
  }
#pragma unroll
  for (int b = 0; b < BPL; ++b) {
    const int j = t + 64 * b;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[b + r * BPL];
    if constexpr (NS > 1) {
      const int jm = j & (NS - 1);
      constexpr int STRIDE = N / (NS * R);
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul_tw<INV>(v[r], tw[jm * r * STRIDE]);
    }
    dft<R, INV>(v);
    if constexpr (TO_LDS) {
      const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lidx(base + r * NS)] = v[r];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) x[b + r * BPL] = v[r];
    }
  }
  if constexpr (TO_LDS) wave_lds_order();
}

// Radix plans.  Every radix divides E = N/64 so that the first pass reads and
// the last pass writes registers; all plans have at most 4 passes.
template <int N>
struct FftPlan;
template <>
struct FftPlan<256> {  // E = 4
  template <bool INV>
  __device__ static void run(float2 (&x)[4], float2* buf, const float2* tw, int t) {
    stockham_pass<256, 4, 4, 1, INV, false, true>(x, buf, tw, t);
    stockham_pass<256, 4, 4, 4, INV, true, true>(x, buf, tw, t);
    stockham_pass<256, 4, 4, 16, INV, true, true>(x, buf, tw, t);
    stockham_pass<256, 4, 4, 64, INV, true, false>(x, buf, tw, t);
  }
};
template <>
struct FftPlan<512> {  // E = 8
  template <bool INV>
  __device__ static void run(float2 (&x)[8], float2* buf, const float2* tw, int t) {
    stockham_pass<512, 8, 8, 1, INV, false, true>(x, buf, tw, t);
    stockham_pass<512, 8, 8, 8, INV, true, true>(x, buf, tw, t);
    stockham_pass<512, 8, 8, 64, INV, true, false>(x, buf, tw, t);
  }
};
template <>
struct FftPlan<1024> {  // E = 16
  template <bool INV>
  __device__ static void run(float2 (&x)[16], float2* buf, const float2* tw, int t) {
    stockham_pass<1024, 16, 16, 1, INV, false, true>(x, buf, tw, t);
    stockham_pass<1024, 16, 4, 16, INV, true, true>(x, buf, tw, t);
    stockham_pass<1024, 16, 16, 64, INV, true, false>(x, buf, tw, t);
  }
};
template <>
struct FftPlan<2048> {  // E = 32
  template <bool INV>
  __device__ static void run(float2 (&x)[32], float2* buf, const float2* tw, int t) {
    stockham_pass<2048, 32, 16, 1, INV, false, true>(x, buf, tw, t);
    stockham_pass<2048, 32, 8, 16, INV, true, true>(x, buf, tw, t);
    stockham_pass<2048, 32, 16, 128, INV, true, false>(x, buf, tw, t);
  }
};
template <>
struct FftPlan<128> {  // E = 2
  template <bool INV>
  __device__ static void run(float2 (&x)[2], float2* buf, const float2* tw, int t) {
    stockham_pass<128, 2, 2, 1, INV, false, true>(x, buf, tw, t);
    stockham_pass<128, 2, 2, 2, INV, true, true>(x, buf, tw, t);
    stockham_pass<128, 2, 2, 4, INV, true, true>(x, buf, tw, t);
    stockham_pass<128, 2, 2, 8, INV, true, true>(x, buf, tw, t);
    stockham_pass<128, 2, 2, 16, INV, true, true>(x, buf, tw, t);
    stockham_pass<128, 2, 2, 32, INV, true, true>(x, buf, tw, t);
    stockham_pass<128, 2, 2, 64, INV, true, false>(x, buf, tw, t);
  }
};

// Unnormalised DFT (forward: exp(-2 pi i jk/N); inverse: exp(+2 pi i jk/N))
// of the row held in x (strided natural layout in and out).
template <int N, bool INV>
__device__ __forceinline__ void fft_row(float2 (&x)[N / 64], float2* buf, const float2* tw, int t) {
  FftPlan<N>::template run<INV>(x, buf, tw, t);
}

}  // namespace wemfwi
