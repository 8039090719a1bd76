// Fused one-way depth step on one wavefront (SSF / PSPI, forward and adjoint).
//
// Row layout: lane t holds padded sample p = t + 64*e (e < E = N/64) of the
// FFT row; model sample x = p - tap lives there when 0 <= x < nx, the rest is
// the zero padding of OneStep (OneStep.h:26-30).
//
// Forward step (SSF.cpp:8-32 / PSPI.cpp:8-29):
//   X = F pad(w_in)                              (one forward FFT)
//   for every reference r selecting some x at this depth:
//     Y = F^H (P_r X)                            (P_r from the phase table)
//     out[x] = Y[x] for x with refid[x] == r     (Selector assigns)
//   out[x] *= taper[x] * exp(-i w phi[x])        (Taper + SplitStep)
// Adjoint step (SSF.cpp:34-59): the same operations transposed, one forward
// FFT per reference and one inverse FFT.
//
// The unitary 1/sqrt(N) factors of both transforms are folded into the phase
// table (entries carry 1/N).
#pragma once

#include "fft_wave.hpp"

#define WEMFWI_GLOBAL __attribute__((address_space(1)))

namespace wemfwi {

// Loads through the global address space (saddr + 32-bit offset addressing
// instead of 64-bit flat addresses per lane).
template <typename T>
__device__ __forceinline__ T ldg(const T* p, int i) {
  return ((const WEMFWI_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ T ldg(const T* p, size_t i) {
  return ((const WEMFWI_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void stg(T* p, int i, T v) {
  ((WEMFWI_GLOBAL T*)p)[i] = v;
}

// Per-model depth tables (device pointers).
struct DepthTables {
  const int* slot_begin;        // [nz+1] first active (depth, reference) slot of a depth
  const int* slot_ref;          // [nslots] reference index of each slot
  const signed char* refid;     // [nz][nx] reference selecting x (-1: none)
  const float2* phi;            // [nz][nx] dz*(S - Sref)/sqrt(Sref) of x's reference (0 for PSPI)
  const float* taper;           // [nx]
};

struct StepGeom {
  int nx;       // propagation width
  int tap;      // FFT zero padding per side
  int ssf;      // 1: split-step correction on
};

// sin and cos of x for |x| < 1e5: Cody-Waite reduction by a three-part pi/2
// (W. J. Cody, W. Waite, "Software Manual for the Elementary Functions", 1980)
// and Taylor polynomials on [-pi/4, pi/4] (truncation < 2e-9).
__device__ __forceinline__ void sincos_cw(float x, float* s, float* c) {
  const float q = rintf(x * 0.63661977236758134f);
  float r = fmaf(-q, 1.5703125f, x);
  r = fmaf(-q, 4.837512969970703125e-4f, r);
  r = fmaf(-q, 7.54978995489188216e-8f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.7557319224e-6f, -1.9841269841e-4f);
  ps = fmaf(r2, ps, 8.3333333333e-3f);
  ps = fmaf(r2, ps, -1.6666666667e-1f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.7557319224e-7f, 2.4801587302e-5f);
  pc = fmaf(r2, pc, -1.3888888889e-3f);
  pc = fmaf(r2, pc, 4.1666666667e-2f);
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  const int iq = (int)q;
  const float ss = (iq & 1) ? cr : sr;
  const float cc = (iq & 1) ? sr : cr;
  *s = (iq & 2) ? -ss : ss;
  *c = ((iq + 1) & 2) ? -cc : cc;
}

__device__ __forceinline__ bool in_row(int x, int nx) { return (unsigned)x < (unsigned)nx; }

// taper[x] * exp(-i w phi[x]) for the sample this lane holds, 0 outside [0,nx)
__device__ __forceinline__ float2 ss_taper(const DepthTables& dt, const StepGeom& g, int d, int x,
                                           float w) {
  const bool ok = in_row(x, g.nx);
  const int xc = ok ? x : 0;
  const float tp = ok ? ldg(dt.taper, xc) : 0.f;
  if (!g.ssf) return make_float2(tp, 0.f);
  const float2 ph = ldg(dt.phi, (size_t)d * g.nx + xc);
  // exp(-i w (pr + i pi)) = exp(w pi) (cos(w pr) - i sin(w pr))
  float sn, cs;
  sincos_cw(w * ph.x, &sn, &cs);
  const float a = tp * exp2f(w * ph.y * 1.4426950408889634f);
  return make_float2(a * cs, -a * sn);
}

// reference ids of the E samples of this lane, packed 4 per 32-bit register
template <int N>
__device__ __forceinline__ void load_refid(const DepthTables& dt, const StepGeom& g, int d, int t,
                                           unsigned (&rid)[(N / 64 + 3) / 4]) {
  constexpr int E = N / 64;
  const signed char* row = dt.refid + (size_t)d * g.nx;
#pragma unroll
  for (int i = 0; i < (E + 3) / 4; ++i) rid[i] = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = t + 64 * e - g.tap;
    const bool ok = in_row(x, g.nx);
    const int v = ok ? (int)ldg(row, ok ? x : 0) : -1;
    rid[e / 4] |= (unsigned)(v & 0xff) << (8 * (e % 4));
  }
}

__device__ __forceinline__ int rid_at(const unsigned* rid, int e) {
  return (int)(signed char)((rid[e / 4] >> (8 * (e % 4))) & 0xff);
}

// out = P_d(in)   (in and out in padded strided layout; out is 0 outside [0,nx))
template <int N>
__device__ __forceinline__ void step_forward(float2 (&st)[N / 64], int d, float w,
                                             const DepthTables& dt, const StepGeom& g,
                                             const float2* __restrict__ ptab, float2* buf,
                                             const float2* tw, int t) {
  constexpr int E = N / 64;
  fft_row<N, false>(st, buf, tw, t);
  unsigned rid[(E + 3) / 4];
  load_refid<N>(dt, g, d, t, rid);
  const int s0 = ldg(dt.slot_begin, d), s1 = ldg(dt.slot_begin, d + 1);
  if (s1 - s0 == 1) {
    // one active reference: transform in place
    const int r = ldg(dt.slot_ref, s0);
    const float2* P = ptab + (size_t)s0 * N;
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmul(st[e], ldg(P, t + 64 * e));
    fft_row<N, true>(st, buf, tw, t);
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (rid_at(rid, e) != r) st[e] = make_float2(0.f, 0.f);
  } else {
    float2 X[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      X[e] = st[e];
      st[e] = make_float2(0.f, 0.f);
    }
    for (int s = s0; s < s1; ++s) {
      const int r = ldg(dt.slot_ref, s);
      const float2* P = ptab + (size_t)s * N;
      float2 Y[E];
#pragma unroll
      for (int e = 0; e < E; ++e) Y[e] = cmul(X[e], ldg(P, t + 64 * e));
      fft_row<N, true>(Y, buf, tw, t);
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (rid_at(rid, e) == r) st[e] = Y[e];
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) st[e] = cmul(st[e], ss_taper(dt, g, d, t + 64 * e - g.tap, w));
}

// out = P_d^H(in)
template <int N>
__device__ __forceinline__ void step_adjoint(float2 (&st)[N / 64], int d, float w,
                                             const DepthTables& dt, const StepGeom& g,
                                             const float2* __restrict__ ptab, float2* buf,
                                             const float2* tw, int t) {
  constexpr int E = N / 64;
  unsigned rid[(E + 3) / 4];
  load_refid<N>(dt, g, d, t, rid);
  // SplitStep adjoint factor exp(+i w conj(phi)) = conj(exp(-i w phi)), taper self-adjoint
#pragma unroll
  for (int e = 0; e < E; ++e) st[e] = cmulc(st[e], ss_taper(dt, g, d, t + 64 * e - g.tap, w));
  const int s0 = ldg(dt.slot_begin, d), s1 = ldg(dt.slot_begin, d + 1);
  if (s1 - s0 == 1) {
    const int r = ldg(dt.slot_ref, s0);
    const float2* P = ptab + (size_t)s0 * N;
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (rid_at(rid, e) != r) st[e] = make_float2(0.f, 0.f);
    fft_row<N, false>(st, buf, tw, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmulc(st[e], ldg(P, t + 64 * e));
  } else {
    float2 acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = make_float2(0.f, 0.f);
    for (int s = s0; s < s1; ++s) {
      const int r = ldg(dt.slot_ref, s);
      const float2* P = ptab + (size_t)s * N;
      float2 Z[E];
#pragma unroll
      for (int e = 0; e < E; ++e) Z[e] = (rid_at(rid, e) == r) ? st[e] : make_float2(0.f, 0.f);
      fft_row<N, false>(Z, buf, tw, t);
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = cadd(acc[e], cmulc(Z[e], ldg(P, t + 64 * e)));
    }
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = acc[e];
  }
  fft_row<N, true>(st, buf, tw, t);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = t + 64 * e - g.tap;
    if (!in_row(x, g.nx)) st[e] = make_float2(0.f, 0.f);
  }
}

}  // namespace wemfwi
