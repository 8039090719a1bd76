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
typedef float wemfwi_f2v __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ T ldg(const T* p, size_t i) {
  return ((const WEMFWI_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ T ldg(const T* p, unsigned i) {
  return ((const WEMFWI_GLOBAL T*)p)[i];
}
__device__ __forceinline__ float2 ldg(const float2* p, size_t i) {
  const wemfwi_f2v v = ((const WEMFWI_GLOBAL wemfwi_f2v*)p)[i];
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 ldg(const float2* p, unsigned i) {
  const wemfwi_f2v v = ((const WEMFWI_GLOBAL wemfwi_f2v*)p)[i];
  return make_float2(v.x, v.y);
}
// Wave-uniform loads through the constant address space: the tables and
// row descriptors are read-only for the whole launch, so these become scalar
// (SMEM) loads into SGPRs instead of per-lane loads into VGPRs.  The index
// must be the same on every lane of the wave.
#define WEMFWI_CONST __attribute__((address_space(4)))
__device__ __forceinline__ int ldu(const int* p, int i) { return ((const WEMFWI_CONST int*)p)[i]; }

// a whole descriptor struct (size a multiple of 4 bytes) by scalar loads
template <class T>
__device__ __forceinline__ T ldu_struct(const T* p, int i) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size must be a multiple of 4 bytes");
  constexpr int n = (int)(sizeof(T) / 4);
  const WEMFWI_CONST unsigned* q = (const WEMFWI_CONST unsigned*)(p + i);
  unsigned w[n];
#pragma unroll
  for (int k = 0; k < n; ++k) w[k] = q[k];
  T v;
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

// index of this wave in its workgroup, as a wave-uniform value
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ void stg(float2* p, unsigned i, float2 v) {
  wemfwi_f2v w;
  w.x = v.x;
  w.y = v.y;
  ((WEMFWI_GLOBAL wemfwi_f2v*)p)[i] = w;
}

// Buffer-resource access: a wave-uniform descriptor (SGPRs) over a range of
// bytes, a per-lane byte offset (VGPR) and a wave-uniform byte offset
// (SGPR / immediate).  Offsets past the range read 0 and drop stores (the
// range check applies to the per-lane offset plus the immediate), which is
// how padding samples of a row are zero.  Word 3 0x00020000: raw buffer,
// 32-bit data format (gfx9 family).
typedef __amdgpu_buffer_rsrc_t brsrc;
// the 64-bit payload type of the b64 buffer builtins (two 32-bit lanes);
// values are moved in and out of it by bit casts only
typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(brsrc(), 0, 0, 0)) wemfwi_b64;

__device__ __forceinline__ brsrc make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
// AUX: cache policy bits of the access; kStream (bit 1, non-temporal) marks
// data read or written once per sweep (wavefield rows), so that it does not
// evict the phase tables and scratch rows that are re-read from L2.
constexpr int kStream = 2;

template <int AUX = 0>
__device__ __forceinline__ float2 bld2(brsrc r, unsigned voff, unsigned soff) {
  const unsigned long long u =
      __builtin_bit_cast(unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
  return make_float2(__builtin_bit_cast(float, (unsigned)u), __builtin_bit_cast(float, (unsigned)(u >> 32)));
}
template <int AUX = 0>
__device__ __forceinline__ void bst2(brsrc r, float2 a, unsigned voff, unsigned soff) {
  const unsigned long long u = ((unsigned long long)__builtin_bit_cast(unsigned, a.y) << 32) |
                               (unsigned long long)__builtin_bit_cast(unsigned, a.x);
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(wemfwi_b64, u), r, (int)voff, (int)soff, AUX);
}
__device__ __forceinline__ unsigned bld1(brsrc r, unsigned voff, unsigned soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}

// Per-model depth tables (device pointers).  The *_lane tables are in the
// padded lane layout of one FFT row: entry t*E + e holds padded sample
// p = t + 64 e (model sample x = p - tap), so each lane reads its E values
// contiguously; padding entries hold refid -1, taper 0 and phi 0.
struct DepthTables {
  const int* slot_begin;        // [nz+1] first active (depth, reference) slot of a depth
  const int* slot_ref;          // [nslots] reference index of each slot
  const signed char* rid_lane;  // [nz][N] reference selecting each sample (-1: none)
  const int* one_ref;           // [nz] 1: a single reference selects every x of the depth
  const float2* phi_lane;       // [nz][N] dz*(S - Sref)/sqrt(Sref) of the sample's reference
  const float* taper_lane;      // [N] taper (Taper.h:11-23)
  const float2* scoef;          // [ntaylor+1][nz][nx] scatterer coefficients C_it
  long long scoef_stride;       // nz*nx
  int nz;                       // depth samples
  int n;                        // FFT length N (lane-table row length)
  float dz;                     // depth step of this model's grid
  float dk;                     // wavenumber step of its scatterer grid
  int nx;                       // propagation width of this model
  int tap;                      // FFT zero padding per side (N = nx + 2 tap)
  int taper;                    // 1: a taper is applied
  int phi_real;                 // 1: every split-step exponent is real (no amplitude term)
  const float* phi_max;         // [nz] largest |Re phi| of the depth's samples
};

struct StepGeom {
  int nx;       // propagation width
  int tap;      // FFT zero padding per side
  int ssf;      // 1: split-step correction on
  int full;     // 1: tap == 0 and nx == N (every lane sample is a model sample)
  int ntaylor;  // Taylor order of the scatterer
  int taper;    // 1: a taper is applied (width > 0)
  FftGen fg;    // FFT length and pass plan when it is not the launch's N
  int pf;       // sweeps: 1 when every row of the launch reads a plain source
                // row at every depth (no src_skip, no smul): full rows then
                // prefetch it by LDS DMA
  int tdma;     // sweeps: 1 when the rows of the launch rarely share a phase
                // table (tables then come from HBM, not L2): the launch runs
                // the kernels whose single-reference depths copy their table
                // row into LDS by DMA
};

// Compile-time variants of the step: FULL (no padding samples), SSF
// (split-step correction), GEN (FFT length g.fg.n below the N lanes of the
// row, transformed by the runtime-planned LDS passes) and K (radix plan of
// the register FFT, FftPlan<N, K>).
template <int N, bool FULL_, bool SSF_, bool GEN_ = false, int K_ = 0>
struct Var {
  static constexpr int n = N;
  static constexpr int E = N / 64;
  static constexpr bool FULL = FULL_;
  static constexpr bool SSF = SSF_;
  static constexpr bool GEN = GEN_;
  static constexpr int K = K_;
  static_assert(!(FULL_ && GEN_), "a row shorter than N has padding samples");
};

// LDS complexes per wave and twiddle-table entries of a variant
template <class V>
struct LdsV {
  static constexpr int value = V::GEN ? 2 * LdsSize<V::n>::value : LdsSize<V::n>::value;
};
template <class V>
__host__ __device__ constexpr int tw_entries() {
  if constexpr (V::GEN) return gen_tw_cap(V::n);
  else return plan_tw_size<V::n, V::K>();
}
template <class V>
struct TwV {
  static constexpr int value = tw_entries<V>();
};

// the row transform of a variant; pre (register-planned lengths only): a
// whole row copied into buf[0..N) by LDS DMA during the last pass
template <class V, bool INV>
__device__ __forceinline__ void fft_v(float2 (&x)[V::E], float2* buf, const float2* tw, int t, const StepGeom& g,
                                      const float2* pre = nullptr, const float2* tab = nullptr) {
  if constexpr (V::GEN && V::K > 0) fft_row_gen_fixed<V::n, INV, V::K>(x, buf, tw, t);
  else if constexpr (V::GEN) fft_row_gen<V::n, INV>(x, buf, tw, t, g.fg);
  else fft_row<V::n, INV, V::K>(x, buf, tw, t, pre, tab);
}

// phase-table entry of lane sample k = t + 64 e from a copy of the table
// (N/2 + 1 entries) in LDS: entry k for k <= N/2, N - k above
template <int N>
__device__ __forceinline__ float2 ptab_lds(const float2* T, int t, int e) {
  if (e < N / 128) return T[t + 64 * e];
  return T[(64 - t) + 64 * (N / 64 - 1 - e)];
}

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

// sincos_cw for |x| <= 0.78: the reduction is the identity there (q = 0),
// so the result is bitwise that of sincos_cw
__device__ __forceinline__ void sincos_small(float x, float* s, float* c) {
  const float r2 = x * x;
  float ps = fmaf(r2, 2.7557319224e-6f, -1.9841269841e-4f);
  ps = fmaf(r2, ps, 8.3333333333e-3f);
  ps = fmaf(r2, ps, -1.6666666667e-1f);
  *s = fmaf(x * r2, ps, x);
  float pc = fmaf(r2, -2.7557319224e-7f, 2.4801587302e-5f);
  pc = fmaf(r2, pc, -1.3888888889e-3f);
  pc = fmaf(r2, pc, 4.1666666667e-2f);
  pc = fmaf(r2, pc, -0.5f);
  *c = fmaf(r2, pc, 1.0f);
}

__device__ __forceinline__ bool in_row(int x, int nx) { return (unsigned)x < (unsigned)nx; }

// Sample x = t + 64 e - tap of a row of nx values starting at p, 0 for
// padding samples.  The load index is x itself or 0 (never negative), so the
// load stays inside the row on every lane.
template <class V>
__device__ __forceinline__ float2 row_sample(const float2* p, const StepGeom& g, int t, int e) {
  const brsrc r = make_rsrc(p, (unsigned)g.nx * 8u);
  if constexpr (V::FULL) return bld2<kStream>(r, (unsigned)t * 8u, 512u * (unsigned)e);
  // x < 0 wraps to a huge offset: out of range like x >= nx, and reads 0
  return bld2<kStream>(r, (unsigned)(t + 64 * e - g.tap) * 8u, 0u);
}

// stores sample x = t + 64 e - tap into a row of nx values (padding samples
// fall outside the range and are dropped)
template <class V>
__device__ __forceinline__ void row_store(float2* p, float2 v, const StepGeom& g, int t, int e) {
  const brsrc r = make_rsrc(p, (unsigned)g.nx * 8u);
  if constexpr (V::FULL) bst2<kStream>(r, v, (unsigned)t * 8u, 512u * (unsigned)e);
  else bst2<kStream>(r, v, (unsigned)(t + 64 * e - g.tap) * 8u, 0u);
}

// taper * exp(-i w phi) of the E samples of this lane (0 on padding samples
// when a taper or the split step is on; 1 otherwise).  pre (or NULL): a whole
// row copied into buf[0..N) by LDS DMA, started right after this depth's
// table loads have been issued (the memory counter is in order: a DMA issued
// first would hold up the table loads behind its HBM round trip).
template <class V>
__device__ __forceinline__ void ss_taper_row(float2 (&f)[V::E], const DepthTables& dt, const StepGeom& g, int d,
                                             float w, int t, const float2* pre = nullptr, float2* buf = nullptr) {
  constexpr int E = V::E;
  float tv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) tv[e] = 1.f;
  if (g.taper) {
    const brsrc r = make_rsrc(dt.taper_lane, (unsigned)V::n * 4u);
#pragma unroll
    for (int e = 0; e < E; ++e) tv[e] = __builtin_bit_cast(float, bld1(r, (unsigned)(t * E + e) * 4u, 0u));
  }
  auto start_dma = [&]() {
    if constexpr (!V::GEN) {
      if (pre) {
        __builtin_amdgcn_sched_barrier(0);
        row_to_lds<V::n>(pre, buf, t);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };
  if constexpr (V::SSF) {
    const brsrc r = make_rsrc(dt.phi_lane + (size_t)d * V::n, (unsigned)V::n * 8u);
    // exp(-i w (pr + i pi)) = exp(w pi) (cos(w pr) - i sin(w pr)).  Two
    // shortcuts, both bitwise the general form and uniform over the wave:
    // w max_x |pr| <= 0.78 at this depth (no argument reduction: smooth
    // models at seismic frequencies), and a real exponent without taper
    // (amplitude 1)
    const bool small = w * __builtin_bit_cast(float, ldu((const int*)dt.phi_max, d)) <= 0.78f;
    if (!g.taper && dt.phi_real) {
      // the real parts only
      float px[E];
#pragma unroll
      for (int e = 0; e < E; ++e) px[e] = __builtin_bit_cast(float, bld1(r, (unsigned)(t * E + e) * 8u, 0u));
      start_dma();
      if (small) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          float sn, cs;
          sincos_small(w * px[e], &sn, &cs);
          f[e] = make_float2(cs, -sn);
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          float sn, cs;
          sincos_cw(w * px[e], &sn, &cs);
          f[e] = make_float2(cs, -sn);
        }
      }
    } else {
      float2 pv[E];
#pragma unroll
      for (int e = 0; e < E; ++e) pv[e] = bld2(r, (unsigned)(t * E + e) * 8u, 0u);
      start_dma();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        float sn, cs;
        if (small) sincos_small(w * pv[e].x, &sn, &cs);
        else sincos_cw(w * pv[e].x, &sn, &cs);
        const float a = tv[e] * exp2f(w * pv[e].y * 1.4426950408889634f);
        f[e] = make_float2(a * cs, -a * sn);
      }
    }
  } else {
    start_dma();
#pragma unroll
    for (int e = 0; e < E; ++e) f[e] = make_float2(tv[e], 0.f);
  }
}

// reference ids of the E samples of this lane, packed 4 per 32-bit register
template <class V>
__device__ __forceinline__ void load_refid(const DepthTables& dt, int d, int t, unsigned (&rid)[(V::E + 3) / 4]) {
  constexpr int E = V::E;
  const signed char* row = dt.rid_lane + (size_t)d * V::n + t * E;
  if constexpr (E >= 4) {
    const brsrc r = make_rsrc(dt.rid_lane + (size_t)d * V::n, (unsigned)V::n);
#pragma unroll
    for (int i = 0; i < E / 4; ++i) rid[i] = bld1(r, (unsigned)(t * E), 4u * (unsigned)i);
  } else {
    rid[0] = 0xffffffffu;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const unsigned v = (unsigned)(unsigned char)ldg(row, (unsigned)e);
      rid[0] = (rid[0] & ~(0xffu << (8 * e))) | (v << (8 * e));
    }
  }
}

__device__ __forceinline__ int rid_at(const unsigned* rid, int e) {
  return (int)(signed char)((rid[e / 4] >> (8 * (e % 4))) & 0xff);
}

// phase-table entry of wavenumber index k = t + 64 e (the table holds
// k <= N/2; entry N-k for the upper half)
template <int N>
__device__ __forceinline__ float2 ptab_at(const float2* P, int t, int e) {
  const brsrc r = make_rsrc(P, (unsigned)(N / 2 + 1) * 8u);
  if (e < N / 128) return bld2(r, (unsigned)t * 8u, 512u * (unsigned)e);
  // N - t - 64 e = (64 - t) + 64 (E - 1 - e), in 1..N/2
  return bld2(r, (unsigned)(64 - t) * 8u, 512u * (unsigned)(N / 64 - 1 - e));
}

// phase-table entry of lane sample k = t + 64 e of a variant: entry k for
// k <= n/2, n - k above (the k grid is symmetric); lanes k >= n of a GEN row
// read entry 0 (their spectrum samples are never used)
template <class V>
__device__ __forceinline__ float2 ptab_v(const float2* P, int t, int e, const StepGeom& g) {
  if constexpr (!V::GEN) {
    return ptab_at<V::n>(P, t, e);
  } else {
    const int n = g.fg.n, k = t + 64 * e;
    const int kh = k <= n / 2 ? k : (k < n ? n - k : 0);
    const brsrc r = make_rsrc(P, (unsigned)(n / 2 + 1) * 8u);
    return bld2(r, (unsigned)kh * 8u, 0u);
  }
}

// row of samples x = t + 64 e - tap (0 on padding samples)
template <class V>
__device__ __forceinline__ void load_samples(float2 (&v)[V::E], const float2* p, const StepGeom& g, int t) {
#pragma unroll
  for (int e = 0; e < V::E; ++e) v[e] = row_sample<V>(p, g, t, e);
}

// out = P_d(in)   (in and out in padded strided layout; out is 0 outside
// [0,nx)).  Loads are issued where their values are first needed: holding
// them across an FFT costs more registers (and waves per SIMD) than their
// latency costs.
template <class V, bool TD>
__device__ __forceinline__ bool step_forward_scr(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                                 const StepGeom& g, const float2* __restrict__ ptab, float2* buf,
                                                 const float2* tw, int t, float2* scr, const float2* pre);

//
// SCR: one code path for every depth (step_forward_scr): at depths with
// several active references the spectrum is parked in the wave's scratch row
// (global memory, written and re-read by the same lane) instead of registers,
// which keeps the step within 128 VGPRs.
// pre (SCR only; FULL rows of register-planned lengths): a whole row copied
// into buf[0..N) by LDS DMA during the step's last inverse transform;
// returns whether that copy was started (buf then holds it once it arrives).
// TD (SCR only): single-reference depths copy their phase-table row into
// buf by LDS DMA during the forward transform (launches whose rows rarely
// share a table).
template <class V, bool SCR = false, bool TD = false>
__device__ __forceinline__ bool step_forward(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                             const StepGeom& g, const float2* __restrict__ ptab, float2* buf,
                                             const float2* tw, int t, float2* scr = nullptr,
                                             const float2* pre = nullptr) {
  if constexpr (SCR) {
    return step_forward_scr<V, TD>(st, d, w, dt, g, ptab, buf, tw, t, scr, pre);
  }
  constexpr int N = V::n;
  constexpr int E = V::E;
  fft_v<V, false>(st, buf, tw, t, g);
  const int s0 = ldu(dt.slot_begin, d), s1 = ldu(dt.slot_begin, d + 1);
  if (s1 - s0 == 1) {
    // one active reference: transform in place
    const int r = ldu(dt.slot_ref, s0);
    const float2* P = ptab + (size_t)s0 * (N / 2 + 1);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmul(st[e], ptab_v<V>(P, t, e, g));
    fft_v<V, true>(st, buf, tw, t, g);
    if (!(V::FULL && ldu(dt.one_ref, d))) {
      unsigned rid[(E + 3) / 4];
      load_refid<V>(dt, d, t, rid);
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (rid_at(rid, e) != r) st[e] = make_float2(0.f, 0.f);
    }
  } else {
    unsigned rid[(E + 3) / 4];
    load_refid<V>(dt, d, t, rid);
    float2 X[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      X[e] = st[e];
      st[e] = make_float2(0.f, 0.f);
    }
#pragma unroll 1
    for (int s = s0; s < s1; ++s) {
      const int r = ldu(dt.slot_ref, s);
      const float2* P = ptab + (size_t)s * (N / 2 + 1);
      float2 Y[E];
#pragma unroll
      for (int e = 0; e < E; ++e) Y[e] = cmul(X[e], ptab_v<V>(P, t, e, g));
      fft_v<V, true>(Y, buf, tw, t, g);
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (rid_at(rid, e) == r) st[e] = Y[e];
    }
  }
  if (V::SSF || g.taper) {
    float2 f[E];
    ss_taper_row<V>(f, dt, g, d, w, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmul(st[e], f[e]);
  }
  return false;
}

// out = P_d^H(in); SCR as in step_forward (the split-step-corrected input row
// is parked in the scratch row at depths with several references)
template <class V, bool TD>
__device__ __forceinline__ bool step_adjoint_scr(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                                 const StepGeom& g, const float2* __restrict__ ptab, float2* buf,
                                                 const float2* tw, int t, float2* scr, const float2* pre);

template <class V, bool SCR = false, bool TD = false>
__device__ __forceinline__ bool step_adjoint(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                             const StepGeom& g, const float2* __restrict__ ptab, float2* buf,
                                             const float2* tw, int t, float2* scr = nullptr,
                                             const float2* pre = nullptr) {
  if constexpr (SCR) {
    return step_adjoint_scr<V, TD>(st, d, w, dt, g, ptab, buf, tw, t, scr, pre);
  }
  constexpr int N = V::n;
  constexpr int E = V::E;
  const int s0 = ldu(dt.slot_begin, d), s1 = ldu(dt.slot_begin, d + 1);
  // SplitStep adjoint factor exp(+i w conj(phi)) = conj(exp(-i w phi)), taper self-adjoint
  if (V::SSF || g.taper) {
    float2 f[E];
    ss_taper_row<V>(f, dt, g, d, w, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmulc(st[e], f[e]);
  }
  if (s1 - s0 == 1) {
    const int r = ldu(dt.slot_ref, s0);
    const float2* P = ptab + (size_t)s0 * (N / 2 + 1);
    if (!(V::FULL && ldu(dt.one_ref, d))) {
      unsigned rid[(E + 3) / 4];
      load_refid<V>(dt, d, t, rid);
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (rid_at(rid, e) != r) st[e] = make_float2(0.f, 0.f);
    }
    fft_v<V, false>(st, buf, tw, t, g);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmulc(st[e], ptab_v<V>(P, t, e, g));
  } else {
    unsigned rid[(E + 3) / 4];
    load_refid<V>(dt, d, t, rid);
    float2 acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = make_float2(0.f, 0.f);
#pragma unroll 1
    for (int s = s0; s < s1; ++s) {
      const int r = ldu(dt.slot_ref, s);
      const float2* P = ptab + (size_t)s * (N / 2 + 1);
      float2 Z[E];
#pragma unroll
      for (int e = 0; e < E; ++e) Z[e] = (rid_at(rid, e) == r) ? st[e] : make_float2(0.f, 0.f);
      fft_v<V, false>(Z, buf, tw, t, g);
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = cadd(acc[e], cmulc(Z[e], ptab_v<V>(P, t, e, g)));
    }
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = acc[e];
  }
  fft_v<V, true>(st, buf, tw, t, g);
  if constexpr (!V::FULL) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (!in_row(t + 64 * e - g.tap, g.nx)) st[e] = make_float2(0.f, 0.f);
  }
  return false;
}

// One loop over the active references of the depth for both the single-
// and the multi-reference case (two separate paths in one sweep loop cost the
// register allocator more than either path alone).
//   forward: X = F in; for each reference r: Y = F^H (P_r X), out[x in r] = Y[x]
template <class V, bool TD>
__device__ __forceinline__ bool step_forward_scr(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                                 const StepGeom& g, const float2* __restrict__ ptab, float2* buf,
                                                 const float2* tw, int t, float2* scr, const float2* pre) {
  constexpr int N = V::n;
  constexpr int E = V::E;
  if constexpr (V::GEN) pre = nullptr;  // the runtime-planned transform has no DMA hook
  const int s0 = ldu(dt.slot_begin, d), s1 = ldu(dt.slot_begin, d + 1);
  if (s0 == s1) {
    // no reference selects any x of this depth: Selector leaves the row zero
    // (RefSampler's stale buffer can leave a depth without a live reference)
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = make_float2(0.f, 0.f);
    return false;
  }
  const bool multi = s1 - s0 > 1;
  // one reference: its phase-table row is copied into the transform buffer
  // by LDS DMA during the forward transform's last pass (read from there,
  // the N/2 + 1 loads do not queue one after another behind the register
  // pressure of the step)
  const float2* tab = (TD && !V::GEN && !multi) ? ptab + (size_t)s0 * (N / 2 + 1) : nullptr;
  fft_v<V, false>(st, buf, tw, t, g, nullptr, tab);
  const bool all = V::FULL && ldu(dt.one_ref, d);
  const brsrc xs = make_rsrc(scr, (unsigned)N * 8u);  // the wave's scratch row, lane index t + 64 e
  if (multi) {
#pragma unroll
    for (int e = 0; e < E; ++e) bst2(xs, st[e], (unsigned)t * 8u, 512u * (unsigned)e);
  }
  unsigned rid[(E + 3) / 4];
  if (!all) load_refid<V>(dt, d, t, rid);
#pragma unroll 1
  for (int s = s0; s < s1; ++s) {
    const int r = ldu(dt.slot_ref, s);
    const float2* P = ptab + (size_t)s * (N / 2 + 1);
    float2 Y[E];
    // in two halves, so that only half of the spectrum / phase loads are in
    // flight next to the merged output
    if (multi) {
#pragma unroll
      for (int e = 0; e < E / 2; ++e) Y[e] = cmul(bld2(xs, (unsigned)t * 8u, 512u * (unsigned)e), ptab_v<V>(P, t, e, g));
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = E / 2; e < E; ++e) Y[e] = cmul(bld2(xs, (unsigned)t * 8u, 512u * (unsigned)e), ptab_v<V>(P, t, e, g));
    } else if (TD && tab) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the table copy has arrived
#pragma unroll
      for (int e = 0; e < E; ++e) Y[e] = cmul(st[e], ptab_lds<N>(buf, t, e));
    } else {
#pragma unroll
      for (int e = 0; e < E / 2; ++e) Y[e] = cmul(st[e], ptab_v<V>(P, t, e, g));
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = E / 2; e < E; ++e) Y[e] = cmul(st[e], ptab_v<V>(P, t, e, g));
    }
    // the last inverse transform of the depth starts the DMA of pre, unless
    // the split step / taper follows (it starts the DMA after its table loads)
    fft_v<V, true>(Y, buf, tw, t, g, (s == s1 - 1 && !(V::SSF || g.taper)) ? pre : nullptr);
    if (all) {
#pragma unroll
      for (int e = 0; e < E; ++e) st[e] = Y[e];
    } else {
      const bool first = s == s0;
#pragma unroll
      for (int e = 0; e < E; ++e)
        st[e] = (rid_at(rid, e) == r) ? Y[e] : (first ? make_float2(0.f, 0.f) : st[e]);
    }
  }
  if (V::SSF || g.taper) {
    float2 f[E];
    ss_taper_row<V>(f, dt, g, d, w, t, pre, buf);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmul(st[e], f[e]);
  }
  return pre != nullptr;
}

//   adjoint: u = conj(f) in; acc = sum_r conj(P_r) F (u restricted to r); out = F^H acc
template <class V, bool TD>
__device__ __forceinline__ bool step_adjoint_scr(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                                 const StepGeom& g, const float2* __restrict__ ptab, float2* buf,
                                                 const float2* tw, int t, float2* scr, const float2* pre) {
  constexpr int N = V::n;
  constexpr int E = V::E;
  if constexpr (V::GEN) pre = nullptr;  // the runtime-planned transform has no DMA hook
  const int s0 = ldu(dt.slot_begin, d), s1 = ldu(dt.slot_begin, d + 1);
  if (s0 == s1) {  // the adjoint of a zero row
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = make_float2(0.f, 0.f);
    return false;
  }
  if (V::SSF || g.taper) {
    float2 f[E];
    ss_taper_row<V>(f, dt, g, d, w, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmulc(st[e], f[e]);
  }
  const bool multi = s1 - s0 > 1;
  const bool all = V::FULL && ldu(dt.one_ref, d);
  const brsrc xs = make_rsrc(scr, (unsigned)N * 8u);  // the wave's scratch row, lane index t + 64 e
  if (multi) {
#pragma unroll
    for (int e = 0; e < E; ++e) bst2(xs, st[e], (unsigned)t * 8u, 512u * (unsigned)e);
  }
  unsigned rid[(E + 3) / 4];
  if (!all) load_refid<V>(dt, d, t, rid);
#pragma unroll 1
  for (int s = s0; s < s1; ++s) {
    const int r = ldu(dt.slot_ref, s);
    const float2* P = ptab + (size_t)s * (N / 2 + 1);
    float2 Z[E];
    // several references: each reads back only the parked samples it
    // selects (each sample read once over the references)
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (multi) {
        Z[e] = make_float2(0.f, 0.f);
        if (rid_at(rid, e) == r) Z[e] = bld2(xs, (unsigned)t * 8u, 512u * (unsigned)e);
      } else {
        Z[e] = (all || rid_at(rid, e) == r) ? st[e] : make_float2(0.f, 0.f);
      }
    }
    // one reference: its phase-table row arrives in the transform buffer by
    // LDS DMA during the transform's last pass (as in the forward step)
    const float2* tab = (TD && !V::GEN && !multi) ? P : nullptr;
    fft_v<V, false>(Z, buf, tw, t, g, nullptr, tab);
    const bool first = s == s0;
    if (TD && tab) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the table copy has arrived
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float2 z = cmulc(Z[e], ptab_lds<N>(buf, t, e));
        st[e] = first ? z : cadd(st[e], z);
      }
    } else {
#pragma unroll
      for (int e = 0; e < E / 2; ++e) {
        const float2 z = cmulc(Z[e], ptab_v<V>(P, t, e, g));
        st[e] = first ? z : cadd(st[e], z);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = E / 2; e < E; ++e) {
        const float2 z = cmulc(Z[e], ptab_v<V>(P, t, e, g));
        st[e] = first ? z : cadd(st[e], z);
      }
    }
  }
  fft_v<V, true>(st, buf, tw, t, g, pre);
  if constexpr (!V::FULL) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (!in_row(t + 64 * e - g.tap, g.nx)) st[e] = make_float2(0.f, 0.f);
  }
  return pre != nullptr;
}

// scatterer coefficient C_it of this lane's sample e (0 on padding samples);
// C is the start of the coefficient row
template <class V>
__device__ __forceinline__ float2 scoef_at(const float2* C, const StepGeom& g, int t, int e) {
  return row_sample<V>(C, g, t, e);
}

// Scatterer (Scatter.cpp:35-111) on one row in padded strided layout:
//   forward:        y = (-i w dz) sum_it C_it * (M_it x)
//   adjoint compat: y = (+i w dz) sum_it conj(C_it) * (M_it d)   (Scatter.cpp:52-66)
//   adjoint exact:  y = (+i w dz) sum_it M_it (conj(C_it) * d)
// with C_it = coef[it] (S - eps Re S)^(-1/2-it) and M_it = Pad^H F^H (k/w)^(2 it)
// F Pad (M_0 = identity on the row).
template <class V, bool ADJ, bool EXACT>
__device__ __forceinline__ void scatter_row(float2 (&st)[V::E], int d, float w, const DepthTables& dt,
                                            const StepGeom& g, float2* buf, const float2* tw, int t) {
  constexpr int N = V::n;
  constexpr int E = V::E;
  const float2* C0 = dt.scoef + (size_t)d * g.nx;
  float2 acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 c = scoef_at<V>(C0, g, t, e);
    acc[e] = ADJ ? cmulc(st[e], c) : cmul(st[e], c);
  }
  const float iw2 = 1.f / (w * w);
  const int n = V::GEN ? g.fg.n : N;  // FFT length
  const float inv_n = 1.f / (float)n;
  for (int it = 1; it <= g.ntaylor; ++it) {
    const float2* Ci = C0 + (size_t)it * dt.scoef_stride;
    float2 Y[E];
#pragma unroll
    for (int e = 0; e < E; ++e) Y[e] = (ADJ && EXACT) ? cmulc(st[e], scoef_at<V>(Ci, g, t, e)) : st[e];
    fft_v<V, false>(Y, buf, tw, t, g);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      // (k/w)^(2 it) / n on the grid of Scatter.cpp:24-31 (lanes k >= n of
      // a GEN row take k = 0; their spectrum samples are never used)
      const int k = t + 64 * e;
      const int kh = k < n / 2 ? k : ((k < n && (n - k) < n / 2) ? n - k : 0);
      const float kk = (float)kh * dt.dk;
      const float q = kk * kk * iw2;
      float qi = q;
      for (int j = 1; j < it; ++j) qi *= q;
      Y[e] = cscale(Y[e], qi * inv_n);
    }
    fft_v<V, true>(Y, buf, tw, t, g);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (ADJ && EXACT) {
        if (V::FULL || in_row(t + 64 * e - g.tap, g.nx)) acc[e] = cadd(acc[e], Y[e]);
      } else {
        const float2 c = scoef_at<V>(Ci, g, t, e);
        acc[e] = cadd(acc[e], ADJ ? cmulc(Y[e], c) : cmul(Y[e], c));
      }
    }
  }
  const float s = w * dt.dz;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    // forward: (-i s) acc ; adjoint: (+i s) acc
    st[e] = ADJ ? make_float2(-acc[e].y * s, acc[e].x * s) : make_float2(acc[e].y * s, -acc[e].x * s);
  }
}

}  // namespace wemfwi
