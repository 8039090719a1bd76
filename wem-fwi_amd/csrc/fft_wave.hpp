// Wavefront-resident Stockham FFT for CDNA4 (gfx950).
//
// One 64-lane wavefront owns one row of N = 64*E complex64 samples.  Lane t
// keeps samples t + 64*e (e = 0..E-1) in registers ("strided natural
// layout").  A transform is a chain of radix-R Stockham passes
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC 2008): pass p with sub-transform length NS reads
// butterfly j's inputs at j + r*N/R and writes its outputs at
// (j/NS)*NS*R + j%NS + r*NS.  Every radix divides E, so the first pass reads
// and the last pass writes the strided natural layout directly and only the
// interior passes exchange data through LDS.  LDS addresses are padded by one
// complex every 16 to keep the scattered first-pass writes bank-conflict
// free.  All LDS addresses of a pass are one per-lane base register plus
// compile-time offsets (ds_read/ds_write immediate offsets).
//
// Twiddles: a per-pass table in LDS, r-major, entry (r-1)*NS + j%NS =
// exp(-2*pi*i*(j%NS)*r/(NS*R)), written once per launch from a
// double-precision table.
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

// a * w (forward) or a * conj(w) (inverse) with w = hi + lo carried as a
// float pair per part (one rounding at the end of each chain)
template <bool INV>
__device__ __forceinline__ float2 cmul_tw_pair(float2 a, float2 hi, float2 lo) {
  const float wi = INV ? -hi.y : hi.y, li = INV ? -lo.y : lo.y;
  return make_float2(fmaf(a.x, hi.x, fmaf(-a.y, wi, fmaf(a.x, lo.x, -a.y * li))),
                     fmaf(a.x, wi, fmaf(a.y, hi.x, fmaf(a.x, li, a.y * lo.x))));
}

// forward: multiply by -i ; inverse: multiply by +i
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  if constexpr (INV) return make_float2(-a.y, a.x);
  else return make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- rounding
//
// Tables and constants applied again at every depth step: the wavefield
// passes through each FFT twiddle and butterfly constant hundreds of times,
// so an error of the constant's modulus compounds coherently instead of
// averaging out.  Rounding cos and sin separately to the nearest float
// leaves a modulus error of up to 2^-24 whose mean over a plan's table is not
// zero (all three constants of the radix-16 butterfly, cos(pi/8), sin(pi/8)
// and sqrt(1/2), round down: about -5e-8 per depth step in rows of 1024).

// the float below (dir < 0) or above (dir > 0) the finite nonzero f
__host__ __device__ inline float float_step(float f, int dir) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  const bool up = (dir > 0) == (f > 0.f);  // away from zero
  u = up ? u + 1u : u - 1u;
  return __builtin_bit_cast(float, u);
}

// the two floats that bracket x (equal when x is a float)
__host__ __device__ inline void float_bracket(double x, float& lo, float& hi) {
  const float f = (float)x;
  if ((double)f == x || f == 0.f) {
    lo = hi = f;
  } else if ((double)f > x) {
    hi = f;
    lo = float_step(f, -1);
  } else {
    lo = f;
    hi = float_step(f, +1);
  }
}

// The complex value z = re + i im rounded to the one of the (up to) four
// bracketing float pairs whose ratio to z is nearest 1 in |log modulus| +
// |phase|, to first order |Re d| + |Im d| with d = zf / z - 1; ties go to
// the per-part nearest.  Modulus and phase errors then stay below 2^-25
// relative and of either sign, so the per-step gain of a chain of such
// factors averages out over a table.
__host__ __device__ inline float2 round_complex(double re, double im) {
  float rl, rh, il, ih;
  float_bracket(re, rl, rh);
  float_bracket(im, il, ih);
  float2 best = make_float2((float)re, (float)im);
  // (zf - z) conj(z) = |z|^2 d: the common factor |z|^2 does not change the
  // order of the candidates
  auto cost = [&](float a, float b) {
    const double dr = (double)a - re, di = (double)b - im;
    const double pr = dr * re + di * im, pi = di * re - dr * im;
    return fabs(pr) + fabs(pi);
  };
  double bc = cost(best.x, best.y);
  const float rs[2] = {rl, rh}, is[2] = {il, ih};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const double c = cost(rs[a], is[b]);
      if (c < bc) {
        bc = c;
        best = make_float2(rs[a], is[b]);
      }
    }
  return best;
}

// A real constant carried as an unevaluated float pair hi + lo (hi the
// nearest float, lo the nearest float of the remainder): x * (hi + lo) =
// fma(x, hi, x * lo) rounds once and carries the constant to ~2^-48.
struct Konst {
  float hi, lo;
};
__host__ __device__ constexpr Konst konst(double c) { return Konst{(float)c, (float)(c - (double)(float)c)}; }
__device__ __forceinline__ float kmul(float x, Konst k) { return fmaf(x, k.hi, x * k.lo); }
// x * a + y * b with both constants as pairs, one rounding at the end of
// each chain
__device__ __forceinline__ float kdot(float x, Konst a, float y, Konst b) {
  return fmaf(x, a.hi, fmaf(y, b.hi, fmaf(x, a.lo, y * b.lo)));
}

// padded LDS index (one pad slot every 16 complex): the layout of the
// runtime-planned passes and of the power-of-two passes unless lay_m says
// otherwise
__host__ __device__ constexpr int lidx(int i) { return i + (i >> 4); }

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

// cos(pi m / 8) and sin(pi m / 8), m = 0..15
constexpr double kCos16[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                               0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
constexpr double kSin16[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                               1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// multiply by W16^m = exp(-2*pi*i*m/16) (forward) or its conjugate
// (inverse); the constants are float pairs (Konst)
template <bool INV, int M>
__device__ __forceinline__ float2 tw16(float2 a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) return a;
  else if constexpr (m == 4) return mul_mi<INV>(a);
  else if constexpr (m == 8) return make_float2(-a.x, -a.y);
  else if constexpr (m == 12) return mul_mi<!INV>(a);
  else if constexpr ((m & 3) == 2) {
    // |cos| = |sin| = sqrt(1/2): w = sqrt(1/2) (sc + i si), sc, si = +-1
    constexpr Konst h = konst(0.70710678118654752440);
    constexpr float sc = kCos16[m] > 0 ? 1.f : -1.f;
    constexpr float si = (kSin16[m] > 0) == INV ? 1.f : -1.f;  // forward: exp(-i theta)
    return make_float2(kmul(sc * a.x - si * a.y, h), kmul(si * a.x + sc * a.y, h));
  } else {
    constexpr Konst wr = konst(kCos16[m]);
    constexpr Konst wi = konst(INV ? kSin16[m] : -kSin16[m]);
    constexpr Konst nwi = Konst{-wi.hi, -wi.lo};
    return make_float2(kdot(a.x, wr, a.y, nwi), kdot(a.x, wi, a.y, wr));
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

// ---------------------------------------------------------------- plans

// Radix plans (every radix divides E = N/64).  K numbers the alternative
// plans of a length that the launch autotuner times (K = 0: the default).
template <int N, int K = 0>
struct FftPlan;
template <>
struct FftPlan<128, 0> {
  static constexpr int P = 7;
  static constexpr int R[P] = {2, 2, 2, 2, 2, 2, 2};
};
template <>
struct FftPlan<256, 0> {
  static constexpr int P = 4;
  static constexpr int R[P] = {4, 4, 4, 4};
};
template <>
struct FftPlan<512, 0> {
  static constexpr int P = 3;
  static constexpr int R[P] = {8, 8, 8};
};
template <>
struct FftPlan<1024, 0> {
  static constexpr int P = 3;
  static constexpr int R[P] = {16, 4, 16};
};
template <>
struct FftPlan<1024, 1> {
  static constexpr int P = 3;
  static constexpr int R[P] = {16, 16, 4};
};
template <>
struct FftPlan<1024, 2> {
  static constexpr int P = 3;
  static constexpr int R[P] = {4, 16, 16};
};
template <>
struct FftPlan<2048, 0> {
  static constexpr int P = 3;
  static constexpr int R[P] = {16, 8, 16};
};
template <>
struct FftPlan<2048, 1> {
  static constexpr int P = 3;
  static constexpr int R[P] = {16, 16, 8};
};
// number of plans of a length (dropped after measurement: 1024 as
// {8, 8, 16}, whose NS = 8 pass needs more LDS than a 1024 row has for a
// conflict-free layout, and 2048 as {8, 16, 16}, which spills registers)
template <int N>
struct FftPlans {
  static constexpr int value = N == 1024 ? 3 : (N == 2048 ? 2 : 1);
};

template <int N, int K = 0>
__host__ __device__ constexpr int plan_ns(int p) {
  int ns = 1;
  for (int q = 0; q < p; ++q) ns *= FftPlan<N, K>::R[q];
  return ns;
}
// Pass p's twiddles are multiples of 2 pi / 16 (NS R <= 16): the few
// values cos, sin (pi m / 8) whose float roundings all shrink the modulus
// (by up to 2^-24) and that carry half of the row or more, so their bias
// adds up over the depth steps.  Such passes hold each twiddle as a float
// pair, hi then lo (Konst), and multiply with the pair; other passes use
// round_complex values.
template <int N, int K = 0>
__host__ __device__ constexpr bool plan_tw_pair(int p) {
  return plan_ns<N, K>(p) > 1 && plan_ns<N, K>(p) * FftPlan<N, K>::R[p] <= 16;
}
// offset of pass p's twiddles in the concatenated table
template <int N, int K = 0>
__host__ __device__ constexpr int plan_tw_off(int p) {
  int off = 0;
  for (int q = 0; q < p; ++q)
    off += (FftPlan<N, K>::R[q] - 1) * plan_ns<N, K>(q) * (plan_tw_pair<N, K>(q) ? 2 : 1);
  return off;
}
template <int N, int K = 0>
__host__ __device__ constexpr int plan_tw_size() {
  return plan_tw_off<N, K>(FftPlan<N, K>::P);
}

// LDS layout between pass p and pass p + 1: element i at
// i + (i >> lay_a) * lay_m.  LDS serves a 64-bit access 16 lanes at a time
// over 32 banks, so 16 consecutive lanes must hit 16 distinct bank pairs.
// The next pass reads 16 consecutive elements per 16 lanes (free of
// conflicts under any such layout).  A pass with NS < 16 writes, per 16
// lanes, 16/NS runs of NS consecutive elements NS R apart: the runs get
// the offsets 0, NS, 2 NS, ... modulo 16 from m = NS with a = log2(NS R)
// (a = 4 when NS R < 16).  Other passes (NS = 1: one run per lane R apart;
// NS >= 16: a run per 16 lanes) are conflict-free with one pad per 16.
template <int N, int K = 0>
__host__ __device__ constexpr int lay_m(int p) {
  const int ns = plan_ns<N, K>(p);
  return ns > 1 && ns < 16 ? ns : 1;
}
template <int N, int K = 0>
__host__ __device__ constexpr int lay_a(int p) {
  const int ns = plan_ns<N, K>(p), w = ns * FftPlan<N, K>::R[p];
  if (!(ns > 1 && ns < 16) || w % 16 != 0) return 4;
  int a = 0;
  while ((1 << a) < w) ++a;
  return a;
}
template <int N, int K, int p>
__host__ __device__ constexpr int lay(int i) {
  return i + (i >> lay_a<N, K>(p)) * lay_m<N, K>(p);
}
// LDS complexes per row: the largest layout over the plans of a length
template <int N, int K = 0>
__host__ __device__ constexpr int plan_lds_size() {
  int m = N + N / 16;
  for (int p = 0; p + 1 < FftPlan<N, K>::P; ++p) {
    const int v = N - 1 + ((N - 1) >> lay_a<N, K>(p)) * lay_m<N, K>(p) + 1;
    m = v > m ? v : m;
  }
  return m;
}
template <int N, int K = 0>
__host__ __device__ constexpr int lds_size_k() {
  if constexpr (K + 1 < FftPlans<N>::value) {
    const int a = plan_lds_size<N, K>(), b = lds_size_k<N, K + 1>();
    return a > b ? a : b;
  } else {
    return plan_lds_size<N, K>();
  }
}
template <int N>
struct LdsSize {
  static constexpr int value = lds_size_k<N>();
};
// reads of pass p (layout of boundary p - 1) split as lane base + constant:
// lay(t + c) == lay(t) + lay(c) for every lane t < 64 and offset c of the pass
template <int N, int K, int p>
__host__ __device__ constexpr bool lay_reads_split() {
  constexpr int R = FftPlan<N, K>::R[p], NB = N / R, BPL = N / 64 / R;
  for (int b = 0; b < BPL; ++b)
    for (int r = 0; r < R; ++r)
      for (int t = 0; t < 64; ++t) {
        const int c = 64 * b + r * NB;
        if (lay<N, K, p - 1>(t + c) != lay<N, K, p - 1>(t) + lay<N, K, p - 1>(c)) return false;
      }
  return true;
}

// ---------------------------------------------------------------- passes

// Copies the N samples of a whole row from global memory into buf[0..N) in
// natural order by LDS DMA (buffer_load_dwordx4 ... lds: 64 lanes x 16 bytes
// per instruction; one lane offset t * 16 and a scalar offset per
// instruction, no VGPRs held for the data).  Marked non-temporal: each row
// is read once per sweep.  The wave must not access buf[0..N) until the
// copy has arrived (the compiler waits for it, vmcnt, before the next LDS
// read of buf).
// N samples (a row) or, with M entries, a table of M complex values (the
// range check of the buffer descriptor leaves the rest of the last 1 KB
// block zero).  AUX: cache policy (2: non-temporal).
template <int N, int M = N, int AUX = 2>
__device__ __forceinline__ void row_to_lds(const float2* src, float2* buf, int t) {
  constexpr int BLOCKS = (M + 127) / 128;
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(src), (short)0, M * 8, 0x00020000);
#pragma unroll
  for (int i = 0; i < BLOCKS; ++i)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)(buf + 128 * i), 16,
                                             t * 16, i * 1024, 0, AUX);
}

__device__ __forceinline__ void wave_lds_order() {
  // LDS instructions of one wavefront execute in order; this keeps the
  // compiler from moving LDS accesses across a pass boundary.
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Pass p of the plan.  FROM_LDS: inputs come from LDS (else from x in the
// strided natural layout); TO_LDS: outputs go to LDS (else to x).
// pre (last pass only, or NULL): a whole row whose copy into buf[0..N) by
// LDS DMA is started as soon as this pass's reads of buf have completed, so
// that it overlaps the pass's arithmetic and whatever follows the transform;
// tab likewise, a phase table of N/2 + 1 entries (read from L2, cached).
template <int N, int K, int p, bool INV, bool FROM_LDS, bool TO_LDS>
__device__ __forceinline__ void stockham_pass(float2 (&x)[N / 64], float2* __restrict__ buf,
                                              const float2* __restrict__ tw, int t,
                                              const float2* pre = nullptr, const float2* tab = nullptr) {
  constexpr int E = N / 64;
  constexpr int R = FftPlan<N, K>::R[p];
  constexpr int NS = plan_ns<N, K>(p);
  constexpr int NB = N / R;
  constexpr int BPL = E / R;
  static_assert(BPL * R == E, "radix must divide the per-lane count");
  static_assert(NS <= 64 || 64 * (BPL - 1) < NS, "twiddle index wraps");
  if constexpr (FROM_LDS) {
    // element t + 64 b + r NB in the layout of the previous boundary: a lane
    // base plus compile-time offsets
    static_assert(lay_reads_split<N, K, p>(), "LDS layout does not split into lane base + offset");
    const float2* base = buf + lay<N, K, p - 1>(t);
#pragma unroll
    for (int b = 0; b < BPL; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) x[b + r * BPL] = base[lay<N, K, p - 1>(64 * b + r * NB)];
    wave_lds_order();
    if constexpr (!TO_LDS) {
      if (pre || tab) {
        // the reads above must have returned before the DMA overwrites buf
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        if (pre) row_to_lds<N>(pre, buf, t);
        else row_to_lds<N, N / 2 + 1, 0>(tab, buf, t);
      }
    }
  }
  // twiddle base: entry (r-1)*NS + (j & (NS-1)), j = t + 64 b (pairs: two
  // float2 per entry)
  constexpr bool PAIR = plan_tw_pair<N, K>(p);
  const float2* twb = tw + plan_tw_off<N, K>(p) + (NS >= 64 ? t : (t & (NS - 1))) * (PAIR ? 2 : 1);
  // output base of butterfly j = t + 64 b
  int obase;
  if constexpr (NS == 1) obase = R * t + ((R * t) >> 4);
  else if constexpr (NS >= 64) obase = t + (t >> 4);
  else if constexpr (NS % 16 == 0) {
    const int u = (t / NS) * NS * R + (t & (NS - 1));
    obase = u + (u >> 4);
  } else {
    obase = 0;
  }
#pragma unroll
  for (int b = 0; b < BPL; ++b) {
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[b + r * BPL];
    if constexpr (NS > 1) {
      constexpr int boff = NS >= 64 ? 64 : 0;  // (j & (NS-1)) advances by 64 per b when NS > 64
#pragma unroll
      for (int r = 1; r < R; ++r) {
        if constexpr (PAIR) v[r] = cmul_tw_pair<INV>(v[r], twb[2 * (r - 1) * NS], twb[2 * (r - 1) * NS + 1]);
        else v[r] = cmul_tw<INV>(v[r], twb[(r - 1) * NS + (NS > 64 ? b * boff : 0)]);
      }
    }
    dft<R, INV>(v);
    if constexpr (TO_LDS) {
      // output index (j/NS)*NS*R + j%NS + r NS, j = t + 64 b
      static_assert((NS > 1 && NS < 16) || (lay_a<N, K>(p) == 4 && lay_m<N, K>(p) == 1),
                    "these write paths address the one-pad-per-16 layout");
      if constexpr (NS == 1) {
        // j R + r = R t + 64 R b + r  ->  lidx = obase + 68 R b + r
#pragma unroll
        for (int r = 0; r < R; ++r) buf[obase + 68 * R * b + r] = v[r];
      } else if constexpr (NS >= 64) {
        // NS multiple of 64: j/NS = (64 b)/NS, j%NS = t + 64 b - NS (j/NS)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int q = (64 * b) / NS;            // compile-time per b
          const int c = q * NS * R + (64 * b - q * NS) + r * NS;
          buf[obase + c + (c >> 4)] = v[r];
        }
      } else if constexpr (NS % 16 == 0) {
        // 16 <= NS < 64: j/NS = t/NS + 64 b/NS, j%NS = t%NS; offsets are
        // multiples of 16, so the padding splits off the lane base
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int c = (64 * b / NS) * NS * R + r * NS;
          buf[obase + c + (c >> 4)] = v[r];
        }
      } else {
        // NS < 16: the boundary's own layout (lay_m, lay_a), per output
        const int u = (t / NS) * NS * R + (t & (NS - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) buf[lay<N, K, p>(u + (64 * b / NS) * NS * R + r * NS)] = v[r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) x[b + r * BPL] = v[r];
    }
  }
  if constexpr (TO_LDS) wave_lds_order();
}

template <int N, int K, bool INV, int p>
__device__ __forceinline__ void fft_passes(float2 (&x)[N / 64], float2* buf, const float2* tw, int t,
                                           const float2* pre, const float2* tab) {
  constexpr int P = FftPlan<N, K>::P;
  static_assert(P >= 2, "the last pass reads buf");
  if constexpr (p < P) {
    stockham_pass<N, K, p, INV, (p > 0), (p < P - 1)>(x, buf, tw, t, p == P - 1 ? pre : nullptr,
                                                        p == P - 1 ? tab : nullptr);
    fft_passes<N, K, INV, p + 1>(x, buf, tw, t, pre, tab);
  }
}

// Unnormalised DFT (forward: exp(-2 pi i jk/N); inverse: exp(+2 pi i jk/N))
// of the row held in x (strided natural layout in and out).  pre: a whole
// row of N samples copied into buf[0..N) by LDS DMA during the last pass;
// tab: a phase table of N/2 + 1 entries copied there instead (NULL: none).
template <int N, bool INV, int K = 0>
__device__ __forceinline__ void fft_row(float2 (&x)[N / 64], float2* buf, const float2* tw, int t,
                                        const float2* pre = nullptr, const float2* tab = nullptr) {
  fft_passes<N, K, INV, 0>(x, buf, tw, t, pre, tab);
}

// ---------------------------------------------------------------- any length
//
// Rows whose FFT length n is not the power of two N of the launch (n <= N,
// e.g. 500 or 600 samples in a 512- or 1024-lane row): the row is staged in
// LDS and transformed by a runtime chain of Stockham passes between two LDS
// buffers (buf and buf + LdsSize<N>), radix 16, 8, 4, 2, 3 and 5 as register
// butterflies and any other prime radix as a direct DFT.  Pass p with radix
// R and sub-transform length NS reads butterfly j's inputs at j + r n/R and
// writes output s at (j/NS) NS R + j%NS + s NS, as in the power-of-two
// passes.  Twiddle table: pass p holds entries (r-1) NS + m = exp(-2 pi i m
// r / (NS R)) (m < NS, 1 <= r < R), followed for a direct-DFT radix by R
// entries exp(-2 pi i q / R); entries are double-precision values rounded
// once.  At most n - 1 + 3 * 61 entries (radices above 61 are not planned).
struct FftGen {
  int n;                   // transform length
  int np;                  // number of passes (at most 10)
  unsigned long long rad;  // radix of pass p in bits 6p .. 6p+5 (a word, not an
                           // array: a runtime-indexed array would live in scratch)
};
__host__ __device__ constexpr int gen_radix(const FftGen& f, int p) { return (int)((f.rad >> (6 * p)) & 63u); }

constexpr int kGenMaxRadix = 61;
// radices transformed as a direct DFT (their own R-entry table follows the
// pass twiddles); 2, 3, 4, 5, 8 and 16 are register butterflies
__host__ __device__ constexpr bool gen_direct(int R) { return R > 5 && R != 8 && R != 16; }
// register-butterfly passes of a runtime plan whose twiddles are few
// distinct values (NS R <= 64) hold them as float pairs, as plan_tw_pair
// passes do (the LDS-bound runtime passes take the extra multiplies for
// free); their pair entries add at most 63 to the table
__host__ __device__ constexpr bool gen_tw_pair(int R, int NS) { return NS > 1 && NS * R <= 64 && !gen_direct(R); }
__host__ __device__ constexpr int gen_tw_cap(int N) { return N + 3 * kGenMaxRadix + 64; }

constexpr Konst kSin3 = konst(0.86602540378443864676);   // sin(2 pi / 3)
constexpr Konst kCos5a = konst(0.30901699437494742410);  // cos(2 pi / 5)
constexpr Konst kCos5b = konst(-0.80901699437494742410); // cos(4 pi / 5)
constexpr Konst kSin5a = konst(0.95105651629515357212);  // sin(2 pi / 5)
constexpr Konst kSin5b = konst(0.58778525229247312917);  // sin(4 pi / 5)
constexpr Konst kSin5aN = konst(-0.95105651629515357212);

template <bool INV>
__device__ __forceinline__ void dft3(float2 (&v)[3]) {
  const float2 t1 = cadd(v[1], v[2]);
  const float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  const float2 d = csub(v[1], v[2]);
  // -i sin(2pi/3) (v1 - v2), conjugated for INV
  const float2 t3 = mul_mi<INV>(make_float2(kmul(d.x, kSin3), kmul(d.y, kSin3)));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2 (&v)[5]) {
  const float2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
  const float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
  const float2 a = v[0];
  const float2 p1 = make_float2(a.x + kdot(b1.x, kCos5a, b2.x, kCos5b), a.y + kdot(b1.y, kCos5a, b2.y, kCos5b));
  const float2 p2 = make_float2(a.x + kdot(b1.x, kCos5b, b2.x, kCos5a), a.y + kdot(b1.y, kCos5b, b2.y, kCos5a));
  // -i (s1 d1 + s2 d2) and -i (s2 d1 - s1 d2) (conjugated for INV)
  const float2 q1 = mul_mi<INV>(make_float2(kdot(d1.x, kSin5a, d2.x, kSin5b), kdot(d1.y, kSin5a, d2.y, kSin5b)));
  const float2 q2 = mul_mi<INV>(make_float2(kdot(d1.x, kSin5b, d2.x, kSin5aN), kdot(d1.y, kSin5b, d2.y, kSin5aN)));
  v[0] = make_float2(a.x + b1.x + b2.x, a.y + b1.y + b2.y);
  v[1] = cadd(p1, q1);
  v[4] = csub(p1, q1);
  v[2] = cadd(p2, q2);
  v[3] = csub(p2, q2);
}

// q = j / NS for j < 2^21 from a per-pass multiplier mg = ceil(2^32 / NS)
// (exact: j (mg NS - 2^32) < 2^32)
__device__ __forceinline__ int div_ns(int j, unsigned long long mg) {
  return (int)(((unsigned long long)(unsigned)j * mg) >> 32);
}

// butterfly j of a pass with a register-butterfly radix R (2, 3, 4, 5, 8,
// 16); MG: j / NS by the multiplier mg (else by integer division)
template <int R, bool INV, bool MG>
__device__ __forceinline__ void gen_bfly(const float2* __restrict__ A, float2* __restrict__ B,
                                         const float2* __restrict__ tw, int NB, int NS, unsigned long long mg,
                                         int j) {
  const int q = MG ? div_ns(j, mg) : j / NS, m = j - q * NS;
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = A[lidx(j + r * NB)];
  if (gen_tw_pair(R, NS)) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const int e = 2 * ((r - 1) * NS + m);
      v[r] = cmul_tw_pair<INV>(v[r], tw[e], tw[e + 1]);
    }
  } else if (NS > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul_tw<INV>(v[r], tw[(r - 1) * NS + m]);
  }
  if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 5) dft5<INV>(v);
  else dft<R, INV>(v);
  const int o = q * NS * R + m;
#pragma unroll
  for (int r = 0; r < R; ++r) B[lidx(o + r * NS)] = v[r];
}

// one pass with a register-butterfly radix; PAIR: two butterflies per lane
// and iteration, so that the LDS loads of the second overlap the first, and
// the quotient j / NS by multiplication (measured: faster in rows of 1024
// and 2048 lanes, slower in rows of 512)
template <int R, bool INV, bool PAIR>
__device__ __forceinline__ void gen_pass_r(const float2* __restrict__ A, float2* __restrict__ B,
                                           const float2* __restrict__ tw, int n, int NS, int t) {
  const int NB = n / R;
  if constexpr (PAIR) {
    const unsigned long long mg = ((1ull << 32) + (unsigned long long)NS - 1) / (unsigned long long)NS;
    for (int j = t; j < NB; j += 128) {
      gen_bfly<R, INV, true>(A, B, tw, NB, NS, mg, j);
      if (j + 64 < NB) gen_bfly<R, INV, true>(A, B, tw, NB, NS, mg, j + 64);
    }
  } else {
    for (int j = t; j < NB; j += 64) gen_bfly<R, INV, false>(A, B, tw, NB, NS, 0ull, j);
  }
}

// largest register radix of runtime plans in rows of N lanes: rows of 1024
// and 2048 lanes run at 1-2 waves per SIMD for their LDS anyway, so radix 8
// and 16 butterflies (and two per iteration) cost them no occupancy; shorter
// rows keep to radix 4 and their register budget
__host__ __device__ constexpr int gen_max_radix(int N) { return N >= 1024 ? 16 : 4; }

// one pass with a prime radix R > 5: direct DFT, twiddled inputs re-read
// from LDS for every output (the transform is rare, registers are not)
template <bool INV>
__device__ __forceinline__ void gen_pass_dft(const float2* __restrict__ A, float2* __restrict__ B,
                                             const float2* __restrict__ tw, const float2* __restrict__ wr, int n,
                                             int R, int NS, int t) {
  const int NB = n / R;
  for (int j = t; j < NB; j += 64) {
    const int q = j / NS, m = j - q * NS;
    const int o = q * NS * R + m;
    for (int s = 0; s < R; ++s) {
      float2 acc = A[lidx(j)];
      int e = 0;  // (r s) mod R
      for (int r = 1; r < R; ++r) {
        e += s;
        if (e >= R) e -= R;
        float2 v = A[lidx(j + r * NB)];
        if (NS > 1) v = cmul_tw<INV>(v, tw[(r - 1) * NS + m]);
        acc = cadd(acc, cmul_tw<INV>(v, wr[e]));
      }
      B[lidx(o + s * NS)] = acc;
    }
  }
}

// Unnormalised DFT of length f.n <= N of the row held in x (strided natural
// layout in and out; samples n.. are zero on output).
template <int N, bool INV>
__device__ __forceinline__ void fft_row_gen(float2 (&x)[N / 64], float2* buf, const float2* tw, int t,
                                            const FftGen& f) {
  constexpr int E = N / 64;
  const int n = f.n;
  float2* A = buf;
  float2* B = buf + LdsSize<N>::value;
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (t + 64 * e < n) A[lidx(t + 64 * e)] = x[e];
  wave_lds_order();
  int NS = 1, off = 0;
  for (int p = 0; p < f.np; ++p) {
    const int R = gen_radix(f, p);
    const float2* twp = tw + off;
    constexpr bool kBig = gen_max_radix(N) == 16;
    switch (R) {
      case 2: gen_pass_r<2, INV, kBig>(A, B, twp, n, NS, t); break;
      case 3: gen_pass_r<3, INV, kBig>(A, B, twp, n, NS, t); break;
      case 4: gen_pass_r<4, INV, kBig>(A, B, twp, n, NS, t); break;
      case 5: gen_pass_r<5, INV, kBig>(A, B, twp, n, NS, t); break;
      case 8:
        if constexpr (kBig) gen_pass_r<8, INV, true>(A, B, twp, n, NS, t);
        break;
      case 16:
        if constexpr (kBig) gen_pass_r<16, INV, true>(A, B, twp, n, NS, t);
        break;
      default: gen_pass_dft<INV>(A, B, twp, twp + (R - 1) * NS, n, R, NS, t); break;
    }
    wave_lds_order();
    off += (R - 1) * NS * (gen_tw_pair(R, NS) ? 2 : 1) + (gen_direct(R) ? R : 0);
    NS *= R;
    float2* T = A;
    A = B;
    B = T;
  }
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = t + 64 * e < n ? A[lidx(t + 64 * e)] : make_float2(0.f, 0.f);
  wave_lds_order();
}

// Pass chains fixed at compile time for FFT lengths that the reference's
// runs and hypercubes use (nk = 500 and 600 in its runs; 100, 200, 300, 750,
// 1000 and 1500 as axis lengths): the chain fft_gen_plan builds for the
// length in its row (the smallest built row of at least nk lanes), over the
// same twiddle table, with every radix, length, sub-length and table offset
// a constant, so that the index arithmetic folds.  Same butterflies in the
// same order as fft_row_gen: bit-identical results, which the tests check
// for every entry.  (2000 = 16 5 5 5 was measured not bit-identical to the
// runtime chain and keeps the runtime chain; no entry has a radix-16 pass.)
// ID 0: none.
constexpr int kGenFixedCount = 9;
constexpr int kGenFixedN[kGenFixedCount] = {0, 500, 600, 100, 200, 300, 750, 1000, 1500};
__host__ __device__ constexpr int gen_row_lanes(int n) {
  int N = 128;
  while (N < n) N *= 2;
  return N;
}
struct GenChain {
  int P;
  int R[10];
};
// fft_gen_plan's chain for length n (largest register radix maxr first,
// then one of 8 / 4 / 2 below it, then the odd primes in increasing order)
__host__ __device__ constexpr GenChain gen_chain(int n, int maxr) {
  GenChain c{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n;
  while (m % maxr == 0 && c.P < 10) {
    c.R[c.P++] = maxr;
    m /= maxr;
  }
  const int small[3] = {8, 4, 2};
  for (int i = 0; i < 3; ++i)
    if (small[i] < maxr && m % small[i] == 0 && c.P < 10) {
      c.R[c.P++] = small[i];
      m /= small[i];
    }
  for (int p = 3; m > 1 && c.P < 10; p += 2)
    while (m % p == 0 && c.P < 10) {
      c.R[c.P++] = p;
      m /= p;
    }
  return c;
}
template <int ID>
struct GenFixed {
  static constexpr int n = kGenFixedN[ID];
  static constexpr int N = gen_row_lanes(n);
  static constexpr GenChain chain = gen_chain(n, gen_max_radix(N));
  static constexpr int P = chain.P;
};
// the chains as documented (500 = 4 5 5 5, 600 = 8 3 5 5, ...)
constexpr bool gen_chain_is(const GenChain& c, const int* r, int p) {
  if (c.P != p) return false;
  for (int i = 0; i < p; ++i)
    if (c.R[i] != r[i]) return false;
  return true;
}
namespace gen_chain_check {
constexpr int c500[] = {4, 5, 5, 5}, c600[] = {8, 3, 5, 5}, c100[] = {4, 5, 5}, c200[] = {4, 2, 5, 5},
              c300[] = {4, 3, 5, 5}, c750[] = {2, 3, 5, 5, 5}, c1000[] = {8, 5, 5, 5}, c1500[] = {4, 3, 5, 5, 5};
static_assert(gen_chain_is(GenFixed<1>::chain, c500, 4) && gen_chain_is(GenFixed<2>::chain, c600, 4) &&
                  gen_chain_is(GenFixed<3>::chain, c100, 3) && gen_chain_is(GenFixed<4>::chain, c200, 4) &&
                  gen_chain_is(GenFixed<5>::chain, c300, 4) && gen_chain_is(GenFixed<6>::chain, c750, 5) &&
                  gen_chain_is(GenFixed<7>::chain, c1000, 4) && gen_chain_is(GenFixed<8>::chain, c1500, 5),
              "compile-time chains");
}  // namespace gen_chain_check

template <int N, bool INV, int ID, int p, int NS, int OFF>
__device__ __forceinline__ void gen_fixed_passes(float2*& A, float2*& B, const float2* tw, int t) {
  if constexpr (p < GenFixed<ID>::P) {
    constexpr int R = GenFixed<ID>::chain.R[p];
    constexpr int n = GenFixed<ID>::n;
    constexpr bool kBig = gen_max_radix(N) == 16;
    static_assert(!gen_direct(R) && (R <= 5 || kBig), "fixed chains use register butterflies");
    gen_pass_r<R, INV, kBig>(A, B, tw + OFF, n, NS, t);
    wave_lds_order();
    float2* T = A;
    A = B;
    B = T;
    gen_fixed_passes<N, INV, ID, p + 1, NS * R, OFF + (R - 1) * NS * (gen_tw_pair(R, NS) ? 2 : 1)>(A, B, tw, t);
  }
}

// fft_row_gen for a fixed chain (f.n == GenFixed<ID>::n)
template <int N, bool INV, int ID>
__device__ __forceinline__ void fft_row_gen_fixed(float2 (&x)[N / 64], float2* buf, const float2* tw, int t) {
  constexpr int E = N / 64;
  constexpr int n = GenFixed<ID>::n;
  static_assert(n <= N, "the chain's length exceeds the row");
  float2* A = buf;
  float2* B = buf + LdsSize<N>::value;
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (t + 64 * e < n) A[lidx(t + 64 * e)] = x[e];
  wave_lds_order();
  gen_fixed_passes<N, INV, ID, 0, 1, 0>(A, B, tw, t);
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = t + 64 * e < n ? A[lidx(t + 64 * e)] : make_float2(0.f, 0.f);
  wave_lds_order();
}

// Fills the concatenated per-pass twiddle table of FftPlan<N, K> (double
// precision, rounded once by round_complex).
template <int N, int K = 0>
__global__ void twiddle_table_kernel(float2* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plan_tw_size<N, K>()) return;
  int p = 0;
  while (p + 1 < FftPlan<N, K>::P && plan_tw_off<N, K>(p + 1) <= i) ++p;
  const int R = FftPlan<N, K>::R[p];
  const int NS = plan_ns<N, K>(p);
  const bool pair = plan_tw_pair<N, K>(p);
  const int k = (i - plan_tw_off<N, K>(p)) >> (pair ? 1 : 0);
  const int r = k / NS + 1;
  const int jm = k % NS;
  const double a = -2.0 * 3.14159265358979323846 * (double)(jm * r) / (double)(NS * R);
  double s, c;
  sincos(a, &s, &c);
  if (!pair) {
    out[i] = round_complex(c, s);
  } else {
    const float2 hi = make_float2((float)c, (float)s);
    out[i] = ((i - plan_tw_off<N, K>(p)) & 1) ? make_float2((float)(c - (double)hi.x), (float)(s - (double)hi.y)) : hi;
  }
}

}  // namespace wemfwi
