// The persistent depth-sweep kernel (one (shot, frequency) row per
// wavefront, all depths of a row in one launch) and its launch templates,
// shared by the translation units that instantiate its workgroup sizes.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/wemfwi.h"
#include "kernels.h"
#include "sweep.hpp"

namespace wemfwi {

// ------------------------------------------------------------------ rows

template <class V>
__device__ __forceinline__ void load_row(float2 (&v)[V::E], const float2* __restrict__ p, const StepGeom& g,
                                         int t) {
  load_samples<V>(v, p, g, t);
}

template <class V>
__device__ __forceinline__ void store_row(float2* __restrict__ p, const float2 (&v)[V::E], const StepGeom& g,
                                          int t) {
#pragma unroll
  for (int e = 0; e < V::E; ++e) row_store<V>(p, v[e], g, t, e);
}

// Waves (rows) per workgroup: the twiddle table in LDS is shared by the
// waves of a workgroup.  For N = 1024 two workgroups of 8 waves (2 x 77.6 KB
// of LDS) give 16 waves per CU, i.e. 4 per SIMD; for N = 2048 one workgroup
// of 8 waves (155 KB) gives 2 per SIMD.  Sweeps of 1024- and 2048-lane rows
// also come in workgroups of half that many waves (the tile), which spread a
// launch of fewer rows than fill the chip over more CUs; the sweep launcher
// picks between the two by timing.
template <int N>
struct WavesPerBlock {
  static constexpr int value = N >= 1024 ? 8 : 4;
};
// GEN rows use two LDS row buffers per wave: 4 waves of 2 x 34.8 KB at
// N = 2048
template <class V>
struct WavesV {
  static constexpr int value = (V::GEN && V::n == 2048) ? 4 : WavesPerBlock<V::n>::value;
};

// In-place or out-of-place Down/Up recursion (OneWay.cpp:32-109) of one
// (shot, frequency) row per wavefront.  step = +1 runs iz = 0..nz-2 (Down
// forward, Up adjoint), step = -1 runs iz = nz-1..1 (Down adjoint, Up forward).
// Forward steps use depth cur, adjoint steps depth next.  The row's width,
// padding, taper and depth count come from its model, so one launch may hold
// rows of different apertures; the launch fixes the FFT length and variant.
// Waves per SIMD the register allocator targets for an FFT length.
template <int N>
struct Occupancy {
  static constexpr int value = N <= 256 ? 8 : (N == 512 ? 6 : (N == 1024 ? 4 : 2));
};

// W: waves (rows) per workgroup.
template <class V, bool ADJ, bool PF, bool TD, int W>
__global__ void __launch_bounds__(64 * W)
    __attribute__((amdgpu_waves_per_eu(V::GEN ? 2 : Occupancy<V::n>::value, 8)))
    oneway_kernel(const SweepRow* __restrict__ rows, int nrows, const DepthTables* __restrict__ models,
                  StepGeom g, int step, const float2* __restrict__ twg) {
  constexpr int N = V::n;
  constexpr int E = V::E;
  constexpr int kWavesPerBlock = W;
  __shared__ float2 tw[TwV<V>::value];
  __shared__ float2 bufs[kWavesPerBlock][LdsV<V>::value];
  const int t = threadIdx.x & 63;
  const int wv = wave_id();
  for (int i = threadIdx.x; i < TwV<V>::value; i += blockDim.x) tw[i] = ldg(twg, (size_t)i);
  __syncthreads();
  const int row = blockIdx.x * kWavesPerBlock + wv;
  if (row >= nrows) return;
  const SweepRow R = ldu_struct(rows, row);
  const DepthTables dt = ldu_struct(models, R.model);
  g.nx = dt.nx;
  g.tap = dt.tap;
  g.taper = dt.taper;
  float2* buf = bufs[wv];
  const int nx = g.nx;
  const bool recursive = R.recursive != 0;
  // a row covers depths first .. first + step * nsteps (a whole sweep, or a
  // segment started from a checkpointed state row)
  const int first = R.first >= 0 ? R.first : (step > 0 ? 0 : dt.nz - 1);
  const int nsteps = R.nsteps >= 0 ? R.nsteps : dt.nz - 1;
  const int zb = R.zbase;
  float2 st[E];
  if (R.init || !R.src_skip || (first >= R.src_lo && first <= R.src_hi)) {
    load_row<V>(st, R.init ? R.init : R.in + (size_t)(first - zb) * nx, g, t);
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = make_float2(0.f, 0.f);
  }
  if (R.smul && !R.init) {
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cmul(st[e], row_sample<V>(R.smul + (size_t)first * nx, g, t, e));
  }
  // the first row is stored unless it already holds the state (in place,
  // read as it is); with src_skip it may not have been initialised
  if (R.in != R.out || R.init || R.smul || R.src_skip) store_row<V>(R.out + (size_t)(first - zb) * nx, st, g, t);
  if (R.out2) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      row_store<V>(R.out2 + (size_t)(first - zb) * nx, cmulc(st[e], row_sample<V>(R.omul + (size_t)first * nx, g, t, e)),
                   g, t, e);
  }
  // PF (full rows whose every depth has a plain source row; the launch
  // guarantees it): the source row is read by LDS DMA into the wave's
  // transform buffer, started once the step's last inverse transform has
  // read that buffer for the last time (in its last pass, or right after the
  // split step's table loads), so that the row's HBM latency overlaps the
  // rest of the step's arithmetic instead of stalling the source add.
  constexpr bool PFX = PF && V::FULL;
  for (int k = 0; k < nsteps; ++k) {
    const int cur = first + k * step;
    const int nxt = cur + step;
    if (!recursive && k > 0) load_row<V>(st, R.in + (size_t)(cur - zb) * nx, g, t);
    const bool src_here = PFX || !R.src_skip || (nxt >= R.src_lo && nxt <= R.src_hi);
    const float2* pre = nullptr;
    if constexpr (PFX) {
      pre = R.in + (size_t)(nxt - zb) * nx;
      __builtin_assume(pre != nullptr);
    }
    bool in_lds;
    if constexpr (ADJ) in_lds = step_adjoint<V, true, TD>(st, nxt, R.w, dt, g, R.ptab, buf, tw, t, R.scratch, pre);
    else in_lds = step_forward<V, true, TD>(st, cur, R.w, dt, g, R.ptab, buf, tw, t, R.scratch, pre);
    if (src_here) {
      float2 src[E];
      if (PFX && in_lds) {
        // the DMA writes LDS behind the memory counter: wait for it
        // explicitly (the compiler does not track the copy through the
        // generic buffer pointer)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
        for (int e = 0; e < E; ++e) src[e] = buf[t + 64 * e];
      } else {
        load_row<V>(src, R.in + (size_t)(nxt - zb) * nx, g, t);
      }
      if (!PFX && R.smul) {
#pragma unroll
        for (int e = 0; e < E; ++e) src[e] = cmul(src[e], row_sample<V>(R.smul + (size_t)nxt * nx, g, t, e));
      }
#pragma unroll
      for (int e = 0; e < E; ++e) st[e] = cadd(st[e], src[e]);
    }
    store_row<V>(R.out + (size_t)(nxt - zb) * nx, st, g, t);
    if (R.out2) {
#pragma unroll
      for (int e = 0; e < E; ++e)
        row_store<V>(R.out2 + (size_t)(nxt - zb) * nx, cmulc(st[e], row_sample<V>(R.omul + (size_t)nxt * nx, g, t, e)),
                     g, t, e);
    }
  }
}

// TD: the phase-table DMA variant (g.tdma); H: the half tile (WavesV / 2
// waves per workgroup)
template <class V, bool TD, bool H>
static hipError_t launch_oneway_w(const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g, int step,
                                  bool adj, const float2* tw, hipStream_t s) {
  constexpr int kWavesPerBlock = H ? WavesV<V>::value / 2 : WavesV<V>::value;
  dim3 grid((nrows + kWavesPerBlock - 1) / kWavesPerBlock), block(64 * kWavesPerBlock);
  // the source-row prefetch only on full rows, when the caller states that
  // every row reads a plain source row at every depth (g.pf)
  if constexpr (V::FULL) {
    if (g.pf) {
      if (adj)
        hipLaunchKernelGGL((oneway_kernel<V, true, true, TD, kWavesPerBlock>), grid, block, 0, s, rows, nrows, models, g, step, tw);
      else
        hipLaunchKernelGGL((oneway_kernel<V, false, true, TD, kWavesPerBlock>), grid, block, 0, s, rows, nrows, models, g, step, tw);
      return hipGetLastError();
    }
  }
  if (adj)
    hipLaunchKernelGGL((oneway_kernel<V, true, false, TD, kWavesPerBlock>), grid, block, 0, s, rows, nrows, models, g, step, tw);
  else
    hipLaunchKernelGGL((oneway_kernel<V, false, false, TD, kWavesPerBlock>), grid, block, 0, s, rows, nrows, models, g, step, tw);
  return hipGetLastError();
}

template <class V, bool H = false>
static hipError_t launch_oneway_v(const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz,
                                  int step, bool adj, const float2* tw, hipStream_t s) {
  if (g.tdma && !V::GEN) return launch_oneway_w<V, true, H>(rows, nrows, models, g, step, adj, tw, s);
  return launch_oneway_w<V, false, H>(rows, nrows, models, g, step, adj, tw, s);
}

// the four register-FFT variants of radix plan K
template <int N, int K, bool H>
static hipError_t launch_oneway_k(const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz,
                                  int step, bool adj, const float2* tw, hipStream_t s) {
  if (g.full && g.ssf)
    return launch_oneway_v<Var<N, true, true, false, K>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
  if (g.full) return launch_oneway_v<Var<N, true, false, false, K>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
  if (g.ssf) return launch_oneway_v<Var<N, false, true, false, K>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
  return launch_oneway_v<Var<N, false, false, false, K>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
}

// rows whose FFT length is not N (GEN variants): the compile-time pass
// chain of their length where there is one (GenFixed; WEMFWI_GEN_FIXED=0
// keeps the runtime chain), else the runtime-planned passes
template <int N, bool H = false>
static hipError_t launch_oneway_gen(const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz,
                                    int step, bool adj, const float2* tw, hipStream_t s) {
  return gen_fixed_dispatch<N>(
      g.fg.n,
      [&](auto id) {
        constexpr int ID = decltype(id)::value;
        if (g.ssf) return launch_oneway_v<Var<N, false, true, true, ID>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
        return launch_oneway_v<Var<N, false, false, true, ID>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
      },
      [&] {
        if (g.ssf) return launch_oneway_v<Var<N, false, true, true>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
        return launch_oneway_v<Var<N, false, false, true>, H>(rows, nrows, models, g, nz, step, adj, tw, s);
      });
}

template <int N, bool H = false, int K = 0>
static hipError_t launch_oneway_plan(int k, const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g,
                                     int nz, int step, bool adj, const float2* tw, hipStream_t s) {
  if constexpr (K + 1 < FftPlans<N>::value) {
    if (k != K) return launch_oneway_plan<N, H, K + 1>(k, rows, nrows, models, g, nz, step, adj, tw, s);
  }
  if (k != K) return hipErrorInvalidValue;
  return launch_oneway_k<N, K, H>(rows, nrows, models, g, nz, step, adj, tw, s);
}

}  // namespace wemfwi
