// Device-side descriptors shared by the kernels and the host layer.
#pragma once

#include <cstdlib>

#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "sweep.hpp"

namespace wemfwi {

// One (shot, frequency) row of a Down/Up sweep.
struct SweepRow {
  const float2* in;      // [nz][nx] source terms (model of OneWay::forward/adjoint)
  float2* out;           // [nz][nx] result; == in for the in-place recursion
  const float2* ptab;    // [nslots][N/2+1] phase table of this row's frequency
  float w;               // angular frequency
  int model;             // index into the DepthTables array
  int recursive;         // 1: W[next] = P W[cur] + in[next] with W the output rows
                         // 0: out[next] = P in[cur] + in[next] (one bidiagonal step)
  float2* scratch;       // [N] row of this wave for depths with several references
  int first;             // depth of the initial row; -1: 0 or nz-1 by the direction
  int nsteps;            // depth steps from it; -1: nz-1 (the whole sweep)
  const float2* init;    // initial state row [nx]; NULL: the row in[first]
  int zbase;             // depth of the first row of in / out (0: whole [nz][nx] arrays)
  const float2* smul;    // [nz][nx] (absolute depths) multiplier of every source row
                         // (also the initial row unless init is set); NULL: none
  float2* out2;          // second output rows (layout of out): conj(omul) * result; NULL: none
  const float2* omul;    // [nz][nx] (absolute depths) multiplier of out2
  int src_skip;          // 1: only the source rows of depths src_lo..src_hi are read (the
  int src_lo, src_hi;    // others are zero and need not be initialised); 0: every row
};

// One independent depth step.
struct StepRow {
  const float2* in;      // [nx]
  float2* out;           // [nx]
  const float2* ptab;    // [nslots][N/2+1]
  float w;
  int d;                 // depth of the step
  int model;
  int add;               // 1: out += step(in)
  float2* scratch;       // [N]
};

bool fft_size_supported(int n);
// FFT radix plans of a built length (the sweep autotuner chooses among them)
int fft_plans(int N);
int twiddle_table_size(int N, int plan);
hipError_t launch_twiddles(int N, int plan, float2* out, hipStream_t s);
// twiddle table of plan `plan` of length N on the current device (cached)
int twiddles(int n, int plan, const float2** out);
// the table of plan `plan` of length n if it was built before (no allocation)
int twiddles_cached(int n, int plan, const float2** out);
// JSON list of the sweep autotuner's decisions
std::string autotune_report();

hipError_t launch_phase_table(float2* ptab, const double* wlist, int nw, const double2* sref,
                              int nslots, int N, int NH, double dk, double dz, double eps,
                              hipStream_t stream);
// tune = false: no autotuning launch and no twiddle-table allocation; the
// plan is the forced or recorded one when its table already exists, else
// plan 0 with tw
hipError_t launch_oneway(int N, const SweepRow* rows, int nrows, const DepthTables* models,
                         StepGeom g, int nz, int step, bool adj, const float2* tw, hipStream_t s,
                         bool tune = true);
// the compile-time pass chains of GenFixed lengths are used unless
// WEMFWI_GEN_FIXED=0 (the runtime chain: same results, slower)
inline bool gen_fixed_enabled() {
  const char* e = std::getenv("WEMFWI_GEN_FIXED");
  return !(e && *e && std::atoi(e) == 0);
}

// Calls fixed(std::integral_constant<int, ID>) for the compile-time chain ID
// of FFT length n in rows of N lanes when there is one (and they are
// enabled), else runtime().
template <int N, int ID = 1, class F, class G>
inline hipError_t gen_fixed_dispatch(int n, F&& fixed, G&& runtime) {
  if constexpr (ID < kGenFixedCount) {
    if constexpr (GenFixed<ID>::N == N) {
      if (n == GenFixed<ID>::n && gen_fixed_enabled()) return fixed(std::integral_constant<int, ID>{});
    }
    return gen_fixed_dispatch<N, ID + 1>(n, fixed, runtime);
  } else {
    return runtime();
  }
}

// sweep launch of radix plan k with the half workgroup tile (half the waves
// per workgroup of launch_oneway's full tile); only where oneway_half_tile
// holds (gen: rows whose FFT length g.fg.n is not N, k unused)
hipError_t launch_oneway_half(int N, int k, const SweepRow* rows, int nrows, const DepthTables* models,
                              StepGeom g, int nz, int step, bool adj, const float2* tw, hipStream_t s);
constexpr bool oneway_half_tile(int N, bool gen) { return gen ? N == 1024 : (N == 1024 || N == 2048); }
hipError_t launch_onestep(int N, const StepRow* rows, int nrows, const DepthTables* models,
                          StepGeom g, bool adj, const float2* tw, hipStream_t s);
// StepGeom.tdma of a sweep launch of `rows` rows that use `tables` distinct
// phase tables: 1 when a table serves fewer than two rows on average (it is
// then read from HBM by one row, and a DMA copy beats the row's serial
// table loads); WEMFWI_TABLE_DMA=0 / 1 forces it off / on
int table_dma(size_t rows, size_t tables);

}  // namespace wemfwi
