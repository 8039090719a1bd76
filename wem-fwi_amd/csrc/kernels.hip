// HIP kernels of the one-way WEM hot path for gfx950, and their launchers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "kernels.h"
#include "sweep.hpp"
#include "sweep_kernel.hpp"

namespace wemfwi {

// ------------------------------------------------------------------ tables

// Phase-shift multipliers exp(-i kz dz)/N per (frequency, active slot, k),
// evaluated in double precision (phase_shift.ispc:8-14 with the k grid of
// Phshift_cmplx.cpp:12-24).  kz = re + i im with im <= 0; Im is formed as
// -|b|/(2 re) (or Re as |b|/(2|im|) when a < 0) -- equal to the reference's
// -sqrt((c-a)/2) since (c-a)(c+a) = b^2, without its cancellation.
// Only k = 0..n/2 is used: the grid satisfies k[n-i] = -k[i], so the
// multiplier of index k is entry min(k, n-k); n is the FFT length, NH the
// entries per (frequency, slot) (N/2 + 1 of the row's N lanes, >= n/2 + 1).
// Rounded by round_complex: the table of a depth repeats wherever the
// reference slowness does.
__global__ void phase_table_kernel(float2* __restrict__ ptab, const double* __restrict__ wlist,
                                   int nw, const double2* __restrict__ sref, int nslots, int N,
                                   int NH, double dk, double dz, double eps) {
  const size_t total = (size_t)nw * nslots * NH;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % NH);
    const int s = (int)((i / NH) % nslots);
    const int iw = (int)(i / ((size_t)NH * nslots));
    const int h = N / 2;
    const double kk = k < h ? k * dk : 0.0;
    const double w = wlist[iw];
    const double2 sr = sref[s];
    const double a = w * w * sr.x - kk * kk;
    const double b = w * w * (sr.y - eps * sr.x);
    const double c = hypot(a, b);
    const double big = sqrt(0.5 * (c + fabs(a)));
    const double small = big > 0.0 ? fabs(b) / (2.0 * big) : 0.0;
    const double re = a >= 0.0 ? big : small;
    const double im = a >= 0.0 ? -small : -big;
    const double amp = exp(im * dz) / (double)N;
    double sn, cs;
    sincos(re * dz, &sn, &cs);
    ptab[i] = round_complex(amp * cs, -amp * sn);
  }
}

// Independent single depth steps (SSF::forward/adjoint as a unit operator).
template <class V, bool ADJ>
__global__ void __launch_bounds__(64 * WavesV<V>::value)
    __attribute__((amdgpu_waves_per_eu(V::GEN ? 2 : Occupancy<V::n>::value, 8)))
    onestep_kernel(const StepRow* __restrict__ rows, int nrows, const DepthTables* __restrict__ models,
                   StepGeom g, const float2* __restrict__ twg) {
  constexpr int N = V::n;
  constexpr int E = V::E;
  constexpr int kWavesPerBlock = WavesV<V>::value;
  __shared__ float2 tw[TwV<V>::value];
  __shared__ float2 bufs[kWavesPerBlock][LdsV<V>::value];
  const int t = threadIdx.x & 63;
  const int wv = wave_id();
  for (int i = threadIdx.x; i < TwV<V>::value; i += blockDim.x) tw[i] = ldg(twg, (size_t)i);
  __syncthreads();
  const int row = blockIdx.x * kWavesPerBlock + wv;
  if (row >= nrows) return;
  const StepRow R = ldu_struct(rows, row);
  const DepthTables dt = ldu_struct(models, R.model);
  float2 st[E];
  load_row<V>(st, R.in, g, t);
  if constexpr (ADJ) step_adjoint<V, true>(st, R.d, R.w, dt, g, R.ptab, bufs[wv], tw, t, R.scratch);
  else step_forward<V, true>(st, R.d, R.w, dt, g, R.ptab, bufs[wv], tw, t, R.scratch);
  if (R.add) {
    float2 o[E];
    load_row<V>(o, R.out, g, t);
#pragma unroll
    for (int e = 0; e < E; ++e) st[e] = cadd(st[e], o[e]);
  }
  store_row<V>(R.out, st, g, t);
}

// ------------------------------------------------------------------ launchers

template <int N, int K = 0>
static int tw_size_k(int k) {
  if constexpr (K + 1 < FftPlans<N>::value) {
    if (k != K) return tw_size_k<N, K + 1>(k);
  }
  return k == K ? plan_tw_size<N, K>() : 0;
}

template <int N, int K = 0>
static hipError_t launch_twiddles_k(int k, float2* out, hipStream_t s) {
  if constexpr (K + 1 < FftPlans<N>::value) {
    if (k != K) return launch_twiddles_k<N, K + 1>(k, out, s);
  }
  if (k != K) return hipErrorInvalidValue;
  const int n = plan_tw_size<N, K>();
  hipLaunchKernelGGL((twiddle_table_kernel<N, K>), dim3((n + 255) / 256), dim3(256), 0, s, out);
  return hipGetLastError();
}

int fft_plans(int N) {
  switch (N) {
    case 128: return FftPlans<128>::value;
    case 256: return FftPlans<256>::value;
    case 512: return FftPlans<512>::value;
    case 1024: return FftPlans<1024>::value;
    case 2048: return FftPlans<2048>::value;
    default: return 0;
  }
}

int twiddle_table_size(int N, int k) {
  switch (N) {
    case 128: return tw_size_k<128>(k);
    case 256: return tw_size_k<256>(k);
    case 512: return tw_size_k<512>(k);
    case 1024: return tw_size_k<1024>(k);
    case 2048: return tw_size_k<2048>(k);
    default: return 0;
  }
}

hipError_t launch_twiddles(int N, int k, float2* out, hipStream_t s) {
  switch (N) {
    case 128: return launch_twiddles_k<128>(k, out, s);
    case 256: return launch_twiddles_k<256>(k, out, s);
    case 512: return launch_twiddles_k<512>(k, out, s);
    case 1024: return launch_twiddles_k<1024>(k, out, s);
    case 2048: return launch_twiddles_k<2048>(k, out, s);
    default: return hipErrorInvalidValue;
  }
}

bool fft_size_supported(int n) {
  return n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048;
}

hipError_t launch_phase_table(float2* ptab, const double* wlist, int nw, const double2* sref,
                              int nslots, int N, int NH, double dk, double dz, double eps,
                              hipStream_t stream) {
  const size_t total = (size_t)nw * nslots * NH;
  if (total == 0) return hipSuccess;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(phase_table_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ptab, wlist, nw,
                     sref, nslots, N, NH, dk, dz, eps);
  return hipGetLastError();
}

int table_dma(size_t rows, size_t tables) {
  const char* e = std::getenv("WEMFWI_TABLE_DMA");
  if (e && *e) return std::atoi(e) ? 1 : 0;
  return tables * 2 > rows ? 1 : 0;
}

// ------------------------------------------------------------------ autotuner
//
// FFT lengths with several radix plans (FftPlans<N> > 1) pick theirs by
// measurement: the first sweep launch of each (N, full row, split step,
// adjoint, device) copies the descriptor of its first row, builds up to 4096
// rows (at most 1 GiB of scratch) with that row's model and frequency, runs
// each plan over the first kTuneSteps depth steps (one warm-up launch, then
// the fastest of three timed with HIP events) and keeps the fastest plan for
// every later launch.  WEMFWI_FFT_PLAN=k forces plan k, WEMFWI_AUTOTUNE=0 keeps plan 0;
// a launch inside stream capture uses the recorded choice or plan 0.
//
// The workgroup tile of 1024- and 2048-lane sweeps (of 1024-lane ones for
// runtime-planned FFT lengths) is picked the same way per row count: the
// first launch of each (N, FFT length, row kind, variant, plan, row count,
// device) times the full tile against the half tile on exactly as
// many sample rows (at most kTileMaxRows) over at most kTileSteps depth
// steps, and keeps the half tile when it is at least 2 % faster.  The tile
// does not change the arithmetic of a row (one wavefront per row either
// way), only how the rows of a launch spread over the CUs.
// WEMFWI_SWEEP_TILE=full|half forces one; launches of more rows, inside
// stream capture before their row count was timed, or with autotuning off
// take the full tile.
constexpr int kTuneSteps = 32;
constexpr int kTileSteps = 16;
constexpr int kTileMaxRows = 16384;

struct TuneEntry {
  int N, full, ssf, adj, dev, plan, rows, steps;
  std::vector<float> ms;  // mean launch time of each plan
};

struct TileEntry {
  int N, nk, full, ssf, adj, pf, tdma, dev, plan, rows, steps, half;
  float ms[2];  // full tile, half tile
};

static std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
static std::vector<TuneEntry>& tune_table() {
  static std::vector<TuneEntry> t;
  return t;
}
static std::vector<TileEntry>& tile_table() {
  static std::vector<TileEntry> t;
  return t;
}

static int env_int(const char* name, int dflt) {
  const char* e = std::getenv(name);
  return (e && *e) ? std::atoi(e) : dflt;
}

// R scratch rows of T + 1 zeroed depths, each a copy of a launch's first
// row descriptor (its model, frequency and phase table), for timing launches
// that leave the caller's rows alone.
struct TuneSample {
  float2* buf = nullptr;
  float2* scr = nullptr;
  SweepRow* drows = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TuneSample() = default;
  TuneSample(const TuneSample&) = delete;
  TuneSample& operator=(const TuneSample&) = delete;
  ~TuneSample() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(drows);
    (void)hipFree(scr);
    (void)hipFree(buf);
  }
  bool make(const SweepRow& r0, int nx, int N, int R, int T, int step, hipStream_t s) {
    const size_t rowlen = (size_t)(T + 1) * nx;
    bool ok = hipMalloc((void**)&buf, (size_t)R * rowlen * sizeof(float2)) == hipSuccess &&
              hipMalloc((void**)&scr, (size_t)R * N * sizeof(float2)) == hipSuccess &&
              hipMalloc((void**)&drows, (size_t)R * sizeof(SweepRow)) == hipSuccess &&
              hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    if (ok) ok = hipMemsetAsync(buf, 0, (size_t)R * rowlen * sizeof(float2), s) == hipSuccess;
    if (!ok) return false;
    std::vector<SweepRow> h(R, r0);
    for (int i = 0; i < R; ++i) {
      h[i].in = buf + (size_t)i * rowlen;
      h[i].out = buf + (size_t)i * rowlen;
      h[i].scratch = scr + (size_t)i * N;
      h[i].recursive = 1;
      h[i].first = step > 0 ? 0 : T;
      h[i].nsteps = T;
      h[i].init = nullptr;
      h[i].zbase = 0;
      h[i].smul = nullptr;
      h[i].out2 = nullptr;
      h[i].omul = nullptr;
      h[i].src_skip = 0;
    }
    return hipMemcpy(drows, h.data(), h.size() * sizeof(SweepRow), hipMemcpyHostToDevice) == hipSuccess;
  }
  // one warm-up launch, then the fastest of three timed ones
  template <class F>
  bool time(F&& launch, hipStream_t s, float* best) {
    if (launch() != hipSuccess) return false;
    for (int it = 0; it < 3; ++it) {
      float m1 = 0.f;
      if (hipEventRecord(e0, s) != hipSuccess || launch() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
          hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&m1, e0, e1) != hipSuccess)
        return false;
      if (it == 0 || m1 < *best) *best = m1;
    }
    return true;
  }
};

// the first row's descriptor and model of a launch, after the whole device
// drains: a tuning launch timed beside other work (a batch plan's groups run
// on several streams) would measure the sharing, not the launch
static bool first_row(const SweepRow* rows, const DepthTables* models, SweepRow* r0, DepthTables* dt) {
  return hipDeviceSynchronize() == hipSuccess &&
         hipMemcpy(r0, rows, sizeof(*r0), hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(dt, models + r0->model, sizeof(*dt), hipMemcpyDeviceToHost) == hipSuccess;
}

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

template <int N>
static int tuned_plan(const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz, int step,
                      bool adj, hipStream_t s, bool tune) {
  const int np = FftPlans<N>::value;
  const int forced = env_int("WEMFWI_FFT_PLAN", -1);
  if (forced >= 0) return forced < np ? forced : 0;
  if (np < 2 || env_int("WEMFWI_AUTOTUNE", 1) == 0) return 0;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(tune_mutex());
  for (const TuneEntry& e : tune_table())
    if (e.N == N && e.full == g.full && e.ssf == g.ssf && e.adj == (int)adj && e.dev == dev) return e.plan;
  if (!tune || capturing(s)) return 0;
  SweepRow r0;
  DepthTables dt;
  if (!first_row(rows, models, &r0, &dt)) return 0;
  const int T = std::min(dt.nz - 1, kTuneSteps);
  const size_t cap_rows = ((size_t)1 << 30) / ((size_t)(T + 1) * std::max(dt.nx, 1) * sizeof(float2) + 1);
  const int R = (int)std::min<size_t>((size_t)std::min(nrows, 4096), std::max<size_t>(cap_rows, 1));
  TuneEntry ent{N, g.full, g.ssf, (int)adj, dev, 0, R, T, {}};
  if (T < 1 || R < 1) {
    tune_table().push_back(ent);
    return 0;
  }
  bool ok;
  {
    TuneSample smp;
    ok = smp.make(r0, dt.nx, N, R, T, step, s);
    float best = 0.f;
    for (int k = 0; ok && k < np; ++k) {
      const float2* tw = nullptr;
      if (twiddles(N, k, &tw) != WEMFWI_OK) {
        ok = false;
        break;
      }
      float ms = 0.f;
      ok = smp.time([&] { return launch_oneway_plan<N>(k, smp.drows, R, models, g, nz, step, adj, tw, s); }, s, &ms);
      if (!ok) break;
      ent.ms.push_back(ms);
      if (k == 0 || ms < best) {
        best = ms;
        ent.plan = k;
      }
    }
  }
  if (!ok) {
    (void)hipGetLastError();
    ent.plan = 0;
    ent.ms.clear();
  }
  tune_table().push_back(ent);
  return ent.plan;
}

// true: launch plan k of the sweep (the runtime-planned passes when the FFT
// length g.fg.n is not N) with the half tile
template <int N>
static bool tuned_half(int k, const float2* tw, const SweepRow* rows, int nrows, const DepthTables* models,
                       StepGeom g, int nz, int step, bool adj, hipStream_t s, bool tune) {
  const bool gen = g.fg.n != N;
  if (!oneway_half_tile(N, gen)) return false;
  if (const char* e = std::getenv("WEMFWI_SWEEP_TILE"); e && *e) return std::string(e) == "half";
  if (env_int("WEMFWI_AUTOTUNE", 1) == 0 || nrows > kTileMaxRows) return false;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(tune_mutex());
  for (const TileEntry& e : tile_table())
    if (e.N == N && e.nk == g.fg.n && e.full == g.full && e.ssf == g.ssf && e.adj == (int)adj && e.pf == g.pf &&
        e.tdma == g.tdma && e.dev == dev && e.plan == k && e.rows == nrows)
      return e.half != 0;
  if (!tune || capturing(s)) return false;
  SweepRow r0;
  DepthTables dt;
  if (!first_row(rows, models, &r0, &dt)) return false;
  const size_t per = (size_t)nrows * std::max(dt.nx, 1) * sizeof(float2);  // one depth of every row
  const size_t cap = ((size_t)1 << 30) / per;                                // depths within 1 GiB
  const int T = std::min(std::min(dt.nz - 1, kTileSteps), (int)std::min<size_t>(cap > 0 ? cap - 1 : 0, kTileSteps));
  TileEntry ent{N, g.fg.n, g.full, g.ssf, (int)adj, g.pf, g.tdma, dev, k, nrows, T, 0, {0.f, 0.f}};
  if (T >= 2) {
    TuneSample smp;
    bool ok = smp.make(r0, dt.nx, N, nrows, T, step, s);
    ok = ok && smp.time(
                   [&] {
                     return gen ? launch_oneway_gen<N>(smp.drows, nrows, models, g, nz, step, adj, tw, s)
                                : launch_oneway_plan<N>(k, smp.drows, nrows, models, g, nz, step, adj, tw, s);
                   },
                   s, &ent.ms[0]);
    ok = ok && smp.time([&] { return launch_oneway_half(N, k, smp.drows, nrows, models, g, nz, step, adj, tw, s); },
                        s, &ent.ms[1]);
    if (ok) ent.half = ent.ms[1] < 0.98f * ent.ms[0] ? 1 : 0;
    else (void)hipGetLastError();
  }
  tile_table().push_back(ent);
  return ent.half != 0;
}

template <int N>
static hipError_t launch_oneway_n(const SweepRow* rows, int nrows, const DepthTables* models, StepGeom g, int nz,
                                  int step, bool adj, const float2* tw, hipStream_t s, bool tune) {
  if (g.fg.n != N) {
    if (tuned_half<N>(0, tw, rows, nrows, models, g, nz, step, adj, s, tune))
      return launch_oneway_half(N, 0, rows, nrows, models, g, nz, step, adj, tw, s);
    return launch_oneway_gen<N>(rows, nrows, models, g, nz, step, adj, tw, s);
  }
  int k = tuned_plan<N>(rows, nrows, models, g, nz, step, adj, s, tune);
  if (k != 0) {
    const float2* twk = nullptr;
    if ((tune ? twiddles(N, k, &twk) : twiddles_cached(N, k, &twk)) == WEMFWI_OK) tw = twk;
    else if (tune) return hipErrorInvalidValue;
    else k = 0;  // plan 0's table is the model's own (tw)
  }
  if (tuned_half<N>(k, tw, rows, nrows, models, g, nz, step, adj, s, tune))
    return launch_oneway_half(N, k, rows, nrows, models, g, nz, step, adj, tw, s);
  return launch_oneway_plan<N>(k, rows, nrows, models, g, nz, step, adj, tw, s);
}

// JSON list of the autotuner's decisions so far
std::string autotune_report() {
  std::lock_guard<std::mutex> lock(tune_mutex());
  std::string out = "[";
  for (size_t i = 0; i < tune_table().size(); ++i) {
    const TuneEntry& e = tune_table()[i];
    char head[256];
    std::snprintf(head, sizeof(head),
                  "%s{\"kernel\": \"oneway_kernel\", \"N\": %d, \"full\": %d, \"ssf\": %d, \"adjoint\": %d, "
                  "\"device\": %d, \"rows\": %d, \"steps\": %d, \"plan\": %d, \"ms\": [",
                  i ? ", " : "", e.N, e.full, e.ssf, e.adj, e.dev, e.rows, e.steps, e.plan);
    out += head;
    for (size_t k = 0; k < e.ms.size(); ++k) {
      char v[32];
      std::snprintf(v, sizeof(v), "%s%.4f", k ? ", " : "", e.ms[k]);
      out += v;
    }
    out += "]}";
  }
  for (size_t i = 0; i < tile_table().size(); ++i) {
    const TileEntry& e = tile_table()[i];
    char line[400];
    std::snprintf(line, sizeof(line),
                  "%s{\"kernel\": \"oneway_tile\", \"N\": %d, \"nk\": %d, \"full\": %d, \"ssf\": %d, \"adjoint\": %d, "
                  "\"pf\": %d, \"tdma\": %d, \"device\": %d, \"rows\": %d, \"steps\": %d, \"plan\": %d, "
                  "\"tile\": \"%s\", \"ms\": [%.4f, %.4f]}",
                  (i || !tune_table().empty()) ? ", " : "", e.N, e.nk, e.full, e.ssf, e.adj, e.pf, e.tdma, e.dev, e.rows,
                  e.steps, e.plan, e.half ? "half" : "full", e.ms[0], e.ms[1]);
    out += line;
  }
  return out + "]";
}

hipError_t launch_oneway(int N, const SweepRow* rows, int nrows, const DepthTables* models,
                         StepGeom g, int nz, int step, bool adj, const float2* tw, hipStream_t s, bool tune) {
  if (nrows <= 0 || nz < 2) return hipSuccess;
  switch (N) {
    case 128: return launch_oneway_n<128>(rows, nrows, models, g, nz, step, adj, tw, s, tune);
    case 256: return launch_oneway_n<256>(rows, nrows, models, g, nz, step, adj, tw, s, tune);
    case 512: return launch_oneway_n<512>(rows, nrows, models, g, nz, step, adj, tw, s, tune);
    case 1024: return launch_oneway_n<1024>(rows, nrows, models, g, nz, step, adj, tw, s, tune);
    case 2048: return launch_oneway_n<2048>(rows, nrows, models, g, nz, step, adj, tw, s, tune);
    default: return hipErrorInvalidValue;
  }
}

// single steps use the general variant (padding and split step allowed);
// FX: the compile-time pass chain of GEN rows (GenFixed id, 0: runtime)
template <int N, bool GEN, int FX = 0>
static hipError_t launch_onestep_v(const StepRow* rows, int nrows, const DepthTables* models,
                                   StepGeom g, bool adj, const float2* tw, hipStream_t s) {
  using V = Var<N, false, true, GEN, FX>;
  constexpr int kWavesPerBlock = WavesV<V>::value;
  dim3 grid((nrows + kWavesPerBlock - 1) / kWavesPerBlock), block(64 * kWavesPerBlock);
  if (adj)
    hipLaunchKernelGGL((onestep_kernel<V, true>), grid, block, 0, s, rows, nrows, models, g, tw);
  else
    hipLaunchKernelGGL((onestep_kernel<V, false>), grid, block, 0, s, rows, nrows, models, g, tw);
  return hipGetLastError();
}

template <int N>
static hipError_t launch_onestep_n(const StepRow* rows, int nrows, const DepthTables* models,
                                   StepGeom g, bool adj, const float2* tw, hipStream_t s) {
  if (g.fg.n != N)
    return gen_fixed_dispatch<N>(
        g.fg.n,
        [&](auto id) { return launch_onestep_v<N, true, decltype(id)::value>(rows, nrows, models, g, adj, tw, s); },
        [&] { return launch_onestep_v<N, true>(rows, nrows, models, g, adj, tw, s); });
  return launch_onestep_v<N, false>(rows, nrows, models, g, adj, tw, s);
}

hipError_t launch_onestep(int N, const StepRow* rows, int nrows, const DepthTables* models,
                          StepGeom g, bool adj, const float2* tw, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  switch (N) {
    case 128: return launch_onestep_n<128>(rows, nrows, models, g, adj, tw, s);
    case 256: return launch_onestep_n<256>(rows, nrows, models, g, adj, tw, s);
    case 512: return launch_onestep_n<512>(rows, nrows, models, g, adj, tw, s);
    case 1024: return launch_onestep_n<1024>(rows, nrows, models, g, adj, tw, s);
    case 2048: return launch_onestep_n<2048>(rows, nrows, models, g, adj, tw, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace wemfwi
