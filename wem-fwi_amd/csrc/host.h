// Host-side model object behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "kernels.h"

namespace wemfwi {

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int refsampler_host(int nz, int nx, const std::complex<float>* slow, int nref, std::complex<float>* sref,
                    int* refid);
int twiddles(int n, const float2** out);

// Per-call device buffers (wavefields, accumulators, descriptors) come from a
// per-device cache of freed blocks: cached_free synchronises the device, as
// hipFree does, and keeps the block for a later cached_malloc of a size it
// fits (at most 1/4 + 2 MiB larger), so repeated calls of the same shape do
// not return tens of GB to the driver and wait for it to clear them again.
// The cache holds at most WEMFWI_CACHE_LIMIT_MB (default half of the device
// memory); a failing hipMalloc empties it and retries.
int cached_malloc(void** p, size_t bytes, const char* what);
void cached_free(void* p);

// Device allocation of n elements (WEMFWI_ERR_ALLOC on failure).
template <typename T>
int dev_alloc(T** p, size_t n, const char* what) {
  *p = nullptr;
  if (n == 0) return WEMFWI_OK;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(WEMFWI_ERR_ALLOC, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
  }
  return WEMFWI_OK;
}

// Device allocation initialised from host memory.
template <typename T>
int dev_upload(T** p, const T* src, size_t n, const char* what) {
  int rc = dev_alloc(p, n, what);
  if (rc) return rc;
  if (n == 0) return WEMFWI_OK;
  hipError_t e = hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, what);
  return WEMFWI_OK;
}

// Grow-only device allocation.
struct Workspace {
  void* ptr = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes, const char* what);
  void release();
};

struct Model {
  wemfwi_model_desc d{};
  int nk = 0;        // FFT length nx + 2*tap
  int N = 0;         // lanes of a row: the built power of two (128 ... 2048) >= nk
  FftGen fg{};       // FFT plan of nk (runtime passes when nk != N)
  bool fft_ok = false;  // nk fits a built row and has a plan (propagation possible)
  int nslots = 0;    // active (depth, reference) pairs
  std::vector<std::complex<float>> slow, den, sref;
  std::vector<int> refid, slot_begin, slot_ref;

  int* d_slot_begin = nullptr;
  int* d_slot_ref = nullptr;
  signed char* d_rid_lane = nullptr;     // [nz][N] lane layout
  int* d_one_ref = nullptr;             // [nz]
  float2* d_phi_lane = nullptr;         // [nz][N] lane layout
  float* d_taper_lane = nullptr;        // [N] lane layout
  double2* d_slot_sref = nullptr;
  float* d_phi_max = nullptr;  // [nz] largest |Re phi| of each depth
  float2* d_scoef = nullptr;   // [ntaylor+1][nz][nx] scatterer coefficients
  float2* d_refl = nullptr;    // [nz][nx] reflection coefficient (Reflect.h:19-26)
  float2* d_drefl = nullptr;   // [4][nz][nx] a1, b1, a2, b2 of dReflect.h:18-27,40-47
  DepthTables* d_tables = nullptr;
  DepthTables h_tables{};
  const float2* d_tw = nullptr;

  Workspace ptab, wlist, rows;
  Workspace scratch;  // [nrows][N] per-row scratch rows of the sweep / step kernels
  // host staging of asynchronous uploads (kept alive until the call syncs)
  std::vector<double> h_wlist;
  std::vector<SweepRow> h_sweep;
  std::vector<StepRow> h_step;

  ~Model();
  // slow_refl: slowness of the Reflect coefficients when it differs from the
  // propagation slowness (NULL: the same)
  int build(const wemfwi_model_desc& desc, const std::complex<float>* slow, const std::complex<float>* den,
            const std::complex<float>* slow_refl = nullptr);
  StepGeom geom() const;
  std::vector<float> freqs;   // frequency set of the current phase table (sorted)
  unsigned table_gen = 0;     // incremented each time the phase table is rebuilt
  int NH() const { return N / 2 + 1; }
  int set_freqs(int n, const float* freq_hz, hipStream_t s);
  // recomputes the phase table of the current frequency set in place (one
  // kernel launch, no host copies: capturable)
  int enqueue_tables(hipStream_t s);
  int require_fft() const;  // WEMFWI_ERR_UNSUPPORTED unless fft_ok
  int phase_tables(int nrows, const float* freq_hz, std::vector<float>& wrow, std::vector<const float2*>& prow,
                   hipStream_t s);
  int oneway(int up, int adjoint, int nrows, const float* freq_hz, const std::complex<float>* in,
             std::complex<float>* out, int add, int recursive, hipStream_t s);
  // oneway on caller-provided device memory: scratch rows [nrows][N] (scr)
  // and nrows row descriptors (drows).  fixed: allocate nothing -- the phase
  // tables of every frequency must exist (set_freqs) and the sweep runs the
  // FFT plan already chosen (no autotuning launch)
  int oneway_on(int up, int adjoint, int nrows, const float* freq_hz, const std::complex<float>* in,
                std::complex<float>* out, int add, int recursive, float2* scr, SweepRow* drows, bool fixed,
                hipStream_t s);
  // frequencies of freq_hz[0..n) missing from the current phase table
  bool has_freqs(int n, const float* freq_hz) const;
  int onestep(int adjoint, int nrows, const float* freq_hz, const int* depth, const std::complex<float>* in,
              std::complex<float>* out, int add, hipStream_t s);
};

// Sweeps of rows with their own models (widths, bands) planned once: rows
// grouped by kernel variant, descriptors and scratch rows uploaded; run()
// launches the groups on side streams in a measured order.
struct BatchPlan {
  struct Group {
    int N, off, n;
    StepGeom g;
    int nz;
    const float2* tw;
  };
  std::vector<Model*> models;          // distinct models
  std::vector<unsigned> gens;          // their phase-table generations at planning
  std::vector<Group> groups;
  Workspace desc, scratch;
  const SweepRow* d_rows = nullptr;
  const DepthTables* d_tabs = nullptr;
  std::vector<hipEvent_t> events;      // fork / join events of enqueue (one more than groups)
  // launch order of the groups: largest FFT length first until the first
  // enqueue outside stream capture has timed the candidate orders (each
  // other group moved to the front) and kept the fastest, or found the
  // decision of an earlier plan of the same group shape.  Only plans whose
  // runs may be repeated are timed: no row reads memory that a row writes.
  std::vector<int> launch_order;
  std::vector<float> order_ms;         // best time of each candidate (empty: not timed)
  bool order_done = false;
  bool rerunnable = false;
  // a plan whose runs cannot be repeated and whose shape has no decision
  // yet: each run() (which waits for its work anyway) launches the candidate
  // order tried least so far for that shape and records its time; after two
  // runs of every candidate the shape has its decision
  bool online = false;
  std::vector<long long> shape;
  ~BatchPlan();
  // Optional per-row products of the sweep kernel (SweepRow smul, out2,
  // omul): arrays of nrows pointers, NULL for none
  struct RowProducts {
    const float2* const* smul = nullptr;
    float2* const* out2 = nullptr;
    const float2* const* omul = nullptr;
  };
  // src_lo / src_hi (optional, per row): only the source rows of those
  // depths are read (the sweep kernel's src_skip)
  int build(int nrows, Model* const* ms, const float* freq_hz, const std::complex<float>* const* in,
            std::complex<float>* const* out, hipStream_t s, const int* src_lo = nullptr,
            const int* src_hi = nullptr, const RowProducts* prod = nullptr);
  // WEMFWI_ERR_ARG if a model's phase table was rebuilt since planning
  int check() const;
  // launches the groups behind the work already on s and makes s wait for
  // them (no host synchronisation, so it can be captured into a graph;
  // the first call outside capture times the launch orders and waits for
  // them); tables = true first recomputes the models' phase tables in place
  int enqueue(int up, int adjoint, hipStream_t s, bool tables);
  int run(int up, int adjoint, hipStream_t s);

 private:
  int launch(int up, int adjoint, hipStream_t s, const std::vector<int>& ord);
  int tune_order(int up, int adjoint, hipStream_t s);
  int run_trial(int up, int adjoint, hipStream_t s);
  std::vector<std::vector<int>> candidates() const;
};

// JSON entries (each starting with ", ") of the launch-order decisions so far
std::string batch_order_report();

// cos^2 source/receiver spread weights of half-width ng (Injection.h:53-65)
std::vector<float> cos2_window(int ng);

}  // namespace wemfwi

struct wemfwi_model {
  wemfwi::Model m;
};
