// Shot x frequency operators of the reference (WEM, BornFull, BornRefl, the
// frequency-extended variants) and the remaining unit operators, built from
// the batched sweep / step / elementwise kernels.  Every (shot, frequency)
// pair is one row; rows are processed in chunks that fit device memory and
// reductions over rows are deterministic (fixed order, double precision).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <complex>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "born.h"
#include "host.h"
#include "kernels.h"

using cf = std::complex<float>;

namespace wemfwi {

// ------------------------------------------------------------------ survey

std::vector<float> cos2_window(int ng) {
  // Injection.h:53-65 (cos^2 spread; float arithmetic as the reference)
  std::vector<float> g(2 * ng + 1, 1.f);
  if (ng == 0) return g;
  const float pi = (float)(4.0 * std::atan(1.0));
  const int size = 2 * ng + 1;
  for (int i = 0; i < size / 2; ++i) g[i] = (-std::cos(pi / ng * i) + 1) / 2;
  for (int i = size / 2; i < size; ++i) g[i] = (std::cos(pi / ng * (i - size + ng + 1)) + 1) / 2;
  return g;
}

struct DevPlan {
  int nx = -1, nz = -1, ng = -1, offset = -1;
  InjPlan inj{};
  RecPlan rec{};
  std::vector<void*> allocs;
  void release() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    nx = -1;
  }
};

struct Survey {
  wemfwi_survey_desc d{};
  std::vector<int> src_iz, src_ix, rec_iz, rec_ix, bin;
  std::vector<float> freq;
  std::vector<cf> wave;
  DevPlan rplan;  // receivers on the current grid
  // per-shot device lists (conformal maps give every shot its own grid):
  // empty unless set; then shot s has sources ssrc_*[s] and receivers srec_*[s]
  std::vector<std::vector<int>> ssrc_iz, ssrc_ix, srec_iz, srec_ix;
  std::vector<DevPlan> ssrc_plan, srec_plan;
  int ckpt = 0;  // Born adjoints: depth steps between kept wavefield rows (0: keep all)
  bool per_shot() const { return !ssrc_iz.empty(); }
  // depths holding sources (receivers = false) or receivers of any shot:
  // [lo, hi], lo > hi when there are none
  void depth_range(bool receivers, int& lo, int& hi) const {
    lo = 1 << 30;
    hi = -1;
    auto take = [&](const std::vector<int>& v) {
      for (int z : v) {
        lo = std::min(lo, z);
        hi = std::max(hi, z);
      }
    };
    take(receivers ? rec_iz : src_iz);
    for (const auto& v : receivers ? srec_iz : ssrc_iz) take(v);
  }
  ~Survey() {
    rplan.release();
    for (auto& p : ssrc_plan) p.release();
    for (auto& p : srec_plan) p.release();
  }
};

template <typename T>
static int upload(std::vector<void*>& allocs, T** p, const std::vector<T>& h, const char* what) {
  *p = nullptr;
  if (h.empty()) return WEMFWI_OK;
  hipError_t e = hipMalloc((void**)p, h.size() * sizeof(T));
  if (e != hipSuccess) return fail(WEMFWI_ERR_ALLOC, std::string("hipMalloc failed for ") + what);
  allocs.push_back(*p);
  e = hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, what);
  return WEMFWI_OK;
}

// Receiver injection / recording plans of a device list on an nx-wide grid.
static int build_plan(DevPlan& P, const std::vector<int>& iz, const std::vector<int>& ix, int ng, int offset,
                      int nx, int nz) {
  if (P.nx == nx && P.nz == nz && P.ng == ng && P.offset == offset) return WEMFWI_OK;
  P.release();
  const std::vector<float> g = cos2_window(ng);
  // recording: contributions by device, in the order of Injection.cpp:69-80
  std::vector<int> rptr(1, 0);
  std::vector<long long> rcell;
  std::vector<float> rw;
  // injection: gather by cell; contributions listed in device order per cell
  std::vector<std::vector<std::pair<int, float>>> bycell;
  std::vector<long long> cells;
  std::vector<int> cell_of;  // temp map
  std::vector<long long> keys;
  for (size_t i = 0; i < ix.size(); ++i) {
    if (iz[i] < 0 || iz[i] >= nz) return fail(WEMFWI_ERR_ARG, "device depth index outside the model");
    for (int ig = -ng; ig <= ng; ++ig) {
      const int off = ix[i] + ig + offset;
      if (off < 0 || off >= nx) continue;
      const long long c = (long long)iz[i] * nx + off;
      rcell.push_back(c);
      rw.push_back(g[ig + ng]);
      keys.push_back(c);
    }
    rptr.push_back((int)rcell.size());
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  bycell.assign(keys.size(), {});
  for (size_t i = 0; i < ix.size(); ++i)
    for (int c = rptr[i]; c < rptr[i + 1]; ++c) {
      const size_t k = std::lower_bound(keys.begin(), keys.end(), rcell[c]) - keys.begin();
      bycell[k].push_back({(int)i, rw[c]});
    }
  std::vector<int> iptr(1, 0), idev;
  std::vector<float> iw;
  for (auto& v : bycell) {
    for (auto& pr : v) {
      idev.push_back(pr.first);
      iw.push_back(pr.second);
    }
    iptr.push_back((int)idev.size());
  }
  int rc;
  long long* d_cells;
  int *d_iptr, *d_idev, *d_rptr;
  float *d_iw, *d_rw;
  long long* d_rcell;
  if ((rc = upload(P.allocs, &d_cells, keys, "injection cells"))) return rc;
  if ((rc = upload(P.allocs, &d_iptr, iptr, "injection ptr"))) return rc;
  if ((rc = upload(P.allocs, &d_idev, idev, "injection devices"))) return rc;
  if ((rc = upload(P.allocs, &d_iw, iw, "injection weights"))) return rc;
  if ((rc = upload(P.allocs, &d_rptr, rptr, "recording ptr"))) return rc;
  if ((rc = upload(P.allocs, &d_rcell, rcell, "recording cells"))) return rc;
  if ((rc = upload(P.allocs, &d_rw, rw, "recording weights"))) return rc;
  P.inj = InjPlan{(int)keys.size(), d_cells, d_iptr, d_idev, d_iw};
  P.rec = RecPlan{(int)ix.size(), d_rptr, d_rcell, d_rw};
  P.nx = nx;
  P.nz = nz;
  P.ng = ng;
  P.offset = offset;
  return WEMFWI_OK;
}

// ------------------------------------------------------------------ batches

#define RCCHK_PUT(expr)    \
  do {                     \
    int _rc = (expr);      \
    if (_rc) return _rc;   \
  } while (0)

// Recording of one operator call's device buffers and host decisions, so
// that the same call can run again under stream capture: the recording pass
// runs eagerly and keeps every buffer it allocates (descriptor tables each
// in a buffer of their own, uploaded once); the replay pass takes the same
// buffers in the same order, uploads nothing, waits for nothing and makes
// the same decisions, so that only kernel launches and device copies reach
// the stream.  Set for the calling thread by the gradient plan while it
// builds (g_arena); null everywhere else.
struct Arena {
  bool replay = false;
  std::vector<std::pair<void*, size_t>> bufs;
  size_t next = 0;
  std::vector<long long> decisions;
  size_t next_decision = 0;
  ~Arena() {
    for (auto& b : bufs) cached_free(b.first);
  }
  int take(void** p, size_t bytes, const char* what) {
    if (!replay) {
      if (int rc = cached_malloc(p, bytes, what)) return rc;
      bufs.push_back({*p, bytes});
      return WEMFWI_OK;
    }
    if (next >= bufs.size() || bufs[next].second != bytes)
      return fail(WEMFWI_ERR_ARG, "recorded call: replay does not match the recording");
    *p = bufs[next++].first;
    return WEMFWI_OK;
  }
  long long decide(long long v) {
    if (!replay) {
      decisions.push_back(v);
      return v;
    }
    return next_decision < decisions.size() ? decisions[next_decision++] : v;
  }
  void rewind() {
    replay = true;
    next = next_decision = 0;
  }
};
static thread_local Arena* g_arena = nullptr;

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  bool recorded = false;  // owned by the arena of a recorded call
  ~DevBuf() {
    if (!recorded) cached_free(p);
  }
  int alloc(size_t bytes, const char* what) {
    if (!recorded) cached_free(p);
    p = nullptr;
    n = bytes;
    recorded = g_arena != nullptr;
    if (int rc = recorded ? g_arena->take(&p, bytes, what) : cached_malloc(&p, bytes, what)) {
      n = 0;
      return rc;
    }
    return WEMFWI_OK;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

template <typename T>
static int put(DevBuf& b, const std::vector<T>& h, hipStream_t s, const char* what) {
  if (g_arena) {
    // a buffer of its own for every table of a recorded call (a replayed
    // launch reads what it read when recorded), uploaded when recorded
    if (h.empty()) return WEMFWI_OK;
    RCCHK_PUT(b.alloc(h.size() * sizeof(T), what));
    if (g_arena->replay) return WEMFWI_OK;
    hipError_t e = hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return e == hipSuccess ? WEMFWI_OK : hip_fail(e, what);
  }
  if (b.n < h.size() * sizeof(T)) {
    int rc = b.alloc(h.size() * sizeof(T), what);
    if (rc) return rc;
  }
  if (h.empty()) return WEMFWI_OK;
  // the copy completes before returning, so the host vector may be a
  // temporary of the caller (descriptor tables are small)
  hipError_t e = hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, what);
}

#define HIPCHK(expr, what)                              \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hip_fail(_e, (what));  \
  } while (0)

#define RCCHK(expr)        \
  do {                     \
    int _rc = (expr);      \
    if (_rc) return _rc;   \
  } while (0)

// How the models of a call are indexed: one per frequency (the
// frequency-extended operators; a single model is nw copies of it) or one per
// shot (conformal maps).
enum class Per { Freq, Shot };

// A chunk of (shot, frequency) rows with their per-row models.
struct Rows {
  size_t first = 0;                    // position of row 0 in the call's row order
  std::vector<int> shot, fi;           // shot and frequency index of each row
  std::vector<int> slice;              // index of the row's model / perturbation slice
  std::vector<Model*> model;           // background model of each row
};

// Rows per chunk: half of the free device memory, or at most
// WEMFWI_CHUNK_ROWS rows when that variable is set (to bound memory, and to
// exercise chunked calls on small problems).
static size_t chunk_rows(size_t per_row_bytes, size_t nrows) {
  if (g_arena && g_arena->replay) return (size_t)g_arena->decide((long long)nrows);
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) fr = (size_t)8 << 30;
  const size_t budget = fr / 2;
  size_t c = per_row_bytes ? budget / per_row_bytes : nrows;
  if (const char* e = std::getenv("WEMFWI_CHUNK_ROWS")) {
    const long long cap = std::atoll(e);
    if (cap > 0) c = std::min(c, (size_t)cap);
  }
  if (c < 1) c = 1;
  c = std::min(c, nrows);
  // a replayed call cuts its rows as when it was recorded (free memory has
  // changed since)
  return g_arena ? (size_t)g_arena->decide((long long)c) : c;
}

// Context of one operator call: geometry shared by all models, device-side
// descriptor tables for the sweep / lin kernels.
struct Ctx {
  Model* m0;            // reference model (geometry)
  Survey* sv;
  hipStream_t s;
  int nz, nx, N;
  size_t nzx;
  std::vector<Model*> models;  // distinct models of the call
  std::vector<DepthTables> htables;  // host staging of d_models (alive until the call ends)
  DevBuf d_models;             // DepthTables[nmodels]
  DevBuf scr;                  // [rows][N] scratch rows of the sweep kernel
  std::vector<MulRow> hmul;    // batched products (host staging)
  DevBuf dmul;
  DevBuf d_rows;
  std::vector<SweepRow> hsw;
  std::vector<LinRow> hlin;
  std::vector<float2*> hp1, hp2, hp3;
  DevBuf dp1, dp2, dp3;
  std::vector<long long> hvb;
  DevBuf dvb;
  std::vector<int> hiz, hix;
  std::vector<float2> hval;
  DevBuf diz, dix, dval, dgw;
  std::vector<float> hgw;

  int init(const std::vector<Model*>& ms, Survey* survey, hipStream_t st) {
    models = ms;
    m0 = ms[0];
    sv = survey;
    s = st;
    nz = m0->d.nz;
    nx = m0->d.nx;
    N = m0->N;
    nzx = (size_t)nz * nx;
    RCCHK(m0->require_fft());
    for (Model* m : ms)
      if (m->d.nz != nz || m->d.nx != nx || m->N != N || m->d.tap != m0->d.tap || m->d.pspi != m0->d.pspi ||
          m->d.taper != m0->d.taper || m->d.ntaylor != m0->d.ntaylor)
        return fail(WEMFWI_ERR_ARG, "models of one call must share grid and parameters");
    htables.clear();
    for (Model* m : ms) htables.push_back(m->h_tables);
    RCCHK(put(d_models, htables, s, "model tables"));
    // frequency sets: every model holds the survey's frequencies
    for (Model* m : ms) {
      bool have = m->freqs.size() == sv->freq.size();
      if (have) {
        std::vector<float> f = sv->freq;
        std::sort(f.begin(), f.end());
        f.erase(std::unique(f.begin(), f.end()), f.end());
        have = f == m->freqs;
      }
      if (!have) RCCHK(m->set_freqs((int)sv->freq.size(), sv->freq.data(), s));
    }
    return WEMFWI_OK;
  }
  // distinct phase tables ((model, frequency) pairs) of a chunk's rows
  static size_t distinct_tables(const Rows& r) {
    std::vector<std::pair<Model*, int>> v;
    for (size_t i = 0; i < r.shot.size(); ++i) v.push_back({r.model[i], r.fi[i]});
    std::sort(v.begin(), v.end());
    return (size_t)(std::unique(v.begin(), v.end()) - v.begin());
  }
  int model_index(Model* m) const {
    return (int)(std::find(models.begin(), models.end(), m) - models.begin());
  }
  const float2* ptab(Model* m, int fi) const {
    const size_t k = std::lower_bound(m->freqs.begin(), m->freqs.end(), sv->freq[fi]) - m->freqs.begin();
    return (const float2*)m->ptab.ptr + k * (size_t)std::max(m->nslots, 1) * m->NH();
  }
  float omega(Model* m, int fi) const {
    const size_t k = std::lower_bound(m->freqs.begin(), m->freqs.end(), sv->freq[fi]) - m->freqs.begin();
    return (float)m->h_wlist[k];
  }

  // in-place / recursive sweep of W (rows x nzx) with sources src;
  // smul: per row, an [nz][nx] table multiplying every source row (and the
  // initial row) as it is read; q2: rows (layout of W) that also receive
  // conj(r) * the result, r the row model's reflection coefficients
  // src_rows: only the source rows of depths [lo, hi] are read (nullptr: all)
  int sweep(const Rows& r, int up, int adjoint, const float2* src, float2* W,
            const std::vector<const float2*>* smul = nullptr, float2* q2 = nullptr,
            const std::pair<int, int>* src_rows = nullptr) {
    const int n = (int)r.shot.size();
    hsw.resize(n);
    for (int i = 0; i < n; ++i) {
      hsw[i].in = src + (size_t)i * nzx;
      hsw[i].out = W + (size_t)i * nzx;
      hsw[i].ptab = ptab(r.model[i], r.fi[i]);
      hsw[i].w = omega(r.model[i], r.fi[i]);
      hsw[i].model = model_index(r.model[i]);
      hsw[i].recursive = 1;
      hsw[i].first = -1;
      hsw[i].nsteps = -1;
      hsw[i].init = nullptr;
      hsw[i].zbase = 0;
      hsw[i].smul = smul ? (*smul)[i] : nullptr;
      hsw[i].out2 = q2 ? q2 + (size_t)i * nzx : nullptr;
      hsw[i].omul = q2 ? r.model[i]->d_refl : nullptr;
      hsw[i].src_skip = src_rows ? 1 : 0;
      hsw[i].src_lo = src_rows ? src_rows->first : 0;
      hsw[i].src_hi = src_rows ? src_rows->second : -1;
    }
    if (scr.n < (size_t)n * N * sizeof(float2)) RCCHK(scr.alloc((size_t)n * N * sizeof(float2), "scratch rows"));
    for (int i = 0; i < n; ++i) hsw[i].scratch = scr.as<float2>() + (size_t)i * N;
    RCCHK(put(d_rows, hsw, s, "sweep rows"));
    const int step = ((!up) != (adjoint != 0)) ? 1 : -1;
    StepGeom g = m0->geom();
    g.pf = (!src_rows && !smul) ? 1 : 0;  // plain source rows at every depth
    g.tdma = table_dma((size_t)n, distinct_tables(r));
    HIPCHK(launch_oneway(N, d_rows.as<SweepRow>(), n, d_models.as<DepthTables>(), g, nz, step, adjoint != 0,
                         m0->d_tw, s),
           "oneway kernel");
    return sync();
  }

  // LinDown (dir=+1) / LinUp (dir=-1); post_refl (adjoint): each bg row,
  // once used, is replaced by r * bg (r the row model's reflection
  // coefficient); assign (adjoint): out rows are stored, not added to
  int lin(const Rows& r, int dir, int adjoint, int exact, float2* bg, const float2* m, size_t m_stride,
          float2* out, bool post_refl = false, bool assign = false) {
    const int n = (int)r.shot.size();
    hlin.resize(n);
    for (int i = 0; i < n; ++i) {
      hlin[i].bg = bg + (size_t)i * nzx;
      hlin[i].m = m + (adjoint ? (size_t)i * nzx : (size_t)0) + (adjoint ? 0 : m_stride * (size_t)r.slice[i]);
      hlin[i].out = out + (size_t)i * nzx;
      hlin[i].ptab = ptab(r.model[i], r.fi[i]);
      hlin[i].w = omega(r.model[i], r.fi[i]);
      hlin[i].model = model_index(r.model[i]);
      hlin[i].zbase = 0;
      hlin[i].post = post_refl ? r.model[i]->d_refl : nullptr;
      hlin[i].assign = assign ? 1 : 0;
    }
    RCCHK(put(d_rows, hlin, s, "lin rows"));
    HIPCHK(launch_lin(N, d_rows.as<LinRow>(), n, d_models.as<DepthTables>(), m0->geom(), nz, dir, adjoint != 0,
                      exact != 0, m0->d_tw, s),
           "lin kernel");
    return sync();
  }

  // In-place sweep of a depth segment: row i of W holds depths zbase..
  // (wstride samples apart from row to row); the sweep runs nsteps steps
  // from depth first, starting from init + i * init_stride (NULL: the
  // buffer's own row first), and adds the source rows already in W.
  // src_rows: only the source rows of depths [lo, hi] are read (nullptr: all)
  int sweep_seg(const Rows& r, int up, int adjoint, float2* W, size_t wstride, int zbase, int first, int nsteps,
                const float2* init, size_t init_stride, float2* q2 = nullptr,
                const std::pair<int, int>* src_rows = nullptr) {
    const int n = (int)r.shot.size();
    hsw.resize(n);
    for (int i = 0; i < n; ++i) {
      hsw[i].in = W + (size_t)i * wstride;
      hsw[i].out = W + (size_t)i * wstride;
      hsw[i].ptab = ptab(r.model[i], r.fi[i]);
      hsw[i].w = omega(r.model[i], r.fi[i]);
      hsw[i].model = model_index(r.model[i]);
      hsw[i].recursive = 1;
      hsw[i].first = first;
      hsw[i].nsteps = nsteps;
      hsw[i].init = init ? init + (size_t)i * init_stride : nullptr;
      hsw[i].zbase = zbase;
      hsw[i].smul = nullptr;
      hsw[i].out2 = q2 ? q2 + (size_t)i * wstride : nullptr;
      hsw[i].omul = q2 ? r.model[i]->d_refl : nullptr;
      hsw[i].src_skip = src_rows ? 1 : 0;
      hsw[i].src_lo = src_rows ? src_rows->first : 0;
      hsw[i].src_hi = src_rows ? src_rows->second : -1;
    }
    if (scr.n < (size_t)n * N * sizeof(float2)) RCCHK(scr.alloc((size_t)n * N * sizeof(float2), "scratch rows"));
    for (int i = 0; i < n; ++i) hsw[i].scratch = scr.as<float2>() + (size_t)i * N;
    RCCHK(put(d_rows, hsw, s, "sweep rows"));
    const int step = ((!up) != (adjoint != 0)) ? 1 : -1;
    StepGeom g = m0->geom();
    g.pf = src_rows ? 0 : 1;  // plain source rows at every depth
    g.tdma = table_dma((size_t)n, distinct_tables(r));
    HIPCHK(launch_oneway(N, d_rows.as<SweepRow>(), n, d_models.as<DepthTables>(), g, nz, step, adjoint != 0,
                         m0->d_tw, s),
           "oneway kernel");
    return WEMFWI_OK;
  }

  // Adjoint LinDown / LinUp over depths [dlo, dlo + dcount) of segment
  // buffers (bg, m and out share the row layout of sweep_seg).
  int lin_seg(const Rows& r, int dir, int exact, float2* bg, const float2* m, float2* out, size_t wstride,
              int zbase, int dlo, int dcount, bool post_refl = false, bool assign = false) {
    if (dcount <= 0) return WEMFWI_OK;
    const int n = (int)r.shot.size();
    hlin.resize(n);
    for (int i = 0; i < n; ++i) {
      hlin[i].bg = bg + (size_t)i * wstride;
      hlin[i].m = m + (size_t)i * wstride;
      hlin[i].out = out + (size_t)i * wstride;
      hlin[i].ptab = ptab(r.model[i], r.fi[i]);
      hlin[i].w = omega(r.model[i], r.fi[i]);
      hlin[i].model = model_index(r.model[i]);
      hlin[i].zbase = zbase;
      hlin[i].post = post_refl ? r.model[i]->d_refl : nullptr;
      hlin[i].assign = assign ? 1 : 0;
    }
    RCCHK(put(d_rows, hlin, s, "lin rows"));
    HIPCHK(launch_lin(N, d_rows.as<LinRow>(), n, d_models.as<DepthTables>(), m0->geom(), nz, dir, true, exact != 0,
                      m0->d_tw, s, dlo, dcount),
           "lin kernel");
    return WEMFWI_OK;
  }

  // n rows of nx samples: dst + i * dpitch <- src + i * spitch
  int copy_rows(float2* dst, size_t dpitch, const float2* src, size_t spitch, int n) {
    HIPCHK(hipMemcpy2DAsync(dst, dpitch * sizeof(float2), src, spitch * sizeof(float2), (size_t)nx * sizeof(float2),
                            (size_t)n, hipMemcpyDeviceToDevice, s),
           "copy rows");
    return WEMFWI_OK;
  }

  // rows: out_i (= or +=) op(A_i) * B_i over len samples, one launch
  int mul_rows(const std::vector<MulRow>& rows, size_t len, bool conjA, bool acc) {
    if (rows.empty()) return WEMFWI_OK;
    hmul = rows;
    RCCHK(put(dmul, hmul, s, "product rows"));
    HIPCHK(launch_mul_rows(dmul.as<MulRow>(), (int)hmul.size(), (long long)len, conjA, acc, s), "mul rows");
    return WEMFWI_OK;
  }

  // Rows of the per-row fields an injection writes: every field holds the
  // depths zbase.. with a stride of wstride samples between rows; sources
  // outside [zlo, zhi] are skipped (the whole field by default).
  struct Win {
    long long wstride;
    int zbase, zlo, zhi;
  };
  Win whole() const { return Win{(long long)nzx, 0, 0, nz - 1}; }

  // consecutive rows of one shot: [i0, i1)
  template <typename F>
  int shot_runs(const Rows& r, F&& f) {
    const int n = (int)r.shot.size();
    for (int i0 = 0; i0 < n;) {
      int i1 = i0 + 1;
      while (i1 < n && r.shot[i1] == r.shot[i0]) ++i1;
      RCCHK(f(i0, i1));
      i0 = i1;
    }
    return WEMFWI_OK;
  }

  // sources of the rows' shots, value = wavelet of the row's frequency
  int inject_sources(const Rows& r, float2* W, int ng, int offset) { return inject_sources(r, W, ng, offset, whole()); }
  int inject_sources(const Rows& r, float2* W, int ng, int offset, const Win& wn) {
    if (sv->per_shot()) return inject_shot_sources(r, W, ng, offset, wn);
    const int n = (int)r.shot.size();
    hiz.resize(n);
    hix.resize(n);
    hval.resize(n);
    for (int i = 0; i < n; ++i) {
      hiz[i] = sv->src_iz[r.shot[i]];
      hix[i] = sv->src_ix[r.shot[i]];
      const cf v = sv->wave[r.fi[i]];
      hval[i] = make_float2(v.real(), v.imag());
      if (hiz[i] < 0 || hiz[i] >= nz) return fail(WEMFWI_ERR_ARG, "source depth index outside the model");
    }
    hgw = cos2_window(ng);
    RCCHK(put(diz, hiz, s, "source iz"));
    RCCHK(put(dix, hix, s, "source ix"));
    RCCHK(put(dval, hval, s, "source values"));
    RCCHK(put(dgw, hgw, s, "source spread"));
    HIPCHK(launch_inject_point(W, wn.wstride, nx, diz.as<int>(), dix.as<int>(), dval.as<float2>(),
                               dgw.as<float>(), ng, offset, n, s, wn.zbase, wn.zlo, wn.zhi),
           "inject sources");
    return WEMFWI_OK;
  }

  // per-shot source lists: gather by cell over each shot's plan, every
  // device of a row carrying the row's wavelet value
  int inject_shot_sources(const Rows& r, float2* W, int ng, int offset, const Win& wn) {
    const int n = (int)r.shot.size();
    hval.resize(n);
    hvb.resize(n);
    for (int i = 0; i < n; ++i) {
      const cf v = sv->wave[r.fi[i]];
      hval[i] = make_float2(v.real(), v.imag());
      hvb[i] = i;
    }
    RCCHK(put(dval, hval, s, "source values"));
    RCCHK(put(dvb, hvb, s, "source value index"));
    return shot_runs(r, [&](int i0, int i1) {
      const int sh = r.shot[i0];
      RCCHK(build_plan(sv->ssrc_plan[sh], sv->ssrc_iz[sh], sv->ssrc_ix[sh], ng, offset, nx, nz));
      HIPCHK(launch_inject_csr(W + (size_t)i0 * wn.wstride, wn.wstride, sv->ssrc_plan[sh].inj,
                               dval.as<float2>() + i0, dvb.as<long long>(), 0, i1 - i0, s,
                               (long long)wn.zbase * nx, (long long)wn.zlo * nx, ((long long)wn.zhi + 1) * nx),
             "inject sources");
      return WEMFWI_OK;
    });
  }

  int vbase(const Rows& r) {
    const int n = (int)r.shot.size();
    hvb.resize(n);
    for (int i = 0; i < n; ++i)
      hvb[i] = (long long)r.shot[i] * sv->d.nr * sv->d.nt + sv->bin[r.fi[i]];
    return put(dvb, hvb, s, "data offsets");
  }

  // W_i += R^H d (receivers of the rows' traces)
  int inject_receivers(const Rows& r, float2* W, const float2* data, int ng, int offset) {
    return inject_receivers(r, W, data, ng, offset, whole());
  }
  int inject_receivers(const Rows& r, float2* W, const float2* data, int ng, int offset, const Win& wn) {
    const long long cb = (long long)wn.zbase * nx, clo = (long long)wn.zlo * nx, chi = ((long long)wn.zhi + 1) * nx;
    if (sv->per_shot()) {
      RCCHK(vbase(r));
      return shot_runs(r, [&](int i0, int i1) {
        const int sh = r.shot[i0];
        RCCHK(build_plan(sv->srec_plan[sh], sv->srec_iz[sh], sv->srec_ix[sh], ng, offset, nx, nz));
        HIPCHK(launch_inject_csr(W + (size_t)i0 * wn.wstride, wn.wstride, sv->srec_plan[sh].inj, data,
                                 dvb.as<long long>() + i0, sv->d.nt, i1 - i0, s, cb, clo, chi),
               "inject receivers");
        return WEMFWI_OK;
      });
    }
    RCCHK(build_plan(sv->rplan, sv->rec_iz, sv->rec_ix, ng, offset, nx, nz));
    RCCHK(vbase(r));
    HIPCHK(launch_inject_csr(W, wn.wstride, sv->rplan.inj, data, dvb.as<long long>(), sv->d.nt,
                             (int)r.shot.size(), s, cb, clo, chi),
           "inject receivers");
    return WEMFWI_OK;
  }

  // data += R W_i
  int record(const Rows& r, float2* data, const float2* W, int ng, int offset) {
    if (sv->per_shot()) {
      RCCHK(vbase(r));
      RCCHK(shot_runs(r, [&](int i0, int i1) {
        const int sh = r.shot[i0];
        RCCHK(build_plan(sv->srec_plan[sh], sv->srec_iz[sh], sv->srec_ix[sh], ng, offset, nx, nz));
        HIPCHK(launch_record(data, dvb.as<long long>() + i0, sv->d.nt, W + (size_t)i0 * nzx, (long long)nzx,
                             sv->srec_plan[sh].rec, i1 - i0, s),
               "record receivers");
        return WEMFWI_OK;
      }));
      return sync();
    }
    RCCHK(build_plan(sv->rplan, sv->rec_iz, sv->rec_ix, ng, offset, nx, nz));
    RCCHK(vbase(r));
    HIPCHK(launch_record(data, dvb.as<long long>(), sv->d.nt, W, (long long)nzx, sv->rplan.rec,
                         (int)r.shot.size(), s),
           "record receivers");
    return sync();
  }

  // the row models' reflection coefficient tables
  static std::vector<const float2*> refl_rows(const Rows& r) {
    std::vector<const float2*> v(r.shot.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = r.model[i]->d_refl;
    return v;
  }

  // zeroes depths [lo, hi] of n rows of nzx samples (nothing when lo > hi);
  // with a window (wstride samples per row, first depth zbase) its rows
  int zero_depths(float2* W, int n, int lo, int hi, size_t wstride = 0, int zbase = 0) {
    if (lo > hi || n <= 0) return WEMFWI_OK;
    const size_t pitch = wstride ? wstride : nzx;
    return zero_rows(W + (size_t)(lo - zbase) * nx, pitch, (size_t)(hi - lo + 1) * nx, n);
  }
  // n rows of width samples, pitch samples apart (kernels, not memsets:
  // a recorded call may be captured in a graph)
  int zero_rows(float2* p, size_t pitch, size_t width, int n) {
    HIPCHK(launch_zero2d(p, (long long)pitch, (long long)width, n, s), "zero rows");
    return WEMFWI_OK;
  }
  // bytes: a multiple of 8
  int zero(void* p, size_t bytes) {
    HIPCHK(launch_zero2d(p, (long long)(bytes / 8), (long long)(bytes / 8), 1, s), "zero");
    return WEMFWI_OK;
  }
  int sync() {
    if (g_arena) return WEMFWI_OK;  // a recorded call is one stream-ordered sequence
    HIPCHK(hipStreamSynchronize(s), "stream sync");
    return WEMFWI_OK;
  }
};

static std::vector<Rows> make_chunks(int ns, int nw, const std::vector<Model*>& fm, Per per, size_t per_row,
                                     bool freq_major) {
  std::vector<std::pair<int, int>> all;
  if (freq_major) {
    for (int i = 0; i < nw; ++i)
      for (int s = 0; s < ns; ++s) all.push_back({s, i});
  } else {
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < nw; ++i) all.push_back({s, i});
  }
  size_t c = chunk_rows(per_row, all.size());
  // whole reduction runs per chunk, so that sums do not depend on the cut
  if (c < all.size() && c >= (size_t)kRunRows) c -= c % kRunRows;
  std::vector<Rows> out;
  for (size_t b = 0; b < all.size(); b += c) {
    Rows r;
    r.first = b;
    for (size_t k = b; k < std::min(all.size(), b + c); ++k) {
      r.shot.push_back(all[k].first);
      r.fi.push_back(all[k].second);
      const int sl = per == Per::Shot ? all[k].first : all[k].second;
      r.slice.push_back(sl);
      r.model.push_back(fm[sl]);
    }
    out.push_back(r);
  }
  return out;
}

// Every source and receiver depth of the survey (shared and per-shot lists)
// must lie inside the model before anything is queued: the depth-only clears
// and the injection windows address wavefield rows by these indices.
static int check_survey(Survey* sv, Model* m) {
  if (!sv || !m) return fail(WEMFWI_ERR_ARG, "null survey or model");
  const int nz = m->d.nz;
  auto inside = [nz](const std::vector<int>& v) {
    for (int z : v)
      if (z < 0 || z >= nz) return false;
    return true;
  };
  bool ok = inside(sv->src_iz) && inside(sv->rec_iz);
  for (const auto& v : sv->ssrc_iz) ok = ok && inside(v);
  for (const auto& v : sv->srec_iz) ok = ok && inside(v);
  if (!ok) return fail(WEMFWI_ERR_ARG, "device depth index outside the model");
  return WEMFWI_OK;
}

// ------------------------------------------------------------------ WEM

// WEM::full_fwd (wem_prop.cpp:52-74) for every (shot, frequency) row:
// inject, Down, record, r*W, Up, record.
static int wem_forward(const std::vector<Model*>& fm, Per per, Survey* sv, float2* data, int add, hipStream_t s) {
  RCCHK(check_survey(sv, fm[0]));
  Ctx c;
  std::vector<Model*> uniq;
  for (Model* m : fm)
    if (std::find(uniq.begin(), uniq.end(), m) == uniq.end()) uniq.push_back(m);
  RCCHK(c.init(uniq, sv, s));
  const size_t dbytes = (size_t)sv->d.ns * sv->d.nr * sv->d.nt * sizeof(float2);
  if (!add) RCCHK(c.zero(data, dbytes));
  auto chunks = make_chunks(sv->d.ns, sv->d.nw, fm, per, c.nzx * sizeof(float2), false);
  DevBuf W;
  RCCHK(W.alloc(chunks[0].shot.size() * c.nzx * sizeof(float2), "wavefield"));
  float2* w = W.as<float2>();
  std::pair<int, int> srows;
  sv->depth_range(false, srows.first, srows.second);
  for (const Rows& r : chunks) {
    const int n = (int)r.shot.size();
    // only the source depths are read as sources: only they are cleared
    RCCHK(c.zero_depths(w, n, srows.first, srows.second));
    RCCHK(c.inject_sources(r, w, sv->d.ngs, sv->d.offset));
    RCCHK(c.sweep(r, 0, 0, w, w, nullptr, nullptr, &srows));
    RCCHK(c.record(r, data, w, sv->d.ngr, sv->d.offset));
    {
      const auto rr = Ctx::refl_rows(r);
      RCCHK(c.sweep(r, 1, 0, w, w, &rr));   // Up with sources r * W
    }
    RCCHK(c.record(r, data, w, sv->d.ngr, sv->d.offset));
  }
  return c.sync();
}

// ------------------------------------------------------------------ Born

// dReflect of the perturbation of each model (one per frequency for the
// extended operator): dr_m = dReflect(dS_m, dR_m).
// BornFull::full_fwd (born_prop.cpp:75-113) per (shot, frequency) row:
//   D  = Down(inj_src)
//   s1 = Down(LinDown(dS; D)),                          record s1
//   s2 = Up(D * dr + r * s1),                           record s2
//        (Born_refl::forward and the down-scattering term share one Up)
//   U  = Up(r * D)
//   s3 = Up(LinUp(dS; U)),                              record s3
static int born_forward(const std::vector<Model*>& fm, Per per, Survey* sv, const float2* dS, const float2* dR,
                        size_t mstride, float2* data, int add, hipStream_t s) {
  RCCHK(check_survey(sv, fm[0]));
  std::vector<Model*> uniq;
  for (Model* m : fm)
    if (std::find(uniq.begin(), uniq.end(), m) == uniq.end()) uniq.push_back(m);
  Ctx c;
  RCCHK(c.init(uniq, sv, s));
  const size_t dbytes = (size_t)sv->d.ns * sv->d.nr * sv->d.nt * sizeof(float2);
  if (!add) RCCHK(c.zero(data, dbytes));
  const int nw = sv->d.nw;
  // reflectivity perturbation of every model slice (frequency or shot)
  const size_t nslice = mstride ? fm.size() : 1;
  DevBuf drb;
  RCCHK(drb.alloc(nslice * c.nzx * sizeof(float2), "dr"));
  for (size_t k = 0; k < nslice; ++k) {
    Model* m = fm[k];
    HIPCHK(launch_dreflect_fwd(drb.as<float2>() + k * c.nzx, m->d_drefl, dS + k * mstride, dR + k * mstride, c.nz,
                               c.nx, false, s),
           "dreflect");
  }
  auto chunks = make_chunks(sv->d.ns, nw, fm, per, 3 * c.nzx * sizeof(float2), false);
  const size_t cap = chunks[0].shot.size() * c.nzx * sizeof(float2);
  DevBuf B, S1, T;
  RCCHK(B.alloc(cap, "background"));
  RCCHK(S1.alloc(cap, "scattered"));
  RCCHK(T.alloc(cap, "scattered"));
  float2 *bg = B.as<float2>(), *s1 = S1.as<float2>(), *tt = T.as<float2>();
  std::pair<int, int> srows;
  sv->depth_range(false, srows.first, srows.second);
  for (const Rows& r : chunks) {
    const int n = (int)r.shot.size();
    const size_t bytes = n * c.nzx * sizeof(float2);
    RCCHK(c.zero_depths(bg, n, srows.first, srows.second));   // only the source depths are read as sources
    RCCHK(c.inject_sources(r, bg, sv->d.ngs, sv->d.offset));
    RCCHK(c.sweep(r, 0, 0, bg, bg, nullptr, nullptr, &srows));
    RCCHK(c.zero(s1, bytes));
    RCCHK(c.lin(r, +1, 0, 0, bg, dS, mstride, s1));
    RCCHK(c.sweep(r, 0, 0, s1, s1));
    RCCHK(c.record(r, data, s1, sv->d.ngr, sv->d.offset));
    {
      std::vector<MulRow> ic(n), rs1(n);
      for (int i = 0; i < n; ++i) {
        const size_t o = (size_t)i * c.nzx;
        const float2* dri = drb.as<float2>() + (mstride ? (size_t)r.slice[i] * c.nzx : 0);
        ic[i] = MulRow{tt + o, bg + o, dri};
        rs1[i] = MulRow{tt + o, r.model[i]->d_refl, s1 + o};
      }
      RCCHK(c.mul_rows(ic, c.nzx, false, false));   // D * dr      (Born_refl::forward)
      RCCHK(c.mul_rows(rs1, c.nzx, false, true));   // + r * s1    (down-scattering)
    }
    RCCHK(c.sweep(r, 1, 0, tt, tt));
    RCCHK(c.record(r, data, tt, sv->d.ngr, sv->d.offset));
    {
      const auto rr = Ctx::refl_rows(r);
      RCCHK(c.sweep(r, 1, 0, bg, bg, &rr));         // U = Up(r * D), the product taken as it is read
    }
    RCCHK(c.zero(s1, bytes));
    RCCHK(c.lin(r, -1, 0, 0, bg, dS, mstride, s1));
    RCCHK(c.sweep(r, 1, 0, s1, s1));
    RCCHK(c.record(r, data, s1, sv->d.ngr, sv->d.offset));
  }
  return c.sync();
}

// Reduction of a chunk's per-row contributions into the FP64 accumulators,
// grouped by model slice (consecutive rows of one slice) or over all rows,
// in runs of the call's row order (launch_reduce_rows), so the sum does not
// depend on how the rows were cut into chunks or depth segments.
//   depths [z0, z1) of row i at in + i * stride + (z0 - zbase) * nx
static int reduce_chunk(Ctx& c, const Rows& r, bool sliced, double2* acc, const float2* in, size_t stride, int zbase,
                        int z0, int z1, const float2* in2 = nullptr) {
  const int n = (int)r.shot.size();
  int i0 = 0;
  while (i0 < n) {
    int i1 = i0 + 1;
    if (sliced)
      while (i1 < n && r.slice[i1] == r.slice[i0]) ++i1;
    else
      i1 = n;
    const size_t slot = sliced ? (size_t)r.slice[i0] : 0;
    const size_t off = (size_t)i0 * stride + (size_t)(z0 - zbase) * c.nx;
    HIPCHK(launch_reduce_rows(acc + slot * c.nzx + (size_t)z0 * c.nx, in + off, (long long)stride,
                              (long long)(z1 - z0) * c.nx, i1 - i0, (int)((r.first + i0) % kRunRows), true, c.s,
                              in2 ? in2 + off : nullptr),
           "reduce");
    i0 = i1;
  }
  return WEMFWI_OK;
}

static int born_adjoint_all(Ctx& c, const std::vector<Model*>& fm, Per per, double2* accS, double2* accDR,
                            bool sliced, const float2* data, int exact, bool ext_rowlast) {
  Survey* sv = c.sv;
  const int nw = sv->d.nw, ns = sv->d.ns;
  // rows of one model slice contiguous: frequency-major for per-frequency
  // models, shot-major for per-shot models
  auto chunks = make_chunks(ns, nw, fm, per, 4 * c.nzx * sizeof(float2), per == Per::Freq);
  const size_t cap = chunks[0].shot.size() * c.nzx * sizeof(float2);
  DevBuf B, Rb, Qb, CS;
  RCCHK(B.alloc(cap, "background"));
  RCCHK(Rb.alloc(cap, "receiver wavefield"));
  RCCHK(Qb.alloc(cap, "receiver wavefield"));
  RCCHK(CS.alloc(cap, "dS contributions"));
  float2 *bg = B.as<float2>(), *R = Rb.as<float2>(), *Q = Qb.as<float2>(), *cs = CS.as<float2>();
  const size_t last = (size_t)(c.nz - 1) * c.nx;
  std::pair<int, int> srows, rrows;
  sv->depth_range(false, srows.first, srows.second);
  sv->depth_range(true, rrows.first, rrows.second);
  for (const Rows& r : chunks) {
    const int n = (int)r.shot.size();
    // the sweeps read only the source / receiver depths as sources: only
    // those rows are cleared before the injection
    RCCHK(c.zero_depths(bg, n, srows.first, srows.second));
    RCCHK(c.inject_sources(r, bg, sv->d.ngs, sv->d.offset));
    RCCHK(c.sweep(r, 0, 0, bg, bg, nullptr, nullptr, &srows));
    RCCHK(c.zero_depths(R, n, rrows.first, rrows.second));
    RCCHK(c.inject_receivers(r, R, data, sv->d.ngr, sv->d.offset));
    RCCHK(c.sweep(r, 1, 1, R, R, nullptr, Q, &rrows));   // R, and Q = conj(r) R on every row
    // dr += conj(D) * R (Born_refl::adjoint), summed straight from D and R
    // before LinDown^H rescales D
    RCCHK(reduce_chunk(c, r, sliced, accDR, bg, c.nzx, 0, 0, c.nz, R));
    RCCHK(c.inject_receivers(r, Q, data, sv->d.ngr, sv->d.offset));   // Q = conj(r) R + R^H d
    RCCHK(c.sweep(r, 0, 1, Q, Q));
    // LinDown^H stores rows 0..nz-2 of cs (no fill); row nz-1 starts at 0
    RCCHK(c.zero_rows(cs + last, c.nzx, c.nx, n));
    // LinDown^H; each D row d < nz-1 it reads becomes r * D (the Up source)
    RCCHK(c.lin(r, +1, 1, exact, bg, Q, 0, cs, true, true));
    {
      std::vector<MulRow> rbg(n);
      for (int i = 0; i < n; ++i) {
        const size_t o = (size_t)i * c.nzx + last;
        rbg[i] = MulRow{bg + o, r.model[i]->d_refl + last, bg + o};
      }
      RCCHK(c.mul_rows(rbg, c.nx, false, false));  // r * D at the last depth
    }
    RCCHK(c.sweep(r, 1, 0, bg, bg));
    RCCHK(c.lin(r, -1, 1, exact, bg, R, 0, cs));
    if (ext_rowlast) {
      for (int i = 0; i < n; ++i)
        if (r.shot[i] != ns - 1)
          RCCHK(c.zero_rows(cs + (size_t)i * c.nzx + (size_t)(c.nz - 1) * c.nx, c.nx, c.nx, 1));
    }
    RCCHK(reduce_chunk(c, r, sliced, accS, cs, c.nzx, 0, 0, c.nz));
    RCCHK(c.sync());
  }
  return WEMFWI_OK;
}

// dReflect^H of the summed dr, then write the model outputs
static int born_adjoint_out(Ctx& c, const std::vector<Model*>& fm, float2* gS, float2* gR, size_t mstride,
                            size_t nslice, const double2* accS, const double2* accDR, int add) {
  hipStream_t s = c.s;
  DevBuf fS, fR, fdr;
  RCCHK(fS.alloc(c.nzx * sizeof(float2), "tmp"));
  RCCHK(fR.alloc(c.nzx * sizeof(float2), "tmp"));
  RCCHK(fdr.alloc(c.nzx * sizeof(float2), "tmp"));
  for (size_t k = 0; k < nslice; ++k) {
    Model* m = fm[k];
    float2* oS = gS + k * mstride;
    float2* oR = gR + k * mstride;
    HIPCHK(launch_d2f(fS.as<float2>(), accS + k * c.nzx, (long long)c.nzx, false, s), "d2f");
    HIPCHK(launch_d2f(fdr.as<float2>(), accDR + k * c.nzx, (long long)c.nzx, false, s), "d2f");
    RCCHK(c.zero(fR.p, c.nzx * sizeof(float2)));
    HIPCHK(launch_dreflect_adj(fS.as<float2>(), fR.as<float2>(), m->d_drefl, fdr.as<float2>(), c.nz, c.nx, s),
           "dreflect adjoint");
    if (!add) {
      HIPCHK(hipMemcpyAsync(oS, fS.p, c.nzx * sizeof(float2), hipMemcpyDeviceToDevice, s), "copy");
      HIPCHK(hipMemcpyAsync(oR, fR.p, c.nzx * sizeof(float2), hipMemcpyDeviceToDevice, s), "copy");
    } else {
      HIPCHK(launch_add(oS, fS.as<float2>(), (long long)c.nzx, s), "add");
      HIPCHK(launch_add(oR, fR.as<float2>(), (long long)c.nzx, s), "add");
    }
  }
  return c.sync();
}

// Checkpointed form of born_adjoint_all: the depth axis is cut into segments
// [a_j, b_j], a_j = j k, b_j = min(a_j + k, nz - 1), each owning the output
// depths [a_j, b_j) (the last one also nz - 1).  A forward pass keeps, per
// row, D at a_j and R at a_j - 1 of every segment; the backward pass visits
// the segments bottom-up, recomputes D and R on the segment from those rows,
// and carries the bottom-up recursions Q = Down^H(.) and U = Up(r D) from one
// segment to the next through their rows at a_j.  Every output sample gets
// the same operations in the same order as with whole fields stored (the
// sweep state outside [0, nx) is zero after a forward step and is discarded
// by the reference restriction at the start of an adjoint step, so a state
// reloaded from a kept row continues bit for bit), hence the same gradient.
// Device memory per row: (4 (k + 2) + 2 nseg + 2) nx samples instead of
// 4 nz nx.
static int born_adjoint_ckpt(Ctx& c, const std::vector<Model*>& fm, Per per, double2* accS, double2* accDR,
                             bool sliced, const float2* data, int exact, bool ext_rowlast, int k) {
  Survey* sv = c.sv;
  const int nw = sv->d.nw, ns = sv->d.ns, nz = c.nz, nx = c.nx;
  const int nseg = (nz - 1 + k - 1) / k;
  const int SR = std::min(k + 2, nz);       // rows of a segment buffer
  const size_t W = (size_t)SR * nx;        // samples per row of a segment buffer
  const size_t per_row = (4 * W + (2 * (size_t)nseg + 2) * nx) * sizeof(float2);
  auto chunks = make_chunks(ns, nw, fm, per, per_row, per == Per::Freq);
  const size_t nmax = chunks[0].shot.size();
  DevBuf B, Rb, Qb, CS, KD, KR, CQ, CU;
  RCCHK(B.alloc(nmax * W * sizeof(float2), "background segment"));
  RCCHK(Rb.alloc(nmax * W * sizeof(float2), "receiver segment"));
  RCCHK(Qb.alloc(nmax * W * sizeof(float2), "receiver segment"));
  RCCHK(CS.alloc(nmax * W * sizeof(float2), "dS contributions"));
  RCCHK(KD.alloc(nmax * nseg * nx * sizeof(float2), "background checkpoints"));
  RCCHK(KR.alloc(nmax * nseg * nx * sizeof(float2), "receiver checkpoints"));
  RCCHK(CQ.alloc(nmax * nx * sizeof(float2), "carried row"));
  RCCHK(CU.alloc(nmax * nx * sizeof(float2), "carried row"));
  float2 *bg = B.as<float2>(), *R = Rb.as<float2>(), *Q = Qb.as<float2>(), *cs = CS.as<float2>(),
         *kd = KD.as<float2>(), *kr = KR.as<float2>(), *cq = CQ.as<float2>(), *cu = CU.as<float2>();
  const size_t kpitch = (size_t)nseg * nx;  // checkpoint rows of one chunk row
  std::pair<int, int> srows, rrows;
  sv->depth_range(false, srows.first, srows.second);
  sv->depth_range(true, rrows.first, rrows.second);
  for (const Rows& r : chunks) {
    const int n = (int)r.shot.size();
    // D on [a, b] and R on [ra, rb] of segment j, from the kept rows
    // (backward pass: with Q = conj(r) R on the rows of R's segment)
    // Only the device depths inside a segment's injection window are read
    // as sources, and only those window rows are cleared.
    auto fields = [&](int j, float2* q2) -> int {
      const int a = j * k, b = std::min(a + k, nz - 1), zb = std::max(a - 1, 0);
      const bool last = j == nseg - 1;
      const std::pair<int, int> sw{std::max(srows.first, j ? a + 1 : 0), std::min(srows.second, b)};
      RCCHK(c.zero_depths(bg, n, sw.first, sw.second, W, zb));
      RCCHK(c.inject_sources(r, bg, sv->d.ngs, sv->d.offset, Ctx::Win{(long long)W, zb, j ? a + 1 : 0, b}));
      RCCHK(c.sweep_seg(r, 0, 0, bg, W, zb, a, b - a, j ? kd + (size_t)j * nx : nullptr, kpitch, nullptr, &sw));
      const int ra = j ? a - 1 : 0, rb = last ? b : b - 1;
      const std::pair<int, int> rw{std::max(rrows.first, j ? a : 0), std::min(rrows.second, rb)};
      RCCHK(c.zero_depths(R, n, rw.first, rw.second, W, zb));
      RCCHK(c.inject_receivers(r, R, data, sv->d.ngr, sv->d.offset, Ctx::Win{(long long)W, zb, j ? a : 0, rb}));
      RCCHK(c.sweep_seg(r, 1, 1, R, W, zb, ra, rb - ra, j ? kr + (size_t)j * nx : nullptr, kpitch, q2, &rw));
      return WEMFWI_OK;
    };
    // forward pass: the kept rows D(a_j) and R(a_j - 1)
    for (int j = 0; j + 1 < nseg; ++j) {
      RCCHK(fields(j, nullptr));
      const int a = j * k, b = a + k, zb = std::max(a - 1, 0);
      RCCHK(c.copy_rows(kd + (size_t)(j + 1) * nx, kpitch, bg + (size_t)(b - zb) * nx, W, n));
      RCCHK(c.copy_rows(kr + (size_t)(j + 1) * nx, kpitch, R + (size_t)(b - 1 - zb) * nx, W, n));
    }
    // backward pass
    for (int j = nseg - 1; j >= 0; --j) {
      const int a = j * k, b = std::min(a + k, nz - 1), zb = std::max(a - 1, 0);
      const bool last = j == nseg - 1;
      const int e = last ? nz : b;  // owned depths [a, e)
      const size_t oa = (size_t)(a - zb) * nx;
      RCCHK(fields(j, Q));   // Q = conj(r) R on R's rows [ra, rb], which hold [a, qe]
      // dr += conj(D) * R over the owned depths, before LinDown^H rescales D
      RCCHK(reduce_chunk(c, r, sliced, accDR, bg, W, zb, a, e, R));
      const int qe = last ? b : b - 1;  // Q source rows [a, qe]
      RCCHK(c.inject_receivers(r, Q, data, sv->d.ngr, sv->d.offset, Ctx::Win{(long long)W, zb, a, qe}));
      RCCHK(c.sweep_seg(r, 0, 1, Q, W, zb, b, b - a, last ? nullptr : cq, nx));
      if (j > 0) RCCHK(c.copy_rows(cq, nx, Q + oa, W, n));
      // LinDown^H stores the rows [a, b) of cs (no fill); the last segment's
      // row nz - 1 = b starts at 0
      if (last) RCCHK(c.zero_rows(cs + (size_t)(b - zb) * nx, W, nx, n));
      // LinDown^H over [a, b); the D rows it reads become r * D (Up sources)
      RCCHK(c.lin_seg(r, +1, exact, bg, Q, cs, W, zb, a, b - a, true, true));
      if (last) {  // the Up recursion starts from r * D at the last depth
        std::vector<MulRow> rbg(n);
        for (int i = 0; i < n; ++i) {
          const size_t o = (size_t)i * W + (size_t)(b - zb) * nx;
          rbg[i] = MulRow{bg + o, r.model[i]->d_refl + (size_t)b * nx, bg + o};
        }
        RCCHK(c.mul_rows(rbg, (size_t)nx, false, false));
      }
      RCCHK(c.sweep_seg(r, 1, 0, bg, W, zb, b, b - a, last ? nullptr : cu, nx));
      if (j > 0) RCCHK(c.copy_rows(cu, nx, bg + oa, W, n));
      const int ulo = std::max(a, 1);
      RCCHK(c.lin_seg(r, -1, exact, bg, R, cs, W, zb, ulo, e - ulo));
      if (ext_rowlast && last) {
        for (int i = 0; i < n; ++i)
          if (r.shot[i] != ns - 1)
            RCCHK(c.zero_rows(cs + (size_t)i * W + (size_t)(nz - 1 - zb) * nx, nx, nx, 1));
      }
      RCCHK(reduce_chunk(c, r, sliced, accS, cs, W, zb, a, e));
    }
    RCCHK(c.sync());
  }
  return WEMFWI_OK;
}

// BornFull::full_adj (born_prop.cpp:178-232) per row, summed over rows in a
// fixed order in double precision:
//   D = Down(inj_src),  R = Up^H(R^H d)
//   dr    += conj(D) * R                              (Born_refl::adjoint)
//   Q     = Down^H(R^H d + conj(r) * R)               (both LinDown^H terms)
//   dS    += LinDown^H(Q; D)
//   U     = Up(r * D),  dS += LinUp^H(R; U)
//   (dS, dR) += dReflect^H(dr)
// ext_rowlast: reproduce ext_born_prop.cpp's serial shot loop, in which each
// LinDown::adjoint zeroes the last depth row (LinOneWay.cpp:64-65) so only
// the last shot's LinUp contribution remains there.
static int born_adjoint(const std::vector<Model*>& fm, Per per, Survey* sv, float2* gS, float2* gR, size_t mstride,
                        const float2* data, int add, int exact, bool ext_rowlast, hipStream_t s) {
  RCCHK(check_survey(sv, fm[0]));
  std::vector<Model*> uniq;
  for (Model* m : fm)
    if (std::find(uniq.begin(), uniq.end(), m) == uniq.end()) uniq.push_back(m);
  Ctx c;
  RCCHK(c.init(uniq, sv, s));
  const size_t nslice = mstride ? fm.size() : 1;
  // FP64 accumulators per model slice: dr, dS
  DevBuf accDR, accS;
  RCCHK(accDR.alloc(nslice * c.nzx * sizeof(double2), "dr accumulator"));
  RCCHK(accS.alloc(nslice * c.nzx * sizeof(double2), "dS accumulator"));
  RCCHK(c.zero(accDR.p, accDR.n));
  RCCHK(c.zero(accS.p, accS.n));
  const int ck = sv->ckpt;
  if (ck > 0 && ck < c.nz - 1) {
    RCCHK(born_adjoint_ckpt(c, fm, per, accS.as<double2>(), accDR.as<double2>(), mstride != 0, data, exact,
                            ext_rowlast, ck));
  } else {
    RCCHK(born_adjoint_all(c, fm, per, accS.as<double2>(), accDR.as<double2>(), mstride != 0, data, exact,
                           ext_rowlast));
  }
  return born_adjoint_out(c, fm, gS, gR, mstride, nslice, accS.as<double2>(), accDR.as<double2>(), add);
}


// ------------------------------------------------------------------ BornRefl

// ext_born_refl.cpp:8-56 (forward) and 60-110 (adjoint) on the unpadded
// grid: source spread ng without offset, receivers without spread.
static int bornrefl_forward(Model* m, Survey* sv, const float2* image, int nh, float2* data, int add,
                            hipStream_t s) {
  RCCHK(check_survey(sv, m));
  Ctx c;
  RCCHK(c.init({m}, sv, s));
  if (!add) RCCHK(c.zero(data, (size_t)sv->d.ns * sv->d.nr * sv->d.nt * sizeof(float2)));
  std::vector<Model*> fm(sv->d.nw, m);
  auto chunks = make_chunks(sv->d.ns, sv->d.nw, fm, Per::Freq, 2 * c.nzx * sizeof(float2), false);
  const size_t cap = chunks[0].shot.size() * c.nzx * sizeof(float2);
  DevBuf B, S;
  RCCHK(B.alloc(cap, "background"));
  RCCHK(S.alloc(cap, "scattered"));
  float2 *bg = B.as<float2>(), *sc = S.as<float2>();
  std::pair<int, int> srows;
  sv->depth_range(false, srows.first, srows.second);
  for (const Rows& r : chunks) {
    const int n = (int)r.shot.size();
    RCCHK(c.zero_depths(bg, n, srows.first, srows.second));   // only the source depths are read as sources
    RCCHK(c.inject_sources(r, bg, sv->d.ngs, 0));
    RCCHK(c.sweep(r, 0, 0, bg, bg, nullptr, nullptr, &srows));
    if (nh <= 0) {
      // sc = Up(D * image[w]): the IC product taken as the Up sweep reads D
      std::vector<const float2*> im(n);
      for (int i = 0; i < n; ++i) im[i] = image + (size_t)r.fi[i] * c.nzx;
      RCCHK(c.sweep(r, 1, 0, bg, sc, &im));
    } else {
      for (int i = 0; i < n; ++i) {
        const size_t o = (size_t)i * c.nzx;
        HIPCHK(hipMemsetAsync(sc + o, 0, c.nzx * sizeof(float2), s), "memset");
        HIPCHK(launch_offset_ic_fwd(sc + o, bg + o, image, nh, c.nz, c.nx, s), "offset ic");
      }
      RCCHK(c.sweep(r, 1, 0, sc, sc));
    }
    RCCHK(c.record(r, data, sc, 0, 0));
  }
  return c.sync();
}

static int bornrefl_adjoint(Model* m, Survey* sv, float2* image, int nh, const float2* data, int add,
                            hipStream_t s) {
  RCCHK(check_survey(sv, m));
  Ctx c;
  RCCHK(c.init({m}, sv, s));
  const int nw = sv->d.nw;
  const size_t nimg = nh > 0 ? (size_t)nh : (size_t)nw;
  DevBuf acc;
  RCCHK(acc.alloc(nimg * c.nzx * sizeof(double2), "image accumulator"));
  RCCHK(c.zero(acc.p, acc.n));
  std::vector<Model*> fm(nw, m);
  auto chunks = make_chunks(sv->d.ns, nw, fm, Per::Freq, 2 * c.nzx * sizeof(float2), true);
  const size_t cap = chunks[0].shot.size() * c.nzx * sizeof(float2);
  DevBuf B, S;
  RCCHK(B.alloc(cap, "background"));
  RCCHK(S.alloc(cap, "receiver wavefield"));
  float2 *bg = B.as<float2>(), *sc = S.as<float2>();
  DevBuf pb, ps;
  std::pair<int, int> srows, rrows;
  sv->depth_range(false, srows.first, srows.second);
  sv->depth_range(true, rrows.first, rrows.second);
  for (const Rows& r : chunks) {
    const int n = (int)r.shot.size();
    // only the source / receiver depths are read as sources and cleared
    RCCHK(c.zero_depths(bg, n, srows.first, srows.second));
    RCCHK(c.inject_sources(r, bg, sv->d.ngs, 0));
    RCCHK(c.sweep(r, 0, 0, bg, bg, nullptr, nullptr, &srows));
    RCCHK(c.zero_depths(sc, n, rrows.first, rrows.second));
    RCCHK(c.inject_receivers(r, sc, data, 0, 0));
    RCCHK(c.sweep(r, 1, 1, sc, sc, nullptr, nullptr, &rrows));
    if (nh <= 0) {
      // image[w] += conj(D) R, the product formed in the reduction
      int i0 = 0;
      while (i0 < n) {
        int i1 = i0 + 1;
        while (i1 < n && r.fi[i1] == r.fi[i0]) ++i1;
        HIPCHK(launch_reduce_rows(acc.as<double2>() + (size_t)r.fi[i0] * c.nzx, bg + (size_t)i0 * c.nzx,
                                  (long long)c.nzx, (long long)c.nzx, i1 - i0, (int)((r.first + i0) % kRunRows),
                                  true, s, sc + (size_t)i0 * c.nzx),
               "reduce");
        i0 = i1;
      }
    } else {
      std::vector<const float2*> hb(n), hs(n);
      for (int i = 0; i < n; ++i) {
        hb[i] = bg + (size_t)i * c.nzx;
        hs[i] = sc + (size_t)i * c.nzx;
      }
      RCCHK(put(pb, hb, s, "row pointers"));
      RCCHK(put(ps, hs, s, "row pointers"));
      HIPCHK(launch_offset_ic_adj(acc.as<double2>(), pb.as<const float2*>(), ps.as<const float2*>(), n, nh, c.nz,
                                  c.nx, s),
             "offset ic adjoint");
    }
    RCCHK(c.sync());
  }
  HIPCHK(launch_d2f(image, acc.as<double2>(), (long long)(nimg * c.nzx), add != 0, s), "d2f");
  return c.sync();
}

// ------------------------------------------------------------------ gradient plan

// The FWI gradient of one model, survey and observed data as a replayable
// hipGraph: res = WEM(m) - obs, J = 0.5 ||res||^2 (FP64, fixed order) and
// (gS, gR) = Born(m)^H res with every (shot, frequency) row batched
// (scripts/fwi.py:70-131 with WEM.py:52-81: the objective and its gradient
// for one model).  Built by running the sequence once eagerly with a
// recording arena (every device buffer kept, every descriptor table in a
// buffer of its own) and once more under stream capture replaying that
// arena, so the graph holds kernel launches and device copies only.  The
// plan keeps every buffer of the call (for store-all gradients 4 wavefields
// per row of a chunk) until it is destroyed; the model, survey, observed data
// and output buffers are bound at creation.
struct GradientPlan {
  Model* m = nullptr;
  Survey* sv = nullptr;
  const float2* obs = nullptr;
  float2 *res = nullptr, *gS = nullptr, *gR = nullptr;
  int exact = 1;
  Arena arena;
  void* d_acc = nullptr;  // [kMisfitParts] partial sums, then J
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GradientPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    cached_free(d_acc);
  }
  double* d_J() const { return (double*)d_acc + kMisfitParts; }
  int run(hipStream_t s) {
    std::vector<Model*> fm(sv->d.nw, m);
    RCCHK(wem_forward(fm, Per::Freq, sv, res, 0, s));
    const long long n = (long long)sv->d.ns * sv->d.nr * sv->d.nt;
    HIPCHK(launch_residual_sumsq(res, obs, n, (double*)d_acc, d_J(), s), "misfit");
    return born_adjoint(fm, Per::Freq, sv, gS, gR, 0, res, 0, exact, false, s);
  }
  int build(hipStream_t s) {
    RCCHK(cached_malloc(&d_acc, (kMisfitParts + 1) * sizeof(double), "misfit partial sums"));
    g_arena = &arena;
    int rc = run(s);  // recording pass (eager: its results are the gradient of this call)
    hipError_t e = hipStreamSynchronize(s);
    if (rc == WEMFWI_OK && e != hipSuccess) rc = hip_fail(e, "gradient plan");
    if (rc == WEMFWI_OK) {
      arena.rewind();
      e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
      if (e != hipSuccess) {
        rc = hip_fail(e, "begin capture");
      } else {
        rc = run(cap);
        hipError_t e2 = hipStreamEndCapture(cap, &graph);
        if (rc == WEMFWI_OK && e2 != hipSuccess) rc = hip_fail(e2, "end capture");
        if (rc == WEMFWI_OK && arena.next != arena.bufs.size())
          rc = fail(WEMFWI_ERR_ARG, "gradient plan: replay does not match the recording");
      }
      if (rc == WEMFWI_OK) {
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "graph instantiate");
      }
    }
    g_arena = nullptr;
    return rc;
  }
};

// ------------------------------------------------------------------ uneven Born

// BornFull::full_fwd / full_adj (born_prop.cpp:75-113, 178-232) of shots that
// each carry their own background on their own aperture (nz x nx_s, placed
// at column x0 of the perturbation grid) and their own frequencies.  The
// rows of a group of consecutive shots (shot order, then frequency order)
// share one set of wavefield buffers; every sweep is one batch plan over all
// of them (one launch per kernel variant) and every LinDown / LinUp one
// launch per FFT length; injection, recording, row products and the FP64
// reductions run per shot with the arguments and in the order of
// born_forward / born_adjoint_all on a one-shot survey, so every shot's data
// and gradient slice are bitwise those of such a call.
// ext (extBornFull, ext_born_prop.cpp:8-194): frequency i of a shot has its
// own background fm[i] and perturbation slice w0 + i; the windows, the
// dReflect products and the FP64 accumulators are then per row, not per
// shot, and each row's contribution goes to its own slice.
struct UShot {
  std::vector<Model*> fm;  // background of each frequency (one entry: all)
  int x0 = 0, nw = 0, w0 = 0;
  const float* freq_hz = nullptr;
  const wemfwi_complex* wavelet = nullptr;
  int src_iz = 0, src_ix = 0, nr = 0;
  const int* rec_iz = nullptr;
  const int* rec_ix = nullptr;
  wemfwi_complex* data = nullptr;
  Model* model(int i) const { return fm.size() > 1 ? fm[i] : fm[0]; }
};

struct UnevenBorn {
  const UShot* shots = nullptr;
  bool ext = false;
  int s0 = 0, s1 = 0, nz = 0, ngs = 0, ngr = 0;
  hipStream_t s = nullptr;
  std::vector<Model*> rm;      // model of each row
  std::vector<float> rf;       // frequency of each row
  std::vector<size_t> roff;    // first sample of each row in a group buffer
  std::vector<int> rshot;      // shot (of the group) of each row
  std::vector<int> rslice;     // perturbation slice of each row
  std::vector<int> first;      // first row of each shot of the group (one more entry)
  size_t total = 0;            // samples of a group buffer
  std::vector<size_t> woff;    // first sample of each shot's [nz][nx_s] slot
  size_t wtotal = 0;
  std::vector<DevPlan> rplan;  // receiver plans, one per shot
  std::vector<Model*> models;  // distinct models of the group
  std::vector<int> midx;       // row -> distinct model
  std::vector<float> wrow;
  std::vector<const float2*> prow;
  DevBuf d_tabs, d_lin, d_mul, d_iz, d_ix, d_val, d_gw, d_vb;

  ~UnevenBorn() {
    for (auto& p : rplan) p.release();
  }
  const UShot& S(int k) const { return shots[s0 + k]; }
  int nshots() const { return s1 - s0; }
  int nx(int k) const { return S(k).fm[0]->d.nx; }
  size_t nzx(int k) const { return (size_t)nz * nx(k); }
  int nw(int k) const { return first[k + 1] - first[k]; }
  // first sample of shot k's rows in a group buffer
  size_t base(int k) const { return nw(k) ? roff[first[k]] : 0; }
  // row r's perturbation window / dReflect slot: its shot's slot, or (ext)
  // the row's own
  size_t wo(int r) const { return ext ? roff[r] : woff[rshot[r]]; }
  size_t wsize() const { return ext ? total : wtotal; }

  int init(const UShot* sh, bool ext_, int a, int b, int nz_, int ngs_, int ngr_, hipStream_t st) {
    shots = sh;
    ext = ext_;
    s0 = a;
    s1 = b;
    nz = nz_;
    ngs = ngs_;
    ngr = ngr_;
    s = st;
    first.assign(1, 0);
    for (int k = 0; k < nshots(); ++k) {
      woff.push_back(wtotal);
      wtotal += nzx(k);
      for (int i = 0; i < S(k).nw; ++i) {
        rm.push_back(S(k).model(i));
        rf.push_back(S(k).freq_hz[i]);
        roff.push_back(total);
        rshot.push_back(k);
        rslice.push_back(ext ? S(k).w0 + i : 0);
        total += nzx(k);
      }
      first.push_back((int)rm.size());
    }
    const int n = (int)rm.size();
    midx.resize(n);
    for (int i = 0; i < n; ++i) {
      auto it = std::find(models.begin(), models.end(), rm[i]);
      midx[i] = (int)(it - models.begin());
      if (it == models.end()) models.push_back(rm[i]);
    }
    // phase tables: every model gets the frequencies of its rows
    wrow.resize(n);
    prow.resize(n);
    for (size_t u = 0; u < models.size(); ++u) {
      std::vector<int> idx;
      std::vector<float> f;
      for (int i = 0; i < n; ++i)
        if (midx[i] == (int)u) {
          idx.push_back(i);
          f.push_back(rf[i]);
        }
      std::vector<float> w;
      std::vector<const float2*> p;
      RCCHK(models[u]->phase_tables((int)f.size(), f.data(), w, p, s));
      for (size_t k = 0; k < idx.size(); ++k) {
        wrow[idx[k]] = w[k];
        prow[idx[k]] = p[k];
      }
    }
    std::vector<DepthTables> ht;
    for (Model* m : models) ht.push_back(m->h_tables);
    RCCHK(put(d_tabs, ht, s, "model tables"));
    // per-row source positions and values, data columns
    std::vector<int> hiz(std::max(n, 1)), hix(std::max(n, 1));
    std::vector<float2> hval(std::max(n, 1));
    std::vector<long long> hvb(std::max(n, 1));
    for (int k = 0; k < nshots(); ++k)
      for (int i = 0; i < nw(k); ++i) {
        const int r = first[k] + i;
        hiz[r] = S(k).src_iz;
        hix[r] = S(k).src_ix;
        hval[r] = make_float2(S(k).wavelet[i].re, S(k).wavelet[i].im);
        hvb[r] = i;  // data [nr][nw]: receiver stride nw, column i
      }
    RCCHK(put(d_iz, hiz, s, "source iz"));
    RCCHK(put(d_ix, hix, s, "source ix"));
    RCCHK(put(d_val, hval, s, "source values"));
    RCCHK(put(d_vb, hvb, s, "data columns"));
    RCCHK(put(d_gw, cos2_window(ngs), s, "source spread"));
    rplan.resize(nshots());
    for (int k = 0; k < nshots(); ++k) {
      const std::vector<int> iz(S(k).rec_iz, S(k).rec_iz + S(k).nr), ix(S(k).rec_ix, S(k).rec_ix + S(k).nr);
      RCCHK(build_plan(rplan[k], iz, ix, ngr, 0, nx(k), nz));
    }
    return WEMFWI_OK;
  }

  // one sweep of every row of the group (batch plan: one launch per kernel
  // variant); lo / hi: per row, the only source depths read (none: all);
  // smul_refl: source rows multiplied by the row model's reflection
  // coefficients as they are read; q2: second output conj(r) * result
  int sweep(int up, int adjoint, const float2* in, float2* out, const std::vector<int>* lo = nullptr,
            const std::vector<int>* hi = nullptr, bool smul_refl = false, float2* q2 = nullptr) {
    const int n = (int)rm.size();
    if (n == 0) return WEMFWI_OK;
    std::vector<const cf*> pin(n);
    std::vector<cf*> pout(n);
    std::vector<const float2*> sm, om;
    std::vector<float2*> o2;
    BatchPlan::RowProducts pr;
    for (int i = 0; i < n; ++i) {
      pin[i] = (const cf*)(in + roff[i]);
      pout[i] = (cf*)(out + roff[i]);
    }
    if (smul_refl) {
      sm.resize(n);
      for (int i = 0; i < n; ++i) sm[i] = rm[i]->d_refl;
      pr.smul = sm.data();
    }
    if (q2) {
      o2.resize(n);
      om.resize(n);
      for (int i = 0; i < n; ++i) {
        o2[i] = q2 + roff[i];
        om[i] = rm[i]->d_refl;
      }
      pr.out2 = o2.data();
      pr.omul = om.data();
    }
    BatchPlan p;
    RCCHK(p.build(n, rm.data(), rf.data(), pin.data(), pout.data(), s, lo ? lo->data() : nullptr,
                  hi ? hi->data() : nullptr, &pr));
    return p.run(up, adjoint, s);
  }

  // LinDown (dir = +1) / LinUp (dir = -1) of every row, one launch per FFT
  // length and Taylor order; m: per row, the perturbation (forward) or the
  // wavefield row (adjoint)
  int lin(int dir, int adjoint, int exact, float2* bg, const std::vector<const float2*>& m, float2* out,
          bool post_refl = false, bool assign = false) {
    const int n = (int)rm.size();
    if (n == 0) return WEMFWI_OK;
    std::vector<long long> key(n);
    for (int i = 0; i < n; ++i) key[i] = ((long long)rm[i]->N * 4096 + rm[i]->nk) * 64 + rm[i]->d.ntaylor;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<LinRow> h(n);
    for (int k = 0; k < n; ++k) {
      const int i = order[k];
      h[k] = LinRow{bg + roff[i], m[i], out + roff[i], prow[i], wrow[i], midx[i], 0,
                    post_refl ? rm[i]->d_refl : nullptr, assign ? 1 : 0};
    }
    RCCHK(put(d_lin, h, s, "lin rows"));
    for (int k0 = 0; k0 < n;) {
      int k1 = k0 + 1;
      while (k1 < n && key[order[k1]] == key[order[k0]]) ++k1;
      const Model* mg = rm[order[k0]];
      HIPCHK(launch_lin(mg->N, d_lin.as<LinRow>() + k0, k1 - k0, d_tabs.as<DepthTables>(), mg->geom(), nz, dir,
                        adjoint != 0, exact != 0, mg->d_tw, s),
             "lin kernel");
      k0 = k1;
    }
    HIPCHK(hipStreamSynchronize(s), "lin");
    return WEMFWI_OK;
  }

  // per-row pointers into a window buffer (each row's window slot), or to
  // the row itself in a group buffer
  std::vector<const float2*> per_window(const float2* w) const {
    std::vector<const float2*> v(rm.size());
    for (size_t r = 0; r < rm.size(); ++r) v[r] = w + wo((int)r);
    return v;
  }
  std::vector<const float2*> per_row(const float2* w) const {
    std::vector<const float2*> v(rm.size());
    for (size_t r = 0; r < rm.size(); ++r) v[r] = w + roff[r];
    return v;
  }
  // depths [lo, hi] of shot k's rows
  int zero_depths(float2* W, int k, int lo, int hi) {
    if (lo > hi || nw(k) == 0) return WEMFWI_OK;
    HIPCHK(launch_zero2d(W + base(k) + (size_t)lo * nx(k), (long long)nzx(k), (long long)(hi - lo + 1) * nx(k),
                         nw(k), s),
           "zero rows");
    return WEMFWI_OK;
  }
  int inject_sources(float2* W) {
    for (int k = 0; k < nshots(); ++k) {
      if (nw(k) == 0) continue;
      const int r = first[k];
      HIPCHK(launch_inject_point(W + base(k), (long long)nzx(k), nx(k), d_iz.as<int>() + r, d_ix.as<int>() + r,
                                 d_val.as<float2>() + r, d_gw.as<float>(), ngs, 0, nw(k), s),
             "inject sources");
    }
    return WEMFWI_OK;
  }
  int inject_receivers(float2* W) {
    for (int k = 0; k < nshots(); ++k) {
      if (nw(k) == 0) continue;
      HIPCHK(launch_inject_csr(W + base(k), (long long)nzx(k), rplan[k].inj, (const float2*)S(k).data,
                               d_vb.as<long long>() + first[k], nw(k), nw(k), s, 0, 0, (long long)nzx(k)),
             "inject receivers");
    }
    return WEMFWI_OK;
  }
  int record(const float2* W) {
    for (int k = 0; k < nshots(); ++k) {
      if (nw(k) == 0) continue;
      HIPCHK(launch_record((float2*)S(k).data, d_vb.as<long long>() + first[k], nw(k), W + base(k),
                           (long long)nzx(k), rplan[k].rec, nw(k), s),
             "record receivers");
    }
    HIPCHK(hipStreamSynchronize(s), "record");
    return WEMFWI_OK;
  }
  int mul_rows(std::vector<MulRow>& rows, size_t len, bool acc) {
    if (rows.empty()) return WEMFWI_OK;
    RCCHK(put(d_mul, rows, s, "product rows"));
    HIPCHK(launch_mul_rows(d_mul.as<MulRow>(), (int)rows.size(), (long long)len, false, acc, s), "mul rows");
    return WEMFWI_OK;
  }
  // receiver depth range of shot k ([lo, hi], lo > hi without receivers)
  void rec_depths(int k, int& lo, int& hi) const {
    lo = 1 << 30;
    hi = -1;
    for (int r = 0; r < S(k).nr; ++r) {
      lo = std::min(lo, S(k).rec_iz[r]);
      hi = std::max(hi, S(k).rec_iz[r]);
    }
  }
  std::vector<int> row_vec(const std::vector<int>& per_shot_v) const {
    std::vector<int> v(rm.size());
    for (int k = 0; k < nshots(); ++k)
      for (int r = first[k]; r < first[k + 1]; ++r) v[r] = per_shot_v[k];
    return v;
  }
};

// born_forward's sequence on the rows of one group of shots; dS, dR: the
// perturbation slices [nslice][nz][nxg]
static int uneven_born_forward_group(UnevenBorn& u, const float2* dS, const float2* dR, int nxg) {
  const int K = u.nshots(), nz = u.nz;
  const size_t npix = (size_t)nz * nxg;
  hipStream_t s = u.s;
  DevBuf WS, WR, DRb, B, S1, T;
  RCCHK(WS.alloc(u.wsize() * sizeof(float2), "perturbation windows"));
  RCCHK(WR.alloc(u.wsize() * sizeof(float2), "perturbation windows"));
  RCCHK(DRb.alloc(u.wsize() * sizeof(float2), "dr"));
  float2 *ws = WS.as<float2>(), *wr = WR.as<float2>(), *dr = DRb.as<float2>();
  for (int k = 0; k < K; ++k) {
    // one window per shot, or (ext) one per row from its slice
    const int r0 = u.ext ? u.first[k] : 0, r1 = u.ext ? u.first[k + 1] : 1;
    for (int r = r0; r < r1; ++r) {
      const size_t o = u.ext ? u.roff[r] : u.woff[k];
      const Model& m = u.ext ? *u.rm[r] : *u.S(k).fm[0];
      const size_t q = u.ext ? (size_t)u.rslice[r] * npix : 0;
      HIPCHK(launch_window(ws + o, dS + q, nz, u.nx(k), nxg, u.S(k).x0, s), "window");
      HIPCHK(launch_window(wr + o, dR + q, nz, u.nx(k), nxg, u.S(k).x0, s), "window");
      HIPCHK(launch_dreflect_fwd(dr + o, m.d_drefl, ws + o, wr + o, nz, u.nx(k), false, s), "dreflect");
    }
  }
  if (u.total == 0) return WEMFWI_OK;
  RCCHK(B.alloc(u.total * sizeof(float2), "background"));
  RCCHK(S1.alloc(u.total * sizeof(float2), "scattered"));
  RCCHK(T.alloc(u.total * sizeof(float2), "scattered"));
  float2 *bg = B.as<float2>(), *s1 = S1.as<float2>(), *tt = T.as<float2>();
  std::vector<int> src(K);
  for (int k = 0; k < K; ++k) src[k] = u.S(k).src_iz;
  const std::vector<int> slo = u.row_vec(src);
  // D = Down(inj_src): only the source depth is cleared and read as a source
  for (int k = 0; k < K; ++k) RCCHK(u.zero_depths(bg, k, src[k], src[k]));
  RCCHK(u.inject_sources(bg));
  RCCHK(u.sweep(0, 0, bg, bg, &slo, &slo));
  const std::vector<const float2*> mS = u.per_window(ws);
  // s1 = Down(LinDown(dS; D)), recorded
  HIPCHK(launch_zero2d(s1, (long long)u.total, (long long)u.total, 1, s), "zero");
  RCCHK(u.lin(+1, 0, 0, bg, mS, s1));
  RCCHK(u.sweep(0, 0, s1, s1));
  RCCHK(u.record(s1));
  // s2 = Up(D * dr + r * s1), recorded
  for (int k = 0; k < K; ++k) {
    if (u.nw(k) == 0) continue;
    std::vector<MulRow> ic, rs1;
    for (int r = u.first[k]; r < u.first[k + 1]; ++r) {
      ic.push_back(MulRow{tt + u.roff[r], bg + u.roff[r], dr + u.wo(r)});
      rs1.push_back(MulRow{tt + u.roff[r], u.rm[r]->d_refl, s1 + u.roff[r]});
    }
    RCCHK(u.mul_rows(ic, u.nzx(k), false));
    RCCHK(u.mul_rows(rs1, u.nzx(k), true));
  }
  RCCHK(u.sweep(1, 0, tt, tt));
  RCCHK(u.record(tt));
  // U = Up(r * D); s3 = Up(LinUp(dS; U)), recorded
  RCCHK(u.sweep(1, 0, bg, bg, nullptr, nullptr, true));
  HIPCHK(launch_zero2d(s1, (long long)u.total, (long long)u.total, 1, s), "zero");
  RCCHK(u.lin(-1, 0, 0, bg, mS, s1));
  RCCHK(u.sweep(1, 0, s1, s1));
  RCCHK(u.record(s1));
  return WEMFWI_OK;
}

// born_adjoint_all and born_adjoint_out on the rows of one group of shots;
// each shot's float gradient slice (ext: each row's, into its own slice) is
// added to the FP64 accumulators gS, gR [nslice][nz][nxg] at its columns,
// shot after shot.  lastcut (ext, mode 0): per row, the columns [a, b) of
// its shot's aperture whose last depth row a later shot overwrites in the
// reference's serial loop
// column ranges [a, b) per shot and frequency
using ColCuts = std::vector<std::vector<std::vector<std::pair<int, int>>>>;

static int uneven_born_adjoint_group(UnevenBorn& u, double2* gS, double2* gR, int nxg, int exact,
                                     const ColCuts* lastcut) {
  const int K = u.nshots(), nz = u.nz;
  const size_t npix = (size_t)nz * nxg;
  hipStream_t s = u.s;
  if (u.total == 0) return WEMFWI_OK;
  DevBuf B, Rb, Qb, CS, AS, AD, FS, FR, FD;
  RCCHK(B.alloc(u.total * sizeof(float2), "background"));
  RCCHK(Rb.alloc(u.total * sizeof(float2), "receiver wavefield"));
  RCCHK(Qb.alloc(u.total * sizeof(float2), "receiver wavefield"));
  RCCHK(CS.alloc(u.total * sizeof(float2), "dS contributions"));
  RCCHK(AS.alloc(u.wsize() * sizeof(double2), "dS accumulators"));
  RCCHK(AD.alloc(u.wsize() * sizeof(double2), "dr accumulators"));
  float2 *bg = B.as<float2>(), *R = Rb.as<float2>(), *Q = Qb.as<float2>(), *cs = CS.as<float2>();
  double2 *as = AS.as<double2>(), *ad = AD.as<double2>();
  HIPCHK(launch_zero2d(as, (long long)(2 * u.wsize()), (long long)(2 * u.wsize()), 1, s), "zero");
  HIPCHK(launch_zero2d(ad, (long long)(2 * u.wsize()), (long long)(2 * u.wsize()), 1, s), "zero");
  std::vector<int> src(K), rlo(K), rhi(K);
  for (int k = 0; k < K; ++k) {
    src[k] = u.S(k).src_iz;
    u.rec_depths(k, rlo[k], rhi[k]);
  }
  const std::vector<int> slo = u.row_vec(src), rl = u.row_vec(rlo), rh = u.row_vec(rhi);
  // D = Down(inj_src); R = Up^H(R^H d) with Q = conj(r) R on every row
  for (int k = 0; k < K; ++k) RCCHK(u.zero_depths(bg, k, src[k], src[k]));
  RCCHK(u.inject_sources(bg));
  RCCHK(u.sweep(0, 0, bg, bg, &slo, &slo));
  for (int k = 0; k < K; ++k) RCCHK(u.zero_depths(R, k, rlo[k], rhi[k]));
  RCCHK(u.inject_receivers(R));
  RCCHK(u.sweep(1, 1, R, R, &rl, &rh, false, Q));
  // dr += conj(D) * R: per shot over its rows (runs from its first row), or
  // (ext) every row into its own accumulator (one row per run: one launch)
  if (u.ext) {
    HIPCHK(launch_reduce_rows(ad, bg, (long long)u.total, (long long)u.total, 1, 0, true, s, R), "reduce");
  } else {
    for (int k = 0; k < K; ++k) {
      if (u.nw(k) == 0) continue;
      HIPCHK(launch_reduce_rows(ad + u.woff[k], bg + u.base(k), (long long)u.nzx(k), (long long)u.nzx(k), u.nw(k),
                                0, true, s, R + u.base(k)),
             "reduce");
    }
  }
  // Q = Down^H(conj(r) R + R^H d)
  RCCHK(u.inject_receivers(Q));
  RCCHK(u.sweep(0, 1, Q, Q));
  // LinDown^H stores rows 0..nz-2 of cs; row nz-1 starts at 0; each D row it
  // reads becomes r * D
  const size_t last = (size_t)(nz - 1);
  for (int k = 0; k < K; ++k)
    if (u.nw(k))
      HIPCHK(launch_zero2d(cs + u.base(k) + last * u.nx(k), (long long)u.nzx(k), u.nx(k), u.nw(k), s), "zero rows");
  RCCHK(u.lin(+1, 1, exact, bg, u.per_row(Q), cs, true, true));
  for (int k = 0; k < K; ++k) {
    if (u.nw(k) == 0) continue;
    std::vector<MulRow> rbg;
    const size_t lo = last * u.nx(k);
    for (int r = u.first[k]; r < u.first[k + 1]; ++r)
      rbg.push_back(MulRow{bg + u.roff[r] + lo, u.rm[r]->d_refl + lo, bg + u.roff[r] + lo});
    RCCHK(u.mul_rows(rbg, (size_t)u.nx(k), false));
  }
  // U = Up(r D); dS += LinUp^H(R; U)
  RCCHK(u.sweep(1, 0, bg, bg));
  RCCHK(u.lin(-1, 1, exact, bg, u.per_row(R), cs));
  if (lastcut) {
    // the reference's serial loop: a later shot's LinDown::adjoint clears
    // the last row of the slice (LinOneWay.cpp:64-65) where it covers it
    for (int k = 0; k < K; ++k)
      for (int r = u.first[k]; r < u.first[k + 1]; ++r)
        for (const auto& c : (*lastcut)[u.s0 + k][r - u.first[k]])
          HIPCHK(launch_zero2d(cs + u.roff[r] + last * u.nx(k) + c.first, c.second - c.first, c.second - c.first, 1,
                               s),
                 "zero row");
  }
  if (u.ext) {
    HIPCHK(launch_reduce_rows(as, cs, (long long)u.total, (long long)u.total, 1, 0, true, s), "reduce");
  } else {
    for (int k = 0; k < K; ++k) {
      if (u.nw(k) == 0) continue;
      const long long n = (long long)u.nzx(k);
      HIPCHK(launch_reduce_rows(as + u.woff[k], cs + u.base(k), n, n, u.nw(k), 0, true, s), "reduce");
    }
  }
  size_t wmax = 0;
  for (int k = 0; k < K; ++k) wmax = std::max(wmax, u.nzx(k));
  RCCHK(FS.alloc(wmax * sizeof(float2), "tmp"));
  RCCHK(FR.alloc(wmax * sizeof(float2), "tmp"));
  RCCHK(FD.alloc(wmax * sizeof(float2), "tmp"));
  for (int k = 0; k < K; ++k) {
    if (u.nw(k) == 0) continue;
    const long long n = (long long)u.nzx(k);
    // dReflect^H of each window's dr, then its slice into the sums over shots
    const int r0 = u.ext ? u.first[k] : 0, r1 = u.ext ? u.first[k + 1] : 1;
    for (int r = r0; r < r1; ++r) {
      const size_t o = u.ext ? u.roff[r] : u.woff[k];
      const Model& m = u.ext ? *u.rm[r] : *u.S(k).fm[0];
      const size_t q = u.ext ? (size_t)u.rslice[r] * npix : 0;
      HIPCHK(launch_d2f(FS.as<float2>(), as + o, n, false, s), "d2f");
      HIPCHK(launch_d2f(FD.as<float2>(), ad + o, n, false, s), "d2f");
      HIPCHK(launch_zero2d(FR.p, n, n, 1, s), "zero");
      HIPCHK(launch_dreflect_adj(FS.as<float2>(), FR.as<float2>(), m.d_drefl, FD.as<float2>(), nz, u.nx(k), s),
             "dreflect adjoint");
      HIPCHK(launch_window_add(gS + q, FS.as<float2>(), nz, u.nx(k), nxg, u.S(k).x0, s), "window add");
      HIPCHK(launch_window_add(gR + q, FR.as<float2>(), nz, u.nx(k), nxg, u.S(k).x0, s), "window add");
    }
  }
  HIPCHK(hipStreamSynchronize(s), "uneven born adjoint");
  return WEMFWI_OK;
}

// Validation and grouping of the entry points (gS / gR null: the forward);
// nslice perturbation slices [nslice][nz][nxg] (1 unless ext)
static int uneven_born(const char* name, const std::vector<UShot>& shots, bool ext, int nslice, int ngs, int ngr,
                       int nz, int nxg, const float2* dS, const float2* dR, float2* gS, float2* gR, int add,
                       int mode, hipStream_t s) {
  const std::string who = name;
  const bool fwd = gS == nullptr;
  const int ns = (int)shots.size();
  if (ngs < 0 || ngr < 0 || nz < 2 || nxg < 1 || nslice < 1) return fail(WEMFWI_ERR_ARG, who + ": invalid argument");
  if (fwd ? (!dS || !dR) : (!gS || !gR)) return fail(WEMFWI_ERR_ARG, who + ": null perturbation");
  // validate every shot before anything is queued
  for (int sh = 0; sh < ns; ++sh) {
    const UShot& S = shots[sh];
    const std::string at = who + ": shot " + std::to_string(sh);
    if (S.fm.empty() || S.nw < 0 || (S.nw > 0 && (!S.freq_hz || !S.wavelet || !S.data)) || S.nr < 0 ||
        (S.nr > 0 && (!S.rec_iz || !S.rec_ix)))
      return fail(WEMFWI_ERR_ARG, at + ": null argument");
    for (Model* mp : S.fm) {
      if (!mp) return fail(WEMFWI_ERR_ARG, at + ": null model");
      if (mp->d.nz != nz) return fail(WEMFWI_ERR_ARG, at + ": model nz differs");
      if (mp->d.nx != S.fm[0]->d.nx) return fail(WEMFWI_ERR_ARG, at + ": the frequencies' models differ in width");
      if (S.nw > 0) RCCHK(mp->require_fft());
    }
    if (ext && (S.w0 < 0 || S.w0 + S.nw > nslice))
      return fail(WEMFWI_ERR_ARG, at + ": band outside the perturbation slices");
    if (S.src_iz < 0 || S.src_iz >= nz) return fail(WEMFWI_ERR_ARG, at + ": source depth outside the model");
    for (int r = 0; r < S.nr; ++r)
      if (S.rec_iz[r] < 0 || S.rec_iz[r] >= nz) return fail(WEMFWI_ERR_ARG, at + ": receiver depth outside the model");
  }
  // ext, mode 0: per shot and frequency, the aperture columns [a, b) of the
  // last depth row that a later shot with the same slice also covers
  ColCuts cut;
  const bool rowlast = ext && !fwd && mode == 0;
  if (rowlast) {
    cut.resize(ns);
    for (int k = 0; k < ns; ++k) {
      const int ax = shots[k].x0, bx = ax + shots[k].fm[0]->d.nx;
      cut[k].resize(shots[k].nw);
      for (int i = 0; i < shots[k].nw; ++i) {
        const int q = shots[k].w0 + i;
        std::vector<std::pair<int, int>> iv;
        for (int k2 = k + 1; k2 < ns; ++k2)
          if (q >= shots[k2].w0 && q < shots[k2].w0 + shots[k2].nw) {
            const int a = std::max(ax, shots[k2].x0), b = std::min(bx, shots[k2].x0 + shots[k2].fm[0]->d.nx);
            if (a < b) iv.push_back({a - ax, b - ax});
          }
        std::sort(iv.begin(), iv.end());
        for (const auto& c : iv) {
          if (!cut[k][i].empty() && c.first <= cut[k][i].back().second)
            cut[k][i].back().second = std::max(cut[k][i].back().second, c.second);
          else
            cut[k][i].push_back(c);
        }
      }
    }
  }
  size_t budget = 0;
  {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) fr = (size_t)8 << 30;
    budget = fr / 2;
  }
  DevBuf GS, GR;
  const size_t npix = (size_t)nz * nxg, nacc = npix * (size_t)nslice;
  if (fwd) {
    for (int sh = 0; sh < ns; ++sh)
      if (shots[sh].nw > 0 && shots[sh].nr > 0 && !add)
        HIPCHK(launch_zero2d(shots[sh].data, (long long)shots[sh].nr * shots[sh].nw,
                             (long long)shots[sh].nr * shots[sh].nw, 1, s),
               "zero data");
  } else {
    RCCHK(GS.alloc(nacc * sizeof(double2), "gradient accumulator"));
    RCCHK(GR.alloc(nacc * sizeof(double2), "gradient accumulator"));
    HIPCHK(launch_zero2d(GS.p, (long long)(2 * nacc), (long long)(2 * nacc), 1, s), "zero");
    HIPCHK(launch_zero2d(GR.p, (long long)(2 * nacc), (long long)(2 * nacc), 1, s), "zero");
  }
  // groups of consecutive shots whose wavefields fit half the free memory
  const size_t nbuf = fwd ? 3 : 4;
  int s0 = 0;
  while (s0 < ns) {
    int s1 = s0;
    size_t bytes = 0;
    while (s1 < ns) {
      const Model& m = *shots[s1].fm[0];
      const size_t per = ext ? (nbuf + (fwd ? 3 : 4)) * (size_t)shots[s1].nw * sizeof(float2)
                             : nbuf * (size_t)shots[s1].nw * sizeof(float2) + 4 * sizeof(double2);
      const size_t b = per * m.d.nz * m.d.nx;
      if (s1 > s0 && bytes + b > budget) break;
      bytes += b;
      ++s1;
    }
    UnevenBorn u;
    RCCHK(u.init(shots.data(), ext, s0, s1, nz, ngs, ngr, s));
    if (fwd) RCCHK(uneven_born_forward_group(u, dS, dR, nxg));
    else RCCHK(uneven_born_adjoint_group(u, GS.as<double2>(), GR.as<double2>(), nxg, mode, rowlast ? &cut : nullptr));
    s0 = s1;
  }
  if (!fwd) {
    HIPCHK(launch_d2f(gS, GS.as<double2>(), (long long)nacc, add != 0, s), "gradient");
    HIPCHK(launch_d2f(gR, GR.as<double2>(), (long long)nacc, add != 0, s), "gradient");
  }
  HIPCHK(hipStreamSynchronize(s), name);
  return WEMFWI_OK;
}

}  // namespace wemfwi

// ------------------------------------------------------------------ C ABI

using namespace wemfwi;

struct wemfwi_survey {
  Survey s;
};

struct wemfwi_gradient {
  GradientPlan p;
};

extern "C" {

int wemfwi_survey_create(const wemfwi_survey_desc* d, wemfwi_survey** out) {
  if (!d || !out) return fail(WEMFWI_ERR_ARG, "survey_create: null argument");
  *out = nullptr;
  if (d->ns < 1 || d->nr < 1 || d->nw < 1 || d->nt < 1 || !d->src_iz || !d->src_ix || !d->rec_iz || !d->rec_ix ||
      !d->freq_hz || !d->bin || !d->wavelet || d->ngs < 0 || d->ngr < 0 || d->ngs > 31 || d->ngr > 31)
    return fail(WEMFWI_ERR_ARG, "survey_create: invalid survey");
  for (int i = 0; i < d->nw; ++i)
    if (d->bin[i] < 0 || d->bin[i] >= d->nt) return fail(WEMFWI_ERR_ARG, "survey_create: frequency bin outside trace");
  wemfwi_survey* h = new (std::nothrow) wemfwi_survey();
  if (!h) return fail(WEMFWI_ERR_ALLOC, "survey_create: out of host memory");
  Survey& s = h->s;
  s.d = *d;
  s.src_iz.assign(d->src_iz, d->src_iz + d->ns);
  s.src_ix.assign(d->src_ix, d->src_ix + d->ns);
  s.rec_iz.assign(d->rec_iz, d->rec_iz + d->nr);
  s.rec_ix.assign(d->rec_ix, d->rec_ix + d->nr);
  s.freq.assign(d->freq_hz, d->freq_hz + d->nw);
  s.bin.assign(d->bin, d->bin + d->nw);
  for (int i = 0; i < d->nw; ++i) s.wave.push_back(cf(d->wavelet[i].re, d->wavelet[i].im));
  s.d.src_iz = s.src_iz.data();
  s.d.src_ix = s.src_ix.data();
  s.d.rec_iz = s.rec_iz.data();
  s.d.rec_ix = s.rec_ix.data();
  s.d.freq_hz = s.freq.data();
  s.d.bin = s.bin.data();
  s.d.wavelet = nullptr;
  *out = h;
  return WEMFWI_OK;
}

int wemfwi_survey_destroy(wemfwi_survey* s) {
  delete s;
  return WEMFWI_OK;
}

int wemfwi_survey_set_checkpoint(wemfwi_survey* survey, int every) {
  if (!survey || every < 0) return fail(WEMFWI_ERR_ARG, "survey_set_checkpoint: invalid argument");
  survey->s.ckpt = every;
  return WEMFWI_OK;
}

int wemfwi_wem_forward(wemfwi_model* model, wemfwi_survey* survey, wemfwi_complex* data, int add, void* stream) {
  if (!model || !survey || !data) return fail(WEMFWI_ERR_ARG, "wem_forward: null argument");
  std::vector<Model*> fm(survey->s.d.nw, &model->m);
  return wem_forward(fm, Per::Freq, &survey->s, (float2*)data, add, (hipStream_t)stream);
}

int wemfwi_extwem_forward(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* data, int add,
                          void* stream) {
  if (!models || !survey || !data) return fail(WEMFWI_ERR_ARG, "extwem_forward: null argument");
  std::vector<Model*> fm;
  for (int i = 0; i < survey->s.d.nw; ++i) {
    if (!models[i]) return fail(WEMFWI_ERR_ARG, "extwem_forward: null model");
    fm.push_back(&models[i]->m);
  }
  return wem_forward(fm, Per::Freq, &survey->s, (float2*)data, add, (hipStream_t)stream);
}

int wemfwi_born_forward(wemfwi_model* model, wemfwi_survey* survey, const wemfwi_complex* dslow,
                        const wemfwi_complex* dden, wemfwi_complex* data, int add, void* stream) {
  if (!model || !survey || !dslow || !dden || !data) return fail(WEMFWI_ERR_ARG, "born_forward: null argument");
  std::vector<Model*> fm(survey->s.d.nw, &model->m);
  return born_forward(fm, Per::Freq, &survey->s, (const float2*)dslow, (const float2*)dden, 0, (float2*)data, add,
                      (hipStream_t)stream);
}

int wemfwi_gradient_create(wemfwi_model* model, wemfwi_survey* survey, const wemfwi_complex* obs,
                           wemfwi_complex* res, wemfwi_complex* gS, wemfwi_complex* gR, int mode, void* stream,
                           wemfwi_gradient** out) {
  if (!model || !survey || !obs || !res || !gS || !gR || !out) return fail(WEMFWI_ERR_ARG, "gradient_create: null argument");
  *out = nullptr;
  wemfwi_gradient* h = new (std::nothrow) wemfwi_gradient();
  if (!h) return fail(WEMFWI_ERR_ALLOC, "gradient_create: out of host memory");
  GradientPlan& p = h->p;
  p.m = &model->m;
  p.sv = &survey->s;
  p.obs = (const float2*)obs;
  p.res = (float2*)res;
  p.gS = (float2*)gS;
  p.gR = (float2*)gR;
  p.exact = mode != 0;
  if (int rc = p.build((hipStream_t)stream)) {
    delete h;
    return rc;
  }
  *out = h;
  return WEMFWI_OK;
}

int wemfwi_gradient_launch(wemfwi_gradient* plan, void* stream, double* misfit) {
  if (!plan || !plan->p.exec) return fail(WEMFWI_ERR_ARG, "gradient_launch: null plan");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipGraphLaunch(plan->p.exec, s);
  if (e == hipSuccess && misfit) e = hipMemcpyAsync(misfit, plan->p.d_J(), sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "gradient launch");
}

int wemfwi_gradient_misfit(wemfwi_gradient* plan, double* misfit) {
  if (!plan || !misfit) return fail(WEMFWI_ERR_ARG, "gradient_misfit: null argument");
  hipError_t e = hipMemcpy(misfit, plan->p.d_J(), sizeof(double), hipMemcpyDeviceToHost);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "gradient misfit");
}

int wemfwi_gradient_destroy(wemfwi_gradient* plan) {
  delete plan;
  return WEMFWI_OK;
}

int wemfwi_born_adjoint(wemfwi_model* model, wemfwi_survey* survey, wemfwi_complex* dslow, wemfwi_complex* dden,
                        const wemfwi_complex* data, int add, int mode, void* stream) {
  if (!model || !survey || !dslow || !dden || !data) return fail(WEMFWI_ERR_ARG, "born_adjoint: null argument");
  std::vector<Model*> fm(survey->s.d.nw, &model->m);
  return born_adjoint(fm, Per::Freq, &survey->s, (float2*)dslow, (float2*)dden, 0, (const float2*)data, add,
                      mode != 0, false, (hipStream_t)stream);
}

int wemfwi_extborn_forward(wemfwi_model* const* models, wemfwi_survey* survey, const wemfwi_complex* dslow,
                           const wemfwi_complex* dden, wemfwi_complex* data, int add, void* stream) {
  if (!models || !survey || !dslow || !dden || !data) return fail(WEMFWI_ERR_ARG, "extborn_forward: null argument");
  std::vector<Model*> fm;
  for (int i = 0; i < survey->s.d.nw; ++i) fm.push_back(&models[i]->m);
  const size_t ms = (size_t)fm[0]->d.nz * fm[0]->d.nx;
  return born_forward(fm, Per::Freq, &survey->s, (const float2*)dslow, (const float2*)dden, ms, (float2*)data, add,
                      (hipStream_t)stream);
}

int wemfwi_extborn_adjoint(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* dslow,
                           wemfwi_complex* dden, const wemfwi_complex* data, int add, int mode, void* stream) {
  if (!models || !survey || !dslow || !dden || !data) return fail(WEMFWI_ERR_ARG, "extborn_adjoint: null argument");
  std::vector<Model*> fm;
  for (int i = 0; i < survey->s.d.nw; ++i) fm.push_back(&models[i]->m);
  const size_t ms = (size_t)fm[0]->d.nz * fm[0]->d.nx;
  return born_adjoint(fm, Per::Freq, &survey->s, (float2*)dslow, (float2*)dden, ms, (const float2*)data, add,
                      mode != 0, mode == 0, (hipStream_t)stream);
}

int wemfwi_survey_set_shot_devices(wemfwi_survey* survey, int shot, int nsrc, const int* src_iz, const int* src_ix,
                                   const int* rec_iz, const int* rec_ix) {
  if (!survey || nsrc < 1 || !src_iz || !src_ix || !rec_iz || !rec_ix)
    return fail(WEMFWI_ERR_ARG, "survey_set_shot_devices: invalid argument");
  Survey& s = survey->s;
  if (shot < 0 || shot >= s.d.ns) return fail(WEMFWI_ERR_ARG, "survey_set_shot_devices: shot out of range");
  if (!s.per_shot()) {
    // every shot starts from its point source and the shared receivers
    s.ssrc_iz.resize(s.d.ns);
    s.ssrc_ix.resize(s.d.ns);
    s.srec_iz.assign(s.d.ns, s.rec_iz);
    s.srec_ix.assign(s.d.ns, s.rec_ix);
    for (int i = 0; i < s.d.ns; ++i) {
      s.ssrc_iz[i] = {s.src_iz[i]};
      s.ssrc_ix[i] = {s.src_ix[i]};
    }
    s.ssrc_plan.resize(s.d.ns);
    s.srec_plan.resize(s.d.ns);
  }
  s.ssrc_iz[shot].assign(src_iz, src_iz + nsrc);
  s.ssrc_ix[shot].assign(src_ix, src_ix + nsrc);
  s.srec_iz[shot].assign(rec_iz, rec_iz + s.d.nr);
  s.srec_ix[shot].assign(rec_ix, rec_ix + s.d.nr);
  s.ssrc_plan[shot].release();
  s.srec_plan[shot].release();
  return WEMFWI_OK;
}

static int shot_models(wemfwi_model* const* models, const Survey& s, std::vector<Model*>& fm, const char* what) {
  if (!models) return fail(WEMFWI_ERR_ARG, std::string(what) + ": null models");
  for (int i = 0; i < s.d.ns; ++i) {
    if (!models[i]) return fail(WEMFWI_ERR_ARG, std::string(what) + ": null model");
    fm.push_back(&models[i]->m);
  }
  return WEMFWI_OK;
}

int wemfwi_shotwem_forward(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* data, int add,
                           void* stream) {
  if (!survey || !data) return fail(WEMFWI_ERR_ARG, "shotwem_forward: null argument");
  std::vector<Model*> fm;
  RCCHK(shot_models(models, survey->s, fm, "shotwem_forward"));
  return wem_forward(fm, Per::Shot, &survey->s, (float2*)data, add, (hipStream_t)stream);
}

int wemfwi_shotborn_forward(wemfwi_model* const* models, wemfwi_survey* survey, const wemfwi_complex* dslow,
                            const wemfwi_complex* dden, wemfwi_complex* data, int add, void* stream) {
  if (!survey || !dslow || !dden || !data) return fail(WEMFWI_ERR_ARG, "shotborn_forward: null argument");
  std::vector<Model*> fm;
  RCCHK(shot_models(models, survey->s, fm, "shotborn_forward"));
  const size_t ms = (size_t)fm[0]->d.nz * fm[0]->d.nx;
  return born_forward(fm, Per::Shot, &survey->s, (const float2*)dslow, (const float2*)dden, ms, (float2*)data, add,
                      (hipStream_t)stream);
}

int wemfwi_shotborn_adjoint(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* dslow,
                            wemfwi_complex* dden, const wemfwi_complex* data, int add, int mode, void* stream) {
  if (!survey || !dslow || !dden || !data) return fail(WEMFWI_ERR_ARG, "shotborn_adjoint: null argument");
  std::vector<Model*> fm;
  RCCHK(shot_models(models, survey->s, fm, "shotborn_adjoint"));
  const size_t ms = (size_t)fm[0]->d.nz * fm[0]->d.nx;
  return born_adjoint(fm, Per::Shot, &survey->s, (float2*)dslow, (float2*)dden, ms, (const float2*)data, add,
                      mode != 0, false, (hipStream_t)stream);
}

int wemfwi_gather_sum(long long nout, const long long* ptr, const long long* src, const wemfwi_complex* in,
                      wemfwi_complex* out, int add, void* stream) {
  if (nout < 0 || (nout > 0 && (!ptr || !src || !in || !out)))
    return fail(WEMFWI_ERR_ARG, "gather_sum: invalid argument");
  if (nout == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(launch_gather_sum((float2*)out, nout, ptr, src, (const float2*)in, add != 0, s), "gather_sum");
  HIPCHK(hipStreamSynchronize(s), "gather_sum");
  return WEMFWI_OK;
}

static int ushots(const wemfwi_born_shot* shots, int ns, std::vector<UShot>& out) {
  if ((!shots && ns > 0) || ns < 0) return fail(WEMFWI_ERR_ARG, "uneven_born: invalid argument");
  out.resize(ns);
  for (int k = 0; k < ns; ++k) {
    const wemfwi_born_shot& S = shots[k];
    UShot& u = out[k];
    u.fm.assign(1, S.model ? &S.model->m : nullptr);
    u.x0 = S.x0;
    u.nw = S.nw;
    u.freq_hz = S.freq_hz;
    u.wavelet = S.wavelet;
    u.src_iz = S.src_iz;
    u.src_ix = S.src_ix;
    u.nr = S.nr;
    u.rec_iz = S.rec_iz;
    u.rec_ix = S.rec_ix;
    u.data = S.data;
  }
  return WEMFWI_OK;
}

static int ushots(const wemfwi_extborn_shot* shots, int ns, std::vector<UShot>& out) {
  if ((!shots && ns > 0) || ns < 0) return fail(WEMFWI_ERR_ARG, "uneven_extborn: invalid argument");
  out.resize(ns);
  for (int k = 0; k < ns; ++k) {
    const wemfwi_extborn_shot& S = shots[k];
    UShot& u = out[k];
    if (!S.models || S.nw < 1) {
      u.fm.assign(1, nullptr);  // rejected by the validation
    } else {
      for (int i = 0; i < S.nw; ++i) u.fm.push_back(S.models[i] ? &S.models[i]->m : nullptr);
    }
    u.x0 = S.x0;
    u.nw = S.nw;
    u.w0 = S.w0;
    u.freq_hz = S.freq_hz;
    u.wavelet = S.wavelet;
    u.src_iz = S.src_iz;
    u.src_ix = S.src_ix;
    u.nr = S.nr;
    u.rec_iz = S.rec_iz;
    u.rec_ix = S.rec_ix;
    u.data = S.data;
  }
  return WEMFWI_OK;
}

int wemfwi_uneven_born_forward(const wemfwi_born_shot* shots, int ns, int ngs, int ngr, int nz, int nx_model,
                               const wemfwi_complex* dslow, const wemfwi_complex* dden, int add, void* stream) {
  std::vector<UShot> u;
  RCCHK(ushots(shots, ns, u));
  return uneven_born("uneven_born_forward", u, false, 1, ngs, ngr, nz, nx_model, (const float2*)dslow,
                     (const float2*)dden, nullptr, nullptr, add, 0, (hipStream_t)stream);
}

int wemfwi_uneven_born_adjoint(const wemfwi_born_shot* shots, int ns, int ngs, int ngr, int nz, int nx_model,
                               wemfwi_complex* dslow, wemfwi_complex* dden, int add, int mode, void* stream) {
  if (mode != 0 && mode != 1) return fail(WEMFWI_ERR_ARG, "uneven_born_adjoint: mode must be 0 or 1");
  std::vector<UShot> u;
  RCCHK(ushots(shots, ns, u));
  return uneven_born("uneven_born_adjoint", u, false, 1, ngs, ngr, nz, nx_model, nullptr, nullptr, (float2*)dslow,
                     (float2*)dden, add, mode, (hipStream_t)stream);
}

int wemfwi_uneven_extborn_forward(const wemfwi_extborn_shot* shots, int ns, int ngs, int ngr, int nz, int nx_model,
                                  int nslice, const wemfwi_complex* dslow, const wemfwi_complex* dden, int add,
                                  void* stream) {
  std::vector<UShot> u;
  RCCHK(ushots(shots, ns, u));
  return uneven_born("uneven_extborn_forward", u, true, nslice, ngs, ngr, nz, nx_model, (const float2*)dslow,
                     (const float2*)dden, nullptr, nullptr, add, 0, (hipStream_t)stream);
}

int wemfwi_uneven_extborn_adjoint(const wemfwi_extborn_shot* shots, int ns, int ngs, int ngr, int nz, int nx_model,
                                  int nslice, wemfwi_complex* dslow, wemfwi_complex* dden, int add, int mode,
                                  void* stream) {
  if (mode != 0 && mode != 1) return fail(WEMFWI_ERR_ARG, "uneven_extborn_adjoint: mode must be 0 or 1");
  std::vector<UShot> u;
  RCCHK(ushots(shots, ns, u));
  return uneven_born("uneven_extborn_adjoint", u, true, nslice, ngs, ngr, nz, nx_model, nullptr, nullptr,
                     (float2*)dslow, (float2*)dden, add, mode, (hipStream_t)stream);
}

int wemfwi_bornrefl_forward(wemfwi_model* model, wemfwi_survey* survey, const wemfwi_complex* image,
                            wemfwi_complex* data, int add, void* stream) {
  if (!model || !survey || !image || !data) return fail(WEMFWI_ERR_ARG, "bornrefl_forward: null argument");
  return bornrefl_forward(&model->m, &survey->s, (const float2*)image, 0, (float2*)data, add, (hipStream_t)stream);
}

int wemfwi_bornrefl_adjoint(wemfwi_model* model, wemfwi_survey* survey, wemfwi_complex* image,
                            const wemfwi_complex* data, int add, void* stream) {
  if (!model || !survey || !image || !data) return fail(WEMFWI_ERR_ARG, "bornrefl_adjoint: null argument");
  return bornrefl_adjoint(&model->m, &survey->s, (float2*)image, 0, (const float2*)data, add, (hipStream_t)stream);
}

int wemfwi_offset_image_forward(wemfwi_model* model, wemfwi_survey* survey, int nh, const wemfwi_complex* image,
                                wemfwi_complex* data, int add, void* stream) {
  if (!model || !survey || !image || !data || nh < 1 || nh % 2 == 0)
    return fail(WEMFWI_ERR_ARG, "offset_image_forward: invalid argument (nh must be odd)");
  return bornrefl_forward(&model->m, &survey->s, (const float2*)image, nh, (float2*)data, add, (hipStream_t)stream);
}

int wemfwi_offset_image_adjoint(wemfwi_model* model, wemfwi_survey* survey, int nh, wemfwi_complex* image,
                                const wemfwi_complex* data, int add, void* stream) {
  if (!model || !survey || !image || !data || nh < 1 || nh % 2 == 0)
    return fail(WEMFWI_ERR_ARG, "offset_image_adjoint: invalid argument (nh must be odd)");
  return bornrefl_adjoint(&model->m, &survey->s, (float2*)image, nh, (const float2*)data, add, (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------ unit operators

namespace wemfwi {

// per-row frequency tables and zero/validation helpers of the unit operators
static int unit_rows(Model& m, int nrows, const float* freq_hz, std::vector<float>& w,
                     std::vector<const float2*>& p, hipStream_t s) {
  if (nrows < 0 || (nrows > 0 && !freq_hz)) return fail(WEMFWI_ERR_ARG, "unit operator: invalid rows");
  if (nrows == 0) return WEMFWI_OK;
  return m.phase_tables(nrows, freq_hz, w, p, s);
}

static int unit_scatter(Model& m, int adjoint, int mode, int nrows, const float* freq_hz, const int* depth,
                        const float2* in, float2* out, int add, hipStream_t s) {
  if (!in || !out || (nrows > 0 && !depth)) return fail(WEMFWI_ERR_ARG, "scatter: null argument");
  for (int i = 0; i < nrows; ++i)
    if (depth[i] < 0 || depth[i] >= m.d.nz) return fail(WEMFWI_ERR_ARG, "scatter: depth out of range");
  std::vector<float> w;
  std::vector<const float2*> p;
  RCCHK(unit_rows(m, nrows, freq_hz, w, p, s));
  if (nrows == 0) return WEMFWI_OK;
  std::vector<StepRow> h(nrows);
  for (int i = 0; i < nrows; ++i)
    h[i] = StepRow{in + (size_t)i * m.d.nx, out + (size_t)i * m.d.nx, p[i], w[i], depth[i], 0, add ? 1 : 0,
                   nullptr};
  DevBuf rows;
  RCCHK(put(rows, h, s, "scatter rows"));
  HIPCHK(launch_scatter(m.N, rows.as<StepRow>(), nrows, m.d_tables, m.geom(), adjoint != 0, mode != 0, m.d_tw, s),
         "scatter kernel");
  HIPCHK(hipStreamSynchronize(s), "scatter");
  return WEMFWI_OK;
}

static int unit_lin(Model& m, int up, int adjoint, int mode, int nrows, const float* freq_hz, const float2* bg,
                    const float2* in, float2* out, int add, hipStream_t s) {
  if (!bg || !in || !out) return fail(WEMFWI_ERR_ARG, "lin_oneway: null argument");
  std::vector<float> w;
  std::vector<const float2*> p;
  RCCHK(unit_rows(m, nrows, freq_hz, w, p, s));
  if (nrows == 0) return WEMFWI_OK;
  const size_t nzx = (size_t)m.d.nz * m.d.nx;
  if (!add) {
    HIPCHK(hipMemsetAsync(out, 0, nrows * nzx * sizeof(float2), s), "memset");
  } else {
    // the row the operator overwrites with zeros even when adding: LinDown
    // forward row 0 (LinOneWay.cpp:33-58), LinDown adjoint and LinUp forward
    // row nz-1 (LinOneWay.cpp:60-115); the kernel writes every other row
    const int zrow = (!up && !adjoint) ? 0 : ((!up || !adjoint) ? m.d.nz - 1 : -1);
    if (zrow >= 0)
      HIPCHK(hipMemset2DAsync(out + (size_t)zrow * m.d.nx, nzx * sizeof(float2), 0, m.d.nx * sizeof(float2),
                              nrows, s),
             "memset");
  }
  std::vector<LinRow> h(nrows);
  for (int i = 0; i < nrows; ++i)
    h[i] = LinRow{const_cast<float2*>(bg) + i * nzx, in + i * nzx, out + i * nzx, p[i], w[i], 0, 0, nullptr, 0};
  DevBuf rows;
  RCCHK(put(rows, h, s, "lin rows"));
  HIPCHK(launch_lin(m.N, rows.as<LinRow>(), nrows, m.d_tables, m.geom(), m.d.nz, up ? -1 : +1, adjoint != 0,
                    mode != 0, m.d_tw, s),
         "lin kernel");
  HIPCHK(hipStreamSynchronize(s), "lin_oneway");
  return WEMFWI_OK;
}

static int unit_injection(int nz, int nx, int ndev, const int* iz, const int* ix, int ng, int tap, int adjoint,
                          int nrows, float2* vals, float2* wfld, int add, hipStream_t s) {
  if (nz < 1 || nx < 1 || ndev < 1 || !iz || !ix || ng < 0 || tap < 0 || nrows < 0 || !vals || !wfld)
    return fail(WEMFWI_ERR_ARG, "injection: invalid argument");
  if (nrows == 0) return WEMFWI_OK;
  DevPlan P;
  std::vector<int> viz(iz, iz + ndev), vix(ix, ix + ndev);
  int rc = build_plan(P, viz, vix, ng, tap, nx, nz);
  if (rc) {
    P.release();
    return rc;
  }
  std::vector<long long> vb(nrows);
  for (int i = 0; i < nrows; ++i) vb[i] = (long long)i * ndev;
  DevBuf dvb;
  const size_t nzx = (size_t)nz * nx;
  hipError_t e = hipSuccess;
  if ((rc = put(dvb, vb, s, "device offsets")) == WEMFWI_OK) {
    if (!adjoint) {
      if (!add) e = hipMemsetAsync(wfld, 0, nrows * nzx * sizeof(float2), s);
      if (e == hipSuccess) e = launch_inject_csr(wfld, (long long)nzx, P.inj, vals, dvb.as<long long>(), 1, nrows, s);
    } else {
      if (!add) e = hipMemsetAsync(vals, 0, (size_t)nrows * ndev * sizeof(float2), s);
      if (e == hipSuccess) e = launch_record(vals, dvb.as<long long>(), 1, wfld, (long long)nzx, P.rec, nrows, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "injection");
  }
  P.release();
  return rc;
}

}  // namespace wemfwi

extern "C" {

int wemfwi_scatter(wemfwi_model* model, int adjoint, int mode, int nrows, const float* freq_hz, const int* depth,
                   const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (!model) return fail(WEMFWI_ERR_ARG, "scatter: null model");
  return unit_scatter(model->m, adjoint, mode, nrows, freq_hz, depth, (const float2*)in, (float2*)out, add,
                      (hipStream_t)stream);
}

int wemfwi_lin_oneway(wemfwi_model* model, int up, int adjoint, int mode, int nrows, const float* freq_hz,
                      const wemfwi_complex* bg, const wemfwi_complex* in, wemfwi_complex* out, int add,
                      void* stream) {
  if (!model) return fail(WEMFWI_ERR_ARG, "lin_oneway: null model");
  return unit_lin(model->m, up, adjoint, mode, nrows, freq_hz, (const float2*)bg, (const float2*)in, (float2*)out,
                  add, (hipStream_t)stream);
}

int wemfwi_reflect(wemfwi_model* model, int adjoint, int nrows, const wemfwi_complex* in, wemfwi_complex* out,
                   int add, void* stream) {
  if (!model || !in || !out || nrows < 0) return fail(WEMFWI_ERR_ARG, "reflect: invalid argument");
  if (nrows == 0) return WEMFWI_OK;
  Model& m = model->m;
  hipStream_t s = (hipStream_t)stream;
  const size_t nzx = (size_t)m.d.nz * m.d.nx;
  std::vector<MulRow> h(nrows);
  for (int i = 0; i < nrows; ++i)
    h[i] = MulRow{(float2*)out + i * nzx, m.d_refl, (const float2*)in + i * nzx};
  DevBuf rows;
  RCCHK(put(rows, h, s, "reflect rows"));
  HIPCHK(launch_mul_rows(rows.as<MulRow>(), nrows, (long long)nzx, adjoint != 0, add != 0, s), "reflect");
  HIPCHK(hipStreamSynchronize(s), "reflect");
  return WEMFWI_OK;
}

int wemfwi_dreflect(wemfwi_model* model, int adjoint, wemfwi_complex* dslow, wemfwi_complex* dden,
                    wemfwi_complex* dr, int add, void* stream) {
  if (!model || !dslow || !dden || !dr) return fail(WEMFWI_ERR_ARG, "dreflect: null argument");
  Model& m = model->m;
  hipStream_t s = (hipStream_t)stream;
  const size_t nzx = (size_t)m.d.nz * m.d.nx;
  if (!adjoint) {
    HIPCHK(launch_dreflect_fwd((float2*)dr, m.d_drefl, (const float2*)dslow, (const float2*)dden, m.d.nz, m.d.nx,
                               add != 0, s),
           "dreflect");
  } else {
    if (!add) {
      HIPCHK(hipMemsetAsync(dslow, 0, nzx * sizeof(float2), s), "memset");
      HIPCHK(hipMemsetAsync(dden, 0, nzx * sizeof(float2), s), "memset");
    }
    HIPCHK(launch_dreflect_adj((float2*)dslow, (float2*)dden, m.d_drefl, (const float2*)dr, m.d.nz, m.d.nx, s),
           "dreflect adjoint");
  }
  HIPCHK(hipStreamSynchronize(s), "dreflect");
  return WEMFWI_OK;
}

int wemfwi_phshift(int nk, float dx, float dz, float eps, int adjoint, int nrows, const float* freq_hz,
                   const wemfwi_complex* sref, const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (nk < 2 || !(dx > 0.f) || nrows < 0 || (nrows > 0 && (!freq_hz || !sref || !in || !out)))
    return fail(WEMFWI_ERR_ARG, "phshift: invalid argument");
  if (nrows == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  // Phshift.h:24 w = 2 pi freq; Phshift_cmplx.cpp:13 dk = 2 pi / (dx (nk - 1))
  std::vector<double> w(nrows);
  std::vector<double2> sr(nrows);
  for (int i = 0; i < nrows; ++i) {
    w[i] = 2.0 * M_PI * (double)freq_hz[i];
    sr[i] = make_double2(sref[i].re, sref[i].im);
  }
  DevBuf dw, dsr;
  RCCHK(put(dw, w, s, "phshift frequencies"));
  RCCHK(put(dsr, sr, s, "phshift slownesses"));
  const double dk = 2.0 * M_PI / ((double)dx * (double)(nk - 1));
  HIPCHK(launch_phshift((float2*)out, (const float2*)in, nk, dk, (double)dz, (double)eps, dw.as<double>(),
                        dsr.as<double2>(), nrows, adjoint != 0, add != 0, s),
         "phshift");
  HIPCHK(hipStreamSynchronize(s), "phshift");
  return WEMFWI_OK;
}

int wemfwi_splitstep(wemfwi_model* model, int adjoint, int nrows, const float* freq_hz, const int* depth,
                     const wemfwi_complex* sref, int nloc, const int* loc, const wemfwi_complex* in,
                     wemfwi_complex* out, int add, void* stream) {
  if (!model || nrows < 0 || nloc < 0 || (nrows > 0 && (!freq_hz || !depth || !sref || !in || !out)) ||
      (nloc > 0 && !loc))
    return fail(WEMFWI_ERR_ARG, "splitstep: invalid argument");
  Model& m = model->m;
  const int nx = m.d.nx, nz = m.d.nz;
  for (int i = 0; i < nrows; ++i)
    if (depth[i] < 0 || depth[i] >= nz) return fail(WEMFWI_ERR_ARG, "splitstep: depth out of range");
  for (int j = 0; j < nloc; ++j)
    if (loc[j] < 0 || loc[j] >= nx) return fail(WEMFWI_ERR_ARG, "splitstep: location out of range");
  hipStream_t s = (hipStream_t)stream;
  if (!add && nrows > 0)
    HIPCHK(hipMemsetAsync(out, 0, (size_t)nrows * nx * sizeof(float2), s), "memset");
  if (nrows == 0 || nloc == 0) {
    HIPCHK(hipStreamSynchronize(s), "splitstep");
    return WEMFWI_OK;
  }
  // SplitStep.cpp:12-13 / 25-26, multipliers in FP64 on the host from the
  // model's slowness: forward exp(-i w dz (S - sref)/sqrt(sref)), adjoint
  // exp(+i w dz conj((S - sref)/sqrt(sref)))
  using cd = std::complex<double>;
  std::vector<float2> mult((size_t)nrows * nloc);
  for (int i = 0; i < nrows; ++i) {
    const double w = 2.0 * M_PI * (double)freq_hz[i];
    const cd r(sref[i].re, sref[i].im);
    for (int j = 0; j < nloc; ++j) {
      const cf sv = m.slow[(size_t)depth[i] * nx + loc[j]];
      const cd q = (cd(sv.real(), sv.imag()) - r) / std::sqrt(r);
      const cd e = adjoint ? std::exp(cd(0.0, 1.0) * w * (double)m.d.dz * std::conj(q))
                           : std::exp(cd(0.0, -1.0) * w * (double)m.d.dz * q);
      mult[(size_t)i * nloc + j] = make_float2((float)e.real(), (float)e.imag());
    }
  }
  std::vector<int> hl(loc, loc + nloc);
  DevBuf dm, dl;
  RCCHK(put(dm, mult, s, "splitstep multipliers"));
  RCCHK(put(dl, hl, s, "splitstep locations"));
  HIPCHK(launch_gather_mul((float2*)out, (const float2*)in, dm.as<float2>(), dl.as<int>(), nloc, nx, nrows, s),
         "splitstep");
  HIPCHK(hipStreamSynchronize(s), "splitstep");
  return WEMFWI_OK;
}

int wemfwi_ic(int adjoint, long long n, const wemfwi_complex* bg, wemfwi_complex* m, wemfwi_complex* d, int add,
              void* stream) {
  if (n < 0 || !bg || !m || !d) return fail(WEMFWI_ERR_ARG, "ic: invalid argument");
  if (n == 0) return WEMFWI_OK;
  hipStream_t s = (hipStream_t)stream;
  if (!adjoint)
    HIPCHK(launch_mul((float2*)d, 0, (const float2*)bg, 0, (const float2*)m, 0, n, 1, false, add != 0, s), "ic");
  else
    HIPCHK(launch_mul((float2*)m, 0, (const float2*)bg, 0, (const float2*)d, 0, n, 1, true, add != 0, s),
           "ic adjoint");
  HIPCHK(hipStreamSynchronize(s), "ic");
  return WEMFWI_OK;
}

int wemfwi_injection(int nz, int nx, int ndev, const int* iz, const int* ix, int ng, int tap, int adjoint, int nrows,
                     wemfwi_complex* vals, wemfwi_complex* wfld, int add, void* stream) {
  return unit_injection(nz, nx, ndev, iz, ix, ng, tap, adjoint, nrows, (float2*)vals, (float2*)wfld, add,
                        (hipStream_t)stream);
}

}  // extern "C"
