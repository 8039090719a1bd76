// Host layer of libwemfwi: RefSampler, background models, phase tables and
// the C ABI declared in include/wemfwi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"
#include "host.h"
#include "kernels.h"

using cf = std::complex<float>;
using cd = std::complex<double>;

namespace wemfwi {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(WEMFWI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ kernel timing

struct TimingEntry {
  double ms = 0.0;
  long count = 0;
};
static bool g_timing = false;
static std::map<std::string, TimingEntry>& timing_table() {
  static std::map<std::string, TimingEntry> t;
  return t;
}

// HIP events around one kernel launch on its stream; folded into the table
// after the stream has been synchronised.
struct KernelTimer {
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  explicit KernelTimer(const char* n) : name(n) {}
  void begin(hipStream_t s) {
    if (!g_timing) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(a, s);
  }
  void end(hipStream_t s) {
    if (a && b) (void)hipEventRecord(b, s);
  }
  ~KernelTimer() {
    if (a && b) {
      float ms = 0.f;
      if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
        auto& e = timing_table()[name];
        e.ms += ms;
        e.count += 1;
      }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// ------------------------------------------------------------------ RefSampler

int refsampler_host(int nz, int nx, const cf* slow, int nref, cf* sref, int* refid) {
  if (nz < 1 || nx < 2 || nref < 1) return fail(WEMFWI_ERR_ARG, "refsampler: need nz>=1, nx>=2, nref>=1");
  // uniform_sample (RefSampler.cpp:28-39)
  const int o = nx / (2 * nref);
  for (int iz = 0; iz < nz; ++iz)
    for (int r = 0; r < nref; ++r) {
      const int off = o + r * nx / nref;
      sref[(size_t)iz * nref + r] = slow[(size_t)iz * nx + off];
      if (sref[(size_t)iz * nref + r].real() <= 0.f)
        return fail(WEMFWI_ERR_REFERENCE, "Ref velocity " + std::to_string(sref[(size_t)iz * nref + r].real()) +
                                              " at the layer " + std::to_string(iz));
    }
  // getIndexes (RefSampler.cpp:41-84).  The reference buffer has nx entries of
  // which only nref are refreshed per depth; the whole buffer is sorted.
  std::vector<float> buf(nx, 0.f), s(nx);
  std::vector<int> is(nx), isr(nref);
  std::vector<std::pair<int, int>> lists(nref);
  for (int iz = 0; iz < nz; ++iz) {
    for (int x = 0; x < nx; ++x) s[x] = slow[(size_t)iz * nx + x].real();
    for (int r = 0; r < nref; ++r) buf[r] = sref[(size_t)iz * nref + r].real();
    std::iota(is.begin(), is.end(), 0);
    std::iota(isr.begin(), isr.end(), 0);
    std::stable_sort(is.begin(), is.end(), [&](int a, int b) { return s[a] < s[b]; });
    std::stable_sort(isr.begin(), isr.end(), [&](int a, int b) { return buf[a] < buf[b]; });
    std::sort(s.begin(), s.end());
    std::sort(buf.begin(), buf.end());
    for (auto& l : lists) l = {0, 0};
    int off = 0;
    while (off < nx && (s[off] - buf[0]) < (buf[1] - s[off])) ++off;
    lists[isr[0]] = {0, off};
    int off_beg = off;
    for (int ir = 1; ir < nref - 1; ++ir) {
      while (off < nx && (s[off] - buf[ir]) < (buf[ir + 1] - s[off])) ++off;
      lists[isr[ir]] = {off_beg, off};
      off_beg = off;
    }
    lists[isr[nref - 1]] = {off, nx};
    int* row = refid + (size_t)iz * nx;
    std::fill(row, row + nx, -1);
    for (int r = 0; r < nref; ++r)
      for (int i = lists[r].first; i < lists[r].second; ++i) row[is[i]] = r;
  }
  return WEMFWI_OK;
}

// ------------------------------------------------------------------ device helpers

namespace {

struct BlockCache {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void*>> free_blocks;  // device -> size -> block
  std::map<int, size_t> held;                               // device -> cached bytes
  std::map<int, size_t> limit;                              // device -> cap on cached bytes
  std::map<void*, std::pair<int, size_t>> live;             // block -> (device, size)
};

BlockCache& block_cache() {
  static BlockCache* c = new BlockCache;  // never destroyed: blocks go back with the process
  return *c;
}

size_t cache_limit(BlockCache& c, int dev) {
  auto it = c.limit.find(dev);
  if (it != c.limit.end()) return it->second;
  size_t lim = 0;
  if (const char* e = std::getenv("WEMFWI_CACHE_LIMIT_MB")) {
    lim = (size_t)std::max(0LL, std::atoll(e)) << 20;
  } else {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) lim = tot / 2;
  }
  c.limit[dev] = lim;
  return lim;
}

// frees the cached blocks of device dev (the caller holds the lock and has
// made dev current)
void flush_device(BlockCache& c, int dev) {
  auto& m = c.free_blocks[dev];
  for (auto& b : m) (void)hipFree(b.second);
  m.clear();
  c.held[dev] = 0;
}

}  // namespace

int cached_malloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) return WEMFWI_OK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto& m = c.free_blocks[dev];
  auto it = m.lower_bound(bytes);
  if (it != m.end() && it->first <= bytes + bytes / 4 + ((size_t)2 << 20)) {
    *p = it->second;
    c.held[dev] -= it->first;
    c.live[*p] = {dev, it->first};
    m.erase(it);
    return WEMFWI_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && !m.empty()) {
    (void)hipGetLastError();
    flush_device(c, dev);
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(WEMFWI_ERR_ALLOC, std::string("hipMalloc failed for ") + what + " (" + std::to_string(bytes) +
                                      " bytes): " + hipGetErrorString(e));
  }
  c.live[*p] = {dev, bytes};
  return WEMFWI_OK;
}

void cached_free(void* p) {
  if (!p) return;
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.live.find(p);
  if (it == c.live.end()) {
    (void)hipFree(p);
    return;
  }
  const int dev = it->second.first;
  const size_t size = it->second.second;
  c.live.erase(it);
  int cur = dev;
  (void)hipGetDevice(&cur);
  if (cur != dev) (void)hipSetDevice(dev);
  // the block may still be read or written by queued work (an error path
  // returns before its stream is drained): wait as hipFree does
  (void)hipDeviceSynchronize();
  if (c.held[dev] + size <= cache_limit(c, dev)) {
    c.free_blocks[dev].emplace(size, p);
    c.held[dev] += size;
  } else {
    (void)hipFree(p);
  }
  if (cur != dev) (void)hipSetDevice(cur);
}

void Workspace::release() {
  if (ptr) cached_free(ptr);
  ptr = nullptr;
  cap = 0;
}

int Workspace::reserve(size_t bytes, const char* what) {
  if (bytes <= cap) return WEMFWI_OK;
  release();
  if (int rc = cached_malloc(&ptr, bytes, what)) return rc;
  cap = bytes;
  return WEMFWI_OK;
}

static std::map<int, float2*>& twiddle_cache() {
  static std::map<int, float2*> c;
  return c;
}

// Runtime FFT plan of a length n that is not a built power of two, in rows
// whose largest register radix is maxr (16 or 4): the power of two as passes
// of the largest radix, then one smaller power-of-two pass, then the odd
// primes in increasing order (3 and 5 as register butterflies, larger
// primes up to kGenMaxRadix as direct DFTs).  False when n has a larger
// prime factor.
bool fft_gen_plan(int n, int maxr, FftGen* f) {
  *f = FftGen{};
  f->n = n;
  if (n < 2) return false;
  int m = n, np = 0;
  std::vector<int> rad;
  while (m % maxr == 0) {
    rad.push_back(maxr);
    m /= maxr;
  }
  for (int r : {8, 4, 2})
    if (r < maxr && m % r == 0) {
      rad.push_back(r);
      m /= r;
    }
  for (int p = 3; m > 1; p += 2) {
    while (m % p == 0) {
      if (p > kGenMaxRadix) return false;
      rad.push_back(p);
      m /= p;
    }
  }
  if (rad.size() > 10) return false;
  for (int r : rad) f->rad |= (unsigned long long)r << (6 * np++);
  f->np = np;
  return true;
}

// Twiddle table of a runtime plan (layout of fft_row_gen), padded with
// zeros to the gen_tw_cap(N) entries a launch stages in LDS; double
// precision rounded once (round_complex), or split into float pairs in
// gen_tw_pair passes.  Cached per (n, N, device).
static std::map<long long, float2*>& gen_twiddle_cache() {
  static std::map<long long, float2*> c;
  return c;
}

int gen_twiddles(const FftGen& f, int N, const float2** out) {
  auto& c = gen_twiddle_cache();
  int dev = 0;
  (void)hipGetDevice(&dev);
  const long long key = ((long long)f.n * 4096 + N) * 64 + dev;
  auto it = c.find(key);
  if (it != c.end()) {
    *out = it->second;
    return WEMFWI_OK;
  }
  size_t total = 0;
  for (int p = 0, NS = 1; p < f.np; NS *= gen_radix(f, p), ++p) {
    const int R = gen_radix(f, p);
    total += (size_t)(R - 1) * NS * (gen_tw_pair(R, NS) ? 2 : 1) + (gen_direct(R) ? R : 0);
  }
  if (total > (size_t)gen_tw_cap(N)) return fail(WEMFWI_ERR_UNSUPPORTED, "FFT plan: twiddle table too large");
  std::vector<float2> h((size_t)gen_tw_cap(N), make_float2(0.f, 0.f));
  size_t off = 0;
  int NS = 1;
  for (int p = 0; p < f.np; ++p) {
    const int R = gen_radix(f, p);
    const bool pair = gen_tw_pair(R, NS);
    for (int r = 1; r < R; ++r)
      for (int m = 0; m < NS; ++m) {
        const double a = -2.0 * M_PI * (double)((long long)m * r) / (double)((long long)NS * R);
        const double c = std::cos(a), s = std::sin(a);
        const size_t e = (size_t)(r - 1) * NS + m;
        if (pair) {
          // hi then lo of each part (cmul_tw_pair)
          const float2 hi = make_float2((float)c, (float)s);
          h[off + 2 * e] = hi;
          h[off + 2 * e + 1] = make_float2((float)(c - (double)hi.x), (float)(s - (double)hi.y));
        } else {
          h[off + e] = round_complex(c, s);
        }
      }
    off += (size_t)(R - 1) * NS * (pair ? 2 : 1);
    if (gen_direct(R)) {
      for (int q = 0; q < R; ++q) {
        const double a = -2.0 * M_PI * (double)q / (double)R;
        h[off + q] = round_complex(std::cos(a), std::sin(a));
      }
      off += R;
    }
    NS *= R;
  }
  float2* d = nullptr;
  int rc = dev_alloc(&d, h.size(), "twiddle table");
  if (rc) return rc;
  hipError_t e = hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "twiddle table");
  c[key] = d;
  *out = d;
  return WEMFWI_OK;
}

int twiddles(int n, const float2** out) { return twiddles(n, 0, out); }

int twiddles_cached(int n, int plan, const float2** out) {
  auto& c = twiddle_cache();
  int dev = 0;
  (void)hipGetDevice(&dev);
  auto it = c.find((n * 8 + plan) * 64 + dev);
  if (it == c.end()) return WEMFWI_ERR_ARG;
  *out = it->second;
  return WEMFWI_OK;
}

int twiddles(int n, int plan, const float2** out) {
  auto& c = twiddle_cache();
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int key = (n * 8 + plan) * 64 + dev;
  auto it = c.find(key);
  if (it != c.end()) {
    *out = it->second;
    return WEMFWI_OK;
  }
  float2* d = nullptr;
  if (twiddle_table_size(n, plan) <= 0) return fail(WEMFWI_ERR_ARG, "twiddles: no such FFT plan");
  int rc = dev_alloc(&d, (size_t)twiddle_table_size(n, plan), "twiddle table");
  if (rc) return rc;
  hipError_t e = launch_twiddles(n, plan, d, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "twiddle table");
  c[key] = d;
  *out = d;
  return WEMFWI_OK;
}

// ------------------------------------------------------------------ model

Model::~Model() {
  void* ptrs[] = {d_slot_begin, d_slot_ref, d_rid_lane, d_one_ref, d_phi_lane, d_taper_lane, d_slot_sref, d_tables,
                  d_scoef, d_refl, d_drefl, d_phi_max};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  ptab.release();
  wlist.release();
  rows.release();
  scratch.release();
}

int Model::build(const wemfwi_model_desc& desc, const cf* slow_h, const cf* den_h, const cf* slow_refl_h) {
  d = desc;
  if (d.nz < 2 || d.nx < 2 || d.nref < 1 || d.tap < 0 || d.taper < 0 || d.ntaylor < 0 ||
      d.ntaylor > 4 || !(d.dx > 0) || !(d.dz > 0))
    return fail(WEMFWI_ERR_ARG, "model: invalid grid or parameters");
  if (d.taper == 1) return fail(WEMFWI_ERR_ARG, "model: taper width pad=1 is undefined (Taper.h divides by pad-1)");
  if (2 * d.taper > d.nx) return fail(WEMFWI_ERR_ARG, "model: taper wider than half the row");
  nk = d.nx + 2 * d.tap;
  N = 128;
  while (N < nk && N < 2048) N *= 2;
  // the background tables (Reflect, dReflect, scatterer coefficients,
  // RefSampler) need no FFT; the propagating entry points check fft_ok
  if (nk == N) {
    fg = FftGen{};
    fg.n = N;
    fft_ok = fft_size_supported(N);
  } else {
    fft_ok = nk <= N && fft_gen_plan(nk, gen_max_radix(N), &fg);
    // the compile-time chain of this length (if any) runs over the twiddle
    // table of this runtime plan: they must be the same chain
    const GenChain c = gen_chain(nk, gen_max_radix(N));
    bool same = fft_ok && c.P == fg.np;
    for (int p = 0; same && p < fg.np; ++p) same = gen_radix(fg, p) == c.R[p];
    if (fft_ok && !same) return fail(WEMFWI_ERR_UNSUPPORTED, "model: FFT pass plans disagree");
  }
  const size_t nzx = (size_t)d.nz * d.nx;
  slow.assign(slow_h, slow_h + nzx);
  if (den_h) den.assign(den_h, den_h + nzx);
  else den.assign(nzx, cf(1.f, 0.f));
  sref.resize((size_t)d.nz * d.nref);
  refid.resize(nzx);
  int rc = refsampler_host(d.nz, d.nx, slow.data(), d.nref, sref.data(), refid.data());
  if (rc) return rc;

  // active (depth, reference) slots: references selecting no x are skipped
  slot_begin.assign(d.nz + 1, 0);
  slot_ref.clear();
  std::vector<double2> ssref;
  std::vector<char> used(d.nref);
  for (int iz = 0; iz < d.nz; ++iz) {
    slot_begin[iz] = (int)slot_ref.size();
    std::fill(used.begin(), used.end(), 0);
    for (int x = 0; x < d.nx; ++x) {
      const int r = refid[(size_t)iz * d.nx + x];
      if (r >= 0) used[r] = 1;
    }
    for (int r = 0; r < d.nref; ++r)
      if (used[r]) {
        slot_ref.push_back(r);
        const cf s = sref[(size_t)iz * d.nref + r];
        double2 v;
        v.x = s.real();
        v.y = s.imag();
        ssref.push_back(v);
      }
  }
  slot_begin[d.nz] = (int)slot_ref.size();
  nslots = (int)slot_ref.size();

  // split-step exponent phi = dz (S - Sref)/sqrt(Sref) of x's reference, so
  // that SplitStep.cpp:13 is exp(-i w phi); zero for PSPI.
  std::vector<float2> phi(nzx, make_float2(0.f, 0.f));
  std::vector<signed char> rid8(nzx);
  for (size_t i = 0; i < nzx; ++i) {
    const int r = refid[i];
    rid8[i] = (signed char)r;
    if (r < 0 || d.pspi) continue;
    const size_t iz = i / d.nx;
    const cd sr(sref[iz * d.nref + r].real(), sref[iz * d.nref + r].imag());
    const cd sv(slow[i].real(), slow[i].imag());
    const cd p = (double)d.dz * (sv - sr) / std::sqrt(sr);
    phi[i] = make_float2((float)p.real(), (float)p.imag());
  }
  // Taper.h:11-23
  std::vector<float> taper(d.nx, 1.f);
  for (int i = 0; i < d.taper; ++i) {
    taper[i] = (float)(0.5 * (1.0 - std::cos(M_PI * i / (double)(d.taper - 1))));
    taper[d.nx - 1 - i] = taper[i];
  }

  if ((rc = dev_upload(&d_slot_begin, slot_begin.data(), slot_begin.size(), "slot_begin"))) return rc;
  if ((rc = dev_upload(&d_slot_ref, slot_ref.data(), slot_ref.size(), "slot_ref"))) return rc;
  // lane layout of the FFT row: entry t*E + e is padded sample t + 64 e
  const int E = N / 64;
  std::vector<signed char> rid_lane((size_t)d.nz * N, (signed char)-1);
  std::vector<float2> phi_lane((size_t)d.nz * N, make_float2(0.f, 0.f));
  std::vector<float> taper_lane(N, 0.f);
  std::vector<int> one_ref(d.nz, 0);
  for (int tl = 0; tl < 64; ++tl)
    for (int e = 0; e < E; ++e) {
      const int x = tl + 64 * e - d.tap;
      if (x < 0 || x >= d.nx) continue;
      taper_lane[tl * E + e] = taper[x];
      for (int iz = 0; iz < d.nz; ++iz) {
        rid_lane[(size_t)iz * N + tl * E + e] = rid8[(size_t)iz * d.nx + x];
        phi_lane[(size_t)iz * N + tl * E + e] = phi[(size_t)iz * d.nx + x];
      }
    }
  for (int iz = 0; iz < d.nz; ++iz) {
    if (slot_begin[iz + 1] - slot_begin[iz] != 1) continue;
    const int r = slot_ref[slot_begin[iz]];
    bool all = true;
    for (int x = 0; x < d.nx && all; ++x) all = refid[(size_t)iz * d.nx + x] == r;
    one_ref[iz] = all ? 1 : 0;
  }
  if ((rc = dev_upload(&d_rid_lane, rid_lane.data(), rid_lane.size(), "refid"))) return rc;
  if ((rc = dev_upload(&d_one_ref, one_ref.data(), one_ref.size(), "one_ref"))) return rc;
  if ((rc = dev_upload(&d_phi_lane, phi_lane.data(), phi_lane.size(), "phi"))) return rc;
  if ((rc = dev_upload(&d_taper_lane, taper_lane.data(), taper_lane.size(), "taper"))) return rc;
  if ((rc = dev_upload(&d_slot_sref, ssref.data(), ssref.size(), "slot_sref"))) return rc;
  // Scatterer coefficients coef[it] (S - eps Re S)^(-1/2-it) (Scatter.cpp:82)
  static const double coef[5] = {1., 1. / 2., 3. / 8., 5. / 16., 35. / 128.};
  std::vector<float2> sc((size_t)(d.ntaylor + 1) * nzx);
  for (int it = 0; it <= d.ntaylor; ++it)
    for (size_t i = 0; i < nzx; ++i) {
      const cd s(slow[i].real(), slow[i].imag());
      const cd v = coef[it] * std::pow(s - (double)d.eps * s.real(), -0.5 - it);
      sc[(size_t)it * nzx + i] = make_float2((float)v.real(), (float)v.imag());
    }
  // Reflect.h:19-26 and dReflect.h:18-27,40-47 (coefficients as written);
  // Reflect reads slow_refl when one is given
  const cf* srf = slow_refl_h ? slow_refl_h : slow.data();
  std::vector<float2> rf(nzx, make_float2(0.f, 0.f)), dr(4 * nzx, make_float2(0.f, 0.f));
  for (int iz = 1; iz < d.nz; ++iz)
    for (int x = 0; x < d.nx; ++x) {
      const size_t i = (size_t)iz * d.nx + x, j = i - d.nx;
      const cd s1(slow[j].real(), slow[j].imag()), s2(slow[i].real(), slow[i].imag());
      const cd r1(den[j].real(), den[j].imag()), r2(den[i].real(), den[i].imag());
      const cd q1(srf[j].real(), srf[j].imag()), q2(srf[i].real(), srf[i].imag());
      const cd ra = r2 * std::sqrt(q1), rb = r1 * std::sqrt(q2);
      const cd r = (ra - rb) / (ra + rb);
      rf[i] = make_float2((float)r.real(), (float)r.imag());
      const cd a = r2 * std::sqrt(s1), b = r1 * std::sqrt(s2);
      const cd v = 1.0 / ((a + b) * (a + b));
      const cd s1s2sq = r1 * r2 / v, r1r2sq = 2.0 * s1 * s2 / v;
      const cd c = std::sqrt(s2 / s1);
      const cd q[4] = {c * s1s2sq, 1.0 / c * s1s2sq, r1 * r1r2sq, r2 * r1r2sq};
      for (int k = 0; k < 4; ++k) dr[(size_t)k * nzx + i] = make_float2((float)q[k].real(), (float)q[k].imag());
    }
  if ((rc = dev_upload(&d_scoef, sc.data(), sc.size(), "scatter coefficients"))) return rc;
  if ((rc = dev_upload(&d_refl, rf.data(), rf.size(), "reflection coefficients"))) return rc;
  if ((rc = dev_upload(&d_drefl, dr.data(), dr.size(), "dReflect coefficients"))) return rc;

  DepthTables& t = h_tables;
  t.slot_begin = d_slot_begin;
  t.slot_ref = d_slot_ref;
  t.rid_lane = d_rid_lane;
  t.one_ref = d_one_ref;
  t.phi_lane = d_phi_lane;
  t.taper_lane = d_taper_lane;
  t.scoef = d_scoef;
  t.scoef_stride = (long long)nzx;
  t.nz = d.nz;
  t.n = N;
  t.dz = d.dz;
  t.dk = (float)(2.0 * M_PI / ((double)d.dx * (double)(nk - 1)));
  t.nx = d.nx;
  t.tap = d.tap;
  t.taper = d.taper > 0 ? 1 : 0;
  t.phi_real = 1;
  for (size_t i = 0; i < phi.size() && t.phi_real; ++i) t.phi_real = phi[i].y == 0.f;
  std::vector<float> phi_max(d.nz, 0.f);
  for (size_t i = 0; i < nzx; ++i) phi_max[i / d.nx] = std::max(phi_max[i / d.nx], std::fabs(phi[i].x));
  if ((rc = dev_upload(&d_phi_max, phi_max.data(), phi_max.size(), "phi bound"))) return rc;
  t.phi_max = d_phi_max;
  if ((rc = dev_upload(&d_tables, &t, 1, "depth tables"))) return rc;
  if (fft_ok) {
    if (nk == N) rc = twiddles(N, &d_tw);
    else rc = gen_twiddles(fg, N, &d_tw);
    if (rc) return rc;
  }
  return WEMFWI_OK;
}

StepGeom Model::geom() const {
  StepGeom g;
  g.nx = d.nx;
  g.tap = d.tap;
  g.ssf = d.pspi ? 0 : 1;
  g.full = (d.tap == 0 && d.nx == N) ? 1 : 0;
  g.fg = fg;
  g.pf = 0;
  g.tdma = 0;
  g.ntaylor = d.ntaylor;
  g.taper = d.taper > 0 ? 1 : 0;
  return g;
}

// Phase table of a set of frequencies: exp(-i kz dz)/N per (frequency,
// active slot, k <= N/2) -- the k grid is symmetric, so P[N-k] = P[k].
int Model::require_fft() const {
  if (fft_ok) return WEMFWI_OK;
  return fail(WEMFWI_ERR_UNSUPPORTED, "model: FFT length nx+2*tap=" + std::to_string(nk) +
                                          " not supported (lengths up to 2048 whose prime factors are at most " +
                                          std::to_string(kGenMaxRadix) + ")");
}

int Model::enqueue_tables(hipStream_t s) {
  if (int rc = require_fft()) return rc;
  if (freqs.empty()) return fail(WEMFWI_ERR_ARG, "model: no frequency set to recompute");
  const double dk = 2.0 * M_PI / ((double)d.dx * (double)(nk - 1));
  hipError_t e = launch_phase_table((float2*)ptab.ptr, (const double*)wlist.ptr, (int)freqs.size(), d_slot_sref,
                                    nslots, nk, NH(), dk, (double)d.dz, (double)d.eps, s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "phase table kernel");
}

int Model::set_freqs(int n, const float* freq_hz, hipStream_t s) {
  if (int rc = require_fft()) return rc;
  std::vector<float> uniq(freq_hz, freq_hz + n);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  std::vector<double>& wl = h_wlist;  // member: outlives the asynchronous upload
  wl.resize(uniq.size());
  for (size_t i = 0; i < uniq.size(); ++i) wl[i] = 2.0 * M_PI * (double)uniq[i];
  const size_t per = (size_t)std::max(nslots, 1) * NH();
  int rc = ptab.reserve(uniq.size() * per * sizeof(float2), "phase table");
  if (rc) return rc;
  if ((rc = wlist.reserve(wl.size() * sizeof(double), "frequency list"))) return rc;
  hipError_t e = hipMemcpyAsync(wlist.ptr, wl.data(), wl.size() * sizeof(double), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "frequency upload");
  const double dk = 2.0 * M_PI / ((double)d.dx * (double)(nk - 1));
  KernelTimer kt("phase_table_kernel");
  kt.begin(s);
  e = launch_phase_table((float2*)ptab.ptr, (const double*)wlist.ptr, (int)wl.size(), d_slot_sref, nslots,
                         nk, NH(), dk, (double)d.dz, (double)d.eps, s);
  kt.end(s);
  if (e != hipSuccess) return hip_fail(e, "phase table kernel");
  freqs = uniq;
  ++table_gen;
  return WEMFWI_OK;
}

// Maps rows to the current frequency set, rebuilding it when a row's
// frequency is not in it.
bool Model::has_freqs(int n, const float* freq_hz) const {
  bool have = !freqs.empty();
  for (int i = 0; i < n && have; ++i) have = std::binary_search(freqs.begin(), freqs.end(), freq_hz[i]);
  return have;
}

int Model::phase_tables(int nrows, const float* freq_hz, std::vector<float>& wrow,
                        std::vector<const float2*>& prow, hipStream_t s) {
  if (int rc = require_fft()) return rc;
  if (!has_freqs(nrows, freq_hz)) {
    int rc = set_freqs(nrows, freq_hz, s);
    if (rc) return rc;
  }
  const size_t per = (size_t)std::max(nslots, 1) * NH();
  wrow.resize(nrows);
  prow.resize(nrows);
  for (int i = 0; i < nrows; ++i) {
    const size_t k = std::lower_bound(freqs.begin(), freqs.end(), freq_hz[i]) - freqs.begin();
    wrow[i] = (float)h_wlist[k];
    prow[i] = (const float2*)ptab.ptr + k * per;
  }
  return WEMFWI_OK;
}

int Model::oneway(int up, int adjoint, int nrows, const float* freq_hz, const cf* in, cf* out, int add,
                  int recursive, hipStream_t s) {
  if (nrows < 0 || !freq_hz || !in || !out) return fail(WEMFWI_ERR_ARG, "oneway: null argument");
  if (nrows == 0 || (in == out && !add)) return oneway_on(up, adjoint, nrows, freq_hz, in, out, add, recursive,
                                                          nullptr, nullptr, false, s);
  int rc = require_fft();
  if (rc) return rc;
  if ((rc = scratch.reserve((size_t)nrows * N * sizeof(float2), "scratch rows"))) return rc;
  if ((rc = rows.reserve((size_t)nrows * sizeof(SweepRow), "row descriptors"))) return rc;
  return oneway_on(up, adjoint, nrows, freq_hz, in, out, add, recursive, (float2*)scratch.ptr,
                   (SweepRow*)rows.ptr, false, s);
}

int Model::oneway_on(int up, int adjoint, int nrows, const float* freq_hz, const cf* in, cf* out, int add,
                     int recursive, float2* scr, SweepRow* drows, bool fixed, hipStream_t s) {
  if (nrows < 0 || !freq_hz || !in || !out) return fail(WEMFWI_ERR_ARG, "oneway: null argument");
  if (nrows == 0) return WEMFWI_OK;
  const size_t nzx = (size_t)d.nz * d.nx;
  hipError_t e;
  if (in == out && !add) {
    // OneWay.cpp:34: the shared vector is zeroed before the recursion reads it
    e = hipMemsetAsync(out, 0, nrows * nzx * sizeof(cf), s);
    if (e != hipSuccess) return hip_fail(e, "memset");
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "oneway");
  }
  if (int rc = require_fft()) return rc;
  if (fixed && !has_freqs(nrows, freq_hz))
    return fail(WEMFWI_ERR_ARG, "oneway: a frequency has no phase table (prepare them with set_freqs first)");
  std::vector<float> wrow;
  std::vector<const float2*> prow;
  int rc = phase_tables(nrows, freq_hz, wrow, prow, s);
  if (rc) return rc;
  h_sweep.resize(nrows);
  std::vector<SweepRow>& h = h_sweep;
  for (int i = 0; i < nrows; ++i) {
    h[i].in = (const float2*)(in + i * nzx);
    h[i].out = (float2*)(out + i * nzx);
    h[i].ptab = prow[i];
    h[i].w = wrow[i];
    h[i].model = 0;
    h[i].recursive = recursive ? 1 : 0;
    h[i].first = -1;
    h[i].nsteps = -1;
    h[i].init = nullptr;
    h[i].zbase = 0;
    h[i].smul = nullptr;
    h[i].out2 = nullptr;
    h[i].omul = nullptr;
    h[i].src_skip = 0;
  }
  for (int i = 0; i < nrows; ++i) h[i].scratch = scr + (size_t)i * N;
  e = hipMemcpyAsync(drows, h.data(), h.size() * sizeof(SweepRow), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "row upload");
  const bool down = !up;
  const int step = (down != (adjoint != 0)) ? 1 : -1;
  {
    KernelTimer kt("oneway_kernel");
    kt.begin(s);
    StepGeom g = geom();
    g.pf = 1;  // every row reads its source row at every depth
    {
      std::vector<float> f(freq_hz, freq_hz + nrows);
      std::sort(f.begin(), f.end());
      g.tdma = table_dma((size_t)nrows, (size_t)(std::unique(f.begin(), f.end()) - f.begin()));
    }
    e = launch_oneway(N, drows, nrows, d_tables, g, d.nz, step, adjoint != 0, d_tw, s, !fixed);
    kt.end(s);
    if (e != hipSuccess) return hip_fail(e, "oneway kernel");
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "oneway");
  return WEMFWI_OK;
}

// Streams of the batched sweep: the launches of different kernel variants
// run concurrently (default priority, created once per device).
static int batch_streams(hipStream_t (&out)[4]) {
  static std::map<int, std::vector<hipStream_t>> pool;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::vector<hipStream_t>& v = pool[dev];
  while (v.size() < 4) {
    hipStream_t st;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "stream create");
    v.push_back(st);
  }
  for (int i = 0; i < 4; ++i) out[i] = v[i];
  return WEMFWI_OK;
}

// Plan of an uneven batch: Down/Up recursions out_i = (I - L_i)^-1 in_i of
// rows that each have their own model (aperture, depth count, FFT length) and
// frequency.  Rows are grouped by kernel variant (FFT length, full row, split
// step); each group is one launch, the groups run on separate streams, at
// first the group with the largest LDS footprint first so that its
// workgroups claim compute units before the smaller ones fill the rest, then
// in the measured order (tune_order).  Descriptors, depth tables and scratch
// rows stay on the device between runs.

int BatchPlan::build(int nrows, Model* const* ms, const float* freq_hz, const cf* const* in, cf* const* out,
                     hipStream_t s, const int* src_lo, const int* src_hi, const RowProducts* prod) {
  if (nrows < 1 || !ms || !freq_hz || !in || !out) return fail(WEMFWI_ERR_ARG, "batch: invalid argument");
  std::vector<int> midx(nrows);
  for (int i = 0; i < nrows; ++i) {
    if (!ms[i] || !in[i] || !out[i]) return fail(WEMFWI_ERR_ARG, "batch: null model or buffer");
    if (int rc = ms[i]->require_fft()) return rc;
    auto it = std::find(models.begin(), models.end(), ms[i]);
    midx[i] = (int)(it - models.begin());
    if (it == models.end()) models.push_back(ms[i]);
  }
  // phase tables: every model gets the frequencies of its rows
  std::vector<std::vector<int>> rows_of(models.size());
  for (int i = 0; i < nrows; ++i) rows_of[midx[i]].push_back(i);
  std::vector<float> wrow(nrows);
  std::vector<const float2*> prow(nrows);
  for (size_t u = 0; u < models.size(); ++u) {
    std::vector<float> f;
    for (int i : rows_of[u]) f.push_back(freq_hz[i]);
    std::vector<float> w;
    std::vector<const float2*> p;
    if (int rc = models[u]->phase_tables((int)f.size(), f.data(), w, p, s)) return rc;
    for (size_t k = 0; k < rows_of[u].size(); ++k) {
      wrow[rows_of[u][k]] = w[k];
      prow[rows_of[u][k]] = p[k];
    }
  }
  for (Model* m : models) gens.push_back(m->table_gen);
  // group rows by kernel variant, largest FFT length first
  std::vector<long long> key(nrows);
  for (int i = 0; i < nrows; ++i) {
    const StepGeom g = ms[i]->geom();
    key[i] = -((((long long)ms[i]->N * 4096 + ms[i]->nk) * 2 + g.full) * 2 + g.ssf);
  }
  std::vector<int> order(nrows);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  size_t scr = 0;
  std::vector<size_t> scr_off(nrows);
  for (int k = 0; k < nrows; ++k) {
    scr_off[k] = scr;
    scr += (size_t)ms[order[k]]->N;
  }
  if (int rc = scratch.reserve(scr * sizeof(float2), "batch scratch rows")) return rc;
  std::vector<SweepRow> h(nrows);
  for (int k = 0; k < nrows; ++k) {
    const int i = order[k];
    h[k].in = (const float2*)in[i];
    h[k].out = (float2*)out[i];
    h[k].ptab = prow[i];
    h[k].w = wrow[i];
    h[k].model = midx[i];
    h[k].recursive = 1;
    h[k].first = -1;
    h[k].nsteps = -1;
    h[k].init = nullptr;
    h[k].zbase = 0;
    h[k].smul = (prod && prod->smul) ? prod->smul[i] : nullptr;
    h[k].out2 = (prod && prod->out2) ? prod->out2[i] : nullptr;
    h[k].omul = (prod && prod->omul) ? prod->omul[i] : nullptr;
    h[k].src_skip = (src_lo && src_hi) ? 1 : 0;
    h[k].src_lo = (src_lo && src_hi) ? src_lo[i] : 0;
    h[k].src_hi = (src_lo && src_hi) ? src_hi[i] : -1;
    h[k].scratch = (float2*)scratch.ptr + scr_off[k];
  }
  for (int k0 = 0; k0 < nrows;) {
    int k1 = k0 + 1;
    while (k1 < nrows && key[order[k1]] == key[order[k0]]) ++k1;
    const Model* mg = ms[order[k0]];
    StepGeom g = mg->geom();
    // every row reads a plain source row at every depth: the full-row
    // kernels may prefetch it
    g.pf = (!(src_lo && src_hi) && !(prod && prod->smul)) ? 1 : 0;
    {
      std::vector<std::pair<int, float>> tabs;
      for (int k = k0; k < k1; ++k) tabs.push_back({midx[order[k]], freq_hz[order[k]]});
      std::sort(tabs.begin(), tabs.end());
      g.tdma = table_dma((size_t)(k1 - k0), (size_t)(std::unique(tabs.begin(), tabs.end()) - tabs.begin()));
    }
    groups.push_back(Group{mg->N, k0, k1 - k0, g, mg->d.nz, mg->d_tw});
    k0 = k1;
  }
  launch_order.resize(groups.size());
  std::iota(launch_order.begin(), launch_order.end(), 0);
  // repeatable runs: no input row overlaps an output row (or an out2 row)
  {
    std::vector<std::pair<uintptr_t, uintptr_t>> outs;
    for (int i = 0; i < nrows; ++i) {
      const size_t bytes = (size_t)ms[i]->d.nz * ms[i]->d.nx * sizeof(float2);
      outs.push_back({(uintptr_t)out[i], (uintptr_t)out[i] + bytes});
      if (prod && prod->out2 && prod->out2[i]) outs.push_back({(uintptr_t)prod->out2[i], (uintptr_t)prod->out2[i] + bytes});
    }
    std::sort(outs.begin(), outs.end());
    std::vector<uintptr_t> reach(outs.size());  // largest end among outs[0..k]
    for (size_t k = 0; k < outs.size(); ++k) reach[k] = std::max(outs[k].second, k ? reach[k - 1] : 0);
    rerunnable = true;
    for (int i = 0; i < nrows && rerunnable; ++i) {
      const uintptr_t a0 = (uintptr_t)in[i], a1 = a0 + (size_t)ms[i]->d.nz * ms[i]->d.nx * sizeof(float2);
      // outs starting before a1: the one reaching furthest overlaps [a0, a1) iff it ends after a0
      const size_t k = (size_t)(std::lower_bound(outs.begin(), outs.end(), std::make_pair(a1, (uintptr_t)0)) - outs.begin());
      if (k > 0 && reach[k - 1] > a0) rerunnable = false;
    }
  }
  std::vector<DepthTables> ht;
  for (Model* m : models) ht.push_back(m->h_tables);
  const size_t rows_bytes = (nrows * sizeof(SweepRow) + 255) / 256 * 256;
  if (int rc = desc.reserve(rows_bytes + ht.size() * sizeof(DepthTables), "batch descriptors")) return rc;
  d_rows = (const SweepRow*)desc.ptr;
  d_tabs = (const DepthTables*)((char*)desc.ptr + rows_bytes);
  hipError_t e = hipMemcpyAsync(desc.ptr, h.data(), nrows * sizeof(SweepRow), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync((void*)d_tabs, ht.data(), ht.size() * sizeof(DepthTables), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host staging vectors end here
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "batch descriptor upload");
}

BatchPlan::~BatchPlan() {
  for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  desc.release();
  scratch.release();
}

int BatchPlan::check() const {
  for (size_t u = 0; u < models.size(); ++u)
    if (models[u]->table_gen != gens[u])
      return fail(WEMFWI_ERR_ARG, "batch: a model's phase table was rebuilt after planning; plan again");
  return WEMFWI_OK;
}

int BatchPlan::enqueue(int up, int adjoint, hipStream_t s, bool tables) {
  if (int rc = check()) return rc;
  if (tables)
    for (Model* m : models)
      if (int rc = m->enqueue_tables(s)) return rc;
  if (!order_done)
    if (int rc = tune_order(up, adjoint, s)) return rc;
  return launch(up, adjoint, s, launch_order);
}

// Launch-order decisions per device and group shape (each group's FFT
// length, variant flags, depth count and row count), shared by every plan
// of that shape: plans built per call (the one-call batch, the uneven
// imaging and gradient drivers) time the orders once per process; plans
// whose runs cannot be repeated (in place) reuse a decision of the same
// shape, or time one candidate per run() until every candidate has run
// twice (run_trial).
struct OrderEntry {
  std::vector<long long> key;
  std::vector<int> order;
  std::vector<float> ms;
};
static std::mutex& order_mutex() {
  static std::mutex m;
  return m;
}
static std::vector<OrderEntry>& order_table() {
  static std::vector<OrderEntry> t;
  return t;
}
// shapes being timed one run at a time (plans whose runs cannot be repeated)
struct TrialEntry {
  std::vector<long long> key;
  std::vector<std::vector<int>> cand;
  std::vector<float> ms;
  std::vector<int> tries;
};
static std::vector<TrialEntry>& trial_table() {
  static std::vector<TrialEntry> t;
  return t;
}

static std::vector<long long> order_key(const std::vector<BatchPlan::Group>& groups) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::vector<long long> k{dev};
  for (const BatchPlan::Group& G : groups)
    for (long long v : {(long long)G.N, (long long)G.g.fg.n, (long long)G.g.full, (long long)G.g.ssf,
                        (long long)G.g.pf, (long long)G.g.tdma, (long long)G.nz, (long long)G.n})
      k.push_back(v);
  return k;
}

std::string batch_order_report() {
  std::lock_guard<std::mutex> lock(order_mutex());
  std::string out;
  for (const OrderEntry& e : order_table()) {
    out += ", {\"kernel\": \"batch_order\", \"groups\": [";
    for (size_t g = 0; (g + 1) * 8 < e.key.size(); ++g) {
      char b[96];
      std::snprintf(b, sizeof(b), "%s{\"N\": %lld, \"rows\": %lld}", g ? ", " : "", e.key[1 + 8 * g],
                    e.key[8 + 8 * g]);
      out += b;
    }
    out += "], \"order\": [";
    for (size_t i = 0; i < e.order.size(); ++i) out += (i ? ", " : "") + std::to_string(e.order[i]);
    out += "], \"ms\": [";
    for (size_t i = 0; i < e.ms.size(); ++i) {
      char b[32];
      std::snprintf(b, sizeof(b), "%s%.3f", i ? ", " : "", e.ms[i]);
      out += b;
    }
    out += "]}";
  }
  return out;
}

// Candidate launch orders, each timed twice (HIP events on s around the
// fork and the joins, the work before them finished first): the current
// order and each other group moved to the front.  Groups whose workgroups
// fill a compute unit's LDS by themselves (rows of 2048 lanes) take whole
// units, so which group starts first decides how the rest share the chip:
// in the 64-shot survey of apertures 256-2048 the 1024-lane group ahead of
// the 2048-lane one runs the batch in 13.4-13.7 ms against 14.6.  A run
// inside stream capture keeps the current order and times nothing.
// WEMFWI_BATCH_ORDER=0 keeps the largest-length-first order.
int BatchPlan::tune_order(int up, int adjoint, hipStream_t s) {
  const char* env = std::getenv("WEMFWI_BATCH_ORDER");
  if (groups.size() < 2 || (env && *env && std::atoi(env) == 0)) {
    order_done = true;
    return WEMFWI_OK;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return WEMFWI_OK;
  const std::vector<long long> key = order_key(groups);
  shape = key;
  {
    std::lock_guard<std::mutex> lock(order_mutex());
    for (const OrderEntry& e : order_table())
      if (e.key == key) {
        launch_order = e.order;
        order_ms = e.ms;
        order_done = true;
        return WEMFWI_OK;
      }
  }
  if (!rerunnable) {
    // timed across runs instead (run_trial); enqueue keeps the current order
    order_done = true;
    online = true;
    return WEMFWI_OK;
  }
  const std::vector<std::vector<int>> cand = candidates();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  std::vector<float> ms(cand.size(), 0.f);
  for (size_t c = 0; c < cand.size() && e == hipSuccess; ++c)
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      e = hipEventRecord(e0, s);
      if (e != hipSuccess) break;
      if (int rc = launch(up, adjoint, s, cand[c])) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
      }
      e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e == hipSuccess && (rep == 0 || t < ms[c])) ms[c] = t;
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(e, "batch order timing");
  size_t best = 0;
  for (size_t c = 1; c < cand.size(); ++c)
    if (ms[c] < ms[best]) best = c;
  launch_order = cand[best];
  order_ms = ms;
  order_done = true;
  std::lock_guard<std::mutex> lock(order_mutex());
  order_table().push_back(OrderEntry{key, launch_order, ms});
  return WEMFWI_OK;
}

// the current order and each other group moved to the front (at most 4)
std::vector<std::vector<int>> BatchPlan::candidates() const {
  std::vector<std::vector<int>> cand;
  for (size_t f = 0; f < groups.size() && cand.size() < 4; ++f) {
    std::vector<int> o{(int)f};
    for (size_t k = 0; k < groups.size(); ++k)
      if (k != f) o.push_back((int)k);
    cand.push_back(o);
  }
  return cand;
}

// One run of an online plan: the candidate order of its shape tried least so
// far (results do not depend on the order), timed with HIP events on s.
int BatchPlan::run_trial(int up, int adjoint, hipStream_t s) {
  std::vector<int> ord;
  size_t ci = 0;
  {
    std::lock_guard<std::mutex> lock(order_mutex());
    for (const OrderEntry& e : order_table())
      if (e.key == shape) {
        ord = e.order;  // decided meanwhile (another plan of the shape)
        break;
      }
    if (ord.empty()) {
      auto& tt = trial_table();
      auto it = std::find_if(tt.begin(), tt.end(), [&](const TrialEntry& e) { return e.key == shape; });
      if (it == tt.end()) {
        const auto cand = candidates();
        tt.push_back(TrialEntry{shape, cand, std::vector<float>(cand.size(), 0.f), std::vector<int>(cand.size(), 0)});
        it = tt.end() - 1;
      }
      for (size_t c = 1; c < it->cand.size(); ++c)
        if (it->tries[c] < it->tries[ci]) ci = c;
      ord = it->cand[ci];
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  int rc = WEMFWI_OK;
  if (e == hipSuccess) rc = launch(up, adjoint, s, ord);
  if (e == hipSuccess && rc == WEMFWI_OK) e = hipEventRecord(e1, s);
  if (e == hipSuccess && rc == WEMFWI_OK) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess && rc == WEMFWI_OK) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "batch run");
  std::lock_guard<std::mutex> lock(order_mutex());
  auto& tt = trial_table();
  auto it = std::find_if(tt.begin(), tt.end(), [&](const TrialEntry& x) { return x.key == shape; });
  if (it == tt.end() || it->cand[ci] != ord) return WEMFWI_OK;
  it->ms[ci] = it->tries[ci] ? std::min(it->ms[ci], t) : t;
  it->tries[ci] += 1;
  if (*std::min_element(it->tries.begin(), it->tries.end()) >= 2) {
    size_t best = 0;
    for (size_t c = 1; c < it->cand.size(); ++c)
      if (it->ms[c] < it->ms[best]) best = c;
    order_table().push_back(OrderEntry{shape, it->cand[best], it->ms});
    tt.erase(it);
  }
  return WEMFWI_OK;
}

int BatchPlan::launch(int up, int adjoint, hipStream_t s, const std::vector<int>& ord) {
  if (ord.size() != groups.size()) return fail(WEMFWI_ERR_ARG, "batch: launch order does not cover the groups");
  hipStream_t st[4];
  if (int rc = batch_streams(st)) return rc;
  while (events.size() < groups.size() + 1) {
    hipEvent_t ev;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "event");
    events.push_back(ev);
  }
  hipError_t e = hipEventRecord(events[0], s);
  const int step = ((!up) != (adjoint != 0)) ? 1 : -1;
  for (size_t k = 0; k < ord.size() && e == hipSuccess; ++k) {
    const Group& G = groups[ord[k]];
    hipStream_t sg = st[k % 4];
    e = hipStreamWaitEvent(sg, events[0], 0);
    if (e == hipSuccess) e = launch_oneway(G.N, d_rows + G.off, G.n, d_tabs, G.g, G.nz, step, adjoint != 0, G.tw, sg);
    if (e == hipSuccess) e = hipEventRecord(events[k + 1], sg);
  }
  // the joins go in after every launch: with more streams than hardware
  // queues a side stream can share the caller's queue, and a launch queued
  // behind a join would wait for the groups before it
  for (size_t k = 0; k < groups.size() && e == hipSuccess; ++k) e = hipStreamWaitEvent(s, events[k + 1], 0);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "batch run");
}

int BatchPlan::run(int up, int adjoint, hipStream_t s) {
  if (int rc = check()) return rc;
  if (!order_done)
    if (int rc = tune_order(up, adjoint, s)) return rc;
  if (online) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
      if (int rc = run_trial(up, adjoint, s)) return rc;
      hipError_t e = hipStreamSynchronize(s);
      return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "batch run");
    }
  }
  if (int rc = enqueue(up, adjoint, s, false)) return rc;
  hipError_t e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "batch run");
}

int Model::onestep(int adjoint, int nrows, const float* freq_hz, const int* depth, const cf* in, cf* out,
                   int add, hipStream_t s) {
  if (nrows < 0 || !freq_hz || !depth || !in || !out) return fail(WEMFWI_ERR_ARG, "onestep: null argument");
  if (nrows == 0) return WEMFWI_OK;
  for (int i = 0; i < nrows; ++i)
    if (depth[i] < 0 || depth[i] >= d.nz) return fail(WEMFWI_ERR_ARG, "onestep: depth out of range");
  std::vector<float> wrow;
  std::vector<const float2*> prow;
  int rc = phase_tables(nrows, freq_hz, wrow, prow, s);
  if (rc) return rc;
  h_step.resize(nrows);
  std::vector<StepRow>& h = h_step;
  for (int i = 0; i < nrows; ++i) {
    h[i].in = (const float2*)(in + (size_t)i * d.nx);
    h[i].out = (float2*)(out + (size_t)i * d.nx);
    h[i].ptab = prow[i];
    h[i].w = wrow[i];
    h[i].d = depth[i];
    h[i].model = 0;
    h[i].add = add ? 1 : 0;
  }
  if ((rc = scratch.reserve((size_t)nrows * N * sizeof(float2), "scratch rows"))) return rc;
  for (int i = 0; i < nrows; ++i) h[i].scratch = (float2*)scratch.ptr + (size_t)i * N;
  if ((rc = rows.reserve(h.size() * sizeof(StepRow), "row descriptors"))) return rc;
  hipError_t e = hipMemcpyAsync(rows.ptr, h.data(), h.size() * sizeof(StepRow), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "row upload");
  e = launch_onestep(N, (const StepRow*)rows.ptr, nrows, d_tables, geom(), adjoint != 0, d_tw, s);
  if (e != hipSuccess) return hip_fail(e, "onestep kernel");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "onestep");
  return WEMFWI_OK;
}

}  // namespace wemfwi

// ------------------------------------------------------------------ C ABI

using namespace wemfwi;


struct wemfwi_batch {
  wemfwi::BatchPlan p;
};

extern "C" {

const char* wemfwi_last_error(void) { return g_err.c_str(); }

int wemfwi_version(void) { return 100; }

int wemfwi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

long long wemfwi_cache_bytes(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  return (long long)c.held[dev];
}

int wemfwi_cache_release(void) {
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return WEMFWI_OK;
  for (auto& kv : c.free_blocks) {
    if (kv.second.empty()) continue;
    (void)hipSetDevice(kv.first);
    flush_device(c, kv.first);
  }
  (void)hipSetDevice(cur);
  return WEMFWI_OK;
}

int wemfwi_refsampler(int nz, int nx, const wemfwi_complex* slow, int nref, wemfwi_complex* sref,
                      int* refid) {
  if (!slow || !sref || !refid) return fail(WEMFWI_ERR_ARG, "refsampler: null argument");
  return refsampler_host(nz, nx, (const cf*)slow, nref, (cf*)sref, refid);
}

int wemfwi_model_create(const wemfwi_model_desc* desc, const wemfwi_complex* slow, const wemfwi_complex* den,
                        wemfwi_model** out) {
  if (!desc || !slow || !out) return fail(WEMFWI_ERR_ARG, "model_create: null argument");
  *out = nullptr;
  wemfwi_model* h = new (std::nothrow) wemfwi_model();
  if (!h) return fail(WEMFWI_ERR_ALLOC, "model_create: out of host memory");
  int rc = h->m.build(*desc, (const cf*)slow, (const cf*)den);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return WEMFWI_OK;
}

int wemfwi_model_create_refl(const wemfwi_model_desc* desc, const wemfwi_complex* slow, const wemfwi_complex* den,
                             const wemfwi_complex* slow_refl, wemfwi_model** out) {
  if (!desc || !slow || !slow_refl || !out) return fail(WEMFWI_ERR_ARG, "model_create_refl: null argument");
  *out = nullptr;
  wemfwi_model* h = new (std::nothrow) wemfwi_model();
  if (!h) return fail(WEMFWI_ERR_ALLOC, "model_create_refl: out of host memory");
  int rc = h->m.build(*desc, (const cf*)slow, (const cf*)den, (const cf*)slow_refl);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return WEMFWI_OK;
}

int wemfwi_model_destroy(wemfwi_model* model) {
  delete model;
  return WEMFWI_OK;
}

int wemfwi_oneway(wemfwi_model* model, int up, int adjoint, int nrows, const float* freq_hz,
                  const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (!model) return fail(WEMFWI_ERR_ARG, "oneway: null model");
  return model->m.oneway(up, adjoint, nrows, freq_hz, (const cf*)in, (cf*)out, add, in == out,
                         (hipStream_t)stream);
}

int wemfwi_oneway_recursive(wemfwi_model* model, int up, int adjoint, int nrows, const float* freq_hz,
                            const wemfwi_complex* src, wemfwi_complex* out, void* stream) {
  if (!model) return fail(WEMFWI_ERR_ARG, "oneway_recursive: null model");
  return model->m.oneway(up, adjoint, nrows, freq_hz, (const cf*)src, (cf*)out, 1, 1, (hipStream_t)stream);
}

int wemfwi_batch_create(int nrows, wemfwi_model* const* models, const float* freq_hz,
                        const wemfwi_complex* const* src, wemfwi_complex* const* out, void* stream,
                        wemfwi_batch** plan) {
  if (!plan || nrows < 1 || !models) return fail(WEMFWI_ERR_ARG, "batch_create: invalid argument");
  *plan = nullptr;
  std::vector<Model*> ms(nrows);
  for (int i = 0; i < nrows; ++i) {
    if (!models[i]) return fail(WEMFWI_ERR_ARG, "batch_create: null model");
    ms[i] = &models[i]->m;
  }
  wemfwi_batch* h = new (std::nothrow) wemfwi_batch();
  if (!h) return fail(WEMFWI_ERR_ALLOC, "batch_create: out of host memory");
  int rc = h->p.build(nrows, ms.data(), freq_hz, (const cf* const*)src, (cf* const*)out, (hipStream_t)stream);
  if (rc) {
    delete h;
    return rc;
  }
  *plan = h;
  return WEMFWI_OK;
}

int wemfwi_batch_run(wemfwi_batch* plan, int up, int adjoint, void* stream) {
  if (!plan) return fail(WEMFWI_ERR_ARG, "batch_run: null plan");
  return plan->p.run(up, adjoint, (hipStream_t)stream);
}

int wemfwi_batch_destroy(wemfwi_batch* plan) {
  delete plan;
  return WEMFWI_OK;
}

struct wemfwi_graph {
  std::vector<wemfwi_batch*> plans;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  ~wemfwi_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

int wemfwi_batch_graph_create(const wemfwi_batch_op* ops, int nops, void* stream, wemfwi_graph** out) {
  if (!ops || nops < 1 || !out) return fail(WEMFWI_ERR_ARG, "batch_graph_create: invalid argument");
  *out = nullptr;
  for (int i = 0; i < nops; ++i)
    if (!ops[i].plan) return fail(WEMFWI_ERR_ARG, "batch_graph_create: null plan");
  hipStream_t s = (hipStream_t)stream;
  // one eager pass: the sweep autotuner decides before capture
  for (int i = 0; i < nops; ++i) {
    if (int rc = ops[i].plan->p.enqueue(ops[i].up, ops[i].adjoint, s, ops[i].tables != 0)) return rc;
  }
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "batch graph warm-up");
  wemfwi_graph* g = new (std::nothrow) wemfwi_graph();
  if (!g) return fail(WEMFWI_ERR_ALLOC, "batch_graph_create: out of host memory");
  e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal);
  int rc = e == hipSuccess ? WEMFWI_OK : hip_fail(e, "begin capture");
  for (int i = 0; i < nops && rc == WEMFWI_OK; ++i) {
    rc = ops[i].plan->p.enqueue(ops[i].up, ops[i].adjoint, g->cap, ops[i].tables != 0);
    g->plans.push_back(ops[i].plan);
  }
  if (e == hipSuccess) {
    hipError_t e2 = hipStreamEndCapture(g->cap, &g->graph);
    if (rc == WEMFWI_OK && e2 != hipSuccess) rc = hip_fail(e2, "end capture");
  }
  if (rc == WEMFWI_OK) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "graph instantiate");
  }
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return WEMFWI_OK;
}

int wemfwi_graph_launch(wemfwi_graph* g, void* stream) {
  if (!g || !g->exec) return fail(WEMFWI_ERR_ARG, "graph_launch: null graph");
  for (wemfwi_batch* p : g->plans)
    if (int rc = p->p.check()) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipGraphLaunch(g->exec, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "graph launch");
}

int wemfwi_graph_destroy(wemfwi_graph* g) {
  delete g;
  return WEMFWI_OK;
}

int wemfwi_oneway_batch(int up, int adjoint, int nrows, wemfwi_model* const* models, const float* freq_hz,
                        const wemfwi_complex* const* src, wemfwi_complex* const* out, void* stream) {
  if (nrows == 0) return WEMFWI_OK;
  wemfwi_batch* plan = nullptr;
  int rc = wemfwi_batch_create(nrows, models, freq_hz, src, out, stream, &plan);
  if (rc) return rc;
  rc = wemfwi_batch_run(plan, up, adjoint, stream);
  wemfwi_batch_destroy(plan);
  return rc;
}

int wemfwi_model_set_freqs(wemfwi_model* model, int nfreq, const float* freq_hz, void* stream) {
  if (!model || !freq_hz || nfreq <= 0) return fail(WEMFWI_ERR_ARG, "set_freqs: invalid argument");
  int rc = model->m.set_freqs(nfreq, freq_hz, (hipStream_t)stream);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? WEMFWI_OK : hip_fail(e, "set_freqs");
}

int wemfwi_timing(int enable) {
  g_timing = enable != 0;
  if (enable) timing_table().clear();
  return WEMFWI_OK;
}

int wemfwi_autotune_report(char* buf, int cap) {
  std::string r = autotune_report();
  const std::string b = batch_order_report();
  if (!b.empty()) r.insert(r.size() - 1, r.size() > 2 ? b : b.substr(2));
  if (buf && cap > 0) {
    const size_t n = std::min((size_t)cap - 1, r.size());
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  return (int)r.size();
}

int wemfwi_timing_get(const char* kernel, double* total_ms, long* count) {
  if (!kernel || !total_ms || !count) return fail(WEMFWI_ERR_ARG, "timing_get: null argument");
  auto it = timing_table().find(kernel);
  *total_ms = it == timing_table().end() ? 0.0 : it->second.ms;
  *count = it == timing_table().end() ? 0 : it->second.count;
  return WEMFWI_OK;
}

int wemfwi_onestep(wemfwi_model* model, int adjoint, int nrows, const float* freq_hz, const int* depth,
                   const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (!model) return fail(WEMFWI_ERR_ARG, "onestep: null model");
  return model->m.onestep(adjoint, nrows, freq_hz, depth, (const cf*)in, (cf*)out, add, (hipStream_t)stream);
}

}  // extern "C"
