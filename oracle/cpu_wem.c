/*
 * CPU restatement (FP32, OpenMP) of the reference's one-way depth stepping --
 * TEST INFRASTRUCTURE ONLY: the CPU baseline timed beside the GPU path and a
 * second checker of the FP64 oracle.  Never linked into the product.
 *
 * Follows the reference source in single precision, as the reference runs:
 *   phase shift      propagator/src/phase_shift.ispc:2-32 (formula as written)
 *   k grid           propagator/src/Phshift_cmplx.cpp:12-24
 *   unitary FFT      operator/src/FFT1_c.cpp:73-103 (origin shift omitted)
 *   SSF / PSPI step  propagator/src/SSF.cpp:8-60, PSPI.cpp:8-53,
 *                    SplitStep.cpp:5-29, Selector.cpp:5-19, Taper.h:11-23
 *   Down / Up        propagator/src/OneWay.cpp:32-109 (in-place recursion)
 *   row parallelism  one (shot, frequency) row per task, static schedule, as
 *                    tbb::blocked_range2d + static_partitioner in
 *                    propagator/src/wem_prop.cpp:11-12,76
 * The FFT is a radix-4 Stockham transform (power-of-two lengths), FFTW
 * being absent on the hosts this runs on.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef float complex cf;

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

/*
 * Unnormalised DFT of x (sign -1 forward, +1 backward, FFTW's convention),
 * a Stockham autosort transform: a radix-2 pass first when log2 n is odd,
 * then radix-4 passes, between x and the scratch row tmp (n entries); the
 * inner loops run over contiguous butterflies.  Pass p of sub-length NS
 * multiplies input r of butterfly j (m = j mod NS) by tw[off + (r-1) NS + m]
 * = exp(-2 pi i m r / (4 NS)) (conjugated backward), then writes output s at
 * (j / NS) 4 NS + m + s NS.
 */
static void fft_s4(cf* x, cf* tmp, int n, int sign, const cf* tw) {
  cf *a = x, *b = tmp;
  int ns = 1;
  if (ilog2(n) & 1) {
    const int h = n / 2;
    for (int j = 0; j < h; ++j) {
      b[2 * j] = a[j] + a[j + h];
      b[2 * j + 1] = a[j] - a[j + h];
    }
    cf* t = a;
    a = b;
    b = t;
    ns = 2;
  }
  size_t off = 0;
  const int q4 = n / 4;
  while (ns < n) {
    const cf* t3 = tw + off;
    for (int q = 0; q < n / (4 * ns); ++q) {
      const cf* in = a + (size_t)q * ns;
      cf* out = b + (size_t)q * 4 * ns;
      for (int m = 0; m < ns; ++m) {
        cf a0 = in[m], a1 = in[m + q4], a2 = in[m + 2 * q4], a3 = in[m + 3 * q4];
        if (ns > 1) {
          cf w1 = t3[m], w2 = t3[ns + m], w3 = t3[2 * ns + m];
          if (sign > 0) {
            w1 = conjf(w1);
            w2 = conjf(w2);
            w3 = conjf(w3);
          }
          a1 *= w1;
          a2 *= w2;
          a3 *= w3;
        }
        const cf s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3;
        const cf d13 = (a1 - a3) * (sign < 0 ? -I : I);
        out[m] = s02 + s13;
        out[m + ns] = d02 + d13;
        out[m + 2 * ns] = s02 - s13;
        out[m + 3 * ns] = d02 - d13;
      }
    }
    off += (size_t)3 * ns;
    ns *= 4;
    cf* t = a;
    a = b;
    b = t;
  }
  if (a != x) memcpy(x, a, sizeof(cf) * n);
}

/* fft_s4's twiddle table for length n (fewer than n entries), each
 * evaluated in double and rounded once */
static cf* s4_twiddles(int n) {
  cf* tw = (cf*)malloc(sizeof(cf) * (size_t)(n > 4 ? n : 4));
  size_t off = 0;
  for (int ns = (ilog2(n) & 1) ? 2 : 1; ns < n; ns *= 4) {
    for (int r = 1; r < 4; ++r)
      for (int m = 0; m < ns; ++m) {
        const double a = -2.0 * M_PI * (double)m * r / (4.0 * ns);
        tw[off + (size_t)(r - 1) * ns + m] = (float)cos(a) + I * (float)sin(a);
      }
    off += (size_t)3 * ns;
  }
  return tw;
}

typedef struct {
  int nz, nx, tap, taper, nref, pspi, nk;
  float dx, dz, eps;
  const cf* slow;   /* [nz][nx] */
  const cf* sref;   /* [nz][nref] */
  const int* refid; /* [nz][nx] */
  float* k;         /* [nk] */
  float* filt;      /* [nx] */
  cf* tw;           /* fft_s4 twiddles of length nk */
} Grid;

/* ps_fwd / ps_adj (phase_shift.ispc:2-32) in the reference's float formula */
static void phase(cf* out, const cf* in, const Grid* g, float w, cf s, int adj, int add) {
  const float sre = crealf(s), sim = cimagf(s);
  for (int ik = 0; ik < g->nk; ++ik) {
    const float a = w * w * sre - g->k[ik] * g->k[ik];
    const float b = w * w * (sim - g->eps * sre);
    const float c = sqrtf(a * a + b * b);
    const float re = sqrtf((c + a) / 2), im = -sqrtf((c - a) / 2);
    const float e = expf(im * g->dz), cs = cosf(re * g->dz), sn = sinf(re * g->dz);
    const cf f = adj ? e * (cs + I * sn) : e * (cs - I * sn);
    out[ik] = add ? out[ik] + in[ik] * f : in[ik] * f;
  }
}

static void ss_factor(cf* out, const Grid* g, int iz, float w, cf s_ref, const int* rid, int r, int adj) {
  const cf sq = csqrtf(s_ref);
  for (int x = 0; x < g->nx; ++x) {
    if (rid[x] != r) continue;
    const cf sv = g->slow[(size_t)iz * g->nx + x];
    cf e = -I * w * g->dz * (sv - s_ref) / sq;
    if (adj) e = conjf(e);
    out[x] *= cexpf(e);
  }
}

static void step_fwd(const Grid* g, int iz, float w, const cf* in, cf* out, cf* X, cf* Y, cf* row, cf* F) {
  const float sc = 1.0f / sqrtf((float)g->nk);
  memset(X, 0, sizeof(cf) * g->nk);
  for (int x = 0; x < g->nx; ++x) X[x + g->tap] = in[x];
  fft_s4(X, F, g->nk, -1, g->tw);
  for (int i = 0; i < g->nk; ++i) X[i] *= sc;
  memset(out, 0, sizeof(cf) * g->nx);
  const int* rid = g->refid + (size_t)iz * g->nx;
  for (int r = 0; r < g->nref; ++r) {
    int any = 0;
    for (int x = 0; x < g->nx && !any; ++x) any = rid[x] == r;
    if (!any) continue;
    const cf s = g->sref[(size_t)iz * g->nref + r];
    phase(Y, X, g, w, s, 0, 0);
    fft_s4(Y, F, g->nk, +1, g->tw);
    for (int x = 0; x < g->nx; ++x) row[x] = Y[x + g->tap] * sc * g->filt[x];
    if (!g->pspi) ss_factor(row, g, iz, w, s, rid, r, 0);
    for (int x = 0; x < g->nx; ++x)
      if (rid[x] == r) out[x] = row[x];
  }
}

static void step_adj(const Grid* g, int iz, float w, const cf* in, cf* out, cf* X, cf* Y, cf* row, cf* F) {
  const float sc = 1.0f / sqrtf((float)g->nk);
  memset(X, 0, sizeof(cf) * g->nk);
  const int* rid = g->refid + (size_t)iz * g->nx;
  for (int r = 0; r < g->nref; ++r) {
    int any = 0;
    for (int x = 0; x < g->nx && !any; ++x) any = rid[x] == r;
    if (!any) continue;
    const cf s = g->sref[(size_t)iz * g->nref + r];
    for (int x = 0; x < g->nx; ++x) row[x] = rid[x] == r ? in[x] : 0;
    if (!g->pspi) ss_factor(row, g, iz, w, s, rid, r, 1);
    memset(Y, 0, sizeof(cf) * g->nk);
    for (int x = 0; x < g->nx; ++x) Y[x + g->tap] = row[x] * g->filt[x];
    fft_s4(Y, F, g->nk, -1, g->tw);
    for (int i = 0; i < g->nk; ++i) Y[i] *= sc;
    phase(X, Y, g, w, s, 1, 1);
  }
  fft_s4(X, F, g->nk, +1, g->tw);
  for (int x = 0; x < g->nx; ++x) out[x] = X[x + g->tap] * sc;
}

/*
 * In-place Down (up=0) / Up (up=1), forward or adjoint, of nrows rows
 * W[nrows][nz][nx]; row i at frequency freq_hz[i].  Returns 0, or -1 for a
 * non power-of-two FFT length.
 */
int cpu_oneway(int nrows, int nz, int nx, int tap, int taper, int nref, int pspi, float dx,
               float dz, float eps, const float* freq_hz, const float* slow, const float* sref,
               const int* refid, float* W, int up, int adjoint, int nthreads) {
  Grid g;
  g.nz = nz;
  g.nx = nx;
  g.tap = tap;
  g.taper = taper;
  g.nref = nref;
  g.pspi = pspi;
  g.nk = nx + 2 * tap;
  g.dx = dx;
  g.dz = dz;
  g.eps = eps;
  g.slow = (const cf*)slow;
  g.sref = (const cf*)sref;
  g.refid = refid;
  if (ilog2(g.nk) < 0) return -1;
  g.k = (float*)calloc(g.nk, sizeof(float));
  g.filt = (float*)malloc(sizeof(float) * nx);
  g.tw = s4_twiddles(g.nk);
  const float dk = (float)(2 * M_PI / (dx * (g.nk - 1)));
  for (int ik = 0; ik < g.nk / 2; ++ik) g.k[ik] = ik * dk;
  for (int ik = 1; ik < g.nk / 2; ++ik) g.k[g.nk - ik] = -g.k[ik];
  for (int x = 0; x < nx; ++x) g.filt[x] = 1.f;
  for (int i = 0; i < taper; ++i) {
    g.filt[i] = 0.5f * (float)(1 - cos(M_PI * i / (taper - 1)));
    g.filt[nx - 1 - i] = g.filt[i];
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const int down = !up;
  const int fwd_dir = down != (adjoint != 0);
#pragma omp parallel
  {
    cf* X = (cf*)malloc(sizeof(cf) * g.nk);
    cf* Y = (cf*)malloc(sizeof(cf) * g.nk);
    cf* row = (cf*)malloc(sizeof(cf) * nx);
    cf* tmp = (cf*)malloc(sizeof(cf) * nx);
    cf* F = (cf*)malloc(sizeof(cf) * g.nk);
#pragma omp for schedule(static)
    for (int i = 0; i < nrows; ++i) {
      cf* w3 = (cf*)W + (size_t)i * nz * nx;
      const float w = (float)(2 * M_PI * freq_hz[i]);
      for (int s = 0; s < nz - 1; ++s) {
        const int cur = fwd_dir ? s : nz - 1 - s;
        const int nxt = fwd_dir ? cur + 1 : cur - 1;
        if (adjoint) step_adj(&g, nxt, w, w3 + (size_t)cur * nx, tmp, X, Y, row, F);
        else step_fwd(&g, cur, w, w3 + (size_t)cur * nx, tmp, X, Y, row, F);
        for (int x = 0; x < nx; ++x) w3[(size_t)nxt * nx + x] += tmp[x];
      }
    }
    free(X);
    free(Y);
    free(row);
    free(tmp);
    free(F);
  }
  free(g.k);
  free(g.filt);
  free(g.tw);
  return 0;
}
