/*
 * wemfwi -- C ABI of the MI355X one-way wave-equation (WEM) propagator.
 *
 * Every entry point returns WEMFWI_OK (0) or an error code; the message of
 * the last failure on the calling thread is wemfwi_last_error().  No C++
 * exception crosses this boundary.  Background models (slowness, density)
 * are passed as host arrays; wavefields, images and data are device
 * pointers owned by the caller.  Complex values are interleaved float pairs
 * (the layout of std::complex<float>).  Arrays are C-ordered with x fastest,
 * as the reference's complex2DReg/complex3DReg (complex/include/complex2DReg.h).
 *
 * A handle is used by one thread at a time; all work of a call is enqueued
 * on the given HIP stream (NULL: the default stream) and the call returns
 * after the work has completed.
 */
#ifndef WEMFWI_H
#define WEMFWI_H

#ifdef __cplusplus
extern "C" {
#endif

#define WEMFWI_OK 0
#define WEMFWI_ERR_ARG 1          /* invalid argument / shape */
#define WEMFWI_ERR_HIP 2          /* HIP runtime error */
#define WEMFWI_ERR_UNSUPPORTED 3  /* FFT length or option not built */
#define WEMFWI_ERR_REFERENCE 4    /* non-positive reference slowness */
#define WEMFWI_ERR_ALLOC 5        /* device allocation failed */

typedef struct wemfwi_complex {
  float re, im;
} wemfwi_complex;

/* Propagation grid and one-way parameters.  Mirrors the paramObj keys read by
 * OneStep (propagator/src/OneStep.h:19-38), OneWay (OneWay.cpp:22-28) and
 * Scatter (Scatter.cpp:6-12). */
typedef struct wemfwi_model_desc {
  int nz;        /* depth samples */
  int nx;        /* propagation width (model nx + 2*pad) */
  float dx, dz;  /* grid spacing */
  int tap;       /* "tap": FFT zero padding per side; FFT length nx + 2*tap */
  int taper;     /* "pad": taper width (Taper.h:11-23); 1 is rejected */
  int nref;      /* "nref": reference slownesses per depth */
  int pspi;      /* "prop": 0 = "ssf", 1 = "pspi" */
  float eps;     /* "eps": damping */
  int ntaylor;   /* "ntaylor": Taylor order of the scatterer */
} wemfwi_model_desc;

typedef struct wemfwi_model wemfwi_model;

/* Last error message of the calling thread ("" if none). */
const char* wemfwi_last_error(void);

/* Library version (major*10000 + minor*100 + patch). */
int wemfwi_version(void);

/* Number of visible devices (0 without a GPU). */
int wemfwi_device_count(void);

/* Device memory cache of the library's per-call buffers (no counterpart in
 * the reference, whose operators allocate host scratch per TBB task,
 * wem_prop.cpp:14-17).  Buffers freed at the end of a call are kept per
 * device for the next call of a similar size instead of being returned to
 * the driver, which clears freed VRAM before handing it out again (hundreds
 * of ms for the tens of GB of a batched call).  At most WEMFWI_CACHE_LIMIT_MB
 * (environment; default half of the device memory) is kept; a failing
 * allocation empties the cache and retries.  wemfwi_cache_bytes: bytes held
 * for the current device; wemfwi_cache_release: returns every cached block
 * of every device to the driver (for callers that allocate elsewhere, e.g.
 * torch, after a large call). */
long long wemfwi_cache_bytes(void);
int wemfwi_cache_release(void);

/* Frequency grid and wavelet spectrum of PropParam::calcPropParam
 * (propagator/src/PropParam.h:45-85), host only, float32 arithmetic as the
 * reference: the float wavelet wave[nt] (time origin ot, step dt) is
 * zero-padded by tpad samples; wave_f[nt + tpad] receives its unitary forward
 * FFT (FFT1_r.cpp:74-83); freq[i] = fmin + i/tmax and index[i] = i + imin with
 * tmax = ot + dt*(nt + tpad - 1), imin = (int)(fmin*tmax),
 * imax = (int)(fmax*tmax + 0.5) -- or, when fmax == 0, imin = 1 and
 * imax = nfreq (nfreq == 0: (nt + tpad)/2 + 1).  *nw receives imax - imin + 1;
 * freq/index must hold cap entries (WEMFWI_ERR_ARG when cap < *nw; freq, index
 * and wave_f may be NULL to query *nw only). */
int wemfwi_prop_param(int nt, float ot, float dt, const float* wave, int tpad, float fmin,
                      float fmax, int nfreq, int cap, int* nw, float* freq, int* index,
                      wemfwi_complex* wave_f);

/* Float time traces <-> complex frequency data, the float-data boundary of
 * WEM::forward (WEMnl.cpp:39-65), BornFull::forward/adjoint
 * (BornFull.cpp:39-62,80-102) and BornRefl (BornRefl.cpp:13-49): traces are
 * zero-padded by tpad samples (n = nt + tpad, PropParam.h:49) and transformed
 * with the unitary FFT1 (operator/src/FFT1_r.cpp:74-96), restricted to the nw
 * survey bins (the only bins the propagation reads or writes).
 * trace_to_freq: data_f[tr][bins[i]] = n^-1/2 sum_t data_t[tr][t] e^{-2 pi i bins[i] t/n};
 *   add == 0 zeroes data_f[ntr][nbins] first.
 * freq_to_trace: data_t[tr][t] (+)= Re(n^-1/2 sum_i data_f[tr][bins[i]] e^{+2 pi i bins[i] t/n}),
 *   t < nt (the crop of Pad::adjoint).
 * data_t, data_f are device pointers; bins is a host array, 0 <= bins[i] < min(nbins, n). */
int wemfwi_trace_to_freq(int ntr, int nt, int tpad, const float* data_t, int nw, const int* bins,
                         int nbins, wemfwi_complex* data_f, int add, void* stream);
int wemfwi_freq_to_trace(int ntr, int nt, int tpad, const wemfwi_complex* data_f, int nbins, int nw,
                         const int* bins, float* data_t, int add, void* stream);

/* Regular device geometry of Device::createRegGeom (propagator/src/Device.h:39-63):
 * device i at x = o_x + i*d_x, z = o_z + i*d_z on the model grid (origin ox, oz,
 * spacing dxv, dzv) gets ix = (int)((x - ox)/dxv + 0.5f), iz likewise, in
 * float32.  WEMFWI_ERR_ARG for n == 0, d_x == d_z == 0, a start outside the
 * model, or a negative index (the reference throws SEPException). */
int wemfwi_device_indices(int n, float o_x, float o_z, float d_x, float d_z, float ox, float oz,
                          float dxv, float dzv, int* iz, int* ix);

/* Irregular device geometry of Device::createIrregGeom (Device.h:65-85): device
 * i at coordinates (x[i], z[i]) -- the "<name>x" / "<name>z" arrays of the JSON
 * file named by the "coord" parameter -- gets ix = (int)((x[i] - ox)/dxv + 0.5f),
 * iz likewise, in float32.  WEMFWI_ERR_ARG for a negative index. */
int wemfwi_device_coords(int n, const float* x, const float* z, float ox, float oz, float dxv,
                         float dzv, int* iz, int* ix);

/* RefSampler (propagator/src/RefSampler.cpp:10-91), host only: reference
 * slownesses sref[nz][nref] and the reference selecting each x,
 * refid[nz][nx] (-1: none), including the stale reference-buffer behaviour of
 * getIndexes.  Returns WEMFWI_ERR_REFERENCE for a non-positive reference
 * (RefSampler.cpp:35-36). */
int wemfwi_refsampler(int nz, int nx, const wemfwi_complex* slow, int nref,
                      wemfwi_complex* sref, int* refid);

/* Background model of the one-way operators: replaces the construction of
 * RefSampler, SSF/PSPI, SplitStep, Taper and Phshift for one slowness grid
 * (OneWay::OneWay, OneWay.cpp:11-29).  slow[nz][nx] is the slowness squared
 * S = 1/v^2 on the propagation grid; den may be NULL (unit density). */
int wemfwi_model_create(const wemfwi_model_desc* desc, const wemfwi_complex* slow,
                        const wemfwi_complex* den, wemfwi_model** out);
int wemfwi_model_destroy(wemfwi_model* model);

/* Background model whose Reflect coefficients (Reflect.h:19-26) come from a
 * different slowness than everything else: the conformal-map path builds
 * Reflect from the mapped, padded slowness before scaling it by the map's
 * Jacobian (wem_prop.cpp:35-38, born_prop.cpp:45-47), while RefSampler, the
 * one-way steps, the scatterer and dReflect (built later, Born_refl.cpp:11)
 * see the scaled one.  slow: the scaled slowness; slow_refl: the unscaled. */
int wemfwi_model_create_refl(const wemfwi_model_desc* desc, const wemfwi_complex* slow,
                             const wemfwi_complex* den, const wemfwi_complex* slow_refl,
                             wemfwi_model** out);

/* Frequency set of the model: evaluates the phase-shift multipliers
 * (phase_shift.ispc:8-14) of every active reference at every depth for the
 * given frequencies, in double precision -- the per-frequency part of
 * Phshift::setFreq (Phshift.h:24-25) hoisted out of the depth loop.
 * wemfwi_oneway / wemfwi_onestep rebuild it themselves when a row's
 * frequency is not in the current set. */
int wemfwi_model_set_freqs(wemfwi_model* model, int nfreq, const float* freq_hz, void* stream);

/* Down / Up (OneWay.cpp:32-109) for nrows independent (shot, frequency) rows
 * of [nz][nx] each, row i at angular frequency 2*pi*freq_hz[i].
 * up = 0: Down, 1: Up.  adjoint = 0: forward, 1: adjoint.
 * in == out is the in-place recursion of every reference call site
 * (W[iz+1] = P W[iz] + W[iz+1] for Down forward); with add = 0 it returns
 * zeros, as the reference zeroes the shared vector first.  in != out applies
 * one bidiagonal step per depth (the reference's out-of-place semantics). */
int wemfwi_oneway(wemfwi_model* model, int up, int adjoint, int nrows, const float* freq_hz,
                  const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);

/* wemfwi_oneway without device memory allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as wemfwi_oneway, with these conditions: the
 * model's phase tables must already hold every freq_hz[i]
 * (wemfwi_model_set_freqs, or an earlier wemfwi_oneway call), else
 * WEMFWI_ERR_ARG; the sweep runs the FFT plan the autotuner already chose
 * for this variant (plan 0 if it has not tuned it: no tuning launch); nrows
 * must not exceed wemfwi_oneway_noalloc_max_rows(model) (the 64 MiB device
 * scratch array holds N complex scratch values and one row descriptor per
 * row), else WEMFWI_ERR_ARG.  Every call on a GPU shares that GPU's scratch
 * array: calls that overlap on one GPU (other host threads, or streams not
 * ordered with each other) race. */
int wemfwi_oneway_noalloc(wemfwi_model* model, int up, int adjoint, int nrows, const float* freq_hz,
                          const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);
/* Largest nrows wemfwi_oneway_noalloc accepts for this model (0 for a null
 * model). */
long long wemfwi_oneway_noalloc_max_rows(const wemfwi_model* model);

/* Recursive Down / Up with separate source and output buffers:
 * out = (I - L)^-1 src, i.e. the in-place recursion of wemfwi_oneway applied
 * to a copy of src, without the copy.  No single reference call corresponds;
 * it is the in-place call of OneWay.cpp:32-109 on a vector that was set to
 * the sources. */
int wemfwi_oneway_recursive(wemfwi_model* model, int up, int adjoint, int nrows,
                            const float* freq_hz, const wemfwi_complex* src, wemfwi_complex* out,
                            void* stream);

/* Uneven batch of recursive Down / Up sweeps: row i has its own model
 * models[i] (aperture nx, depth count, FFT length may all differ) and
 * frequency freq_hz[i]; out[i] = (I - L_i)^-1 src[i] over its [nz_i][nx_i]
 * buffers (device pointers; src[i] == out[i] is the in-place recursion).
 * Rows are grouped by kernel variant, one launch per group, the groups on
 * separate streams; every row gets exactly the result of a wemfwi_oneway_recursive
 * call on its own model.  No reference counterpart: the reference runs one
 * (shot, frequency) task at a time (wem_prop.cpp:11-12). */
int wemfwi_oneway_batch(int up, int adjoint, int nrows, wemfwi_model* const* models,
                        const float* freq_hz, const wemfwi_complex* const* src,
                        wemfwi_complex* const* out, void* stream);

/* The same as a plan: rows, models, frequencies and buffers are fixed at
 * creation (phase tables built, descriptors and scratch rows uploaded); each
 * run launches the groups only.  The models and buffers must outlive the
 * plan; a run fails with WEMFWI_ERR_ARG if a model's phase table was rebuilt
 * since planning (e.g. by a call with other frequencies). */
typedef struct wemfwi_batch wemfwi_batch;
int wemfwi_batch_create(int nrows, wemfwi_model* const* models, const float* freq_hz,
                        const wemfwi_complex* const* src, wemfwi_complex* const* out, void* stream,
                        wemfwi_batch** plan);
int wemfwi_batch_run(wemfwi_batch* plan, int up, int adjoint, void* stream);
int wemfwi_batch_destroy(wemfwi_batch* plan);

/* A sequence of batch sweeps captured once into a hipGraph and replayed
 * (no counterpart in the reference, which runs one TBB task per (shot,
 * frequency), wem_prop.cpp:11-12).  Op i runs plan i's rows as Down/Up
 * (up), forward/adjoint; tables = 1 first recomputes the phase tables of
 * the plan's models in place (Phshift::setFreq for the plan's frequencies,
 * Phshift_cmplx.cpp:47-113).  Creation runs the sequence once eagerly (the
 * sweep autotuner decides then), captures it on a private stream and
 * instantiates it; the graph holds kernel launches only.  A launch replays
 * it on `stream` and waits; it fails with WEMFWI_ERR_ARG if a plan's model
 * had its phase table rebuilt since planning.  Plans must outlive the
 * graph. */
typedef struct wemfwi_batch_op {
  wemfwi_batch* plan;
  int up, adjoint, tables;
} wemfwi_batch_op;
typedef struct wemfwi_graph wemfwi_graph;
int wemfwi_batch_graph_create(const wemfwi_batch_op* ops, int nops, void* stream, wemfwi_graph** graph);
int wemfwi_graph_launch(wemfwi_graph* graph, void* stream);
int wemfwi_graph_destroy(wemfwi_graph* graph);

/* SSF / PSPI single depth step (SSF.cpp:8-60, PSPI.cpp:8-53) on nrows
 * independent rows of nx samples: row i at depth depth[i] and frequency
 * freq_hz[i].  add = 1 accumulates into out. */
int wemfwi_onestep(wemfwi_model* model, int adjoint, int nrows, const float* freq_hz,
                   const int* depth, const wemfwi_complex* in, wemfwi_complex* out, int add,
                   void* stream);

/* Acquisition and frequency grid of the shot x frequency operators: what
 * PropParam (propagator/src/PropParam.h:45-85), Device (Device.h:39-63) and
 * the Injection objects (Injection.h:19-65) derive from the wavelet, the
 * model hypercube and the parameter object.  Every (shot, frequency) pair is
 * one row of the batched operators. */
typedef struct wemfwi_survey_desc {
  int ns;                         /* shots */
  const int* src_iz;              /* [ns] source depth index */
  const int* src_ix;              /* [ns] source x index (model grid) */
  int nr;                         /* receivers */
  const int* rec_iz;              /* [nr] */
  const int* rec_ix;              /* [nr] */
  int nw;                         /* frequencies */
  const float* freq_hz;           /* [nw] freq[i] (PropParam.h:76) */
  const int* bin;                 /* [nw] data bin index[i] (PropParam.h:77) */
  const wemfwi_complex* wavelet;  /* [nw] wavelet spectrum at bin[i] */
  int nt;                         /* bins per trace of the data [ns][nr][nt] */
  int ngs, ngr;                   /* cos^2 spread half-widths (par "ngs", "ngr" / "ng") */
  int offset;                     /* device x offset on the propagation grid (par "pad") */
} wemfwi_survey_desc;

typedef struct wemfwi_survey wemfwi_survey;

int wemfwi_survey_create(const wemfwi_survey_desc* desc, wemfwi_survey** out);
int wemfwi_survey_destroy(wemfwi_survey* survey);

/* Wavefield checkpointing of the Born adjoints called with this survey
 * (born_adjoint, extborn_adjoint, shotborn_adjoint; no counterpart in the
 * reference, which keeps whole fields, born_prop.cpp:178-232): every >= 1
 * keeps the background and receiver rows every `every` depth steps and
 * recomputes the fields segment by segment, the per-row device memory
 * dropping from 4 nz nx to (4 (every + 2) + 2 ceil((nz-1)/every) + 2) nx
 * complex samples at the cost of one more pass of the two forward
 * recursions.  The gradient is bitwise the same as with whole fields.
 * 0 (the default), or every >= nz - 1, keeps whole fields. */
int wemfwi_survey_set_checkpoint(wemfwi_survey* survey, int every);

/* WEM::forward into complex data (propagator/src/WEMnl.cpp:67-82,
 * wem_prop.cpp:9-78).  model: background slowness and density on the
 * propagation grid (edge padded by "pad", wem_prop.cpp:26-33).  data
 * [ns][nr][nt] device; add = 0 zeroes data first. */
int wemfwi_wem_forward(wemfwi_model* model, wemfwi_survey* survey, wemfwi_complex* data, int add,
                       void* stream);

/* BornFull::forward / adjoint into / from complex data (BornFull.cpp:64-117,
 * born_prop.cpp:11-235).  dslow, dden: perturbations [nz][nx] on the
 * propagation grid (zero padded, born_prop.cpp:52-58); the adjoint sums
 * every (shot, frequency) contribution in a fixed order in double precision.
 * mode 0 applies Scatter::adjoint in the reference's operator order
 * (Scatter.cpp:52-66); mode 1 applies the exact adjoint of Scatter::forward. */
int wemfwi_born_forward(wemfwi_model* model, wemfwi_survey* survey, const wemfwi_complex* dslow,
                        const wemfwi_complex* dden, wemfwi_complex* data, int add, void* stream);
int wemfwi_born_adjoint(wemfwi_model* model, wemfwi_survey* survey, wemfwi_complex* dslow,
                        wemfwi_complex* dden, const wemfwi_complex* data, int add, int mode,
                        void* stream);

/* The L2 data-misfit gradient of one model as a replayable graph
 * (scripts/fwi.py:70-131 with WEM.py:52-81: J = 1/2 ||WEM(m) - d||^2 and its
 * gradient Born(m)^H (WEM(m) - d), which the reference computes with one
 * WEM::forward and one BornFull::adjoint call).  create runs the whole
 * sequence once (res = WEM(m) - obs, the FP64 misfit, gS / gR =
 * wemfwi_born_adjoint(res) with the given mode, honouring the survey's
 * checkpoint setting) and captures it as a hipGraph; launch replays it and
 * returns J.  obs, res [ns][nr][nt] and gS, gR [nz][nx] are device buffers
 * bound at creation; the plan holds every device buffer of the call until
 * destroyed; a new model needs a new plan.  Results are bitwise those of the
 * eager calls. */
typedef struct wemfwi_gradient wemfwi_gradient;
int wemfwi_gradient_create(wemfwi_model* model, wemfwi_survey* survey, const wemfwi_complex* obs,
                           wemfwi_complex* res, wemfwi_complex* gS, wemfwi_complex* gR, int mode, void* stream,
                           wemfwi_gradient** plan);
int wemfwi_gradient_launch(wemfwi_gradient* plan, void* stream, double* misfit);
/* J of the last run (creation or launch) */
int wemfwi_gradient_misfit(wemfwi_gradient* plan, double* misfit);
int wemfwi_gradient_destroy(wemfwi_gradient* plan);

/* extWEM / extBornFull (ext_wem_prop.cpp:8-59, ext_born_prop.cpp:8-194): one
 * background model per frequency (models[nw]); perturbations
 * [nw][nz][nx].  extborn_adjoint mode 0 reproduces the reference's serial
 * shot loop, in which LinDown::adjoint clears the last depth row of the
 * shared slice (LinOneWay.cpp:64-65): that row keeps the last shot's LinUp
 * term only.  mode 1 is the exact adjoint. */
int wemfwi_extwem_forward(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* data,
                          int add, void* stream);
int wemfwi_extborn_forward(wemfwi_model* const* models, wemfwi_survey* survey,
                           const wemfwi_complex* dslow, const wemfwi_complex* dden,
                           wemfwi_complex* data, int add, void* stream);
int wemfwi_extborn_adjoint(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* dslow,
                           wemfwi_complex* dden, const wemfwi_complex* data, int add, int mode,
                           void* stream);

/* ---------------------------------------------------------------- conformal maps
 * The per-shot grids of PropParam::prepareConformalMaps (PropParam.h:89-131,
 * conformal/src) and the shot x frequency operators on them
 * (wem_prop.cpp:20-45, born_prop.cpp:28-70,136-174,219-231). */

/* Map tables of every shot, host only.  kind 0: identity; 1: ShiftAndScale
 * (source to the origin, farthest corner to |z| = 1; the scale carries over
 * from shot to shot as PropParam.h:91 declares it outside the loop) followed
 * by ComplexLog(eps_log); 2: Rotation(angles[s]) (the chain built there is
 * unused, PropParam.h:123).  The model grid has nz x nx samples, origin
 * (oz, ox), spacing (dz, dx); sources src_*[ns], receivers rec_*[nr].
 * Outputs per shot s (n = nz*nx):
 *   axes[4s..4s+3]  = o_x, d_x, o_z, d_z of the propagation grid;
 *   fwd[s][n]       model -> grid gather (-1: zero) of cmap->forward(model, grid),
 *                   including ComplexLog's uncleared row buffer;
 *   inv[s][n]       grid -> model gather (-1: zero) of cmap->inverse (Born adjoint);
 *   jac[s][2][n]    the two Jacobian factors of scaleByJacobian, in order;
 *   nsrc[s], siz/six[s][nx]: mapped source devices (cmap 1: the first row of
 *                   the log grid, ComplexLog.cpp:167-190);
 *   riz/rix[s][nr]: mapped receivers.
 * A mapped device outside the grid is WEMFWI_ERR_ARG (the reference would
 * index outside its arrays). */
int wemfwi_cmap_plan(int kind, int nz, int nx, float oz, float ox, float dz, float dx, int ns,
                     const int* src_iz, const int* src_ix, int nr, const int* rec_iz, const int* rec_ix,
                     float eps_log, const float* angles, float* axes, int* fwd, int* inv, float* jac,
                     int* nsrc, int* siz, int* six, int* riz, int* rix);

/* Per-shot device lists: shot `shot` injects its source at the nsrc points
 * (src_iz, src_ix) and records its nr traces at (rec_iz, rec_ix); shots not
 * set keep the survey's point source and shared receivers. */
int wemfwi_survey_set_shot_devices(wemfwi_survey* survey, int shot, int nsrc, const int* src_iz,
                                   const int* src_ix, const int* rec_iz, const int* rec_ix);

/* WEM / BornFull with one background model per shot (models[ns], all of one
 * size; grid spacing may differ): WEM::full_fwd and BornFull::full_fwd /
 * full_adj on the mapped grids.  Perturbations are per shot,
 * dslow/dden[ns][nz][nx] on the padded grids (the mapped model perturbation
 * zero-padded, born_prop.cpp:50-58); the adjoint writes each shot's slice
 * (before the inverse map of born_prop.cpp:217-231), summed over
 * frequencies in a fixed order in double precision; mode as born_adjoint. */
int wemfwi_shotwem_forward(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* data,
                           int add, void* stream);
int wemfwi_shotborn_forward(wemfwi_model* const* models, wemfwi_survey* survey,
                            const wemfwi_complex* dslow, const wemfwi_complex* dden,
                            wemfwi_complex* data, int add, void* stream);
int wemfwi_shotborn_adjoint(wemfwi_model* const* models, wemfwi_survey* survey, wemfwi_complex* dslow,
                            wemfwi_complex* dden, const wemfwi_complex* data, int add, int mode,
                            void* stream);

/* out[j] (+)= sum_{c = ptr[j]}^{ptr[j+1]-1} in[src[c]] (device arrays; double
 * accumulation, list order; empty lists give 0 or leave out[j] with add): the
 * model <-> grid gathers of the conformal maps and their sums over shots. */
int wemfwi_gather_sum(long long nout, const long long* ptr, const long long* src,
                      const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);

/* BornRefl (ext_born_refl.cpp:8-110): frequency-extended reflectivity Born
 * operator / migration on the unpadded grid; image [nw][nz][nx] device.  The
 * adjoint is the zero-lag imaging condition image[w] += sum_s conj(D) U
 * (ic.cpp:17-28). */
int wemfwi_bornrefl_forward(wemfwi_model* model, wemfwi_survey* survey, const wemfwi_complex* image,
                            wemfwi_complex* data, int add, void* stream);
int wemfwi_bornrefl_adjoint(wemfwi_model* model, wemfwi_survey* survey, wemfwi_complex* image,
                            const wemfwi_complex* data, int add, void* stream);

/* Subsurface-offset extended imaging (no counterpart in the reference; its
 * h = 0 slice is the zero-lag image summed over frequencies):
 * I[h][z][x] = sum_{w,s} conj(D(z, x-h)) U(z, x+h), h = -(nh/2)..nh/2, nh odd;
 * image [nh][nz][nx].  The forward is its adjoint with respect to U. */
int wemfwi_offset_image_forward(wemfwi_model* model, wemfwi_survey* survey, int nh,
                                const wemfwi_complex* image, wemfwi_complex* data, int add,
                                void* stream);
int wemfwi_offset_image_adjoint(wemfwi_model* model, wemfwi_survey* survey, int nh,
                                wemfwi_complex* image, const wemfwi_complex* data, int add,
                                void* stream);

/* Subsurface-offset migration of shots with uneven apertures and bands (no
 * single counterpart in the reference, which migrates one (shot, frequency)
 * task at a time on one grid, ext_born_refl.cpp:60-110 with the image of
 * wemfwi_offset_image_adjoint): shot s propagates on its own background
 * `model` (nz x nx_s, its aperture), placed at image column x0; its rows are
 * its nw frequencies.  Per row: D = Down(source spread ngs at (src_iz,
 * src_ix)), R = Up^H(receivers at (rec_iz, rec_ix) <- data[r][i]), and
 * image[h][z][x0 + x] += conj(D(z, x-h)) R(z, x+h) over h = -(nh/2)..nh/2,
 * nh odd <= 41.  All rows of the call are swept in batched launches; the
 * image (device [nh][nz][nx_image]) is summed in FP64, shot by shot in
 * order, rows in order.  Frequencies, wavelet and receiver positions are
 * host arrays; data is a device array [nr][nw] per shot. */
typedef struct wemfwi_image_shot {
  wemfwi_model* model;
  int x0;
  int nw;
  const float* freq_hz;
  const wemfwi_complex* wavelet;
  int src_iz, src_ix;
  int nr;
  const int* rec_iz;
  const int* rec_ix;
  const wemfwi_complex* data;
} wemfwi_image_shot;
int wemfwi_uneven_offset_image_adjoint(const wemfwi_image_shot* shots, int ns, int ngs, int nh, int nz,
                                       int nx_image, wemfwi_complex* image, int add, void* stream);

/* Extended least-squares objective of the same shots and its gradient:
 * J = 1/2 sum_s ||L_s I - d_s||^2 and grad = sum_s L_s^H (L_s I - d_s), where
 * I = model is a device image [nh][nz][nx_image], d_s = shots[s].data the
 * observed data [nr][nw], and L_s I models each (shot, frequency) row as
 * BornRefl::fwd_propagate (ext_born_refl.cpp:8-56: D = Down(source), Up,
 * receivers) with the offset scatter U(z, y) = sum_h I(h, z, x0 + y - h)
 * D(z, y - 2h) (the adjoint of the offset imaging condition) in place of IC.
 * L_s^H is wemfwi_uneven_offset_image_adjoint of the residual; D is computed
 * once per row for both.  grad (+)= the gradient (device, distinct from
 * model); *misfit (host) = J, summed in FP64 in a fixed order.  Arguments
 * and errors as wemfwi_uneven_offset_image_adjoint. */
int wemfwi_uneven_offset_gradient(const wemfwi_image_shot* shots, int ns, int ngs, int nh, int nz, int nx_image,
                                  const wemfwi_complex* model, wemfwi_complex* grad, double* misfit, int add,
                                  void* stream);

/* BornFull::forward / adjoint (BornFull.cpp:64-117, born_prop.cpp:11-235) of
 * shots with uneven apertures and bands: shot s propagates on its own
 * background `model` (nz x nx_s, its aperture; any FFT length the models
 * support) placed at column x0 of the perturbation grid [nz][nx_model]
 * (columns x0 + a outside [0, nx_model) are zero, as the reference's
 * zero-padded perturbation, born_prop.cpp:52-58); its rows are its nw
 * frequencies; the source (spread ngs) and the receivers (spread ngr) sit on
 * the aperture grid; data is the shot's device array [nr][nw] (written by the
 * forward, read by the adjoint).  All rows of a call are swept in batched
 * launches (one per kernel variant: FFT length x full row x split step) and
 * LinDown / LinUp likewise; injection, recording and the FP64 reduction run
 * per shot in the order of a one-shot call, so every shot's data and every
 * shot's gradient slice equal those of wemfwi_born_forward /
 * wemfwi_born_adjoint on that shot alone (a one-shot survey with bins
 * 0..nw-1, nt = nw, offset 0) bitwise.  The adjoint sums the shots' slices
 * in shot order in double precision into dslow / dden [nz][nx_model]
 * (device; add = 0 overwrites them); mode as born_adjoint.  Whole
 * wavefields are kept (the survey checkpoint setting does not apply);
 * consecutive shots are grouped so that a group's wavefields fit half of
 * the free device memory.  Frequencies, wavelets and device positions are
 * host arrays. */
typedef struct wemfwi_born_shot {
  wemfwi_model* model;
  int x0;
  int nw;
  const float* freq_hz;
  const wemfwi_complex* wavelet;
  int src_iz, src_ix;
  int nr;
  const int* rec_iz;
  const int* rec_ix;
  wemfwi_complex* data;
} wemfwi_born_shot;
int wemfwi_uneven_born_forward(const wemfwi_born_shot* shots, int ns, int ngs, int ngr, int nz, int nx_model,
                               const wemfwi_complex* dslow, const wemfwi_complex* dden, int add, void* stream);
int wemfwi_uneven_born_adjoint(const wemfwi_born_shot* shots, int ns, int ngs, int ngr, int nz, int nx_model,
                               wemfwi_complex* dslow, wemfwi_complex* dden, int add, int mode, void* stream);

/* extBornFull::forward / adjoint (extBornFull.cpp:11-121,
 * ext_born_prop.cpp:8-194) of shots with uneven apertures and bands: as
 * wemfwi_uneven_born_*, but frequency i of shot s propagates on its own
 * background models[i] (nz x nx_s, the shot's aperture) and meets the
 * perturbation slice w0 + i of dslow / dden [nslice][nz][nx_model] (the
 * reference's one model slice per frequency).  Every shot's data, and every
 * shot's contribution to each slice before the sum over shots (mode 1),
 * equal wemfwi_extborn_forward / wemfwi_extborn_adjoint on that shot alone
 * (models[0..nw), slices w0..w0+nw) bitwise.  The adjoint sums in shot
 * order in double precision.  mode 0 reproduces the reference's serial shot
 * loop (LinOneWay.cpp:64-65): the last depth row of a slice keeps, at each
 * column, the LinUp term of the last shot whose aperture and band cover it
 * (shots without such a successor keep theirs); mode 1 is the exact
 * adjoint. */
typedef struct wemfwi_extborn_shot {
  wemfwi_model* const* models;
  int x0;
  int nw;
  int w0;
  const float* freq_hz;
  const wemfwi_complex* wavelet;
  int src_iz, src_ix;
  int nr;
  const int* rec_iz;
  const int* rec_ix;
  wemfwi_complex* data;
} wemfwi_extborn_shot;
int wemfwi_uneven_extborn_forward(const wemfwi_extborn_shot* shots, int ns, int ngs, int ngr, int nz,
                                  int nx_model, int nslice, const wemfwi_complex* dslow,
                                  const wemfwi_complex* dden, int add, void* stream);
int wemfwi_uneven_extborn_adjoint(const wemfwi_extborn_shot* shots, int ns, int ngs, int ngr, int nz,
                                  int nx_model, int nslice, wemfwi_complex* dslow, wemfwi_complex* dden,
                                  int add, int mode, void* stream);

/* ---------------------------------------------------------------- extended FWI
 * Operators of the frequency-extended objective of scripts/efwi.py:160-190
 * (operator/src/python/Operator.py:228-292,619-775), on device models
 * [nw][nz][nx]. */

/* Matrix M[n][n] of scipy.ndimage.correlate1d (convolve = 0) or convolve1d
 * (convolve = 1) with taps w[L], origin 0, along a line of n samples, so that
 * out[i] = sum_j M[i][j] in[j]; boundary mode 0 reflect, 1 constant (0),
 * 2 nearest, 3 mirror, 4 wrap (extended periodically for L > n).  Host. */
int wemfwi_filter_matrix(int n, int L, const double* w, int mode, int convolve, double* M);

/* out[i][p] (+)= c * sum_j M[i][j] in[j][p] + i * alpha * mean_j Im in[j][p]
 * over ncol columns p, c = c_re + i c_im, M a device float matrix [nw][nw]:
 * Derivative::forward is c = i*beta with the correlate matrix,
 * Derivative::adjoint c = -i*beta with the convolve matrix (Operator.py:272-292).
 * in and out are distinct device arrays [nw][ncol]. */
int wemfwi_freq_filter(int nw, long long ncol, const float* M, float c_re, float c_im, float alpha,
                       const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);

/* Soft clip of one component (comp 0: real, 1: imaginary) of n device
 * samples with h(u) = l u - sqrt(l^2 u^2 + tau^2) (Operator.py:619-623):
 * jacobian = 0: y = sign*h(sign*(x - x0)) + x0 (sign = +1 Softmax /
 * HyperbolicPenalty with x0 = 0, -1 Softmin); jacobian = 1: y = x * h'(sign*(b - x0))
 * at the background bg (Softclip, dSoftmax, dSoftmin; self-adjoint).  The
 * other component passes through; add = 1 accumulates into out. */
int wemfwi_soft_clip(long long n, int comp, int sign, float x0, float l, float tau, int jacobian,
                     const wemfwi_complex* bg, const wemfwi_complex* in, wemfwi_complex* out, int add,
                     void* stream);

/* ---------------------------------------------------------------- model interpolation
 * Preconditioning of the FWI model (scripts/fwi.py:46-68, efwi.py:48-80). */

/* Filter F[nd][nm] from a coarse axis (nm samples, spacing dm) to a fine axis
 * (nd, dd), host: kind 0 Lanczos with a window of a coarse cells
 * (LanczosInterpolation3D.h:29-55; the 2-D operator is taper 0), kind 1 the
 * cubic (a, b) spline (Spline3D.h:22-52; Spline2D.h:97-127 is taper 0);
 * taper widens the kernel to (1 + taper) dm.  Both fold the taps that fall
 * outside the axis back onto it, as written there. */
int wemfwi_interp_filter(int kind, int nd, float dd, int nm, float dm, float a, float b, float taper,
                         float* F);

/* out[(a n_out + i) inner + b] (+)= sum_j F[i][j] in[(a n_in + j) inner + b]
 * for a device float matrix F[n_out][n_in] (transpose = 1: F stored
 * [n_in][n_out], applied transposed): one axis of the separable
 * interpolation (Interpolation3D.h:21-95).  in != out.  Each output row
 * sums only over its band of nonzero taps (found on the device per call). */
int wemfwi_axis_apply(const float* F, int n_out, int n_in, long long outer, long long inner, int transpose,
                      const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);

/* ---------------------------------------------------------------- unit operators
 * Row-batched forms of the reference's unit operators (py_wem_ops.cpp:25-206),
 * for operator tests and custom pipelines.  Wavefields are device arrays;
 * add == 0 overwrites the output (the reference's add=false). */

/* Scatter (Scatter.cpp:35-111) on nrows rows of nx samples, row i at depth[i]
 * and frequency freq_hz[i]; mode 0 keeps Scatter::adjoint's operator order
 * (Scatter.cpp:52-66), mode 1 is the exact adjoint. */
int wemfwi_scatter(wemfwi_model* model, int adjoint, int mode, int nrows, const float* freq_hz,
                   const int* depth, const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);

/* LinDown (up = 0) / LinUp (up = 1) (LinOneWay.cpp:33-142) on [nrows][nz][nx]
 * with background wavefields bg[nrows][nz][nx]: forward
 * out[iz+-1] += P_iz Scatter_iz(bg[iz] * in[iz]); adjoint the transpose. */
int wemfwi_lin_oneway(wemfwi_model* model, int up, int adjoint, int mode, int nrows,
                      const float* freq_hz, const wemfwi_complex* bg, const wemfwi_complex* in,
                      wemfwi_complex* out, int add, void* stream);

/* Reflect (Reflect.h:12-48) on [nrows][nz][nx]: out = r * in (adjoint conj(r) * in). */
int wemfwi_reflect(wemfwi_model* model, int adjoint, int nrows, const wemfwi_complex* in,
                   wemfwi_complex* out, int add, void* stream);

/* dReflect (dReflect.h:11-84) of the model's background on [nz][nx]:
 * forward dr = dReflect(dslow, dden); adjoint (dslow, dden) = dReflect^H(dr). */
int wemfwi_dreflect(wemfwi_model* model, int adjoint, wemfwi_complex* dslow, wemfwi_complex* dden,
                    wemfwi_complex* dr, int add, void* stream);

/* Phshift::forward/adjoint (Phshift_cmplx.cpp:47-113, phase_shift.ispc:2-32,
 * py_wem_ops.cpp:95-113) on nrows independent wavenumber rows of nk samples:
 * out[k] (+)= in[k] exp(-i kz dz) (adjoint: the conjugate factor), kz the
 * damped vertical wavenumber of the reference slowness sref[i] at frequency
 * freq_hz[i] (host arrays) on the k grid of Phshift_cmplx.cpp:12-24
 * (dk = 2 pi / (dx (nk - 1))). */
int wemfwi_phshift(int nk, float dx, float dz, float eps, int adjoint, int nrows, const float* freq_hz,
                   const wemfwi_complex* sref, const wemfwi_complex* in, wemfwi_complex* out, int add,
                   void* stream);

/* SplitStep::forward/adjoint (SplitStep.cpp:5-29, py_wem_ops.cpp:25-46) on
 * nrows rows of the model's width: at the nloc positions loc (setLocation),
 * row i at depth depth[i] (setDepth), frequency freq_hz[i] (setFreq) and
 * reference slowness sref[i] (setSlow): forward out[x] = in[x] exp(-i w dz
 * (S - sref)/sqrt(sref)), adjoint out[x] = in[x] exp(+i w dz conj((S -
 * sref)/sqrt(sref))), assigned at loc; add = 0 zeroes the rows of out first
 * (host arrays freq_hz, depth, sref, loc). */
int wemfwi_splitstep(wemfwi_model* model, int adjoint, int nrows, const float* freq_hz, const int* depth,
                     const wemfwi_complex* sref, int nloc, const int* loc, const wemfwi_complex* in,
                     wemfwi_complex* out, int add, void* stream);

/* IC (ic.cpp:5-45) on n samples: forward d = bg * m; adjoint m = conj(bg) * d. */
int wemfwi_ic(int adjoint, long long n, const wemfwi_complex* bg, wemfwi_complex* m, wemfwi_complex* d,
              int add, void* stream);

/* Injection (Injection.h:19-65, Injection.cpp:8-80) of ndev devices at (iz, ix)
 * (host arrays) with cos^2 spread ng and x offset tap, on nrows wavefields
 * [nrows][nz][nx]: forward wfld[iz][ix+ig+tap] += g[ig] vals[row][dev];
 * adjoint vals[row][dev] = sum g[ig] wfld[iz][ix+ig+tap]. */
int wemfwi_injection(int nz, int nx, int ndev, const int* iz, const int* ix, int ng, int tap,
                     int adjoint, int nrows, wemfwi_complex* vals, wemfwi_complex* wfld, int add,
                     void* stream);

/* Kernel timing with HIP events on the launch stream (off by default).
 * wemfwi_timing(1) clears and enables; wemfwi_timing_get returns the summed
 * duration (ms) and launch count of a kernel by name ("oneway_kernel",
 * "phase_table_kernel", ...). */
int wemfwi_timing(int enable);
int wemfwi_timing_get(const char* kernel, double* total_ms, long* count);

/* Decisions of the sweep autotuner so far (no counterpart in the reference,
 * whose FFTW plans are chosen by FFTW_ESTIMATE, FFT1_c.cpp:34-38): a JSON
 * list with, per (FFT length, full row, split step, adjoint, device), the
 * plan chosen and the measured mean launch time of every radix plan, and
 * per batch group shape ("kernel": "batch_order") the launch order of the
 * groups chosen and the best time of every candidate order.
 * Writes at most cap bytes (NUL-terminated) and returns the full length. */
int wemfwi_autotune_report(char* buf, int cap);

#ifdef __cplusplus
}
#endif

#endif /* WEMFWI_H */
