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

/* Down / Up (OneWay.cpp:32-109) for nrows independent (shot, frequency) rows
 * of [nz][nx] each, row i at angular frequency 2*pi*freq_hz[i].
 * up = 0: Down, 1: Up.  adjoint = 0: forward, 1: adjoint.
 * in == out is the in-place recursion of every reference call site
 * (W[iz+1] = P W[iz] + W[iz+1] for Down forward); with add = 0 it returns
 * zeros, as the reference zeroes the shared vector first.  in != out applies
 * one bidiagonal step per depth (the reference's out-of-place semantics). */
int wemfwi_oneway(wemfwi_model* model, int up, int adjoint, int nrows, const float* freq_hz,
                  const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream);

/* SSF / PSPI single depth step (SSF.cpp:8-60, PSPI.cpp:8-53) on nrows
 * independent rows of nx samples: row i at depth depth[i] and frequency
 * freq_hz[i].  add = 1 accumulates into out. */
int wemfwi_onestep(wemfwi_model* model, int adjoint, int nrows, const float* freq_hz,
                   const int* depth, const wemfwi_complex* in, wemfwi_complex* out, int add,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WEMFWI_H */
