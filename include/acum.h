/*
 * acum.h -- C-ABI of the MI355X acoustic hot path (Farassat 1A tonal + BPM broadband).
 *
 * The library libacum_hip.so replaces two call sites of the reference ACUM-3 program:
 *   - main.cu:156-247, the device allocations, host->device copies, the launch of
 *     findPressureTimeHistory (core.cu:18, declared core.cuh:1) and the device->host copies;
 *   - main.cu:150, the host call of findBBnoise (bb_core.cu:11, declared bb_core.cuh:1).
 * The host-side functions at the end of this header restate the reference's file contract and
 * host preprocessing (io.cu, io.cuh:1-10, preprocess.cu:228, main.cu:100-142) so that a driver
 * can read the same input files and write the same output files.
 *
 * Conventions (all entry points):
 *   - 0 is success; a negative ACUM_ERR_* code is failure.  The library never calls exit();
 *     the text of the last failure is returned by acum_last_error() / acum_host_error().
 *   - The caller owns every host buffer.  The context owns the device buffers it allocates.
 *   - Outputs are overwritten, never accumulated into.  Values equal what the reference leaves
 *     in its caller-initialised buffers (pT/pL/pA start at 0, main.cu:123-130; BBspl/BBoaspl
 *     start at -100 dB, main.cu:132-142).
 *   - One context per host thread; calls are synchronous on the context's stream.
 *   - Index arrays are 64-bit (the reference's int offsets overflow beyond 2^31 floats).
 */
#ifndef ACUM_H
#define ACUM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ACUM_OK = 0,
  ACUM_ERR_ARG = -1,     /* invalid argument, inconsistent sizes, unknown keyword */
  ACUM_ERR_IO = -2,      /* missing, truncated or malformed input file */
  ACUM_ERR_DEVICE = -3,  /* no GPU, or a HIP runtime call failed */
  ACUM_ERR_NOMEM = -4,   /* host or device allocation failed */
  ACUM_ERR_NUMERIC = -5, /* an observer within 1e-6 m of a source position, 1 - Mr <= 0 (or a
                            non-increasing observer time), a non-finite result                 */
  ACUM_ERR_PLAN = -6     /* a tile's observer-time window exceeded its planned bound */
};

/* Arithmetic of the tonal path. */
enum {
  ACUM_FP64 = 0,  /* every operation in FP64, restructured like the mixed path (the reference's
                     own structure in FP64 is acum_tonal_plain)                                */
  ACUM_MIXED = 1  /* FP64 observer time, bin index and bin deltas; FP32 geometry, Doppler and
                     integrands; FP32 per-tile partials summed in FP64 with Kahan compensation */
};

typedef struct acum_ctx acum_ctx;

/* ----------------------------------------------------------------------------------------
 * Tonal case: the arguments of findPressureTimeHistory (core.cu:18) after preprocess_imper.
 * Element data stays in the reference AoS layout: value w of element j of surface i at time
 * step k lives at  X[XPtr[i] + (j*nTime + k)*4 + w]  (core.cu:63-66), with
 *   X  = (x, y, z, dS)               4 floats, m and m^2
 *   N  = (nx, ny, nz, dnx, dny, dnz) 6 floats, unit normal and its time derivative
 *   V  = (vx, vy, vz, ax, ay, az)    6 floats, surface velocity and acceleration
 *   PR = (P, dP/dt)                  2 floats, gauge pressure (Pa) and its time derivative
 * ---------------------------------------------------------------------------------------- */
typedef struct acum_tonal_case {
  int64_t nSurf, nTime, nObs;
  const int64_t* E;                      /* [nSurf] elements per surface                    */
  const float *X, *N, *V, *PR;
  const int64_t *XPtr, *NPtr, *VPtr, *PRPtr; /* [nSurf] float offsets of each surface        */
  const float* rObs;                     /* [nObs*3] observer positions, m                   */
  double dPsi;                           /* azimuth step, radians (io.cu:75)                  */
  double dTau;                           /* time step, s (io.cu:74)                           */
  double a0, rhoRef, oM;                 /* speed of sound, density, rotor speed (rad/s)      */
  double vInf[3];                        /* free-stream velocity, m/s (io.cu:81-83)           */
  int32_t precision;                     /* ACUM_FP64 or ACUM_MIXED                           */
  int32_t reserved;
} acum_tonal_case;

/* ----------------------------------------------------------------------------------------
 * BPM case: the 39 arguments of findBBnoise (bb_core.cu:11, bb_core.cuh:1), with 64-bit
 * offsets.  cn and cc in bbdata are as read_BB_inputs leaves them: already divided by (r/R)^2
 * when omega != 0 (io.cu:509-512).  ff[] carries the band centre frequencies, so both the
 * reference 42-band table (io.cu:693-699) and other band sets work.
 * ---------------------------------------------------------------------------------------- */
typedef struct acum_bpm_case {
  int32_t nSurf_BB, nTime, nObs, nf2, t_del, reserved;
  double dTau, a0;
  double vInf[3];
  const float* rObs;                     /* [nObs*3]                                         */
  const float* ff;                       /* [nf2] band centre frequencies, Hz                */
  const float *span, *omega, *psi_offset; /* [nSurf_BB]                                      */
  const int32_t *ccw, *trip;             /* [nSurf_BB]                                       */
  const float *rotate, *translate;       /* [3*nSurf_BB], degrees and m                      */
  const int32_t* nsect;                  /* [nSurf_BB] span sections per surface             */
  const int64_t *xsPtr, *dsPtr;          /* [nSurf_BB] offsets into xyzcR and bbdata         */
  const float* xyzcR;                    /* [sum nsect*4]  r/R, y/R, z/R, c/R                */
  const float* bbdata;                   /* [sum nsect*nTime*7] d_s, d_p, d*_s, d*_p, cn, cc, theta(deg) */
  const int32_t *nsect_t, *nmach_t, *naoa_t, *nre_t; /* [nSurf_BB] table dimensions        */
  const int64_t *aPtr, *mPtr, *rPtr, *sPtr, *dPtr;  /* [nSurf_BB] table offsets             */
  const float *sect_t, *mach_t, *aoa_t, *re_t;       /* table axes                          */
  const float* bbdata_t;                 /* [sect][mach][re][aoa][7] table data (io.cu:614)  */
} acum_bpm_case;

/* Timing of the last run, measured with HIP events on the context's stream. */
typedef struct acum_timing {
  double prep_ms;    /* AoS -> device record conversion                                     */
  double eval_ms;    /* tonal eval + binning kernel (or BPM spectra kernels)                 */
  double reduce_ms;  /* ordered FP64/Kahan partial reduction                                 */
  double total_ms;   /* everything between the first and the last event                      */
  int64_t evals;     /* source-observer-time evaluations of the run                          */
  int64_t partial_bytes; /* bytes of per-tile partial slabs read by the reduction            */
  int32_t n_partials;    /* partial slabs per observer row (P)                               */
  int32_t n_items;       /* work items of the persistent launch                              */
  int32_t launches;      /* kernel launches of the run                                       */
  int32_t motion_widened_cases; /* cases planned with a wider source-motion bound than tuned */
                                /* after a window overflow (see acum_tune)                    */
  int32_t eval_launches; /* eval kernel launches of a pass: one per kernel variant in use     */
                         /* (still air or moving medium, window size)                         */
  /* after acum_autotune: */
  int32_t tune_cases;    /* cases whose bucket was tuned by this call                          */
  int32_t tune_hits;     /* cases whose bucket was already in the context's cache             */
  int32_t tune_passes;   /* timed passes the tuning took                                       */
} acum_timing;

/* Tile shapes of the tonal kernel (chosen by acum_autotune or set by the caller). */
typedef struct acum_tune {
  int32_t window;      /* observer-time bins per lane in LDS: 16, 20, 24, 32 (+ 48, 64 mixed) */
  int32_t max_spread;  /* largest spread (bins) of a tile's elements around its first one     */
  int32_t max_tile;    /* most elements per tile                                              */
  int32_t groups;      /* partial slabs per observer row (0 = automatic: four work items per  */
                       /* workgroup slot; either way at most what fits a quarter of the       */
                       /* case's element records or 8 GiB, whichever is larger)               */
  int32_t max_chunk;   /* most source time steps per window pass                              */
  int32_t waves;       /* waves per workgroup: always 1 (one 64-observer wave per workgroup;  */
                       /* the LDS window, not the workgroup size, sets the occupancy)         */
  int32_t motion_permille; /* planning bound on observer-time change per metre of source    */
                       /* motion, per mille of the worst case (0 = default 600); a case whose */
                       /* window overflows under it is planned again with the next of 600,    */
                       /* 800 and the full bound that widens it, and run again within the      */
                       /* same call                                                            */
} acum_tune;

/* --- context ----------------------------------------------------------------------------- */
/* A context owns a stream on its device, the staged cases and one device workspace that the
 * broadband and spectra calls carve their buffers from; the workspace grows to the largest such
 * call and is freed by acum_ctx_destroy.                                                      */
int acum_device_count(void);
int acum_ctx_create(acum_ctx** ctx, int device);
void acum_ctx_destroy(acum_ctx* ctx);
const char* acum_last_error(const acum_ctx* ctx);

/* --- tonal ------------------------------------------------------------------------------- */
/* One case, host in / host out.  Replaces main.cu:156-247 (launch at main.cu:236).           */
int acum_tonal(acum_ctx* ctx, const acum_tonal_case* c, float* pT, float* pL, float* pA);
/* The plain FP64 path, kept as the measurement baseline of the restructured paths: the
 * reference's kernel structure (findPressureTimeHistory, core.cu:18-259, with the launch shape
 * of main.cu:194-236) in FP64 -- one thread per (step, observer) striding over elements, three
 * retarded-time solves per eval and FP64 atomic adds into the bins.  Host in / host out like
 * acum_tonal, same errors; c->precision is not used.  acum_get_timing reports the kernel time
 * (eval_ms) and the conversion of the FP64 bins to the float outputs (reduce_ms).  Bin sums
 * depend on the order the atomics are served in, so repeated calls agree to rounding only.    */
int acum_tonal_plain(acum_ctx* ctx, const acum_tonal_case* c, float* pT, float* pL, float* pA);
/* Many cases of uneven size in one persistent launch.  Per-case results equal acum_tonal's.  */
int acum_tonal_batch(acum_ctx* ctx, const acum_tonal_case* cases, int32_t nCases,
                     float* const* pT, float* const* pL, float* const* pA);
/* Staged form: upload once (element data resident in HBM), run many times, fetch results.   */
int acum_tonal_stage(acum_ctx* ctx, const acum_tonal_case* cases, int32_t nCases, int32_t* handle);
int acum_tonal_run(acum_ctx* ctx, int32_t handle, int32_t precision, int32_t repeats, int32_t use_graph);
int acum_tonal_fetch(acum_ctx* ctx, int32_t handle, int32_t icase, float* pT, float* pL, float* pA);
int acum_tonal_release(acum_ctx* ctx, int32_t handle);
/* Diagnostics: for observer iobs and every (surface, element, step) in case order, the bin
 * lookup of the eval -- first bin ipts and periodic shift npi (core.cu:169, 290-355) and the
 * number of bins of the down-sweep and up-sweep (core.cu:199-231).  Arrays hold sum(E)*nTime. */
int acum_tonal_bins(acum_ctx* ctx, const acum_tonal_case* c, int64_t iobs,
                    int32_t* ipts, int32_t* npi, int32_t* ndown, int32_t* nup);
int acum_get_timing(const acum_ctx* ctx, acum_timing* t);
/* The context's tile shape per precision: what a case plans with unless the autotuner has a
 * shape for its bucket.  Replaces the compile-time launch tiles of main.cu:194-200.            */
int acum_get_tuning(const acum_ctx* ctx, int32_t precision, acum_tune* t);
int acum_set_tuning(acum_ctx* ctx, int32_t precision, const acum_tune* t);
/* Autotuner.  Each case of the staged batch falls in a shape bucket: precision, still air or
 * moving medium, the rounded-up base-2 logarithms of its element, step and observer counts, and
 * the device's compute units.  A bucket not yet in the context's cache is tuned on its case:
 * successive halving over window (16, 20, 24) x spread (2, 3, 4) x source-motion bound (0.5,
 * 0.6, 0.8 of the worst case), then max_tile (64, 256), each candidate
 * timed on the case's first 512 observers (all of them in a smaller case), keeping the best
 * quarter per round with twice the passes.  The chosen shape is cached by bucket, and every
 * later plan of a case in that bucket (alone or in any batch) uses it, so batched and single
 * runs of a case stay bit-identical.
 * The winner then runs once on the whole case and keeps its source-motion bound only if that
 * holds for every observer (else the bound that held replaces it).
 * best / best_ms: the shape and time per pass (on the observer subset) of case 0's bucket.
 * acum_get_timing then reports tune_cases, tune_hits and tune_passes of the call.  The tuning
 * passes overwrite the staged cases' histories: run the stage again before fetching them.     */
int acum_autotune(acum_ctx* ctx, int32_t handle, int32_t precision, acum_tune* best, double* best_ms);
/* The shape case icase of a staged batch plans with for a precision (from the cache or the
 * context's shape).                                                                          */
int acum_get_case_tuning(acum_ctx* ctx, int32_t handle, int32_t icase, int32_t precision, acum_tune* t);
/* The autotuner's cache as a text file (one bucket per line), to keep tuned shapes across
 * processes; loading adds to (and overrides within) the context's cache; clear empties it.   */
int acum_tune_cache_save(const acum_ctx* ctx, const char* path);
int acum_tune_cache_load(acum_ctx* ctx, const char* path);
int acum_tune_cache_clear(acum_ctx* ctx);

/* --- broadband --------------------------------------------------------------------------- */
/* Replaces main.cu:150 (findBBnoise, bb_core.cu:11).  BBspl[nObs*nf2*7], BBoaspl[nObs].      */
int acum_bpm(acum_ctx* ctx, const acum_bpm_case* c, float* BBspl, float* BBoaspl);

/* ======================= host side: file contract and preprocessing ====================== */
/* Scalars of the `input` file (io.cu:12-101) and derived constants. */
typedef struct acum_inputs {
  double a0, rhoRef, pRef, oM, dPsi_deg, Minf, Mref, xy_angle, xz_angle, CFDscale, OBSscale;
  int32_t nSurf, nTime, periodic, lowpass, impermeable, BB_noise;
  double dTau, dPsi;   /* derived: s per step, radians per step (io.cu:74-75)               */
  double vInf[3];      /* derived (io.cu:81-83)                                              */
} acum_inputs;

typedef struct acum_run acum_run;  /* everything one run directory holds */

const char* acum_host_error(void);  /* text of the last host-side failure (this thread)     */
/* read_inputs + read_BB_inputs + read_observers + read_surf (main.cu:56-78). */
int acum_run_load(const char* dir, acum_run** run);
void acum_run_free(acum_run* run);
/* preprocess_imper (preprocess.cu:228-413): scaling, periodic fix, derivatives in FP64.      */
int acum_run_preprocess(acum_run* run);
/* The same preprocessing on the context's GPU (preprocess.cu:228-413, called at main.cu:88):
 * the run's element arrays are copied to the device, processed in FP64 with the operations of
 * acum_run_preprocess (bit-identical results) and copied back.  nTime <= 4096.  Timing:
 * prep_ms = upload, eval_ms = kernel, reduce_ms = download.                                 */
int acum_preprocess_gpu(acum_ctx* ctx, acum_run* run);
/* Stages the tonal case of a run for acum_tonal_run / _fetch / _release (one case, handle out).
 * A run that is not preprocessed is uploaded raw and preprocessed on the GPU in place
 * (nTime <= 4096, as acum_preprocess_gpu, bit-identical to acum_run_preprocess), so the
 * preprocessed element data never come back to the host; the run itself is left as it was.
 * Replaces the host preprocess_imper call (main.cu:88) and the device copies (main.cu:156-234). */
int acum_run_stage(acum_ctx* ctx, const acum_run* run, int32_t* handle);
int acum_run_inputs(const acum_run* run, acum_inputs* in);
/* The run's tonal case (pointers into the run).  *c is filled in either way, so its sizes can be
 * read from a run that is not preprocessed, but then the call returns ACUM_ERR_ARG: its element
 * arrays are not yet the kernel's inputs (preprocess first, or stage with acum_run_stage).    */
int acum_run_tonal_case(const acum_run* run, int32_t precision, acum_tonal_case* c);
int acum_run_bpm_case(const acum_run* run, acum_bpm_case* c);
/* pT.out, pL.out, pA.out (io.cu:137-203) and bb_oaspl.out, bb_spl.dat (io.cu:702-796).      */
int acum_run_write_tonal(const acum_run* run, const char* dir, const float* pT, const float* pL, const float* pA);
int acum_run_write_bb(const acum_run* run, const char* dir, const float* BBoaspl, const float* BBspl);
/* Writes the input files of the run (input, obs.dat, surfaces/, input_BB, BB_inputs/).       */
int acum_run_save(const acum_run* run, const char* dir);
/* Binary cache of the parsed surface files: surfaces/surface_k.acumc beside each
 * surface_k.dat (a 64-byte header with E, nt and the text file's size and modification time,
 * then the parsed floats).  acum_run_load copies a matching cache instead of parsing the text;
 * a cache that does not match the text file (or ACUM_NO_SURFACE_CACHE=1) is ignored.  No
 * counterpart in the reference, which parses the text each run (io.cu:206-333).  The run must
 * be as read (not preprocessed).                                                             */
int acum_run_save_cache(const acum_run* run, const char* dir);
/* Raw surface arrays as read (before preprocess): x,y,z,CP,A,nx,ny,nz per (surface, k, j).   */
int acum_run_surface_raw(const acum_run* run, int32_t isurf, float* out /* [nTime*E*8] */);

/* Synthetic rotor cases written in the reference's file contract. */
typedef struct acum_synth_params {
  uint64_t seed;
  int32_t nBlades, nChord, nSpan, nTime;   /* panels per side: nChord x nSpan               */
  int32_t coaxial;                         /* 1: two counter-rotating rotors of nBlades     */
  int32_t obs_layout;                      /* 0: arc in x-z plane, 1: hemisphere az x el, 2: Fibonacci sphere */
  int32_t nObsAz, nObsEl, nObs;            /* layout sizes (nObs for arc / Fibonacci)        */
  int32_t cp_model;                        /* 1: -0.6 (r/R)^2 (1+0.1 sin psi); 2: thin-airfoil x (1+0.2 cos psi); 3: model 2 + pulses */
  int32_t bb;                              /* 1: also write broadband inputs                  */
  int32_t nSect, t_del, trip, nBands;      /* broadband: sections per blade, BL model, trip, bands (34 or 42) */
  double R, chord, rootCut, oM, a0, rhoRef, pRef, Minf, xy_angle, xz_angle;
  double obsRadius, obsElMin, obsElMax;    /* m, degrees                                     */
  double cpNoise, coaxSep;                 /* N(0, cpNoise) per (element, step); coaxial gap (m) */
  double twistRoot, twistTip;              /* broadband pitch, degrees                       */
} acum_synth_params;
int acum_synth_defaults(int32_t config, acum_synth_params* p);
/* Case icase of the batched workload: sizes and speeds drawn from seed.                      */
int acum_synth_batch_case(uint64_t seed, int32_t icase, acum_synth_params* p);
int acum_synth(const acum_synth_params* p, acum_run** run);

/* Tonal overall level per observer: 10 log10(mean_t (pA - mean pA)^2 / (2e-5)^2), dB.       */
int acum_oaspl(int64_t nTime, int64_t nObs, const float* pA, double* oaspl);

/* Tonal band spectra on the GPU.  No counterpart in the reference, which writes the pressure
 * histories only (io.cu:137-203); this is the post-processing step that turns them into levels.
 * Per observer, the one-sided mean-square of pA [nTime][nObs] (mean removed; DFT bins
 * k = 1 .. nTime/2 at k / (nTime dt) Hz, each weighted 2 except the Nyquist bin) summed into the
 * bands of centre frequencies fc[0 .. nBands) (strictly increasing, Hz).  Band b holds the bins
 * with lo_b <= f < hi_b, where the edge between two bands is the geometric mean of their
 * centres and the outer edges are fc[0] 2^(-1/6) and fc[nBands-1] 2^(1/6).
 *   spl[o * nBands + b] = 10 log10(max(E_b, 1e-10 (2e-5)^2) / (2e-5)^2)   (-100 dB floor)
 *   oaspl[o]            = 10 log10(variance / (2e-5)^2), as acum_oaspl
 * dt is the observer-time step in seconds (dPsi / oM for the tonal outputs).  Over bands that
 * cover every bin the band energies sum to the variance.                                      */
int acum_tonal_spectra(acum_ctx* ctx, int64_t nTime, int64_t nObs, const float* pA, double dt,
                       int32_t nBands, const float* fc, float* spl, float* oaspl);
/* The same for case icase of a staged batch after acum_tonal_run, from its device-resident pA
 * (dt = dPsi / oM of the case).                                                              */
int acum_tonal_fetch_spectra(acum_ctx* ctx, int32_t handle, int32_t icase, int32_t nBands,
                             const float* fc, float* spl, float* oaspl);

/* Allocation-free form of acum_tonal_spectra.  No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  pA [nTime][nObs], spl [nObs][nBands] and oaspl [nObs] are DEVICE pointers; fc is a
 * host array.  The call allocates nothing, host or device: the bin energies and observer means go
 * to a fixed device scratch array of 64 MiB per GPU, held from the first context on that GPU until
 * the process ends, and the band table travels as a kernel argument.  Limits: nBands <= 128 and
 * (nTime/2 + 1) * nObs * 8 bytes <= 64 MiB (e.g. 2048 steps x 8000 observers); beyond them the
 * call returns ACUM_ERR_ARG.  Calls that overlap on one GPU (from several contexts or threads)
 * race on the shared scratch; calls on one context are serialised.  Results are bit-identical
 * to acum_tonal_spectra's.                                                                    */
int acum_tonal_spectra_noalloc(acum_ctx* ctx, int64_t nTime, int64_t nObs, const float* pA, double dt,
                               int32_t nBands, const float* fc, float* spl, float* oaspl);

#ifdef __cplusplus
}
#endif
#endif /* ACUM_H */
