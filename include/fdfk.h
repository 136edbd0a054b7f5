/*
 * fdfk.h -- C ABI of the MI355X (gfx950) FDFK2D time-stepping library.
 *
 * Plain C types only (no torch, no HIP types).  Every function returns 0 on
 * success or a negative errno-style code; fdfk_last_error() gives the text of
 * the last failure on the calling thread.  A solver handle is not
 * thread-safe; separate handles may be used concurrently.
 *
 * Array layouts are the reference's Fortran ones (column-major, z fastest):
 *   c     (1:3, izt:nz+npml, 1-npml:nx+npml)   rho, lambda, mu       (FP32)
 *   ux/uz (1:2, izt:nz+npml, 1-npml:nx+npml)   level 1 = time n, level 2 = n-1
 *   seisx/seisz (0:nsis, 1:nrcv)               time fastest per receiver
 *   the twelve background arrays uxl_reg..uzb_pml with the shapes of
 *   src/FDFK2D.f90:330-355 (time slowest, 0:nt).
 * izt = 1 with a free surface (is_PML_top = .false.), izt = 1-npml with an
 * absorbing top layer (is_PML_top = .true., src/Input.f90:179-183).  The top
 * layer needs a hybrid ring across the top as well, four arrays beyond the
 * reference's twelve (fdfk_set_top_background), and a background without a
 * free surface (fdfk_compute_background makes both).
 */
#ifndef FDFK_H
#define FDFK_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FDFK_OK 0
#define FDFK_EINVAL (-22)
#define FDFK_ENOMEM (-12)
#define FDFK_EDEVICE (-5)
#define FDFK_ERANGE (-34)
#define FDFK_ENOSYS (-38)

/* storage precision of wavefield and material arrays (arithmetic is FP32,
 * FP64 for FDFK_FP64) */
#define FDFK_FP32 0
#define FDFK_BF16 1
#define FDFK_FP16 2
/* every array and all arithmetic in FP64 (per-cell kernels; the reference
 * of the FP32 and 16-bit storage modes at full size; no counterpart in the
 * reference, which is FP32) */
#define FDFK_FP64 3

/* motion of a solver: the in-plane P-SV motion of the reference (ux, uz), or
 * the transverse SH motion uy (this library only: an incident SH wave, one
 * component, on the general per-cell kernels; uy takes the place of ux in
 * the background arrays, fields and seismograms, uz stays zero) */
#define FDFK_PSV 0
#define FDFK_SH 1

typedef struct fdfk_solver fdfk_solver;

/* One simulation: the scalars of src/Input.f90:get_parameter plus the arrays
 * of read_data, i.e. the state simulate_wavefield (src/FDFK2D.f90:159) sets up. */
typedef struct fdfk_job {
    int32_t nx, nz, npml, norder;   /* norder = FD order / 2 (Input.f90:130-131) */
    int32_t is_pml_top;             /* 1: absorbing top layer, rows izt = 1-npml.. (FP32 only, not batched) */
    int32_t npower;                 /* PML profile power (FDFK2D.f90:417: 2) */
    float x0, z0, dx, dz;
    double dt;
    double R0;                      /* PML reflection coefficient (FDFK2D.f90:421) */
    float vpml;                     /* <= 0: max Vp of c (FDFK2D.f90:416) */
    int32_t nt;                     /* iterations it = 0..nt (FDFK2D.f90:311) */
    int32_t nstp;                   /* record every nstp iterations (FDFK2D.f90:313) */
    int32_t nrcv;
    const float *rx, *rz;           /* receiver coordinates (m) */
    const float *c;                 /* rho, lambda, mu (reference layout) */
    int32_t precision;              /* FDFK_FP32; FDFK_FP16 / FDFK_BF16: lam and mu stored for
                                       the interior tiles as per-row FP32 profiles plus FP16 /
                                       BF16 residuals with power-of-two scales (single solves,
                                       fused step; the wavefields stay FP32); FDFK_FP64:
                                       everything in FP64 (per-cell kernels, single solves) */
    int32_t wave;                   /* FDFK_PSV or FDFK_SH (FP32 or FP64, single solves) */
} fdfk_job;

/* FK background description: the inputs of comput_background_wavefield
 * (src/FDFK2D.f90:432-434). */
typedef struct fdfk_fk {
    int32_t nlayer_left, nlayer_right;
    const float *fk_left, *fk_right; /* (4, nlayer): ztop, vp, vs, rho per layer */
    int32_t src_is_p;                /* P (1) or SV (0) incidence */
    float ds, f0;                    /* source strength (m) and frequency (Hz) */
    double rayp;                     /* s/m (|rayp| >= 1e-9) */
    const double *bazs;              /* back-azimuth per padded column (rad), NX values */
    /* 0: keep the background of all nt+1 steps (the reference's twelve
     * arrays).  B > 0: keep the FP64 time traces and 2B steps of the arrays,
     * synthesising each next block of B steps in stream order as the run
     * advances (same values; the background of a 20001-step solve shrinks
     * from 17.9 GiB to the traces).  Not in the reference (it stores all
     * steps, src/FDFK2D.f90:320-355). */
    int32_t block_steps;
} fdfk_fk;

/* --- handle API: state stays resident on the GPU between calls --- */
/* replaces simulate_wavefield's allocation + setup (src/FDFK2D.f90:253-421) */
int fdfk_create(const fdfk_job *job, fdfk_solver **out);
/* upload a host background (twelve reference-layout arrays, nt+1 steps each) */
int fdfk_set_background(fdfk_solver *s, const float *const bg[12]);
/* compute the FK background on the GPU (FP64), replaces
 * comput_background_wavefield (src/FDFK2D.f90:432-434); with an absorbing
 * top layer the model has no free surface and the top ring is filled too */
int fdfk_compute_background(fdfk_solver *s, const fdfk_fk *fk);
/* the top ring of a solver with an absorbing top layer (no counterpart in
 * the reference, whose twelve arrays leave the hybrid contour open at the
 * top): uxt_reg, uzt_reg (nbd, nx-2nbd, 0:nt) on rows 1..nbd x columns
 * nbd+1..nx-nbd, uxt_pml, uzt_pml (nbd, nx+2nbd, 0:nt) on rows 1-nbd..0 x
 * columns 1-nbd..nx+nbd, z fastest.  Without it the contour stays open. */
int fdfk_set_top_background(fdfk_solver *s, const float *const top[4]);
int fdfk_get_top_background(fdfk_solver *s, float *const top[4]);
/* FDFK_FP64 solvers: the background in double (the FP32 setters widen),
 * seismograms and fields in double (the FP32 getters round); no counterpart
 * in the reference */
int fdfk_set_background64(fdfk_solver *s, const double *const bg[12]);
int fdfk_set_top_background64(fdfk_solver *s, const double *const top[4]);
int fdfk_seismograms64(fdfk_solver *s, double *seisx, double *seisz);
int fdfk_get_state64(fdfk_solver *s, double *ux, double *uz);
/* copy the device background back in the reference layout (twelve arrays);
 * FDFK_EINVAL when it is synthesised in blocks */
int fdfk_get_background(fdfk_solver *s, float *const bg[12]);
/* advance nsteps iterations of the loop src/FDFK2D.f90:465-532 (record,
 * then updatedispl); recording stays on the device */
int fdfk_run(fdfk_solver *s, int32_t nsteps);
/* next iteration index it (0..nt+1) */
int fdfk_iteration(const fdfk_solver *s);
/* seismograms (0:nsis, 1:nrcv), before the cos(baz) projection of
 * src/FDFK2D.f90:537-543 */
int fdfk_seismograms(fdfk_solver *s, float *seisx, float *seisz);
/* current two-level fields in the reference layout */
int fdfk_get_state(fdfk_solver *s, float *ux, float *uz);
int fdfk_set_state(fdfk_solver *s, const float *ux, const float *uz);
/* wait for all queued work of this handle */
int fdfk_sync(fdfk_solver *s);
/* without waiting: 1 when all queued work of this handle has completed, 0
 * while some is still queued or running (pacing of callers that queue
 * chunks of iterations; no counterpart in the reference) */
int fdfk_query(fdfk_solver *s);
/* kernel timing of the stencil launch of each step (HIP events on the
 * handle's stream): k_step, the interior and the frame in one launch (or
 * k_interior2 alone when the environment sets FDFK_FUSED=0); enable before
 * fdfk_run, read after fdfk_sync */
int fdfk_profile(fdfk_solver *s, int32_t enable);
int fdfk_kernel_stats(fdfk_solver *s, double *interior_ms_total, int64_t *interior_launches,
                      double *step_ms_total, int64_t *steps);
/* bytes of device memory held by the handle */
int64_t fdfk_device_bytes(const fdfk_solver *s);
void fdfk_destroy(fdfk_solver *s);
const char *fdfk_last_error(void);
/* fdfk_destroy keeps a solver's device blocks for later solvers instead of
 * freeing them (a free waits for the whole device, which would stall other
 * solvers and batches running meanwhile); this frees the kept blocks of the
 * current device and the idle inverse-FFT plans of the FK background (up to
 * 8 are kept for reuse).  An allocation that finds the device full does the
 * same and tries once more.  No counterpart in the reference. */
int fdfk_trim_memory(void);
/* Tile-width autotuner: builds a scratch solver of the job's geometry (zero
 * background, same kernels), times `steps` steps at interior tile widths
 * from 0.5x to 2x the occupancy model's choice, and keeps the fastest for
 * every later fdfk_create of that geometry in this process (FDFK_LX still
 * overrides).  Returns the chosen width and its ms per step, and the
 * model's.  No counterpart in the reference. */
int fdfk_autotune(const fdfk_job *job, int32_t steps, int32_t *best_lx, float *best_ms, int32_t *model_lx,
                  float *model_ms);

/* --- batched execution: many resident solvers advanced together with one
 * launch per kernel type per step (tile work-list over jobs).  Members keep
 * their own device state; they must share the FD order, and while they are
 * in a batch only fdfk_batch_run may advance them.  No counterpart in the
 * reference (one simulation per process there). --- */
typedef struct fdfk_batch fdfk_batch;
int fdfk_batch_create(int32_t norder, fdfk_batch **out);
/* the batch's later steps are ordered after the solver's queued work (e.g.
 * its background) on the GPU; neither call waits on the host */
int fdfk_batch_add(fdfk_batch *b, fdfk_solver *s);
/* later work on the solver's stream (fdfk_seismograms, fdfk_run,
 * fdfk_destroy) is ordered after the batch steps already queued */
int fdfk_batch_remove(fdfk_batch *b, fdfk_solver *s);
/* nsteps batched steps; members past their last iteration are skipped */
int fdfk_batch_run(fdfk_batch *b, int32_t nsteps);
/* number of members with iterations left */
int fdfk_batch_active(const fdfk_batch *b);
int fdfk_batch_sync(fdfk_batch *b);
/* Online tile-width tuner of the batched steps (no counterpart in the
 * reference, whose steps are not batched): with enable != 0, an active set
 * of a shape (the members' grids) not seen before by this batch runs its
 * first batched steps on candidate interior tile widths around the
 * occupancy model's, timing each over 4 steps after one untimed step, and
 * then keeps the fastest; a set of a known shape takes its tuned width at
 * once.  Results do not depend on the width.  Off by default: a server
 * whose active set changes every few steps would spend its steps tuning. */
int fdfk_batch_autotune(fdfk_batch *b, int32_t enable);
/* interior tile width of the batch's current tables (0 before its first step) */
int fdfk_batch_tile_width(const fdfk_batch *b);
int fdfk_batch_profile(fdfk_batch *b, int32_t enable);
int fdfk_batch_kernel_stats(fdfk_batch *b, double *interior_ms_total, int64_t *interior_launches,
                            double *step_ms_total, int64_t *steps);
void fdfk_batch_destroy(fdfk_batch *b);

/* --- model smoother: replaces the program utils/gauss_smooth.f90 (weights
 * :310-360, window sum :366-397).  vel and out hold nx traces of nz FP32
 * samples (one trace per column x, the SU model order) in host memory; out
 * may alias vel.  Rows iz0..iz1 of columns ix0..ix1 (1-based, inclusive) are
 * smoothed, the other cells keep vel's values.  method 0: Gaussian window,
 * 1: triangular window (equal half-widths in x and z only), both in the
 * reference's FP32 arithmetic; 2: the Gaussian as two 1-D FP64 passes over
 * the whole model (same result as 0 up to rounding, 2(2n+1) instead of
 * (2nx+1)(2nz+1) products per cell).  FDFK_EINVAL when a window half-width
 * reaches the model's extent (the mirror folds once). --- */
int fdfk_smooth_model(const float *vel, float *out, int32_t nx, int32_t nz, float lambdax,
                      float lambdaz, float dx, float dz, int32_t iz0, int32_t iz1, int32_t ix0,
                      int32_t ix1, int32_t method);
/* Allocation-free variant of fdfk_smooth_model on device buffers.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops).  vel and
 * out are device arrays of nx*nz FP32 values (out must not alias vel); the
 * work is queued on `stream` (a hipStream_t, NULL for the null stream) and
 * the call returns once the host-side weight tables are on the device.  The
 * scratch memory (weight tables and the intermediate model of method 2) is a
 * fixed 64 MiB array per GPU held for the life of the process: a method-2
 * model may hold up to about 16.7 M cells (the 8081 x 2040 padded model of
 * the large solve fits), larger requests fail with FDFK_EINVAL.  No device
 * memory is allocated or freed.  Calls that overlap on one GPU share the
 * scratch array and race. */
int fdfk_smooth_model_dev(const float *vel, float *out, int32_t nx, int32_t nz, float lambdax,
                          float lambdaz, float dx, float dz, int32_t iz0, int32_t iz1, int32_t ix0,
                          int32_t ix1, int32_t method, void *stream);

/* --- drop-in compatibility: the exact external procedures the reference
 * driver links (all arguments by reference, gfortran/flang mangling).  Device
 * mirrors are keyed by the host ux pointer; every call copies level 1/2 back
 * because the reference driver reads ux(1,...) on the host each iteration. --- */
/* replaces updatedispl (called at src/FDFK2D.f90:516-522) */
void updatedispl_(const int32_t *norder, const int32_t *nx, const int32_t *nz,
    const int32_t *npml, const int32_t *izt, const int32_t *is_pml_top, const int32_t *npower,
    const int32_t *nbd, const int32_t *it, const int32_t *nt, const float *dx, const float *dz,
    const double *dt, const float *vpml, const double *R0,
    const float *uxl_reg, const float *uzl_reg, const float *uxr_reg, const float *uzr_reg,
    const float *uxb_reg, const float *uzb_reg, const float *uxl_pml, const float *uzl_pml,
    const float *uxr_pml, const float *uzr_pml, const float *uxb_pml, const float *uzb_pml,
    const float *c, const double *coef0, const double *coef1, float *ux, float *uz,
    float *bpxlx, float *bpzlx, float *bpxrx, float *bpzrx, float *bpxtx, float *bpxtz,
    float *bpztx, float *bpztz, float *bpxbx, float *bpxbz, float *bpzbx, float *bpzbz,
    float *bux1l, float *bux2l, float *bux3l, float *buz1l, float *buz2l, float *buz3l,
    float *bux1r, float *bux2r, float *bux3r, float *buz1r, float *buz2r, float *buz3r,
    float *bux1b, float *bux2b, float *bux3b, float *buz1b, float *buz2b, float *buz3b,
    float *bux1t, float *bux2t, float *bux3t, float *buz1t, float *buz2t, float *buz3t);
/* replaces comput_background_wavefield (called at src/FDFK2D.f90:432-434) */
void comput_background_wavefield_(const int32_t *nx, const int32_t *nz, const int32_t *npml,
    const int32_t *izt, const int32_t *nbd, const int32_t *nt, const int32_t *nlayer_left,
    const int32_t *nlayer_right, const int32_t *src_type, const float *x0, const float *z0,
    const float *dx, const float *dz, const double *dt, const float *ds, const float *f0,
    const double *rayp, const double *bazs, const float *fk_left, const float *fk_right,
    float *uxl_reg, float *uzl_reg, float *uxr_reg, float *uzr_reg, float *uxb_reg,
    float *uzb_reg, float *uxl_pml, float *uzl_pml, float *uxr_pml, float *uzr_pml,
    float *uxb_pml, float *uzb_pml);
/* releases the compatibility mirrors (optional) */
void fdfk_compat_release(void);

#ifdef __cplusplus
}
#endif
#endif
