/*
 * pywave_amd.h -- C ABI of the MI355X (gfx950) acoustic finite-difference engine.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a negative
 * PW_ERR_* code; pw_last_error() then holds a message.  Host arrays are caller-owned.
 * Arrays follow the reference's layouts: models are z-fastest ("Fortran order" of
 * [nz,nx(,ny)]), receiver data are time-fastest.
 *
 * Reference interfaces replaced (pywave = github.com/chenyk1990/pywave):
 *   afdcfun.afd2dc            pywave/src/afd.c:316-578   (module table afd.c:585-607)
 *   afdcfun.afd3dc            pywave/src/afd.c:23-314
 *   pfwicfun.forward_modeling_ac  pywave/src/pfwi.c:27-495
 *   pfwicfun.fwic             pywave/src/pfwi.c:497-1028
 */
#ifndef PYWAVE_AMD_H
#define PYWAVE_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PW_OK 0
#define PW_ERR_ARG -1      /* invalid argument / dimension mismatch */
#define PW_ERR_HIP -2      /* HIP runtime error */
#define PW_ERR_NOMEM -3    /* device or host allocation failed */
#define PW_ERR_STATE -4    /* call order (e.g. run before velocity was set) */
#define PW_ERR_NONFINITE -5 /* a result (traces, wavefield, image, gradient, misfit) holds Inf or NaN:
                               the run blew up, usually a time step beyond the CFL limit */

typedef struct pw_ctx pw_ctx; /* one device + one stream; owns no global state */
typedef struct pw_afd pw_afd; /* one afd modeling / time-reversal problem resident in HBM */

/* ---------------------------------------------------------------- context */
pw_ctx *pw_create(int device);
void pw_destroy(pw_ctx *ctx);
const char *pw_last_error(pw_ctx *ctx);
int pw_device_count(void);
/* Run subsequent work on a caller stream (hipStream_t passed as void*; NULL is the HIP null
   stream, which is what torch reports for its default stream). */
int pw_set_stream(pw_ctx *ctx, void *stream);
/* Return to the context's own (non-blocking) stream. */
int pw_use_own_stream(pw_ctx *ctx);
int pw_synchronize(pw_ctx *ctx);
/* streaming bandwidth of the device, for reading the step kernels' rooflines: `reads` input
 * streams (1-3) and one output stream of n floats each (n a positive multiple of 4), reps timed
 * launches on the context's stream; cached 0 uses non-temporal loads and stores, cached 1 the
 * default policy (buffers that together fit the Infinity Cache then stay in it between launches);
 * *gbs = bytes moved / s / 1e9.  Allocates (reads + 1) * n * 4 bytes for the call.  No counterpart
 * in the reference. */
int pw_stream_bandwidth(pw_ctx *ctx, long long n, int reads, int reps, int cached, double *gbs);

/* ---------------------------------------------------------------- host setup helpers */
/* kiss_fft_next_fast_size (pywave/src/wave_kissfft.c:412-424) */
int pw_next_fast_size(int n);
/* fdm/fdm2d padded extents (pywave/src/wave_fft3.c:45-60, wave_fft2.c:45-51); ny2 = 1 for ny <= 1 */
void pw_padded_sizes(int nz, int nx, int ny, int *nz2, int *nx2, int *ny2);
/* abc_cal weights (pywave/src/wave_abc.c:127-145) */
void pw_abc_weights(int nb, float c, float *w);
/* the zero-phase Ricker fdm/fdm2d build per source (pywave/src/wave_fdm.c:502-516) */
int pw_ricker(int nt, float dt, float f0, float t0, float A, float *out);

/* ---------------------------------------------------------------- afd: one-shot drop-ins */
/* afdcfun.afd2dc (pywave/src/afd.c:316-578).  Same argument meaning and order; vel is nz*nx
 * z-fastest; src is [sx(ns) sz(ns) f(ns) t(ns) A(ns)] or, when tri != 0, the nt*nx receiver data
 * (time fastest).  out receives exactly the reference's return vector:
 * nt*nx traces (time fastest) followed by nz*nx*ntsnap snapshots when jsnap > 0.
 * verb, ifsnaps, abc, ox, oz are accepted and ignored as in the reference (abc is forced on). */
int pw_afd2d(pw_ctx *ctx, const float *vel, const float *src, int tri, int nt, int nx, int nz,
             int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct, float dt,
             float ox, float dx, float oz, float dz, float *out, long long out_len);

/* Allocation-free 2D modelling (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same run as pw_afd2d's modelling (tri = 0, jsnap = 0) on device
 * buffers: d_vel nz*nx (z fastest), src_xz host [sx.., sz..] (2*ns grid indices, as afd2dc
 * truncates them), d_wav ns*nt wavelets (row per source, e.g. from pw_ricker), d_traces nx*nt
 * (the reference's [ix][it]).  Every field of the run lives in one file-scope device array of
 * pw_afd2d_noalloc_capacity() bytes (64 MiB) per GPU, held until the process ends; a run that
 * needs more returns PW_ERR_ARG.  Calls by any thread on one GPU share that array: calls that
 * overlap in time on one GPU race.  Enqueued on the context's stream; nothing is allocated,
 * freed or synchronised. */
size_t pw_afd2d_noalloc_capacity(void);
int pw_afd2d_dev_noalloc(pw_ctx *ctx, const float *d_vel, int nz, int nx, float dz, float dx, int nt,
                         float dt, int nbt, float ct, int ns, const int *src_xz, const float *d_wav,
                         float *d_traces);

/* afdcfun.afd3dc (pywave/src/afd.c:23-314); src is [sx sy sz f t A] (6*ns) or nt*nx*ny data. */
int pw_afd3d(pw_ctx *ctx, const float *vel, const float *src, int tri, int nt, int nx, int ny,
             int nz, int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct,
             float dt, float ox, float dx, float oy, float dy, float oz, float dz, float *out,
             long long out_len);

/* apscfun.aps2dc / aps3dc (pywave/src/aps.c; pywave/aps2d.py:64, pywave/aps3d.py): the afd
 * arguments and return with the pseudo-spectral engine psm2d / psm (pywave/src/wave_psp.c:367,
 * :25), whose Laplacian is IFFT(-(2 pi |k|)^2 FFT(u)) over the padded grid (the reference's KISS
 * FFT arithmetic reproduced on the device). */
int pw_aps2d(pw_ctx *ctx, const float *vel, const float *src, int tri, int nt, int nx, int nz,
             int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct, float dt,
             float ox, float dx, float oz, float dz, float *out, long long out_len);
int pw_aps3d(pw_ctx *ctx, const float *vel, const float *src, int tri, int nt, int nx, int ny,
             int nz, int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct,
             float dt, float ox, float dx, float oy, float dy, float oz, float dz, float *out,
             long long out_len);

/* length of the afd return vector: nt*nx*ny + nz*nx*ny*ntsnap (ny = 1 in 2D) */
long long pw_afd_output_len(int nt, int nz, int nx, int ny, int jsnap);

/* ---------------------------------------------------------------- afd: resident plans
 * The same engine split into set-up, run and read-back so that models stay in HBM and the
 * step loop can be timed on its own.  ndim is 2 or 3 (ny, dy ignored in 2D).
 * Sources: src_pack as in pw_afd2d/pw_afd3d; wavelet (ns*nt, row per source) overrides the
 * built-in Ricker when non-NULL. */
pw_afd *pw_afd_create(pw_ctx *ctx, int ndim, int nz, int nx, int ny, float dz, float dx,
                      float dy, int nt, float dt, int nbt, float ct, int jsnap);
int pw_afd_set_velocity(pw_afd *p, const float *vel);
int pw_afd_set_sources(pw_afd *p, int ns, const float *src_pack, const float *wavelet);
int pw_afd_set_tri_data(pw_afd *p, const float *dat);
/* Enqueue the whole time loop (forward modelling, or backward time-reversal when tri != 0). */
int pw_afd_run(pw_afd *p, int tri);
/* As pw_afd_run, bracketed by HIP events on the launch stream: total milliseconds and the
 * average duration of one step kernel (ms), measured by events around every step launch
 * when per_step != 0 (adds event overhead), else total/nt. Synchronises. */
int pw_afd_run_timed(pw_afd *p, int tri, int per_step, double *total_ms, double *step_ms);
/* Copy the result in the reference's return layout (see pw_afd2d). Synchronises. */
int pw_afd_output(pw_afd *p, float *out, long long out_len);
/* Final wavefield interior after a run (nz*nx*ny, z fastest): the reference's 'img'. */
int pw_afd_final_field(pw_afd *p, float *out, long long out_len);
/* Kernel tile choice for the 3D step (0 = default); returns PW_ERR_ARG for unknown ids. */
int pw_afd_set_variant(pw_afd *p, int variant, int ny_chunk);
/* on != 0: the plan's time loop uses the pseudo-spectral Laplacian of psm2d/psm
 * (pywave/src/wave_psp.c) instead of the finite-difference stencil (set before running) */
int pw_afd_set_spectral(pw_afd *p, int on);
int pw_afd_variant_count(void);
/* info[4] = {rows per block (step-1 rows for two-step kernels), waves per block, kind (0 general,
 * 1 lean, 2 pipelined, 3 pipelined non-temporal, 4 pipelined x-major, 5 two steps per launch),
 * time steps per launch}. No counterpart in the reference (tooling for the kernel choice). */
int pw_afd_variant_info(int variant, int *info);
/* Autotuner: times each 3D kernel variant and y-chunk on `steps` steps of this plan, keeps the
 * fastest (cached per padded grid shape in the context); no-op for 2D plans. */
int pw_afd_autotune(pw_afd *p, int steps, int *variant_out, int *chunk_out);
void pw_afd_destroy(pw_afd *p);

/* ---------------------------------------------------------------- FWI (pfwi discretisation)
 * pararray: the 20 floats of pywave/pfwi.py:86-107 (nz,nx,dz,dx,oz,ox,nt,dt,ot,inv,ns,ds,os,sz,
 * nb,coef,f0,acqui_type,interval,niter); as in pfwic (pywave/src/pfwi.c:796-806) sz=5, rz=1,
 * nr=nx, dr=dx, r0=ox are fixed.  batch = shots advanced per launch (<= 0: all);
 * precision = 4 (f32, reference arithmetic) or 8 (f64, for finite-difference checks).
 * Data arrays are (nt, nr, ns) Fortran order unless noted; device variants use [ns][nt][nr]. */
typedef struct pw_fwi pw_fwi;
pw_fwi *pw_fwi_create(pw_ctx *ctx, const float *pararray, int batch, int precision);
void pw_fwi_destroy(pw_fwi *p);
/* checkpoint segment length in steps (0 = one segment when it fits half the free HBM, else 50) */
int pw_fwi_set_checkpointing(pw_fwi *p, int steps);
/* padnz, padnx, batch, K, nseg, ns, nt, nr */
int pw_fwi_info(pw_fwi *p, int *ints8);
int pw_fwi_set_wavelet(pw_fwi *p, const float *wav);
int pw_fwi_set_observed(pw_fwi *p, const float *data);
/* passive source field (pfwi mode 4: gradient_pas_init / gradient_pas_av,
 * pywave/src/wave_fwigradient.c:528-870): src (nz,nx,nt,ns) Fortran order, added to the Laplacian
 * on the model interior at every step in place of the point sources; every shot then records and
 * injects at r0_v[0] + ir*dr_v (ir < nr).  NULL switches back to the point sources.  Applies to
 * forward, gradient (both modes), device variants and L-BFGS (which then resets its trial step to
 * 1 every 5 iterations as pfwi does, pywave/src/wave_fwi.c:282). */
int pw_fwi_set_source_field(pw_fwi *p, const float *src);
/* the same from device memory of the plan's precision, [ns][nt][nx][nz] */
int pw_fwi_set_source_field_dev(pw_fwi *p, const void *d_src);
/* visco-acoustic plan: memory variable with tau/taus from q (nz,nx Fortran) and the pararray's f0
 * (preparation, pywave/src/wave_fwiutil.c:91-97).  pw_fwi_forward then follows forward_modeling
 * (pywave/src/wave_fwimodeling.c:116-220); pw_fwi_gradient mode 0, pw_fwi_misfit_f64 and
 * pw_fwi_lbfgs follow gradient_v (pywave/src/wave_fwigradient.c:264-452), the gradient fwi()
 * (pywave/src/wave_fwi.c:40-41) takes for media != 1.  f32 plans, point sources only; the exact
 * gradient (mode 1) and the device entry points refuse a visco plan.  NULL returns to acoustic. */
int pw_fwi_set_visco(pw_fwi *p, const float *q);
/* forward_modeling_a (pywave/src/wave_fwimodeling.c:21-114) */
int pw_fwi_forward(pw_fwi *p, const float *vel, float *data_out);
/* one gradient over all shots: mode 0 = gradient_av (pywave/src/wave_fwigradient.c:93-262:
 * imaging condition, windowed residual, first-call 0.1/max scaling (scaling_io <= 0 on input ->
 * computed; NULL -> remembered by the plan), triangle smoothing); mode 1 = exact adjoint of
 * 0.5*||F(v)-d||^2.  fcost = 0.5*sum (d-F(v))^2 reduced in f64. */
int pw_fwi_gradient(pw_fwi *p, const float *vel, int mode, float *grad, double *fcost, float *scaling_io);
/* f64 plans only: exact-adjoint gradient / misfit with f64 model and output */
int pw_fwi_gradient_f64(pw_fwi *p, const double *vel, int mode, double *grad, double *fcost);
int pw_fwi_misfit_f64(pw_fwi *p, const double *vel, double *fcost);
/* device pointers (element type = plan precision) on the context's stream: model (nz*nx),
 * traces / cotangent [ns][nt][nr], gradient (nz*nx).  When one batch holds every shot,
 * pw_fwi_forward_dev keeps its checkpoints, and the next pw_fwi_adjoint_dev reuses them if it is
 * given the same model pointer holding the same values (a 64-bit fingerprint of the model is
 * compared, so a model changed in place between the two calls is swept forward again). */
int pw_fwi_forward_dev(pw_fwi *p, const void *d_vel, void *d_traces);
int pw_fwi_adjoint_dev(pw_fwi *p, const void *d_vel, const void *d_gtraces, void *d_gvel);
/* tile widths (x columns per wave) of the step kernels: w[0] forward, w[1] visco-acoustic forward,
 * w[2] exact adjoint, w[3] reference adjoint, w[4] visco-acoustic reference adjoint; w[5] = 1 once
 * the plan has timed the widths on its grid and batch (its first sweep does), 0 before; w[6] the
 * two-step forward's width, 0 when forward sweeps run one step per launch; w[7] the same for the
 * exact adjoint; w[8] the width of forward steps that store their term for the exact adjoint
 * (9 ints) */
int pw_fwi_widths(pw_fwi *p, int *w);
/* forward time steps one launch may advance: 2 (default; the plan's timing sweep keeps two steps
 * per launch where they are faster), 1 (one kernel per time step, the plain method) or -2 (two
 * steps per launch whatever the timing; acoustic point-source plans) */
int pw_fwi_set_steps_per_launch(pw_fwi *p, int n);
/* force the step kernels' tile widths instead of timing them: w[0] forward, w[1] visco forward,
 * w[2] exact adjoint, w[3] reference adjoint, w[4] visco reference adjoint, each 4, 8 or 16 (one
 * step per launch; forward steps that store their term take w[0]); NULL returns to the timing sweep.  Identical results at every width except the
 * exact adjoint's FMA contraction (f32 rounding).  No counterpart in the reference. */
int pw_fwi_set_widths(pw_fwi *p, const int *w);
/* the plan's last width-timing sweep: triples (kernel, width, ms per launch) written to out (up to
 * cap of them); kernel 0 forward, 1 visco forward, 2 exact adjoint, 3 reference adjoint, 4 visco
 * reference adjoint, 5 two-step forward, 6 two-step exact adjoint, 7 forward storing its term.
 * Returns the number of triples. */
int pw_fwi_tuning_log(pw_fwi *p, float *out, int cap);
/* event-timed milliseconds of the last gradient (forward sweeps, backward sweeps) or of the last
 * pw_fwi_forward (its sweeps, 0) */
int pw_fwi_timing(pw_fwi *p, double *forward_ms, double *backward_ms);
/* fwi() L-BFGS (pywave/src/wave_fwi.c:24-119, wave_fwilbfgs.c:9-165) with f64 dot products;
 * vinv: niter*nz*nx, *nstored = iterates written (the rest untouched) */
int pw_fwi_lbfgs(pw_fwi *p, const float *vel0, int niter, float *vinv, int *nstored);

/* pfwicfun.forward_modeling_ac (pywave/src/pfwi.c:27-495): out = nt*nx*ns, (nt,nx,ns) Fortran */
int pw_forward_modeling_ac(pw_ctx *ctx, const float *vel, const float *q, const float *wav,
                           const float *pararray, float *out, long long out_len);
/* the forward_modelingc that pywave/pfwi.py:72 calls for media=2 (undefined in the reference's
 * extension): visco-acoustic forward_modeling, arguments and output as pw_forward_modeling_ac */
int pw_forward_modeling_c(pw_ctx *ctx, const float *vel, const float *q, const float *wav,
                          const float *pararray, float *out, long long out_len);
/* pfwicfun.fwic (pywave/src/pfwi.c:497-1028): out = nz*nx*niter, (nz,nx,niter) Fortran */
int pw_fwic(pw_ctx *ctx, const float *vel, const float *q, const float *wav, const float *data,
            const float *pararray, float *out, long long out_len);
/* pfwicfun.pfwic (pywave/src/pfwi.c:1507-2034; called by pywave/pfwi.py:196 for mode 4 with
 * inv and not onlysrc): passive FWI for velocity with a known source field.  datasrc = data
 * (nt*nx*ns) followed by src (nz*nx*nt*ns) when pararray[9] (inv) != 0, else src alone (zero
 * data); datasrc_len its length in floats.  out = nz*nx*niter, (nz,nx,niter) Fortran. */
int pw_pfwic(pw_ctx *ctx, const float *vel, const float *q, const float *wav, const float *datasrc,
             const float *pararray, float *out, long long out_len, long long datasrc_len);

/* ---------------------------------------------------------------- passive modelling and LS-TRI
 * pfwi mode 4's lstric (pywave/src/pfwi.c:1031-1505 -> lstri, pywave/src/wave_fwi.c:121-210 ->
 * lstri_op, pywave/src/wave_fwigradient.c:454-525): the time-reversal operator timerev_lop
 * (pywave/src/wave_triutil.c:48-180) maps a space-time source field (nz,nx,nt) per shot to data
 * (nt,nx) recorded at every column one cell below the top of the model; its adjoint maps data to
 * an image field; the inversion is the reference's conjugate-gradient solver (np_solver with
 * np_cgstep, pywave/src/wave_bigsolver.c:1170-1450, wave_cgstep.c:17-67).  Geometry from the
 * 20-float pfwi pararray (nz, nx, dz, dx, nt, dt, ns, nb, coef; niter for pw_lstric). */
typedef struct pw_lstri pw_lstri;
pw_lstri *pw_lstri_create(pw_ctx *ctx, const float *pararray);
void pw_lstri_destroy(pw_lstri *p);
/* vel (nz,nx) Fortran order */
int pw_lstri_set_model(pw_lstri *p, const float *vel);
/* modelling of every shot (all shots advance in one launch per step):
 * src (nz,nx,nt,ns) Fortran -> data (nt,nx,ns) Fortran */
int pw_lstri_model(pw_lstri *p, const float *src, float *data);
/* adjoint (time-reversal image) of every shot: data (nt,nx,ns) -> field (nz,nx,nt,ns) */
int pw_lstri_image(pw_lstri *p, const float *data, float *img);
/* niter CG iterations per shot from a zero field; the CG step vectors carry over from one shot to
 * the next as the reference's np_cgstep state does (restart at the first shot of the call).
 * iters_out: ns ints or NULL (iterations run per shot). */
int pw_lstri_invert(pw_lstri *p, const float *data, int niter, float *src, int *iters_out);
/* on != 0 (default): the inversion replays its two sweeps from hipGraphs captured on the first
 * call; 0 launches them step by step.  Results are identical.  No counterpart in the reference. */
int pw_lstri_use_graph(pw_lstri *p, int on);
/* time steps per sweep launch: 1, 2, 4, or 0 for the plan's choice (timed before its first sweep
 * of each shot count: single-step tile widths 1-16, two-step launches and four-step launches of
 * 4- and 8-column tiles on the plan's own grid); identical results either way.  No counterpart in the reference (a scheduling choice of this
 * engine). */
int pw_lstri_steps_per_launch(pw_lstri *p, int n);
/* single-step tile width (x columns per 64-cell tile): 1, 2, 4, 8, 16, or 0 for the plan's timed
 * choice; identical results either way.  No counterpart in the reference. */
int pw_lstri_set_tile_width(pw_lstri *p, int w);
/* output columns of the four-step launches' tiles: 4, 8, or 0 for the timed choice (identical
 * results).  No counterpart in the reference. */
int pw_lstri_set_four_step_columns(pw_lstri *p, int c);
/* the last sweep's launch choice: w[0] single-step tile width, w[1] steps per launch, w[2] the
 * four-step tiles' output columns (3 ints); the timing sweep's (kernel, width, ms per launch)
 * triples into log[3*cap] (kernel 0 modelling step, 1 adjoint step, 2 two-step modelling launch,
 * 3 two-step adjoint launch, 4 four-step modelling launch, 5 four-step adjoint launch); returns
 * their number.  No counterpart in the reference. */
int pw_lstri_launch_choice(pw_lstri *p, int *w, float *log, int cap);
/* device milliseconds of the last pw_lstri_model / pw_lstri_invert call */
int pw_lstri_timing(pw_lstri *p, double *ms);
/* pfwicfun.lstric (called by pywave/pfwi.py:160 and :247): pararray[9] (inv) != 0 -> datasrc =
 * data (nt*nx*ns), out = source field (nz*nx*nt*ns) after pararray[19] CG iterations per shot;
 * else datasrc = source field, out = data.  in_len / out_len in floats. */
int pw_lstric(pw_ctx *ctx, const float *vel, const float *q, const float *wav, const float *datasrc,
              const float *pararray, float *out, long long out_len, long long in_len);

/* ---------------------------------------------------------------- time-reversal imaging of groups
 * Back-propagates ngroups (1..64) receiver groups of one data set, up to 4 together (one launch per
 * step, the model read once for those groups), and accumulates the source-location image
 * sum over snapshots (it % jsnap == 0) of prod_g u_g^2 in f64 -- the image the reference's demo
 * forms from per-group afd3d(tri=True) runs (demos/test_third_tri.py:57-77).  With more than 4
 * groups the groups run in passes of 4 and the plan keeps each snapshot's running product in HBM
 * (8 bytes per model cell and snapshot). */
typedef struct pw_tri pw_tri;
pw_tri *pw_tri_create(pw_ctx *ctx, int nz, int nx, int ny, float dz, float dx, float dy, int nt,
                      float dt, int nbt, float ct, int ngroups, int jsnap);
int pw_tri_set_velocity(pw_tri *p, const float *vel);
/* dat: (nt, nx, ny) Fortran order; masks: ngroups*nx*ny bytes [g][iy][ix], non-zero = member */
int pw_tri_set_data(pw_tri *p, const float *dat, const unsigned char *masks);
int pw_tri_run(pw_tri *p);
int pw_tri_run_timed(pw_tri *p, double *total_ms);
/* launch shape of the back-propagation kernel, timed before the first run (rows per block 4, 8 or
 * 16; y chunks of 4 to 24 planes; default or non-temporal cache policy on the streamed operands):
 * w[0] rows per block, w[1] y-chunk planes, w[2] non-temporal (3 ints); the timing sweep's (rows,
 * chunk, non-temporal, ms per launch) quadruples into log[4*cap]; returns their number.  No
 * counterpart in the reference. */
int pw_tri_launch_choice(pw_tri *p, int *w, float *log, int cap);
/* force the launch shape (rows 4, 8 or 16, y chunk >= 1, non_temporal 0 or 1; identical results),
 * or rows = 0 for the timed choice.  No counterpart in the reference. */
int pw_tri_set_launch(pw_tri *p, int rows, int y_chunk, int non_temporal);
/* unnormalised f64 image, nz*nx*ny z fastest */
int pw_tri_image(pw_tri *p, double *img, long long len);
void pw_tri_destroy(pw_tri *p);

/* ---------------------------------------------------------------- batched 2D modelling
 * n shots of uneven size and length (each one afd2d modelling problem, pywave/src/afd.c:316 ->
 * pywave/src/wave_fdm.c:397) advanced together: one launch per time step covers every running
 * shot; with use_graph the step loop is captured once into a hipGraph and replayed.
 * dims: n*(nz, nx, nt); d: n*(dz, dx, dt); nbt, ct shared. */
typedef struct pw_batch pw_batch;
pw_batch *pw_batch2d_create(pw_ctx *ctx, int n, const int *dims, const float *d, int nbt, float ct);
/* shot i: vel nz*nx (z fastest), src_pack [sx sz f t A] (5*ns), wavelet ns*nt or NULL */
int pw_batch2d_set_shot(pw_batch *p, int i, const float *vel, int ns, const float *src_pack,
                        const float *wavelet);
int pw_batch2d_run(pw_batch *p, int use_graph, double *ms);
/* traces of shot i, (nt, nx) Fortran order (afd2d's dout) */
int pw_batch2d_traces(pw_batch *p, int i, float *out, long long len);
long long pw_batch2d_cell_updates(pw_batch *p);
/* tile width (x columns per 64-cell tile): 0 (default) times widths 2, 4, 8 and 16 on the batch
 * before its next run and keeps the fastest (once per batch geometry and context); 2, 4, 8 or 16
 * uses that width.  Results are bit-identical at every width.  No counterpart in the reference. */
int pw_batch2d_set_width(pw_batch *p, int xc);
/* the width in use (returned); the timing sweep's (width, ms per launch) pairs into log[2*cap]
 * and their number into *nlog (either may be NULL) */
int pw_batch2d_width(pw_batch *p, float *log, int cap, int *nlog);
void pw_batch2d_destroy(pw_batch *p);

#ifdef __cplusplus
}
#endif
#endif
