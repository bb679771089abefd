/*
 * orc.h -- the oracle: a plain-C CPU restatement of pywave's acoustic FD engine and
 * FWI gradient.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU-baseline
 * leg of bench.py may call it, as the checker; the product never does.
 *
 * Arrays are z-fastest ("Fortran order" of [nz,nx(,ny)]) as in the reference.
 */
#ifndef ORC_H
#define ORC_H

#ifdef __cplusplus
extern "C" {
#endif

/* kiss_fft_next_fast_size (pywave/src/wave_kissfft.c:412-424): smallest m >= n whose only
 * prime factors are 2, 3 and 5. */
int orc_next_fast_size(int n);

/* Padded extents used by fdm2d/fdm (pywave/src/wave_fft2.c:45-51, wave_fft3.c:45-60):
 * first axis 2*nfs((n+1)/2), the others nfs(n). */
void orc_padded_sizes(int nz, int nx, int ny, int *nz2, int *nx2, int *ny2);

/* abc_cal (pywave/src/wave_abc.c:127-145): w[i] = exp(-c*c*(nb-1-i)^2), f32 products,
 * exp in double. */
void orc_abc_weights(int nb, float c, float *w);

/* afd2d (pywave/src/afd.c:316-578 + pywave/src/wave_fdm.c:397-686).
 *   vel    : nz1*nx1 model velocity
 *   spx/spz: ns source grid indices (already truncated to int)
 *   rick   : ns*nt wavelets (row per source), used when tri==0
 *   dat    : nx1*nt receiver data [ix][it]; input when tri!=0, output otherwise
 *   snaps  : ntsnap*nz1*nx1 or NULL;  img: nz1*nx1 or NULL (tri only)
 * Returns 0. */
int orc_afd2d(const float *vel, int nz1, int nx1, int nt, float dt, float dx, float dz,
              int nbt, float ct, int ns, const int *spx, const int *spz, const float *rick,
              int tri, float *dat, int jsnap, float *snaps, float *img);

/* afd3d (pywave/src/afd.c:23-314 + pywave/src/wave_fdm.c:28-381).
 *   dat : nx1*ny1*nt receiver plane data [iy][ix][it]. */
int orc_afd3d(const float *vel, int nz1, int nx1, int ny1, int nt, float dt, float dx,
              float dy, float dz, int nbt, float ct, int ns, const int *spx, const int *spy,
              const int *spz, const float *rick, int tri, float *dat, int jsnap, float *snaps,
              float *img);

/* aps2d / aps3d: the same set-up and time loop with the pseudo-spectral Laplacian of
 * psm2d / psm (pywave/src/wave_psp.c:367-635, :25-365); arguments as orc_afd2d / orc_afd3d. */
int orc_aps2d(const float *vel, int nz1, int nx1, int nt, float dt, float dx, float dz,
              int nbt, float ct, int ns, const int *spx, const int *spz, const float *rick,
              int tri, float *dat, int jsnap, float *snaps, float *img);
int orc_aps3d(const float *vel, int nz1, int nx1, int ny1, int nt, float dt, float dx,
              float dy, float dz, int nbt, float ct, int ns, const int *spx, const int *spy,
              const int *spz, const float *rick, int tri, float *dat, int jsnap, float *snaps,
              float *img);

/* FP64 twin of orc_afd2d / orc_afd3d (modelling only): the same scheme in double precision from
 * the same f32 inputs (rick: ns*nt f32 wavelets).  is3d = 0 gives the 2D engine (ny1 = 1).
 * dat: ny1*nx1*nt [iy][ix][it]; snaps: ntsnap*ny1*nx1*nz1 or NULL. */
int orc_afd_f64(const float *vel, int nz1, int nx1, int ny1, int is3d, int nt, float dt, float dx,
                float dy, float dz, int nbt, float ct, int ns, const int *spx, const int *spy,
                const int *spz, const float *rick, double *dat, int jsnap, double *snaps);
/* its 4th-order Laplacian of a padded [ny2][nx2][nz2] field (ny2 = 1: 2D), zero on the rim */
void orc_lap_f64(const double *u, int nz2, int nx2, int ny2, double dz, double dx, double dy, double *out);

/* ---- transforms (pywave/src/wave_kissfft.c, wave_fft2.c, wave_fft3.c) ---- */

/* complex transform of n points (interleaved re, im); inverse = 1 for the positive exponent
 * (unscaled) */
int orc_cfft(int n, int inverse, const float *in, float *out);
/* real transform of even n: forward n reals -> n/2+1 bins; inverse n/2+1 bins -> n reals
 * (unscaled) */
int orc_rfft(int n, int inverse, const float *in, float *out);

/* the spectral Laplacian of a padded [n3][n2][n1] field (n3 = 1 in 2D) */
typedef struct orc_psl orc_psl;
orc_psl *orc_psl_new(int n1, int n2, int n3, float dz, float dx, float dy);
void orc_psl_free(orc_psl *P);
void orc_psl_apply(orc_psl *P, const float *in, float *out);

/* ---- FWI discretisation (pywave/src/wave_fwiutil.c, wave_fwimodeling.c,
 *      wave_fwigradient.c, wave_fwi.c, wave_fwilbfgs.c) ---- */

typedef struct {
    int nz, nx, nt, ns, nb;
    float dz, dx, dt, z0, x0, t0;
    float coef;
    int acqui_type;
    float ds, s0, r0, dr;
    int nr;
    int interval;
    /* derived by orc_fwi_prepare (wave_fwiutil.c:19-121) */
    int padnz, padnx, sz, rz, ds_v, s0_v, dr_v;
    int *r0_v, *r02, *nr2;  /* per shot, allocated by orc_fwi_prepare */
    float *bc;              /* nb sponge weights */
    /* passive path (pywave/src/wave_fwigradient.c:528-870, wave_fwi.c:212-322): when set, a
     * space-time source field [is][it][ix][iz] (nz*nx*nt*ns) replaces the point sources, all
     * shots record at r0_v[0] + ir*dr_v, ir < nr, and fwi resets the step every 5 iterations */
    const float *srcf;
    /* visco-acoustic path (gradient_v, pywave/src/wave_fwigradient.c:264-452; fwi() with media != 1):
     * when set, the quality factor q (nz*nx, z fastest) and the reference frequency f0 give tau and
     * taus (preparation, wave_fwiutil.c:91-97) and orc_gradient_av / orc_fwi follow gradient_v */
    const float *q;
    float f0;
} orc_acq;

int orc_fwi_prepare(orc_acq *a);
void orc_fwi_release(orc_acq *a);

/* forward_modeling_a: data[is][ir][it] (ns*nr*nt) */
int orc_fwi_forward(const orc_acq *a, const float *vel, const float *wav, float *data);

/* visco-acoustic forward_modeling (pywave/src/wave_fwimodeling.c:116-220): data[is][ir][it];
 * q: quality factor (nz*nx), f0: reference frequency of the tau model */
int orc_fwi_forward_visco(const orc_acq *a, const float *vel, const float *q, float f0, const float *wav,
                          float *data);

/* One gradient_av call.  'state' carries the first-call scaling across calls:
 * *scaling <= 0 on entry means "first call" (compute 0.1/max, store it). */
int orc_gradient_av(const orc_acq *a, const float *vel, const float *wav, const float *data,
                    float *grad, float *fcost, float *scaling);

/* fwi(): L-BFGS with Wolfe line search over orc_gradient_av; vinv is niter*nz*nx,
 * slots after an early stop are left untouched. Returns the number of stored iterates. */
int orc_fwi(const orc_acq *a, const float *vel0, const float *wav, const float *data,
            int niter, float *vinv);

/* gradient_av's f32 misfit accumulation over n residual samples in the reference's order
 * (shot, receiver, sample: the F order of data (nt, nr, ns)) */
float orc_fcost_f32(const float *d, const float *p, long long n);

/* Triangle smoothing pieces used by source_map / gradient_smooth2 */
void orc_source_map(int sx, int sz, int rectx, int rectz, int padnx, int padnz, float *rr);
void orc_gradient_smooth2(int rectx, int rectz, int nx, int nz, int waterz, float scaling,
                          float *grad);

/* ---- least-squares time-reversal imaging (pfwi mode 4 lstric: pywave/src/wave_triutil.c,
 *      wave_fwi.c:121-210, wave_fwigradient.c:454-525, wave_bigsolver.c, wave_cgstep.c) ---- */

/* timerev_lop: adj = 0 models dat[ix][it] (nx*nt) from mod[it][ix][iz] (nt*nx*nz); adj = 1 the
 * adjoint; add = 0 clears the output first.  Receivers at every column, one cell below the top
 * of the model; decay boundary with coefficient coef over nb cells. */
int orc_timerev(int nz, int nx, int nt, int nb, float dt, float dx, float dz, float coef,
                const float *vel, int adj, int add, float *mod, float *dat);

/* lstri over ns shots: inv = 0 models data[is][ix][it] from src[is][it][ix][iz]; inv = 1 solves
 * for src with niter conjugate-gradient iterations per shot (the step state carries from one
 * shot to the next, restarting at the first shot). */
int orc_lstri(int nz, int nx, int nt, int ns, int nb, float dt, float dx, float dz, float coef,
              const float *vel, int inv, int niter, float *data, float *src);

#ifdef __cplusplus
}
#endif
#endif
