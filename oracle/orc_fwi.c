/*
 * orc_fwi.c -- oracle restatement of pywave's FWI discretisation (TEST INFRASTRUCTURE ONLY).
 *
 *   preparation        pywave/src/wave_fwiutil.c:19-121
 *   pad2d              pywave/src/wave_fwiutil.c:123-147  (clamp-to-edge replication)
 *   source_map         pywave/src/wave_fwiutil.c:149-177
 *   laplace (8th)      pywave/src/wave_fwiutil.c:17,179-199
 *   apply_sponge       pywave/src/wave_fwiutil.c:201-223
 *   residual_weighting pywave/src/wave_fwiutil.c:225-296  (oreo = false branch)
 *   gradient_smooth2   pywave/src/wave_fwiutil.c:298-335
 *   triangle smoothing pywave/src/wave_triangle.c:18-41,71-97,121-142,177-193,229-238
 *   forward_modeling_a pywave/src/wave_fwimodeling.c:21-114
 *   gradient_av        pywave/src/wave_fwigradient.c:93-262
 *   forward_modeling   pywave/src/wave_fwimodeling.c:116-220  (visco-acoustic)
 *   gradient_v         pywave/src/wave_fwigradient.c:264-452  (visco-acoustic, fwi() with media != 1)
 *   fwi + L-BFGS       pywave/src/wave_fwi.c:24-119, pywave/src/wave_fwilbfgs.c:9-165
 * Per-value arithmetic (float vs double sub-expressions, association, accumulation order)
 * follows the reference so that the results agree bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static const float K0 = -205. / 72, K1 = 8. / 5, K2 = -1. / 5, K3 = 8. / 315, K4 = -1. / 560;

int orc_fwi_prepare(orc_acq *a)
{
    int nb = a->nb;
    a->bc = malloc(sizeof(float) * (nb > 0 ? nb : 1));
    for (int i = 0; i < nb; i++) {
        float tmp = a->coef * (nb - i);
        a->bc[i] = expf(-tmp * tmp);
    }
    a->padnx = a->nx + 2 * nb;
    a->padnz = a->nz + 2 * nb;
    a->ds_v = a->ds / a->dx + 0.5;
    a->s0_v = a->s0 / a->dx + 0.5 + nb;
    a->sz = 5 + nb;  /* pfwi.c hard-codes sz = 5 before preparation adds nb */
    a->dr_v = a->dr / a->dx + 0.5;
    a->rz = 1 + nb;  /* pfwi.c hard-codes rz = 1 */
    a->r0_v = malloc(sizeof(int) * a->ns);
    a->r02 = malloc(sizeof(int) * a->ns);
    a->nr2 = malloc(sizeof(int) * a->ns);
    float xend = a->x0 + (a->nx - 1) * a->dx;
    for (int i = 0; i < a->ns; i++) {
        if (a->acqui_type == 1) {
            a->r0_v[i] = (a->r0 - a->x0) / a->dx + 0.5 + nb;
            a->r02[i] = 0;
            a->nr2[i] = a->nr;
        } else {
            float sx = a->s0 + a->ds * i;
            float rbegin = (sx + a->r0 < a->x0) ? a->x0 : sx + a->r0;
            float rend = sx + a->r0 + (a->nr - 1) * a->dr;
            rend = (rend < xend) ? rend : xend;
            a->r0_v[i] = rbegin / a->dx + 0.5 + nb;
            a->r02[i] = (rbegin - sx - a->r0) / a->dx + 0.5;
            a->nr2[i] = (rend - rbegin) / a->dx + 1.5;
        }
    }
    return 0;
}

void orc_fwi_release(orc_acq *a)
{
    free(a->bc); free(a->r0_v); free(a->r02); free(a->nr2);
    a->bc = NULL; a->r0_v = a->r02 = a->nr2 = NULL;
}

/* ---------------- triangle smoothing (adjoint form, np_smooth2) ---------------- */

static void tri_smooth2_line(int o, int d, int nx, int nb, float *x, float *tmp)
{
    int np = nx + 2 * nb;
    float wt = 1.0 / (nb * nb);
    /* triple2: -wt at 0, 2wt at nb, -wt at 2nb; saxpy's alpha is a float */
    for (int i = 0; i < np; i++) tmp[i] = 0;
    float a1 = -wt, a2 = 2. * wt, a3 = -wt;
    for (int i = 0; i < nx; i++) tmp[i] += a1 * x[o + i * d];
    for (int i = 0; i < nx; i++) tmp[nb + i] += a2 * x[o + i * d];
    for (int i = 0; i < nx; i++) tmp[2 * nb + i] += a3 * x[o + i * d];
    /* doubint2: forward then backward running sums */
    float t = 0.;
    for (int i = 0; i < np; i++) { t += tmp[i]; tmp[i] = t; }
    t = 0.;
    for (int i = np - 1; i >= 0; i--) { t += tmp[i]; tmp[i] = t; }
    /* fold2: middle copy, then reflections accumulated from the right, then the left */
    for (int i = 0; i < nx; i++) x[o + i * d] = tmp[i + nb];
    for (int j = nb + nx; j < np; j += nx) {
        for (int i = 0; i < nx && i < np - j; i++) x[o + (nx - 1 - i) * d] += tmp[j + i];
        j += nx;
        for (int i = 0; i < nx && i < np - j; i++) x[o + i * d] += tmp[j + i];
    }
    for (int j = nb; j >= 0; j -= nx) {
        for (int i = 0; i < nx && i < j; i++) x[o + i * d] += tmp[j - 1 - i];
        j -= nx;
        for (int i = 0; i < nx && i < j; i++) x[o + (nx - 1 - i) * d] += tmp[j - 1 - i];
    }
}

/* smooth a z-fastest n0 x n1 field along axis 0 (rect0) then axis 1 (rect1) */
static void smooth2d(float *f, int n0, int n1, int rect0, int rect1)
{
    if (rect0 > 1) {
        float *tmp = malloc(sizeof(float) * (n0 + 2 * rect0));
        for (int j = 0; j < n1; j++) tri_smooth2_line(j * n0, 1, n0, rect0, f, tmp);
        free(tmp);
    }
    if (rect1 > 1) {
        float *tmp = malloc(sizeof(float) * (n1 + 2 * rect1));
        for (int j = 0; j < n0; j++) tri_smooth2_line(j, n0, n1, rect1, f, tmp);
        free(tmp);
    }
}

void orc_source_map(int sx, int sz, int rectx, int rectz, int padnx, int padnz, float *rr)
{
    memset(rr, 0, sizeof(float) * (size_t)padnx * padnz);
    rr[(size_t)sx * padnz + sz] = 1.;
    smooth2d(rr, padnz, padnx, rectz, rectx);
}

void orc_gradient_smooth2(int rectx, int rectz, int nx, int nz, int waterz, float scaling,
                          float *grad)
{
    smooth2d(grad, nz, nx, rectz, rectx);
    if (waterz >= 0) {
        for (int i = 0; i < waterz; i++)
            for (int j = 0; j < nx; j++) grad[i + j * nz] = 0.;
    } else {
        for (int i = nz + waterz; i < nz; i++)
            for (int j = 0; j < nx; j++) grad[i + j * nz] = 0.;
    }
    for (int i = 0; i < nz * nx; i++) grad[i] *= scaling;
}

/* ---------------- propagation pieces ---------------- */

static void pad2d(const float *vec, float *arr, int nz, int nx, int nb)
{
    int pnz = nz + 2 * nb, pnx = nx + 2 * nb;
    for (int ix = 0; ix < pnx; ix++) {
        int sx = ix - nb < 0 ? 0 : (ix - nb > nx - 1 ? nx - 1 : ix - nb);
        for (int iz = 0; iz < pnz; iz++) {
            int sz = iz - nb < 0 ? 0 : (iz - nb > nz - 1 ? nz - 1 : iz - nb);
            arr[(size_t)ix * pnz + iz] = vec[(size_t)sx * nz + sz];
        }
    }
}

static void laplace8(const float *p, float *term, int pnx, int pnz, float dx2, float dz2)
{
    for (int ix = 4; ix < pnx - 4; ix++)
        for (int iz = 4; iz < pnz - 4; iz++) {
            size_t j = (size_t)ix * pnz + iz;
            size_t s = pnz;
            term[j] = (K0 * p[j] + K1 * (p[j + s] + p[j - s]) + K2 * (p[j + 2 * s] + p[j - 2 * s])
                       + K3 * (p[j + 3 * s] + p[j - 3 * s]) + K4 * (p[j + 4 * s] + p[j - 4 * s])) / dx2
                    + (K0 * p[j] + K1 * (p[j + 1] + p[j - 1]) + K2 * (p[j + 2] + p[j - 2])
                       + K3 * (p[j + 3] + p[j - 3]) + K4 * (p[j + 4] + p[j - 4])) / dz2;
        }
}

static void sponge(float *p, const float *bc, int pnx, int pnz, int nb)
{
    for (int ix = 0; ix < pnx; ix++) {
        for (int iz = 0; iz < nb; iz++) p[(size_t)ix * pnz + iz] = bc[iz] * p[(size_t)ix * pnz + iz];
        for (int iz = pnz - nb; iz < pnz; iz++)
            p[(size_t)ix * pnz + iz] = bc[pnz - iz - 1] * p[(size_t)ix * pnz + iz];
    }
    for (int iz = 0; iz < pnz; iz++) {
        for (int ix = 0; ix < nb; ix++) p[(size_t)ix * pnz + iz] = bc[ix] * p[(size_t)ix * pnz + iz];
        for (int ix = pnx - nb; ix < pnx; ix++)
            p[(size_t)ix * pnz + iz] = bc[pnx - ix - 1] * p[(size_t)ix * pnz + iz];
    }
}

/* one leapfrog step: p2 = 2 p1 - p0 + v*v*dt2*(lap p1 + rr*w) over the whole padded grid */
/* fsrc: the passive source field of this step on the model interior (nx*nz) or NULL
 * (wave_fwigradient.c:733-737: term[ix+nb][iz+nb] += ww3[is][it][ix][iz]) */
static void fwi_step(const float *p0, const float *p1, float *p2, float *term, const float *vv,
                     const float *rr, float wsrc, const float *fsrc, int nz, int nx, int nb,
                     int pnx, int pnz, float dx2, float dz2, float dt2)
{
    size_t n = (size_t)pnx * pnz;
    laplace8(p1, term, pnx, pnz, dx2, dz2);
    if (rr)
        for (size_t j = 0; j < n; j++) term[j] += rr[j] * wsrc;
    if (fsrc)
        for (int ix = 0; ix < nx; ix++)
            for (int iz = 0; iz < nz; iz++)
                term[(size_t)(ix + nb) * pnz + iz + nb] += fsrc[(size_t)ix * nz + iz];
    for (size_t j = 0; j < n; j++) p2[j] = 2 * p1[j] - p0[j] + vv[j] * vv[j] * dt2 * term[j];
}

/* receivers of shot is: (first index in the trace panel, count, first column); the passive
 * path records every shot at r0_v[0] + ir*dr_v, ir < nr (wave_fwigradient.c:710-713) */
static void receivers(const orc_acq *a, int is, int *r02, int *nr2, int *r0v)
{
    if (a->srcf) { *r02 = 0; *nr2 = a->nr; *r0v = a->r0_v[0]; }
    else { *r02 = a->r02[is]; *nr2 = a->nr2[is]; *r0v = a->r0_v[is]; }
}

int orc_fwi_forward(const orc_acq *a, const float *vel, const float *wav, float *data)
{
    int pnz = a->padnz, pnx = a->padnx, nt = a->nt, nr = a->nr;
    size_t n = (size_t)pnz * pnx;
    float dx2 = a->dx * a->dx, dz2 = a->dz * a->dz, dt2 = a->dt * a->dt;
    float *vv = malloc(sizeof(float) * n), *rr = malloc(sizeof(float) * n);
    float *p0 = malloc(sizeof(float) * n), *p1 = malloc(sizeof(float) * n);
    float *p2 = malloc(sizeof(float) * n), *term = calloc(n, sizeof(float));
    pad2d(vel, vv, a->nz, a->nx, a->nb);
    for (int is = 0; is < a->ns; is++) {
        float *dd = data + (size_t)is * nr * nt;
        memset(dd, 0, sizeof(float) * (size_t)nr * nt);
        memset(p0, 0, sizeof(float) * n); memset(p1, 0, sizeof(float) * n); memset(p2, 0, sizeof(float) * n);
        int sx = a->s0_v + is * a->ds_v;
        orc_source_map(sx, a->sz, 2, 2, pnx, pnz, rr);
        int r02, nr2, r0v;
        receivers(a, is, &r02, &nr2, &r0v);
        for (int it = 0; it < nt; it++) {
            for (int ir = 0; ir < nr2; ir++) {
                int rx = r0v + ir * a->dr_v;
                dd[(size_t)(r02 + ir) * nt + it] = p1[(size_t)rx * pnz + a->rz];
            }
            const float *fs = a->srcf ? a->srcf + ((size_t)is * nt + it) * a->nz * a->nx : NULL;
            fwi_step(p0, p1, p2, term, vv, a->srcf ? NULL : rr, wav[it], fs, a->nz, a->nx, a->nb, pnx, pnz, dx2,
                     dz2, dt2);
            float *t = p0; p0 = p1; p1 = p2; p2 = t;
            sponge(p0, a->bc, pnx, pnz, a->nb);
            sponge(p1, a->bc, pnx, pnz, a->nb);
        }
    }
    free(vv); free(rr); free(p0); free(p1); free(p2); free(term);
    return 0;
}

/* tau and taus (preparation, pywave/src/wave_fwiutil.c:91-97) of the quality factor q, padded by
 * edge replication onto the (padnx, padnz) grid */
static void visco_tables(const orc_acq *a, const float *q, float f0, float *tau, float *taus)
{
    const double PI = 3.14159265358979323846264338328;
    size_t nzx = (size_t)a->nz * a->nx;
    float *tau1 = malloc(sizeof(float) * nzx), *taus1 = malloc(sizeof(float) * nzx);
    for (size_t i = 0; i < nzx; i++) {
        float tmp = sqrtf(q[i] * q[i] + 1);
        float taue = (tmp + 1) / (2. * PI * f0 * q[i]);
        taus1[i] = (tmp - 1) / (2. * PI * f0 * q[i]);
        tau1[i] = taue / taus1[i] - 1;
    }
    pad2d(tau1, tau, a->nz, a->nx, a->nb);
    pad2d(taus1, taus, a->nz, a->nx, a->nb);
    free(tau1); free(taus1);
}

/* visco-acoustic modelling (pywave/src/wave_fwimodeling.c:116-220): a memory variable r with
 * tau, taus from Q and f0 (preparation, wave_fwiutil.c:91-97); r and the source term are formed in
 * double where the reference's expressions promote, then stored in f32; r is damped once per step,
 * the pressure levels as in the acoustic case. */
int orc_fwi_forward_visco(const orc_acq *a, const float *vel, const float *q, float f0, const float *wav,
                          float *data)
{
    int pnz = a->padnz, pnx = a->padnx, nt = a->nt, nr = a->nr;
    size_t n = (size_t)pnz * pnx;
    float dx2 = a->dx * a->dx, dz2 = a->dz * a->dz, dt2 = a->dt * a->dt;
    float idt = 1. / a->dt;
    float *vv = malloc(sizeof(float) * n), *tau = malloc(sizeof(float) * n), *taus = malloc(sizeof(float) * n);
    float *rr = malloc(sizeof(float) * n);
    float *p0 = malloc(sizeof(float) * n), *p1 = malloc(sizeof(float) * n), *p2 = malloc(sizeof(float) * n);
    float *r1 = malloc(sizeof(float) * n), *r2 = malloc(sizeof(float) * n), *term = calloc(n, sizeof(float));
    pad2d(vel, vv, a->nz, a->nx, a->nb);
    visco_tables(a, q, f0, tau, taus);
    for (int is = 0; is < a->ns; is++) {
        float *dd = data + (size_t)is * nr * nt;
        memset(dd, 0, sizeof(float) * (size_t)nr * nt);
        memset(p0, 0, sizeof(float) * n); memset(p1, 0, sizeof(float) * n); memset(p2, 0, sizeof(float) * n);
        memset(r1, 0, sizeof(float) * n); memset(r2, 0, sizeof(float) * n);
        int sx = a->s0_v + is * a->ds_v;
        orc_source_map(sx, a->sz, 2, 2, pnx, pnz, rr);
        for (int it = 0; it < nt; it++) {
            for (int ir = 0; ir < a->nr2[is]; ir++) {
                int rx = a->r0_v[is] + ir * a->dr_v;
                dd[(size_t)(a->r02[is] + ir) * nt + it] = p1[(size_t)rx * pnz + a->rz];
            }
            laplace8(p1, term, pnx, pnz, dx2, dz2);
            for (size_t j = 0; j < n; j++) {
                double num = (double)(tau[j] / taus[j] * term[j]) + ((double)idt - 0.5 / (double)taus[j]) * r1[j];
                r2[j] = (float)(num / ((double)idt + 0.5 / (double)taus[j]));
                double t = (double)term[j] * (1. + (double)tau[j]) - (double)(r2[j] + r1[j]) * 0.5;
                term[j] = (float)(t + (double)(rr[j] * wav[it]));
                p2[j] = 2 * p1[j] - p0[j] + vv[j] * vv[j] * dt2 * term[j];
            }
            float *t = p0; p0 = p1; p1 = p2; p2 = t;
            t = r1; r1 = r2; r2 = t;
            sponge(p0, a->bc, pnx, pnz, a->nb);
            sponge(p1, a->bc, pnx, pnz, a->nb);
            sponge(r1, a->bc, pnx, pnz, a->nb);
        }
    }
    free(vv); free(tau); free(taus); free(rr);
    free(p0); free(p1); free(p2); free(r1); free(r2); free(term);
    return 0;
}

static void residual_weights(float *w, int nt, int nr, int wtn1, int wtn2, int woffn1, int woffn2)
{
    float sw[10];
    for (int it = 0; it < 10; it++) sw[it] = sinf(0.5 * 3.14159265358979323846264338328 * (it + 1) / 11.);
    memset(w, 0, sizeof(float) * (size_t)nt * nr);
    for (int ir = woffn1; ir <= woffn2; ir++)
        for (int it = wtn1; it <= wtn2; it++) w[(size_t)ir * nt + it] = 1.;
    for (int ir = 0; ir < 10; ir++)
        for (int it = wtn1; it <= wtn2; it++) {
            w[(size_t)(woffn1 + ir) * nt + it] *= sw[ir];
            w[(size_t)(woffn2 - ir) * nt + it] *= sw[ir];
        }
    for (int it = 0; it < 10; it++)
        for (int ir = woffn1; ir <= woffn2; ir++) {
            w[(size_t)ir * nt + wtn1 + it] *= sw[it];
            w[(size_t)ir * nt + wtn2 - it] *= sw[it];
        }
}

int orc_gradient_av(const orc_acq *a, const float *vel, const float *wav, const float *data,
                    float *grad, float *fcost, float *scaling)
{
    int pnz = a->padnz, pnx = a->padnx, nt = a->nt, nr = a->nr, nz = a->nz, nx = a->nx, nb = a->nb;
    int interval = a->interval, wnt = (nt - 1) / interval + 1;
    size_t n = (size_t)pnz * pnx, nzx = (size_t)nz * nx;
    float dt = a->dt, dt2 = dt * dt, dx2 = a->dx * a->dx, dz2 = a->dz * a->dz;
    float wdt = dt * interval, wdt2 = wdt * wdt;

    /* residual window (gradient_init, wave_fwigradient.c:74-80), fwic's full windows */
    float wt1 = a->t0, wt2 = a->t0 + (nt - 1) * dt;
    float woff1 = a->r0, woff2 = a->r0 + (nr - 1) * a->dr;
    int wtn1 = (wt1 - a->t0) / dt + 0.5, wtn2 = (wt2 - a->t0) / dt + 0.5;
    int woffn1 = (woff1 - a->r0) / a->dr + 0.5, woffn2 = (woff2 - a->r0) / a->dr + 0.5;
    float *weight = malloc(sizeof(float) * (size_t)nt * nr);
    residual_weights(weight, nt, nr, wtn1, wtn2, woffn1, woffn2);

    float *vv = malloc(sizeof(float) * n), *rr = malloc(sizeof(float) * n);
    float *p0 = malloc(sizeof(float) * n), *p1 = malloc(sizeof(float) * n);
    float *p2 = malloc(sizeof(float) * n), *term = calloc(n, sizeof(float));
    float *wave = malloc(sizeof(float) * nzx * wnt);
    float *pp = malloc(sizeof(float) * (size_t)nr * nt);
    if (!wave) return -1;
    pad2d(vel, vv, nz, nx, nb);
    /* gradient_v: memory variable r1/r2, its scratch tmp, tau/taus; updates on [4, pad-4) only */
    const int visco = a->q != NULL;
    float idt = 1. / dt;
    float *tau = NULL, *taus = NULL, *r1 = NULL, *r2 = NULL, *tmp = NULL;
    if (visco) {
        tau = malloc(sizeof(float) * n); taus = malloc(sizeof(float) * n);
        r1 = malloc(sizeof(float) * n); r2 = malloc(sizeof(float) * n); tmp = calloc(n, sizeof(float));
        visco_tables(a, a->q, a->f0, tau, taus);
    }
    *fcost = 0.;
    memset(grad, 0, sizeof(float) * nzx);

    for (int is = 0; is < a->ns; is++) {
        memset(p0, 0, sizeof(float) * n); memset(p1, 0, sizeof(float) * n); memset(p2, 0, sizeof(float) * n);
        memset(pp, 0, sizeof(float) * (size_t)nr * nt);
        if (visco) { memset(r1, 0, sizeof(float) * n); memset(r2, 0, sizeof(float) * n); }
        int sx = a->s0_v + is * a->ds_v;
        orc_source_map(sx, a->sz, 2, 2, pnx, pnz, rr);
        int r02, nr2, r0v;
        receivers(a, is, &r02, &nr2, &r0v);
        int wit = 0;
        for (int it = 0; it < nt; it++) {
            for (int ir = 0; ir < nr2; ir++) {
                int rx = r0v + ir * a->dr_v;
                pp[(size_t)(r02 + ir) * nt + it] = p1[(size_t)rx * pnz + a->rz];
            }
            if (it % interval == 0) {
                for (int ix = 0; ix < nx; ix++)
                    for (int iz = 0; iz < nz; iz++)
                        wave[(size_t)wit * nzx + (size_t)ix * nz + iz] = p1[(size_t)(ix + nb) * pnz + iz + nb];
                wit++;
            }
            if (visco) {  /* wave_fwigradient.c:318-341 */
                laplace8(p1, term, pnx, pnz, dx2, dz2);
                for (int ix = 4; ix < pnx - 4; ix++)
                    for (int iz = 4; iz < pnz - 4; iz++) {
                        size_t j = (size_t)ix * pnz + iz;
                        double num = (double)(tau[j] / taus[j] * term[j]) + ((double)idt - 0.5 / (double)taus[j]) * r1[j];
                        r2[j] = (float)(num / ((double)idt + 0.5 / (double)taus[j]));
                        double t = (double)term[j] * (1. + (double)tau[j]) - (double)(r2[j] + r1[j]) * 0.5;
                        term[j] = (float)(t + (double)(rr[j] * wav[it]));
                        p2[j] = 2 * p1[j] - p0[j] + vv[j] * vv[j] * dt2 * term[j];
                    }
                float *t = r1; r1 = r2; r2 = t;
            } else {
                const float *fs = a->srcf ? a->srcf + ((size_t)is * nt + it) * nzx : NULL;
                fwi_step(p0, p1, p2, term, vv, a->srcf ? NULL : rr, wav[it], fs, nz, nx, nb, pnx, pnz, dx2, dz2, dt2);
            }
            float *t = p0; p0 = p1; p1 = p2; p2 = t;
            sponge(p0, a->bc, pnx, pnz, nb);
            sponge(p1, a->bc, pnx, pnz, nb);
            if (visco) sponge(r1, a->bc, pnx, pnz, nb);
        }
        wit--;
        const float *dd = data + (size_t)is * nr * nt;
        for (int ir = 0; ir < nr; ir++)
            for (int it = 0; it < nt; it++) {
                size_t k = (size_t)ir * nt + it;
                pp[k] = dd[k] - pp[k];
                *fcost += 0.5 * pp[k] * pp[k];
            }
        for (size_t k = 0; k < (size_t)nr * nt; k++) pp[k] *= weight[k];

        memset(p0, 0, sizeof(float) * n); memset(p1, 0, sizeof(float) * n);
        memset(p2, 0, sizeof(float) * n); memset(term, 0, sizeof(float) * n);
        if (visco) {
            memset(r1, 0, sizeof(float) * n); memset(r2, 0, sizeof(float) * n); memset(tmp, 0, sizeof(float) * n);
        }
        for (int it = nt - 1; it >= 0; it--) {
            if (visco) {  /* wave_fwigradient.c:375-388: the r term, then the Laplacian of tmp */
                for (int ix = 4; ix < pnx - 4; ix++)
                    for (int iz = 4; iz < pnz - 4; iz++) {
                        size_t j = (size_t)ix * pnz + iz;
                        double num = (double)(-tau[j] / taus[j] * p1[j]) + (-(double)idt + 0.5 / (double)taus[j]) * r1[j];
                        r2[j] = (float)(num / (-(double)idt - 0.5 / (double)taus[j]));
                        tmp[j] = (float)((double)p1[j] * (1. + (double)tau[j]) - 0.5 * (double)(r2[j] + r1[j]));
                    }
                laplace8(tmp, term, pnx, pnz, dx2, dz2);
            } else {
                laplace8(p1, term, pnx, pnz, dx2, dz2);
            }
            for (int ir = 0; ir < nr2; ir++) {
                int rx = r0v + ir * a->dr_v;
                term[(size_t)rx * pnz + a->rz] += pp[(size_t)(r02 + ir) * nt + it];
            }
            if (visco) {
                for (int ix = 4; ix < pnx - 4; ix++)
                    for (int iz = 4; iz < pnz - 4; iz++) {
                        size_t j = (size_t)ix * pnz + iz;
                        p2[j] = 2 * p1[j] - p0[j] + vv[j] * vv[j] * dt2 * term[j];
                    }
            } else {
                for (size_t j = 0; j < n; j++) p2[j] = 2 * p1[j] - p0[j] + vv[j] * vv[j] * dt2 * term[j];
            }
            if (it % interval == 0) {
                if (wit != wnt - 1 && wit != 0) {
                    for (int ix = 0; ix < nx; ix++)
                        for (int iz = 0; iz < nz; iz++) {
                            float temp = vv[(size_t)(ix + nb) * pnz + iz + nb];
                            temp = temp * temp * temp;
                            temp = -2. / temp;
                            size_t q = (size_t)ix * nz + iz;
                            grad[q] += (wave[(size_t)(wit + 1) * nzx + q] - 2. * wave[(size_t)wit * nzx + q]
                                        + wave[(size_t)(wit - 1) * nzx + q]) / wdt2
                                       * p1[(size_t)(ix + nb) * pnz + iz + nb] * temp;
                        }
                }
                wit--;
            }
            float *t = p0; p0 = p1; p1 = p2; p2 = t;
            sponge(p0, a->bc, pnx, pnz, nb);
            sponge(p1, a->bc, pnx, pnz, nb);
            if (visco) {
                t = r1; r1 = r2; r2 = t;
                sponge(r1, a->bc, pnx, pnz, nb);
            }
        }
    }
    if (*scaling <= 0) {
        float dmax = 0.;
        for (size_t i = 0; i < nzx; i++)
            if (fabsf(grad[i]) > dmax) dmax = fabsf(grad[i]);
        *scaling = 0.1 / dmax;
    }
    orc_gradient_smooth2(3, 3, nx, nz, 0, *scaling, grad);
    free(weight); free(vv); free(rr); free(p0); free(p1); free(p2); free(term); free(wave); free(pp);
    free(tau); free(taus); free(r1); free(r2); free(tmp);
    return 0;
}

/* ---------------- L-BFGS (float arithmetic as in wave_fwilbfgs.c) ---------------- */

static float fdot(int n, const float *a, const float *b)
{
    float s = 0.;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static float fnorm(int n, const float *a) { return sqrtf(fdot(n, a, a)); }

typedef struct {
    int npair, ipair, nls;
    float c1, c2, factor, alpha, f0, fk, v1, v2, conv;
    float *sk, *yk;
} lbfgs_state;

static void lb_save(int n, const float *x, const float *g, lbfgs_state *o)
{
    if (o->ipair < o->npair) {
        memcpy(o->sk + (size_t)o->ipair * n, x, sizeof(float) * n);
        memcpy(o->yk + (size_t)o->ipair * n, g, sizeof(float) * n);
        o->ipair += 1;
    } else {
        for (int i = 0; i < o->npair - 1; i++) {
            memcpy(o->sk + (size_t)i * n, o->sk + (size_t)(i + 1) * n, sizeof(float) * n);
            memcpy(o->yk + (size_t)i * n, o->yk + (size_t)(i + 1) * n, sizeof(float) * n);
        }
        memcpy(o->sk + (size_t)(o->npair - 1) * n, x, sizeof(float) * n);
        memcpy(o->yk + (size_t)(o->npair - 1) * n, g, sizeof(float) * n);
    }
}

static void lb_update(int n, const float *x, const float *g, lbfgs_state *o)
{
    float *s = o->sk + (size_t)(o->ipair - 1) * n, *y = o->yk + (size_t)(o->ipair - 1) * n;
    for (int i = 0; i < n; i++) { s[i] = x[i] - s[i]; y[i] = g[i] - y[i]; }
}

static void lb_direction(int n, const float *g, float *r, lbfgs_state *o)
{
    int m = o->ipair;
    float *sl = o->sk + (size_t)(m - 1) * n, *yl = o->yk + (size_t)(m - 1) * n;
    if (fnorm(n, sl) == 0. || fnorm(n, yl) == 0.) {
        for (int i = 0; i < n; i++) r[i] = -g[i];
        return;
    }
    float *q = malloc(sizeof(float) * n), *rho = malloc(sizeof(float) * m), *al = malloc(sizeof(float) * m);
    memcpy(q, g, sizeof(float) * n);
    for (int i = m - 1; i >= 0; i--) {
        float *s = o->sk + (size_t)i * n, *y = o->yk + (size_t)i * n;
        float tmp = fdot(n, y, s);
        rho[i] = 1. / tmp;
        tmp = fdot(n, s, q);
        al[i] = rho[i] * tmp;
        for (int j = 0; j < n; j++) q[j] -= al[i] * y[j];
    }
    float tmp = fdot(n, yl, yl);
    float gamma = 1. / tmp / rho[m - 1];
    for (int j = 0; j < n; j++) r[j] = gamma * q[j];
    for (int i = 0; i < m; i++) {
        float *s = o->sk + (size_t)i * n, *y = o->yk + (size_t)i * n;
        tmp = fdot(n, y, r);
        float beta = tmp * rho[i];
        tmp = al[i] - beta;
        for (int j = 0; j < n; j++) r[j] += tmp * s[j];
    }
    for (int j = 0; j < n; j++) r[j] = -r[j];
    free(q); free(rho); free(al);
}

int orc_fwi(const orc_acq *a, const float *vel0, const float *wav, const float *data,
            int niter, float *vinv)
{
    int n = a->nz * a->nx;
    float *x = malloc(sizeof(float) * n), *g = malloc(sizeof(float) * n);
    float *dir = malloc(sizeof(float) * n), *xk = malloc(sizeof(float) * n);
    lbfgs_state o;
    o.npair = 20; o.nls = 20; o.c1 = 1e-4; o.c2 = 0.9; o.factor = 10; o.v1 = 0.; o.v2 = 10.;
    o.conv = 0.01;
    o.sk = malloc(sizeof(float) * (size_t)n * o.npair);
    o.yk = malloc(sizeof(float) * (size_t)n * o.npair);
    memcpy(x, vel0, sizeof(float) * n);
    float fcost, scaling = 0;
    orc_gradient_av(a, x, wav, data, g, &fcost, &scaling);
    o.ipair = 0; o.alpha = 1.; o.fk = fcost; o.f0 = fcost;
    int stored = 0, iter;
    for (iter = 0; iter < niter; iter++) {
        if (a->srcf && iter % 5 == 0) o.alpha = 1.;  /* pfwi: optpar->repeat, pywave/src/wave_fwi.c:282 */
        if (iter == 0) {
            for (int i = 0; i < n; i++) dir[i] = -g[i];
        } else {
            lb_update(n, x, g, &o);
            lb_direction(n, g, dir, &o);
        }
        lb_save(n, x, g, &o);
        /* Wolfe line search */
        int flag = 0, i;
        float alpha1 = 0., alpha2 = 0.;
        memcpy(xk, x, sizeof(float) * n);
        float m1 = fdot(n, g, dir);
        float m2 = m1 * o.c2;
        m1 = m1 * o.c1;
        for (i = 0; i < o.nls; i++) {
            for (int j = 0; j < n; j++) x[j] = xk[j] + o.alpha * dir[j];
            for (int j = 0; j < n; j++) {
                if (x[j] < o.v1) x[j] = o.v1;
                if (x[j] > o.v2) x[j] = o.v2;
            }
            orc_gradient_av(a, x, wav, data, g, &fcost, &scaling);
            float m3 = fdot(n, g, dir);
            if (fcost <= o.fk + o.alpha * m1 && m3 >= m2) {
                o.fk = fcost; flag = 0; break;
            } else if (fcost > o.fk + o.alpha * m1) {
                alpha2 = o.alpha;
                o.alpha = 0.5 * (alpha1 + alpha2);
            } else {
                alpha1 = o.alpha;
                if (alpha2 == 0.) o.alpha *= o.factor;
                else o.alpha = 0.5 * (alpha1 + alpha2);
            }
        }
        if (i == o.nls) flag = (fcost <= o.fk) ? 1 : 2;
        if (flag == 2) break;
        if (o.fk / o.f0 < o.conv) break;
        memcpy(vinv + (size_t)iter * n, x, sizeof(float) * n);
        stored++;
    }
    free(x); free(g); free(dir); free(xk); free(o.sk); free(o.yk);
    return stored;
}

float orc_fcost_f32(const float *d, const float *p, long long n)
{
    /* gradient_av's misfit (pywave/src/wave_fwigradient.c:179-184): the residual is stored in
     * f32, 0.5*r*r is formed in double and added to the f32 accumulator, shot by shot, receiver
     * by receiver, sample by sample */
    float f = 0.f;
    for (long long i = 0; i < n; i++) {
        float r = d[i] - p[i];
        f += 0.5 * r * r;
    }
    return f;
}
