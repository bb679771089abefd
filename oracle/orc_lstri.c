/*
 * orc_lstri.c -- oracle for pfwi mode 4's least-squares time-reversal imaging (lstric).
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 *   operator  pywave/src/wave_triutil.c:14-180 (tri2d_make, timerev_init, timerev_lop) with the
 *             8th-order Lap/LapT of wave_triutil.h and the decay boundary pfwiabc
 *             (wave_triutil.c:264-317)
 *   driver    pywave/src/wave_fwi.c:121-210 (lstri), wave_fwigradient.c:454-525 (lstri_op)
 *   solver    pywave/src/wave_bigsolver.c:1170-1450 (np_solver with a unit model weight) and
 *             wave_cgstep.c:17-67 (np_cgstep), dot products of wave_blas.c
 *
 * Single precision with the reference's operation order (built without FP contraction);
 * dot products accumulate in double in index order like cblas_dsdot, norms in float like the
 * reference's cblas_snrm2 (a plain sum of squares).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* float-divided stencil weights, as wave_triutil.h writes them */
static const float W0 = -205.0f / 72.0f, W1 = 8.0f / 5.0f, W2 = -1.0f / 5.0f, W3 = 8.0f / 315.0f,
                   W4 = -1.0f / 560.0f;

typedef struct {
    int nz, nx, nt, nb, pz, px, depth;
    float sx, sz;   /* 1/dx^2, 1/dz^2 */
    float *v2;      /* v*v*dt^2 on the padded grid, edge-replicated [px][pz] */
    float *decay;   /* nb */
    float *u[3];
} tri_op;

static void tri_setup(tri_op *t, int nz, int nx, int nt, int nb, float dt, float dx, float dz,
                      float cb, const float *vel)
{
    t->nz = nz; t->nx = nx; t->nt = nt; t->nb = nb;
    t->pz = nz + 2 * nb; t->px = nx + 2 * nb;
    t->depth = 1 + nb;  /* receivers one cell below the top of the model (rz = 1) */
    t->sx = 1.0f / (dx * dx);
    t->sz = 1.0f / (dz * dz);
    const float dt2 = dt * dt;
    const size_t np = (size_t)t->pz * t->px;
    t->v2 = (float *)malloc(np * sizeof(float));
    for (int ix = 0; ix < t->px; ++ix) {
        int jx = ix < nb ? nb : (ix >= nb + nx ? nb + nx - 1 : ix);
        for (int iz = 0; iz < t->pz; ++iz) {
            int jz = iz < nb ? nb : (iz >= nb + nz ? nb + nz - 1 : iz);
            float v = vel[(size_t)(jx - nb) * nz + (jz - nb)];
            t->v2[(size_t)ix * t->pz + iz] = v * v * dt2;
        }
    }
    t->decay = (float *)malloc((nb ? nb : 1) * sizeof(float));
    for (int i = 0; i < nb; ++i) {
        int k = nb - 1 - i;
        t->decay[i] = exp(-cb * cb * k * k);
    }
    for (int k = 0; k < 3; ++k) t->u[k] = (float *)calloc(np, sizeof(float));
}

static void tri_release(tri_op *t)
{
    free(t->v2); free(t->decay);
    for (int k = 0; k < 3; ++k) free(t->u[k]);
}

/* decay of the outer nb cells: top/bottom rows first, then left/right columns */
static void tri_decay(const tri_op *t, float *a)
{
    const int pz = t->pz, px = t->px, nb = t->nb;
    for (int iz = 0; iz < nb; ++iz)
        for (int ix = 0; ix < px; ++ix) {
            a[(size_t)ix * pz + iz] *= t->decay[iz];
            a[(size_t)ix * pz + pz - 1 - iz] *= t->decay[iz];
        }
    for (int iz = 0; iz < pz; ++iz)
        for (int ix = 0; ix < nb; ++ix) {
            a[(size_t)ix * pz + iz] *= t->decay[ix];
            a[(size_t)(px - 1 - ix) * pz + iz] *= t->decay[ix];
        }
}

#define AT(a, x, z) (a)[(size_t)(x) * pz + (z)]

/* modelling: dat[ix][it] (nx*nt) from mod[it][ix][iz]; adjoint: mod from dat.  add = 0 clears the
 * output first (np_adjnull). */
static void tri_apply(tri_op *t, int adj, int add, float *mod, float *dat)
{
    const int pz = t->pz, px = t->px, nb = t->nb, nz = t->nz, nx = t->nx, nt = t->nt;
    const size_t np = (size_t)pz * px, nzx = (size_t)nz * nx;
    if (!add) {
        if (adj) memset(mod, 0, nzx * nt * sizeof(float));
        else memset(dat, 0, (size_t)nx * nt * sizeof(float));
    }
    for (int k = 0; k < 3; ++k) memset(t->u[k], 0, np * sizeof(float));
    float *u0 = t->u[0], *u1 = t->u[1], *u2 = t->u[2], *tmp;
    const float *v = t->v2;
    if (!adj) {
        for (int it = 0; it < nt; ++it) {
            for (int ix = nb; ix < nb + nx; ++ix) dat[(size_t)(ix - nb) * nt + it] += AT(u1, ix, t->depth);
            for (int ix = 4; ix < px - 4; ++ix)
                for (int iz = 4; iz < pz - 4; ++iz) {
                    float lx = W4 * (AT(u1, ix + 4, iz) + AT(u1, ix - 4, iz)) +
                               W3 * (AT(u1, ix + 3, iz) + AT(u1, ix - 3, iz)) +
                               W2 * (AT(u1, ix + 2, iz) + AT(u1, ix - 2, iz)) +
                               W1 * (AT(u1, ix + 1, iz) + AT(u1, ix - 1, iz)) + W0 * AT(u1, ix, iz);
                    float lz = W4 * (AT(u1, ix, iz + 4) + AT(u1, ix, iz - 4)) +
                               W3 * (AT(u1, ix, iz + 3) + AT(u1, ix, iz - 3)) +
                               W2 * (AT(u1, ix, iz + 2) + AT(u1, ix, iz - 2)) +
                               W1 * (AT(u1, ix, iz + 1) + AT(u1, ix, iz - 1)) + W0 * AT(u1, ix, iz);
                    float lap = (lx * t->sx + lz * t->sz) * AT(v, ix, iz);
                    AT(u2, ix, iz) = lap + 2.0f * AT(u1, ix, iz) - AT(u0, ix, iz);
                }
            tmp = u0; u0 = u1; u1 = u2; u2 = tmp;
            const float *w = mod + (size_t)it * nzx;
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = 0; iz < nz; ++iz)
                    AT(u1, ix + nb, iz + nb) += w[(size_t)ix * nz + iz] * AT(v, ix + nb, iz + nb);
            tri_decay(t, u0);
            tri_decay(t, u1);
        }
    } else {
        for (int it = nt - 1; it >= 0; --it) {
            tri_decay(t, u1);
            tri_decay(t, u0);
            float *w = mod + (size_t)it * nzx;
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = 0; iz < nz; ++iz)
                    w[(size_t)ix * nz + iz] += AT(u1, ix + nb, iz + nb) * AT(v, ix + nb, iz + nb);
            for (int ix = 4; ix < px - 4; ++ix)
                for (int iz = 4; iz < pz - 4; ++iz) {
#define P(x, z) (AT(u1, x, z) * AT(v, x, z))
                    float lx = W4 * (P(ix + 4, iz) + P(ix - 4, iz)) + W3 * (P(ix + 3, iz) + P(ix - 3, iz)) +
                               W2 * (P(ix + 2, iz) + P(ix - 2, iz)) + W1 * (P(ix + 1, iz) + P(ix - 1, iz)) +
                               W0 * P(ix, iz);
                    float lz = W4 * (P(ix, iz + 4) + P(ix, iz - 4)) + W3 * (P(ix, iz + 3) + P(ix, iz - 3)) +
                               W2 * (P(ix, iz + 2) + P(ix, iz - 2)) + W1 * (P(ix, iz + 1) + P(ix, iz - 1)) +
                               W0 * P(ix, iz);
#undef P
                    AT(u2, ix, iz) = (lx * t->sx + lz * t->sz) + 2.0f * AT(u1, ix, iz) - AT(u0, ix, iz);
                }
            tmp = u0; u0 = u1; u1 = u2; u2 = tmp;
            for (int ix = nb; ix < nb + nx; ++ix) AT(u1, ix, t->depth) += dat[(size_t)(ix - nb) * nt + it];
        }
    }
    t->u[0] = u0; t->u[1] = u1; t->u[2] = u2;
}

#undef AT

int orc_timerev(int nz, int nx, int nt, int nb, float dt, float dx, float dz, float coef,
                const float *vel, int adj, int add, float *mod, float *dat)
{
    tri_op t;
    tri_setup(&t, nz, nx, nt, nb, dt, dx, dz, coef, vel);
    tri_apply(&t, adj, add, mod, dat);
    tri_release(&t);
    return 0;
}

static double dot_d(size_t n, const float *a, const float *b)
{
    double s = 0.;
    for (size_t i = 0; i < n; ++i) s += (double)a[i] * b[i];
    return s;
}

static float sumsq_f(size_t n, const float *a)
{
    float s = 0.f;
    for (size_t i = 0; i < n; ++i) s += a[i] * a[i];
    return s;
}

/* conjugate-gradient state that persists over the shots of one call */
typedef struct {
    int started;
    float *S, *Ss;
} cg_state;

/* one CG step (plane search in the span of the new gradient and the previous step) */
static void cg_step(cg_state *c, size_t nm, size_t nd, float *x, const float *g, float *rr, const float *gg)
{
    double alfa, beta;
    int restart = 0;
    if (!c->started) {
        c->started = 1;
        restart = 1;
    }
    if (restart) {
        memset(c->S, 0, nm * sizeof(float));
        memset(c->Ss, 0, nd * sizeof(float));
        beta = 0.0;
        double gdg = dot_d(nd, gg, gg);
        if (gdg <= 0.) return;
        alfa = -dot_d(nd, gg, rr) / gdg;
    } else {
        double gdg = dot_d(nd, gg, gg), sds = dot_d(nd, c->Ss, c->Ss), gds = dot_d(nd, gg, c->Ss);
        if (gdg == 0. || sds == 0.) return;
        double det = 1.0 - (gds / gdg) * (gds / sds);
        if (det > 1.e-12f) det *= gdg * sds;
        else det = gdg * sds * 1.e-12f;
        double gdr = -dot_d(nd, gg, rr), sdr = -dot_d(nd, c->Ss, rr);
        alfa = (sds * gdr - gds * sdr) / det;
        beta = (-gds * gdr + gdg * sdr) / det;
    }
    const float fa = (float)alfa, fb = (float)beta;
    for (size_t i = 0; i < nm; ++i) c->S[i] *= fb;
    for (size_t i = 0; i < nm; ++i) c->S[i] += fa * g[i];
    for (size_t i = 0; i < nd; ++i) c->Ss[i] *= fb;
    for (size_t i = 0; i < nd; ++i) c->Ss[i] += fa * gg[i];
    for (size_t i = 0; i < nm; ++i) x[i] += c->S[i];
    for (size_t i = 0; i < nd; ++i) rr[i] += c->Ss[i];
}

int orc_lstri(int nz, int nx, int nt, int ns, int nb, float dt, float dx, float dz, float coef,
              const float *vel, int inv, int niter, float *data, float *src)
{
    tri_op t;
    tri_setup(&t, nz, nx, nt, nb, dt, dx, dz, coef, vel);
    const size_t nm = (size_t)nz * nx * nt, nd = (size_t)nx * nt;
    cg_state c = {0, NULL, NULL};
    float *g = NULL, *rr = NULL, *gg = NULL;
    if (inv) {
        c.S = (float *)malloc(nm * sizeof(float));
        c.Ss = (float *)malloc(nd * sizeof(float));
        g = (float *)malloc(nm * sizeof(float));
        rr = (float *)malloc(nd * sizeof(float));
        gg = (float *)malloc(nd * sizeof(float));
    }
    for (int is = 0; is < ns; ++is) {
        float *dd = data + (size_t)is * nd, *ww = src + (size_t)is * nm;
        if (!inv) {
            tri_apply(&t, 0, 0, ww, dd);
            continue;
        }
        /* solve timerev(ww) ~= dd from ww = 0; the model weight is one everywhere */
        memset(ww, 0, nm * sizeof(float));
        for (size_t i = 0; i < nd; ++i) rr[i] = -dd[i];
        const float r0 = sumsq_f(nd, rr);
        float g0 = 1.f;
        for (int iter = 0; iter < niter; ++iter) {
            tri_apply(&t, 1, 0, g, rr);
            tri_apply(&t, 0, 0, g, gg);
            float dpr, dpg;
            if (iter == 0) {
                g0 = sumsq_f(nm, g);
                dpr = 1.f;
                dpg = 1.f;
            } else {
                dpr = sumsq_f(nd, rr) / r0;
                dpg = sumsq_f(nm, g) / g0;
            }
            if (dpr < 1.e-12f || dpg < 1.e-12f) break;
            cg_step(&c, nm, nd, ww, g, rr, gg);
        }
    }
    free(c.S); free(c.Ss); free(g); free(rr); free(gg);
    tri_release(&t);
    return 0;
}
