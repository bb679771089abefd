/*
 * orc_fft.c -- oracle for the transforms behind pywave's pseudo-spectral Laplacian.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * A single-precision mixed-radix FFT following the KISS FFT algorithm the reference vendors
 * (pywave/src/wave_kissfft.c): factorisation radix 4 first, then 2, 3, 5 and odd numbers up to
 * sqrt(n); recursive decimation in time with the radix-p butterflies applied on the way back up;
 * twiddles exp(-+2 pi i k/n) evaluated in double and rounded to float; real transforms of even
 * length through a half-length complex transform and split twiddles.  The 2D/3D wrappers follow
 * pywave/src/wave_fft2.c:72-151 and wave_fft3.c:96-200: (-1)^(i2) / (-1)^(i2+i3) centering on the
 * input, a real transform along the first axis, complex transforms along the others, and the
 * inverse scaled by 1/(n1 n2 (n3)) with the same centering sign.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

typedef struct {
    float r, i;
} cpx;

typedef struct {
    int n, inverse;
    int fac[64];  /* p0, m0, p1, m1, ... */
    cpx *tw;
} cfft;

static cpx cmul(cpx a, cpx b) { cpx c = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; return c; }
static cpx cadd(cpx a, cpx b) { cpx c = {a.r + b.r, a.i + b.i}; return c; }
static cpx csub(cpx a, cpx b) { cpx c = {a.r - b.r, a.i - b.i}; return c; }

static void cfft_init(cfft *c, int n, int inverse)
{
    const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
    c->n = n;
    c->inverse = inverse;
    c->tw = (cpx *)malloc((n ? n : 1) * sizeof(cpx));
    for (int i = 0; i < n; ++i) {
        double ph = -2 * pi * i / n;
        if (inverse) ph = -ph;
        c->tw[i].r = (float)cos(ph);
        c->tw[i].i = (float)sin(ph);
    }
    int m = n, p = 4, k = 0;
    double fsq = floor(sqrt((double)m));
    do {
        while (m % p) {
            p = p == 4 ? 2 : (p == 2 ? 3 : p + 2);
            if (p > fsq) p = m;
        }
        m /= p;
        c->fac[k++] = p;
        c->fac[k++] = m;
    } while (m > 1);
}

static void cfft_free(cfft *c) { free(c->tw); }

static void bfly2(const cfft *c, cpx *F, size_t fs, int m)
{
    for (int k = 0; k < m; ++k) {
        cpx t = cmul(F[m + k], c->tw[k * fs]);
        F[m + k] = csub(F[k], t);
        F[k] = cadd(F[k], t);
    }
}

static void bfly3(const cfft *c, cpx *F, size_t fs, size_t m)
{
    const cpx e = c->tw[fs * m];
    for (size_t k = 0; k < m; ++k) {
        cpx s1 = cmul(F[k + m], c->tw[k * fs]);
        cpx s2 = cmul(F[k + 2 * m], c->tw[2 * k * fs]);
        cpx s3 = cadd(s1, s2), s0 = csub(s1, s2);
        F[k + m].r = (float)((double)F[k].r - (double)s3.r * .5);  /* half taken in double */
        F[k + m].i = (float)((double)F[k].i - (double)s3.i * .5);
        s0.r *= e.i;
        s0.i *= e.i;
        F[k] = cadd(F[k], s3);
        F[k + 2 * m].r = F[k + m].r + s0.i;
        F[k + 2 * m].i = F[k + m].i - s0.r;
        F[k + m].r -= s0.i;
        F[k + m].i += s0.r;
    }
}

static void bfly4(const cfft *c, cpx *F, size_t fs, size_t m)
{
    for (size_t k = 0; k < m; ++k) {
        cpx s0 = cmul(F[k + m], c->tw[k * fs]);
        cpx s1 = cmul(F[k + 2 * m], c->tw[2 * k * fs]);
        cpx s2 = cmul(F[k + 3 * m], c->tw[3 * k * fs]);
        cpx s5 = csub(F[k], s1);
        F[k] = cadd(F[k], s1);
        cpx s3 = cadd(s0, s2), s4 = csub(s0, s2);
        F[k + 2 * m] = csub(F[k], s3);
        F[k] = cadd(F[k], s3);
        cpx a, b;
        if (c->inverse) {
            a.r = s5.r - s4.i; a.i = s5.i + s4.r;
            b.r = s5.r + s4.i; b.i = s5.i - s4.r;
        } else {
            a.r = s5.r + s4.i; a.i = s5.i - s4.r;
            b.r = s5.r - s4.i; b.i = s5.i + s4.r;
        }
        F[k + m] = a;
        F[k + 3 * m] = b;
    }
}

static void bfly5(const cfft *c, cpx *F, size_t fs, int m)
{
    const cpx ya = c->tw[fs * m], yb = c->tw[fs * 2 * m];
    cpx *F0 = F, *F1 = F + m, *F2 = F + 2 * m, *F3 = F + 3 * m, *F4 = F + 4 * m;
    for (int u = 0; u < m; ++u) {
        cpx s0 = F0[u];
        cpx s1 = cmul(F1[u], c->tw[u * fs]);
        cpx s2 = cmul(F2[u], c->tw[2 * u * fs]);
        cpx s3 = cmul(F3[u], c->tw[3 * u * fs]);
        cpx s4 = cmul(F4[u], c->tw[4 * u * fs]);
        cpx s7 = cadd(s1, s4), s10 = csub(s1, s4), s8 = cadd(s2, s3), s9 = csub(s2, s3);
        F0[u].r += s7.r + s8.r;
        F0[u].i += s7.i + s8.i;
        cpx s5 = {s0.r + s7.r * ya.r + s8.r * yb.r, s0.i + s7.i * ya.r + s8.i * yb.r};
        cpx s6 = {s10.i * ya.i + s9.i * yb.i, -(s10.r * ya.i) - s9.r * yb.i};
        F1[u] = csub(s5, s6);
        F4[u] = cadd(s5, s6);
        cpx s11 = {s0.r + s7.r * yb.r + s8.r * ya.r, s0.i + s7.i * yb.r + s8.i * ya.r};
        cpx s12 = {-(s10.i * yb.i) + s9.i * ya.i, s10.r * yb.i - s9.r * ya.i};
        F2[u] = cadd(s11, s12);
        F3[u] = csub(s11, s12);
    }
}

static void bfly_gen(const cfft *c, cpx *F, size_t fs, int m, int p)
{
    cpx *s = (cpx *)malloc(p * sizeof(cpx));
    for (int u = 0; u < m; ++u) {
        for (int q1 = 0, k = u; q1 < p; ++q1, k += m) s[q1] = F[k];
        for (int q1 = 0, k = u; q1 < p; ++q1, k += m) {
            size_t tx = 0;
            F[k] = s[0];
            for (int q = 1; q < p; ++q) {
                tx += fs * k;
                if (tx >= (size_t)c->n) tx -= c->n;
                F[k] = cadd(F[k], cmul(s[q], c->tw[tx]));
            }
        }
    }
    free(s);
}

/* f: input with element stride 'is' (in complex units); fs: the decimation stride */
static void work(const cfft *c, cpx *out, const cpx *f, size_t fs, size_t is, const int *fp)
{
    const int p = fp[0], m = fp[1];
    if (m == 1) {
        for (int j = 0; j < p; ++j) out[j] = f[j * fs * is];
    } else {
        for (int j = 0; j < p; ++j) work(c, out + j * m, f + j * fs * is, fs * p, is, fp + 2);
    }
    switch (p) {
        case 2: bfly2(c, out, fs, m); break;
        case 3: bfly3(c, out, fs, m); break;
        case 4: bfly4(c, out, fs, m); break;
        case 5: bfly5(c, out, fs, m); break;
        default: bfly_gen(c, out, fs, m, p); break;
    }
}

static void cfft_run(const cfft *c, const cpx *in, size_t is, cpx *out) { work(c, out, in, 1, is, c->fac); }

typedef struct {
    int half;
    cfft sub;
    cpx *sup;
} rfft;

static void rfft_init(rfft *r, int nfft, int inverse)
{
    r->half = nfft / 2;
    cfft_init(&r->sub, r->half, inverse);
    r->sup = (cpx *)malloc((nfft / 4 ? nfft / 4 : 1) * sizeof(cpx));
    for (int i = 0; i < r->half / 2; ++i) {
        double ph = -3.14159265358979323846264338327 * ((double)(i + 1) / r->half + .5);
        if (inverse) ph = -ph;
        r->sup[i].r = (float)cos(ph);
        r->sup[i].i = (float)sin(ph);
    }
}

static void rfft_free(rfft *r) { cfft_free(&r->sub); free(r->sup); }

/* x: 2*half reals -> X: half+1 bins */
static void rfft_fwd(const rfft *r, const float *x, cpx *X, cpx *t)
{
    const int h = r->half;
    cfft_run(&r->sub, (const cpx *)x, 1, t);
    X[0].r = t[0].r + t[0].i;
    X[h].r = t[0].r - t[0].i;
    X[0].i = X[h].i = 0;
    for (int k = 1; k <= h / 2; ++k) {
        cpx fpk = t[k], fpnk = {t[h - k].r, -t[h - k].i};
        cpx f1k = cadd(fpk, fpnk), f2k = csub(fpk, fpnk);
        cpx w = cmul(f2k, r->sup[k - 1]);
        X[k].r = (float)((double)(f1k.r + w.r) * .5);
        X[k].i = (float)((double)(f1k.i + w.i) * .5);
        X[h - k].r = (float)((double)(f1k.r - w.r) * .5);
        X[h - k].i = (float)((double)(w.i - f1k.i) * .5);
    }
}

/* X: half+1 bins -> x: 2*half reals */
static void rfft_inv(const rfft *r, const cpx *X, float *x, cpx *t)
{
    const int h = r->half;
    t[0].r = X[0].r + X[h].r;
    t[0].i = X[0].r - X[h].r;
    for (int k = 1; k <= h / 2; ++k) {
        cpx fk = X[k], fnkc = {X[h - k].r, -X[h - k].i};
        cpx fek = cadd(fk, fnkc), d = csub(fk, fnkc);
        cpx fok = cmul(d, r->sup[k - 1]);
        t[k] = cadd(fek, fok);
        t[h - k] = csub(fek, fok);
        t[h - k].i *= -1;
    }
    cfft_run(&r->sub, t, 1, (cpx *)x);
}

int orc_cfft(int n, int inverse, const float *in, float *out)
{
    cfft c;
    cfft_init(&c, n, inverse);
    cfft_run(&c, (const cpx *)in, 1, (cpx *)out);
    cfft_free(&c);
    return 0;
}

int orc_rfft(int n, int inverse, const float *in, float *out)
{
    rfft r;
    rfft_init(&r, n, inverse);
    cpx *t = (cpx *)malloc((n / 2 + 1) * sizeof(cpx));
    if (inverse) rfft_inv(&r, (const cpx *)in, out, t);
    else rfft_fwd(&r, in, (cpx *)out, t);
    free(t);
    rfft_free(&r);
    return 0;
}

/* ---- pseudo-spectral Laplacian of a padded field: out = IFFT(lapl * FFT(in)) ------------- */

struct orc_psl {
    int n1, n2, n3, nk;      /* padded sizes (z real, x, y complex), nk = n1/2 + 1 */
    float wwt;               /* 1/(n1 n2 n3) */
    float *lapl;             /* [n3][n2][nk] */
    rfft r, ri;
    cfft c2, ic2, c3, ic3;
    cpx *tmp;                /* [n3][n2][nk] */
    cpx *line, *t;
    float *ff;
};

/* the wavenumber operator of psm/psm2d (pywave/src/wave_psp.c:115-126, :471-480):
 * -(2 pi |k|)^2 with centred kx, ky and non-negative kz, f32 arithmetic with hypot in double */
orc_psl *orc_psl_new(int n1, int n2, int n3, float dz, float dx, float dy)
{
    orc_psl *P = (orc_psl *)calloc(1, sizeof(orc_psl));
    P->n1 = n1; P->n2 = n2; P->n3 = n3; P->nk = n1 / 2 + 1;
    P->wwt = 1.0 / (n1 * n2 * n3);
    const int nk = P->nk;
    P->lapl = (float *)malloc((size_t)nk * n2 * n3 * sizeof(float));
    const float dkz = 1. / (n1 * dz), dkx = 1. / (n2 * dx), kx0 = -0.5 / dx;
    const float dky = n3 > 1 ? 1. / (n3 * dy) : 0.f, ky0 = n3 > 1 ? -0.5 / dy : 0.f;
    for (int iy = 0; iy < n3; ++iy) {
        float ky = ky0 + iy * dky;
        for (int ix = 0; ix < n2; ++ix) {
            float kx = kx0 + ix * dkx;
            for (int iz = 0; iz < nk; ++iz) {
                float kz = 0.f + iz * dkz;
                float k = n3 > 1 ? 2 * 3.14159265358979323846264338328 * hypot(ky, hypot(kx, kz))
                                 : 2 * 3.14159265358979323846264338328 * hypot(kx, kz);
                P->lapl[((size_t)iy * n2 + ix) * nk + iz] = -k * k;
            }
        }
    }
    rfft_init(&P->r, n1, 0);
    rfft_init(&P->ri, n1, 1);
    cfft_init(&P->c2, n2, 0);
    cfft_init(&P->ic2, n2, 1);
    cfft_init(&P->c3, n3, 0);
    cfft_init(&P->ic3, n3, 1);
    P->tmp = (cpx *)malloc((size_t)nk * n2 * n3 * sizeof(cpx));
    int mx = n2 > n3 ? n2 : n3;
    P->line = (cpx *)malloc((size_t)(mx > nk ? mx : nk) * sizeof(cpx));
    P->t = (cpx *)malloc((size_t)nk * sizeof(cpx));
    P->ff = (float *)malloc((size_t)n1 * sizeof(float));
    return P;
}

void orc_psl_free(orc_psl *P)
{
    if (!P) return;
    free(P->lapl); free(P->tmp); free(P->line); free(P->t); free(P->ff);
    rfft_free(&P->r); rfft_free(&P->ri);
    cfft_free(&P->c2); cfft_free(&P->ic2); cfft_free(&P->c3); cfft_free(&P->ic3);
    free(P);
}

/* in/out: [n3][n2][n1] padded fields (z fastest) */
void orc_psl_apply(orc_psl *P, const float *in, float *out)
{
    const int n1 = P->n1, n2 = P->n2, n3 = P->n3, nk = P->nk;
    /* centering, real transform along z */
    for (int i3 = 0; i3 < n3; ++i3)
        for (int i2 = 0; i2 < n2; ++i2) {
            const float *col = in + ((size_t)i3 * n2 + i2) * n1;
            const int neg = n3 > 1 ? ((i3 % 2 == 0) != (i2 % 2 == 0)) : (i2 % 2);
            for (int i1 = 0; i1 < n1; ++i1) P->ff[i1] = neg ? -col[i1] : col[i1];
            rfft_fwd(&P->r, P->ff, P->tmp + ((size_t)i3 * n2 + i2) * nk, P->t);
        }
    /* along x (stride nk) */
    for (int i3 = 0; i3 < n3; ++i3)
        for (int i1 = 0; i1 < nk; ++i1) {
            cpx *base = P->tmp + (size_t)i3 * n2 * nk + i1;
            cfft_run(&P->c2, base, nk, P->line);
            for (int i2 = 0; i2 < n2; ++i2) base[(size_t)i2 * nk] = P->line[i2];
        }
    /* along y (stride nk n2) */
    if (n3 > 1)
        for (int i2 = 0; i2 < n2; ++i2)
            for (int i1 = 0; i1 < nk; ++i1) {
                cpx *base = P->tmp + (size_t)i2 * nk + i1;
                cfft_run(&P->c3, base, (size_t)nk * n2, P->line);
                for (int i3 = 0; i3 < n3; ++i3) base[(size_t)i3 * n2 * nk] = P->line[i3];
            }
    /* wavenumber multiply */
    const size_t nkk = (size_t)nk * n2 * n3;
    for (size_t ik = 0; ik < nkk; ++ik) {
        P->tmp[ik].r = P->tmp[ik].r * P->lapl[ik];
        P->tmp[ik].i = P->tmp[ik].i * P->lapl[ik];
    }
    /* inverse along y, then x */
    if (n3 > 1)
        for (int i2 = 0; i2 < n2; ++i2)
            for (int i1 = 0; i1 < nk; ++i1) {
                cpx *base = P->tmp + (size_t)i2 * nk + i1;
                cfft_run(&P->ic3, base, (size_t)nk * n2, P->line);
                for (int i3 = 0; i3 < n3; ++i3) base[(size_t)i3 * n2 * nk] = P->line[i3];
            }
    for (int i3 = 0; i3 < n3; ++i3)
        for (int i1 = 0; i1 < nk; ++i1) {
            cpx *base = P->tmp + (size_t)i3 * n2 * nk + i1;
            cfft_run(&P->ic2, base, nk, P->line);
            for (int i2 = 0; i2 < n2; ++i2) base[(size_t)i2 * nk] = P->line[i2];
        }
    /* inverse real transform along z, centering and 1/(n1 n2 n3) */
    for (int i3 = 0; i3 < n3; ++i3)
        for (int i2 = 0; i2 < n2; ++i2) {
            rfft_inv(&P->ri, P->tmp + ((size_t)i3 * n2 + i2) * nk, P->ff, P->t);
            const int neg = n3 > 1 ? ((i3 % 2 == 0) != (i2 % 2 == 0)) : (i2 % 2);
            const float w = neg ? -P->wwt : P->wwt;
            float *col = out + ((size_t)i3 * n2 + i2) * n1;
            for (int i1 = 0; i1 < n1; ++i1) col[i1] = w * P->ff[i1];
        }
}
