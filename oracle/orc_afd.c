/*
 * orc_afd.c -- oracle restatement of pywave's afd2d/afd3d engine (TEST INFRASTRUCTURE ONLY).
 *
 * Follows, per grid cell, the reference's floating-point operation order so that a build
 * without FP contraction reproduces it bit for bit:
 *   padded sizes      pywave/src/wave_fft2.c:45-51, wave_fft3.c:45-60, wave_kissfft.c:412-424
 *   velocity expand   pywave/src/wave_abc.c:14-125   (clamp-to-edge replication)
 *   damping weights   pywave/src/wave_abc.c:127-145
 *   damping apply     pywave/src/wave_abc.c:203-293  (faces in order top,bottom,lx,rx,ly,ry)
 *   2D step           pywave/src/wave_fdm.c:551-662
 *   3D step           pywave/src/wave_fdm.c:209-354
 * Loops are written cache-friendly (z innermost); every cell is computed independently, so
 * loop order does not change any value.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static const float C0 = -30. / 12., C1 = +16. / 12., C2 = -1. / 12.;

int orc_next_fast_size(int n)
{
    for (;; n++) {
        int m = n;
        while (m % 2 == 0) m /= 2;
        while (m % 3 == 0) m /= 3;
        while (m % 5 == 0) m /= 5;
        if (m <= 1) return n;
    }
}

void orc_padded_sizes(int nz, int nx, int ny, int *nz2, int *nx2, int *ny2)
{
    /* real-to-complex first axis: nk = nfs((n+1)/2)+1, n2 = 2*(nk-1) */
    if (nz2) *nz2 = 2 * orc_next_fast_size((nz + 1) / 2);
    if (nx2) *nx2 = orc_next_fast_size(nx);
    if (ny2) *ny2 = orc_next_fast_size(ny);
}

void orc_abc_weights(int nb, float c, float *w)
{
    for (int ib = 0; ib < nb; ib++) {
        float e = -c * c * (nb - 1 - ib) * (nb - 1 - ib); /* float products, left to right */
        w[ib] = exp((double)e);
    }
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* damping of one padded field: faces applied one after another (corner cells get the
 * product in this order) */
static void damp3(float *a, int nz2, int nx2, int ny2, int nb, int nby, const float *w)
{
    size_t pl = (size_t)nz2 * nx2;
    /* top */
    for (int iy = 0; iy < ny2; iy++)
        for (int ix = 0; ix < nx2; ix++)
            for (int iz = 0; iz < nb; iz++) a[pl * iy + (size_t)nz2 * ix + iz] *= w[iz];
    /* bottom */
    for (int iy = 0; iy < ny2; iy++)
        for (int ix = 0; ix < nx2; ix++)
            for (int iz = 0; iz < nb; iz++) a[pl * iy + (size_t)nz2 * ix + nz2 - 1 - iz] *= w[iz];
    /* left x */
    for (int iy = 0; iy < ny2; iy++)
        for (int ix = 0; ix < nb; ix++)
            for (int iz = 0; iz < nz2; iz++) a[pl * iy + (size_t)nz2 * ix + iz] *= w[ix];
    /* right x */
    for (int iy = 0; iy < ny2; iy++)
        for (int ix = 0; ix < nb; ix++)
            for (int iz = 0; iz < nz2; iz++) a[pl * iy + (size_t)nz2 * (nx2 - 1 - ix) + iz] *= w[ix];
    /* left y / right y (absent in 2D: nby == 0) */
    for (int iy = 0; iy < nby; iy++)
        for (int ix = 0; ix < nx2; ix++)
            for (int iz = 0; iz < nz2; iz++) a[pl * iy + (size_t)nz2 * ix + iz] *= w[iy];
    for (int iy = 0; iy < nby; iy++)
        for (int ix = 0; ix < nx2; ix++)
            for (int iz = 0; iz < nz2; iz++) a[pl * (ny2 - 1 - iy) + (size_t)nz2 * ix + iz] *= w[iy];
}

/* shared 2D/3D engine; 2D is ny1 == 1 with no y boundary.  ps != 0: the Laplacian is the
 * pseudo-spectral one of psm/psm2d (pywave/src/wave_psp.c:25-365, 367-635) over the whole padded
 * grid; everything else (set-up, leapfrog, injection, boundary, receivers) is shared. */
static int afd_engine(const float *vel, int nz1, int nx1, int ny1, int is3d, int nt, float dt,
                      float dx, float dy, float dz, int nbt, float ct, int ns, const int *spx,
                      const int *spy, const int *spz, const float *rick, int tri, float *dat,
                      int jsnap, float *snaps, float *img, int ps)
{
    int nby = is3d ? nbt : 0;
    int nz = nz1 + 2 * nbt, nx = nx1 + 2 * nbt, ny = ny1 + 2 * nby;
    int nz2, nx2, ny2;
    orc_padded_sizes(nz, nx, ny, &nz2, &nx2, &ny2);
    if (!is3d) ny2 = 1;
    size_t pl = (size_t)nz2 * nx2, n2 = pl * ny2, n = (size_t)nz * nx * ny;

    float *vv = malloc(n * sizeof(float));
    float *curr = calloc(n2, sizeof(float));
    float *prev = calloc(n2, sizeof(float));
    float *lap = calloc(n2, sizeof(float)); /* rim cells never written: stay 0 */
    float *w = malloc((nbt > 0 ? nbt : 1) * sizeof(float));
    if (!vv || !curr || !prev || !lap || !w) return -1;
    orc_abc_weights(nbt, ct, w);

    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++)
            for (int iz = 0; iz < nz; iz++) {
                int sy = is3d ? clampi(iy - nby, 0, ny1 - 1) : 0;
                int sx = clampi(ix - nbt, 0, nx1 - 1), sz = clampi(iz - nbt, 0, nz1 - 1);
                float v = vel[(size_t)nz1 * nx1 * sy + (size_t)nz1 * sx + sz];
                vv[(size_t)nz * nx * iy + (size_t)nz * ix + iz] = v * v * dt * dt;
            }

    orc_psl *P = ps ? orc_psl_new(nz2, nx2, ny2, dz, dx, dy) : NULL;
    float idz = 1 / (dz * dz), idx = 1 / (dx * dx), idy = 1 / (dy * dy);
    float cc3 = idy + idx + idz, cc2 = idx + idz;
    int it0 = tri ? nt - 1 : 0, it1 = tri ? -1 : nt, its = tri ? -1 : 1;

    for (int it = it0; it != it1; it += its) {
        /* Laplacian on [2,n2-2) per axis (finite differences) or spectral on the padded grid */
        if (ps) {
            orc_psl_apply(P, curr, lap);
        } else if (is3d) {
            for (int iy = 2; iy < ny2 - 2; iy++)
                for (int ix = 2; ix < nx2 - 2; ix++)
                    for (int iz = 2; iz < nz2 - 2; iz++) {
                        size_t j = pl * iy + (size_t)nz2 * ix + iz;
                        lap[j] = C0 * curr[j] * cc3
                               + C1 * (curr[j - pl] + curr[j + pl]) * idy
                               + C2 * (curr[j - 2 * pl] + curr[j + 2 * pl]) * idy
                               + C1 * (curr[j - nz2] + curr[j + nz2]) * idx
                               + C2 * (curr[j - 2 * nz2] + curr[j + 2 * nz2]) * idx
                               + C1 * (curr[j - 1] + curr[j + 1]) * idz
                               + C2 * (curr[j - 2] + curr[j + 2]) * idz;
                    }
        } else {
            for (int ix = 2; ix < nx2 - 2; ix++)
                for (int iz = 2; iz < nz2 - 2; iz++) {
                    size_t j = (size_t)nz2 * ix + iz;
                    lap[j] = C0 * curr[j] * cc2
                           + C1 * (curr[j - nz2] + curr[j + nz2]) * idx
                           + C2 * (curr[j - 2 * nz2] + curr[j + 2 * nz2]) * idx
                           + C1 * (curr[j - 1] + curr[j + 1]) * idz
                           + C2 * (curr[j - 2] + curr[j + 2]) * idz;
                }
        }
        /* leapfrog on the updated domain: c + (c - p), then + lap*vv */
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++)
                for (int iz = 0; iz < nz; iz++) {
                    size_t i = (size_t)nz * nx * iy + (size_t)nz * ix + iz;
                    size_t j = pl * iy + (size_t)nz2 * ix + iz;
                    float old = curr[j], c = old;
                    c += c - prev[j];
                    prev[j] = old;
                    c += lap[j] * vv[i];
                    curr[j] = c;
                }
        /* injection */
        if (tri) {
            for (int iy = 0; iy < ny1; iy++)
                for (int ix = 0; ix < nx1; ix++) {
                    int gy = is3d ? iy + nby : 0;
                    size_t j = pl * gy + (size_t)nz2 * (ix + nbt) + nbt;
                    size_t i = (size_t)nz * nx * gy + (size_t)nz * (ix + nbt) + nbt;
                    curr[j] += vv[i] * dat[((size_t)iy * nx1 + ix) * nt + it];
                }
        } else {
            for (int s = 0; s < ns; s++) {
                int ylo = is3d ? -1 : 0, yhi = is3d ? 1 : 0;
                for (int ddy = ylo; ddy <= yhi; ddy++)
                    for (int ddx = -1; ddx <= 1; ddx++)
                        for (int ddz = -1; ddz <= 1; ddz++) {
                            int gz = spz[s] + nbt + ddz, gx = spx[s] + nbt + ddx;
                            int gy = is3d ? spy[s] + nby + ddy : 0;
                            if (gz < 0 || gz >= nz || gx < 0 || gx >= nx || gy < 0 || gy >= ny)
                                continue; /* the reference would write out of bounds */
                            size_t i = (size_t)nz * nx * gy + (size_t)nz * gx + gz;
                            size_t j = pl * gy + (size_t)nz2 * gx + gz;
                            int k = abs(ddx) + abs(ddy) + abs(ddz) + 1;
                            curr[j] += vv[i] * rick[(size_t)s * nt + it] / k;
                        }
            }
        }
        /* absorbing boundary on both time levels */
        damp3(curr, nz2, nx2, ny2, nbt, nby, w);
        damp3(prev, nz2, nx2, ny2, nbt, nby, w);
        /* receivers: the gpz = nbt row/plane over the model columns */
        if (!tri) {
            for (int iy = 0; iy < ny1; iy++)
                for (int ix = 0; ix < nx1; ix++) {
                    int gy = is3d ? iy + nby : 0;
                    dat[((size_t)iy * nx1 + ix) * nt + it] = curr[pl * gy + (size_t)nz2 * (ix + nbt) + nbt];
                }
        }
        if (jsnap > 0 && it % jsnap == 0 && snaps) {
            float *snap = snaps + (size_t)(it / jsnap) * nz1 * nx1 * ny1;
            for (int iy = 0; iy < ny1; iy++)
                for (int ix = 0; ix < nx1; ix++)
                    for (int iz = 0; iz < nz1; iz++) {
                        int gy = is3d ? iy + nby : 0;
                        snap[(size_t)nz1 * nx1 * iy + (size_t)nz1 * ix + iz] =
                            curr[pl * gy + (size_t)nz2 * (ix + nbt) + iz + nbt];
                    }
        }
    }
    if (tri && img) {
        for (int iy = 0; iy < ny1; iy++)
            for (int ix = 0; ix < nx1; ix++)
                for (int iz = 0; iz < nz1; iz++) {
                    int gy = is3d ? iy + nby : 0;
                    img[(size_t)nz1 * nx1 * iy + (size_t)nz1 * ix + iz] =
                        curr[pl * gy + (size_t)nz2 * (ix + nbt) + iz + nbt];
                }
    }
    free(vv); free(curr); free(prev); free(lap); free(w);
    orc_psl_free(P);
    return 0;
}

int orc_afd2d(const float *vel, int nz1, int nx1, int nt, float dt, float dx, float dz,
              int nbt, float ct, int ns, const int *spx, const int *spz, const float *rick,
              int tri, float *dat, int jsnap, float *snaps, float *img)
{
    return afd_engine(vel, nz1, nx1, 1, 0, nt, dt, dx, 1.0f, dz, nbt, ct, ns, spx, NULL, spz, rick,
                      tri, dat, jsnap, snaps, img, 0);
}

int orc_aps2d(const float *vel, int nz1, int nx1, int nt, float dt, float dx, float dz,
              int nbt, float ct, int ns, const int *spx, const int *spz, const float *rick,
              int tri, float *dat, int jsnap, float *snaps, float *img)
{
    return afd_engine(vel, nz1, nx1, 1, 0, nt, dt, dx, 1.0f, dz, nbt, ct, ns, spx, NULL, spz, rick,
                      tri, dat, jsnap, snaps, img, 1);
}

int orc_afd3d(const float *vel, int nz1, int nx1, int ny1, int nt, float dt, float dx,
              float dy, float dz, int nbt, float ct, int ns, const int *spx, const int *spy,
              const int *spz, const float *rick, int tri, float *dat, int jsnap, float *snaps,
              float *img)
{
    return afd_engine(vel, nz1, nx1, ny1, 1, nt, dt, dx, dy, dz, nbt, ct, ns, spx, spy, spz, rick,
                      tri, dat, jsnap, snaps, img, 0);
}

int orc_aps3d(const float *vel, int nz1, int nx1, int ny1, int nt, float dt, float dx,
              float dy, float dz, int nbt, float ct, int ns, const int *spx, const int *spy,
              const int *spz, const float *rick, int tri, float *dat, int jsnap, float *snaps,
              float *img)
{
    return afd_engine(vel, nz1, nx1, ny1, 1, nt, dt, dx, dy, dz, nbt, ct, ns, spx, spy, spz, rick,
                      tri, dat, jsnap, snaps, img, 1);
}
