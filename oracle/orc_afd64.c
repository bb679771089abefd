/*
 * orc_afd64.c -- FP64 twin of the afd2d/afd3d modelling engine (TEST INFRASTRUCTURE ONLY).
 *
 * The same discretisation as orc_afd.c and the reference C, evaluated in double precision, so
 * that the f32 results (the reference's, the oracle's and the GPU's) can be given an error
 * against the exact arithmetic of the scheme:
 *   padded sizes      pywave/src/wave_fft2.c:45-51, wave_fft3.c:45-60
 *   velocity expand   pywave/src/wave_abc.c:14-125   (clamp-to-edge replication)
 *   damping weights   pywave/src/wave_abc.c:127-145  (exp(-c^2 (nb-1-i)^2), here all in double)
 *   damping apply     pywave/src/wave_abc.c:203-293  (every face multiplies; corners get products)
 *   4th-order Laplacian, leapfrog, 3x3(x3) source stencil, surface receivers, snapshots
 *                     pywave/src/wave_fdm.c:209-354 (3D), :551-662 (2D)
 * Inputs are the f32 values the f32 engines see (model, wavelet samples, dt, spacings, ct),
 * promoted exactly; every operation after that is in double.  Modelling only (no time reversal).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_lap_f64(const double *u, int nz2, int nx2, int ny2, double dz, double dx, double dy, double *out)
{
    const double c0 = -30. / 12., c1 = 16. / 12., c2 = -1. / 12.;
    const double idz = 1 / (dz * dz), idx = 1 / (dx * dx), idy = ny2 > 1 ? 1 / (dy * dy) : 0;
    const size_t pl = (size_t)nz2 * nx2;
    const int y0 = ny2 > 1 ? 2 : 0, y1 = ny2 > 1 ? ny2 - 2 : 1;
    memset(out, 0, sizeof(double) * pl * ny2);
    for (int iy = y0; iy < y1; iy++)
        for (int ix = 2; ix < nx2 - 2; ix++)
            for (int iz = 2; iz < nz2 - 2; iz++) {
                size_t j = pl * iy + (size_t)nz2 * ix + iz;
                double s = c0 * u[j] * (idz + idx + idy)
                         + (c1 * (u[j - nz2] + u[j + nz2]) + c2 * (u[j - 2 * nz2] + u[j + 2 * nz2])) * idx
                         + (c1 * (u[j - 1] + u[j + 1]) + c2 * (u[j - 2] + u[j + 2])) * idz;
                if (ny2 > 1)
                    s += (c1 * (u[j - pl] + u[j + pl]) + c2 * (u[j - 2 * pl] + u[j + 2 * pl])) * idy;
                out[j] = s;
            }
}

static void damp_f64(double *a, int nz2, int nx2, int ny2, int nb, int nby, const double *w)
{
    const size_t pl = (size_t)nz2 * nx2;
    for (int iy = 0; iy < ny2; iy++)
        for (int ix = 0; ix < nx2; ix++) {
            double *c = a + pl * iy + (size_t)nz2 * ix;
            for (int iz = 0; iz < nb; iz++) c[iz] *= w[iz], c[nz2 - 1 - iz] *= w[iz];
            double f = 1.0;
            if (ix < nb) f *= w[ix];
            if (nx2 - 1 - ix < nb) f *= w[nx2 - 1 - ix];
            if (iy < nby) f *= w[iy];
            if (ny2 - 1 - iy < nby) f *= w[ny2 - 1 - iy];
            if (f != 1.0)
                for (int iz = 0; iz < nz2; iz++) c[iz] *= f;
        }
}

int orc_afd_f64(const float *vel, int nz1, int nx1, int ny1, int is3d, int nt, float dt, float dx,
                float dy, float dz, int nbt, float ct, int ns, const int *spx, const int *spy,
                const int *spz, const float *rick, double *dat, int jsnap, double *snaps)
{
    const int nby = is3d ? nbt : 0;
    const int nz = nz1 + 2 * nbt, nx = nx1 + 2 * nbt, ny = ny1 + 2 * nby;
    int nz2, nx2, ny2;
    orc_padded_sizes(nz, nx, ny, &nz2, &nx2, &ny2);
    if (!is3d) ny2 = 1;
    const size_t pl = (size_t)nz2 * nx2, n2 = pl * ny2;
    double *vv = calloc(n2, sizeof(double));
    double *curr = calloc(n2, sizeof(double));
    double *prev = calloc(n2, sizeof(double));
    double *lap = calloc(n2, sizeof(double));
    double *w = calloc(nbt > 0 ? nbt : 1, sizeof(double));
    if (!vv || !curr || !prev || !lap || !w) {
        free(vv); free(curr); free(prev); free(lap); free(w);
        return -1;
    }
    const double c = ct, ddt = dt;
    for (int ib = 0; ib < nbt; ib++) w[ib] = exp(-c * c * (double)(nbt - 1 - ib) * (nbt - 1 - ib));
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++)
            for (int iz = 0; iz < nz; iz++) {
                int sy = is3d ? clampi(iy - nby, 0, ny1 - 1) : 0;
                int sx = clampi(ix - nbt, 0, nx1 - 1), sz = clampi(iz - nbt, 0, nz1 - 1);
                double v = vel[(size_t)nz1 * nx1 * sy + (size_t)nz1 * sx + sz];
                vv[pl * iy + (size_t)nz2 * ix + iz] = v * v * ddt * ddt;
            }
    for (int it = 0; it < nt; it++) {
        orc_lap_f64(curr, nz2, nx2, ny2, dz, dx, is3d ? dy : 1.0, lap);
        /* leapfrog on the updated domain [0,nz) x [0,nx) x [0,ny) */
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++)
                for (int iz = 0; iz < nz; iz++) {
                    size_t j = pl * iy + (size_t)nz2 * ix + iz;
                    double old = curr[j];
                    curr[j] = 2 * old - prev[j] + lap[j] * vv[j];
                    prev[j] = old;
                }
        for (int s = 0; s < ns; s++) {
            int ylo = is3d ? -1 : 0, yhi = is3d ? 1 : 0;
            for (int ddy = ylo; ddy <= yhi; ddy++)
                for (int ddx = -1; ddx <= 1; ddx++)
                    for (int ddz = -1; ddz <= 1; ddz++) {
                        int gz = spz[s] + nbt + ddz, gx = spx[s] + nbt + ddx;
                        int gy = is3d ? spy[s] + nby + ddy : 0;
                        if (gz < 0 || gz >= nz || gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
                        size_t j = pl * gy + (size_t)nz2 * gx + gz;
                        curr[j] += vv[j] * (double)rick[(size_t)s * nt + it] / (abs(ddx) + abs(ddy) + abs(ddz) + 1);
                    }
        }
        damp_f64(curr, nz2, nx2, ny2, nbt, nby, w);
        damp_f64(prev, nz2, nx2, ny2, nbt, nby, w);
        for (int iy = 0; iy < ny1; iy++)
            for (int ix = 0; ix < nx1; ix++) {
                int gy = is3d ? iy + nby : 0;
                dat[((size_t)iy * nx1 + ix) * nt + it] = curr[pl * gy + (size_t)nz2 * (ix + nbt) + nbt];
            }
        if (jsnap > 0 && it % jsnap == 0 && snaps) {
            double *snap = snaps + (size_t)(it / jsnap) * nz1 * nx1 * ny1;
            for (int iy = 0; iy < ny1; iy++)
                for (int ix = 0; ix < nx1; ix++) {
                    int gy = is3d ? iy + nby : 0;
                    memcpy(snap + (size_t)nz1 * nx1 * iy + (size_t)nz1 * ix,
                           curr + pl * gy + (size_t)nz2 * (ix + nbt) + nbt, sizeof(double) * nz1);
                }
        }
    }
    free(vv); free(curr); free(prev); free(lap); free(w);
    return 0;
}
