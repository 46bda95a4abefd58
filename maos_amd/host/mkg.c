/*
 * Average-gradient (G-tilt) operator of a Shack-Hartmann powfs: GS0 of
 * setup_shwfs_grad (maos/powfs.c:603-630), built by mkg (lib/mkg.c:108-468)
 * with the OPD grid and the pupil grid the same, no scaling or offset, and
 * partial cells used (do_partial = 1).
 *
 * The OPD is bilinearly interpolated between the points of all subaperture
 * grids (one common grid: a subaperture's edge cells reach the next
 * subaperture's points), and so is the amplitude.  Gradient x of
 * subaperture S is
 *     g_x = (1/dx) int_S A dPhi/du du dv / int_S A du dv
 * over the cells whose four corner points exist with amplitude > 1e-14, in
 * cell units u, v (the OPD grid step dx); g_y likewise.  A row whose
 * amplitude integral is at most 1% of the subaperture area is zero, a
 * subaperture with saa < saat has an empty row, and entries of magnitude at
 * most 1e-14 are dropped.  Each cell's integrals are exact polynomial
 * integrals of products of the bilinear basis over the cell's part inside
 * the square.
 */
#include "../../include/maos_setup.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MKG_TOL 1e-14

/* integrals over [a, b] of 1-t, t, (1-t)^2, t(1-t), t^2 */
static void ints1(double a, double b, double r[5]) {
    const double d1 = b - a, d2 = (b * b - a * a) * 0.5, d3 = (b * b * b - a * a * a) / 3.;
    r[0] = d1 - d2;
    r[1] = d2;
    r[2] = d1 - 2 * d2 + d3;
    r[3] = d2 - d3;
    r[4] = d3;
}

long mxs_mkg(long *rowp, long *coli, double *val, long nzmax, const double *ptsx, const double *ptsy, long nsa,
             long nxsa, double dx, const double *amp, const double *saa, double saat) {
    if (!rowp || !coli || !val || !ptsx || !ptsy || nsa < 1 || nxsa < 1 || !(dx > 0)) return -1;
    const long npsa = nxsa * nxsa;
    /* the common grid: integer coordinates of every point */
    double x0 = ptsx[0], y0 = ptsy[0];
    for (long isa = 1; isa < nsa; isa++) {
        x0 = fmin(x0, ptsx[isa]);
        y0 = fmin(y0, ptsy[isa]);
    }
    long gx1 = 0, gy1 = 0;
    long *ox = (long *)malloc(sizeof(long) * nsa), *oy = (long *)malloc(sizeof(long) * nsa);
    for (long isa = 0; isa < nsa; isa++) {
        ox[isa] = lround((ptsx[isa] - x0) / dx) + 1; /* one empty column/row of margin */
        oy[isa] = lround((ptsy[isa] - y0) / dx) + 1;
        if (ox[isa] + nxsa + 1 > gx1) gx1 = ox[isa] + nxsa + 1;
        if (oy[isa] + nxsa + 1 > gy1) gy1 = oy[isa] + nxsa + 1;
    }
    const long mx = gx1 + 1, my = gy1 + 1;
    long *map = (long *)malloc(sizeof(long) * mx * my);
    for (long i = 0; i < mx * my; i++) map[i] = -1;
    for (long isa = 0; isa < nsa; isa++)
        for (long iy = 0; iy < nxsa; iy++)
            for (long ix = 0; ix < nxsa; ix++) map[(ox[isa] + ix) + (oy[isa] + iy) * mx] = isa * npsa + ix + iy * nxsa;
    const long nl = nxsa + 2; /* neighbourhood of a subaperture: its grid and a ring */
    double *wx = (double *)malloc(sizeof(double) * nl * nl), *wy = (double *)malloc(sizeof(double) * nl * nl);
    long nnz = 0;
    int overflow = 0;
    for (int pass = 0; pass < 2; pass++) {
        for (long isa = 0; isa < nsa; isa++) {
            const long row = pass * nsa + isa;
            rowp[row] = nnz;
            if (saa && saat > 0 && saa[isa] < saat) continue;
            memset(wx, 0, sizeof(double) * nl * nl);
            memset(wy, 0, sizeof(double) * nl * nl);
            double asum = 0;
            /* the square in grid coordinates: points at ox .. ox + nxsa - 1,
             * edges half a step outside them */
            const double s1x = ox[isa] - 0.5, s2x = ox[isa] + nxsa - 0.5;
            const double s1y = oy[isa] - 0.5, s2y = oy[isa] + nxsa - 0.5;
            for (long cy = oy[isa] - 1; cy < oy[isa] + nxsa; cy++) {
                const double v1 = fmax(s1y - cy, 0.), v2 = fmin(s2y - cy, 1.);
                if (v2 - v1 <= 0) continue;
                for (long cx = ox[isa] - 1; cx < ox[isa] + nxsa; cx++) {
                    const double u1 = fmax(s1x - cx, 0.), u2 = fmin(s2x - cx, 1.);
                    if (u2 - u1 <= 0) continue;
                    /* corners 00, 10, 01, 11 */
                    long q[4];
                    double a[4];
                    int ok = 1;
                    for (int c = 0; c < 4; c++) {
                        q[c] = map[(cx + (c & 1)) + (cy + (c >> 1)) * mx];
                        a[c] = q[c] < 0 ? 0. : (amp ? amp[q[c]] : 1.);
                        ok = ok && fabs(a[c]) > MKG_TOL;
                    }
                    if (!ok) continue;
                    double iu[5], iv[5];
                    ints1(u1, u2, iu);
                    ints1(v1, v2, iv);
                    /* A = a00 (1-u)(1-v) + a10 u(1-v) + a01 (1-u)v + a11 uv */
                    /* int A (1-v): x weight of the lower corners (- left, + right) */
                    const double Iu0 = iu[0], Iu1 = iu[1];
                    const double Av0 = (a[0] * Iu0 + a[1] * Iu1) * iv[2] + (a[2] * Iu0 + a[3] * Iu1) * iv[3];
                    /* int A v: x weight of the upper corners */
                    const double Av1 = (a[0] * Iu0 + a[1] * Iu1) * iv[3] + (a[2] * Iu0 + a[3] * Iu1) * iv[4];
                    /* int A (1-u), int A u: y weights of the left and right corners */
                    const double Iv0 = iv[0], Iv1 = iv[1];
                    const double Au0 = (a[0] * Iv0 + a[2] * Iv1) * iu[2] + (a[1] * Iv0 + a[3] * Iv1) * iu[3];
                    const double Au1 = (a[0] * Iv0 + a[2] * Iv1) * iu[3] + (a[1] * Iv0 + a[3] * Iv1) * iu[4];
                    asum += (a[0] * Iu0 + a[1] * Iu1) * Iv0 + (a[2] * Iu0 + a[3] * Iu1) * Iv1;
                    const long lx = cx - (ox[isa] - 1), ly = cy - (oy[isa] - 1);
                    const long k00 = lx + ly * nl, k10 = k00 + 1, k01 = k00 + nl, k11 = k01 + 1;
                    wx[k00] -= Av0;
                    wx[k10] += Av0;
                    wx[k01] -= Av1;
                    wx[k11] += Av1;
                    wy[k00] -= Au0;
                    wy[k01] += Au0;
                    wy[k10] -= Au1;
                    wy[k11] += Au1;
                }
            }
            const double sc = asum > 0.01 * (double)npsa ? 1. / (asum * dx) : 0.;
            const double *w = pass == 0 ? wx : wy;
            for (long ly = 0; ly < nl; ly++)
                for (long lx = 0; lx < nl; lx++) {
                    const double v = w[lx + ly * nl] * sc;
                    if (!(fabs(v) > MKG_TOL)) continue;
                    const long pt = map[(ox[isa] - 1 + lx) + (oy[isa] - 1 + ly) * mx];
                    if (pt < 0) continue; /* cannot happen: weights live on existing points */
                    if (nnz >= nzmax) {
                        overflow = 1;
                        continue;
                    }
                    coli[nnz] = pt;
                    val[nnz] = v;
                    nnz++;
                }
        }
    }
    rowp[2 * nsa] = nnz;
    free(map);
    free(ox);
    free(oy);
    free(wx);
    free(wy);
    return overflow ? -2 : nnz;
}
