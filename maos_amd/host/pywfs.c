/*
 * Pyramid WFS setup tables (host, FP64): the focal-plane phase mask of the
 * pyramid, the detector transfer function of the pupil images, and the
 * detector subaperture grid with the offset of each pupil image.  These are
 * the tables pywfs_new builds (lib/pywfs.c:110-370); the per-step imaging
 * runs on the GPU (pywfs.hip).
 */
#include "../../include/maos_setup.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* A facet of the pyramid deflects the light by a quarter of the grid: the
 * phase rises by pi/2 per focal-plane sample away from the apex, which
 * displaces each pupil image by notf/4 samples (lib/pywfs.c:196-232).  For
 * four sides the phase is pi/2 (|x| + |y|); three sides keep one facet of
 * slope 1 along +y and the two others rotated by 120 degrees; a roof is
 * pi/2 |x|; one side (the Zernike sensor) is a pi phase dot of nstop samples.
 * The mask is zero outside the clear aperture of diameter nstop samples
 * centred on the apex (notf/2, notf/2); flat edges and a flat vertex of the
 * given half-widths (in samples) keep zero phase. */
int mxs_pywfs_pyramid(double *mask_, long notf, int nside, long nstop, double eskip, double vskip) {
    if (notf < 4 || (notf & 1) || nside < 1 || nside > 4 || nstop < 1) return -1;
    cplx *mask = (cplx *)mask_;
    memset(mask, 0, sizeof(cplx) * notf * notf);
    const long c = notf / 2;
    const long skip = notf > nstop ? (notf - nstop) / 2 : 0;
    const double r2 = 0.25 * (double)nstop * (double)nstop;
    const double s3 = sqrt(3.);
    for (long iy = skip; iy < notf - skip; iy++) {
        for (long ix = skip; ix < notf - skip; ix++) {
            const double x = fabs((double)(ix - c)), y = (double)(iy - c), ay = fabs(y);
            if (x * x + ay * ay >= r2) continue;
            double h = 0; /* facet height in quarter waves */
            if (nside == 1) {
                h = 2.;
            } else if (nside == 2) {
                h = x;
            } else if (nside == 3) {
                h = y * s3 > x ? y : 0.5 * fabs(y - x * s3);
            } else if (!(x < eskip || ay < eskip || (x < vskip && ay < vskip))) {
                h = x + ay;
            }
            mask[ix + iy * notf] = cexp(I * (0.5 * M_PI * h));
        }
    }
    return 0;
}

/* The pupil images are sampled by detector pixels of side dsa (in the pupil
 * plane, sampled at dx2): the image is convolved with the pixel box and a
 * Gaussian charge diffusion of rms pixblur*dsa, i.e. its transform is
 * multiplied by sinc(kx dsa) sinc(ky dsa) (dsa/dx2)^2
 * exp(-2 (pi pixblur dsa)^2 |k|^2) at k = j / (notf dx2), j centred, then
 * moved to FFT (corner) order (lib/pywfs.c:236-262). */
void mxs_pywfs_nominal(double *nominal, long notf, double dsa, double dx2, double pixblur) {
    const long c = notf / 2;
    const double du = 1. / (dx2 * (double)notf);
    const double dupix = dsa * du;
    const double area = (dsa / dx2) * (dsa / dx2);
    const double blur = pixblur * dsa;
    const double e0 = -2. * (M_PI * blur * du) * (M_PI * blur * du);
    for (long iy = 0; iy < notf; iy++) {
        const long jy = iy - c;
        const double ay = jy ? sin(M_PI * jy * dupix) / (M_PI * jy * dupix) : 1.;
        /* corner order: centred sample (jx, jy) goes to ((jx + notf) % notf, ...) */
        const long oy = (jy + notf) % notf;
        for (long ix = 0; ix < notf; ix++) {
            const long jx = ix - c;
            const double ax = jx ? sin(M_PI * jx * dupix) / (M_PI * jx * dupix) : 1.;
            double v = ax * ay * area;
            if (blur != 0) v *= exp(e0 * (double)(jx * jx + jy * jy));
            nominal[(jx + notf) % notf + oy * notf] = v;
        }
    }
}

/* Detector subapertures: an order x order grid of pitch dsa centred on the
 * optical axis (the centres, lib/pywfs.c:281-296), and the centre of each
 * pupil image relative to the grid centre in units of notf/2 samples
 * (pywfs_mksi, lib/pywfs.c:34-107): four sides (ind = 2 iy + ix) at
 * (ix - 1/2, iy - 1/2); three sides at (0, -1/2) and ((ind - 3/2) sqrt(3)/2,
 * 1/4); a roof at (ix - 1/2, 0); one side at the centre. */
long mxs_pywfs_saloc(double *x, double *y, long maxn, long order, double dsa) {
    if (order < 1 || order * order > maxn) return -1;
    const double o = (-0.5 * (double)order + 0.5) * dsa;
    for (long iy = 0; iy < order; iy++)
        for (long ix = 0; ix < order; ix++) {
            x[ix + iy * order] = o + ix * dsa;
            y[ix + iy * order] = o + iy * dsa;
        }
    return order * order;
}

int mxs_pywfs_sioff(double *sioff, int nside) {
    if (nside < 1 || nside > 4) return -1;
    for (int ind = 0; ind < nside; ind++) {
        const int ix = ind % 2, iy = ind / 2;
        double ox = 0, oy = 0;
        if (nside == 2) {
            ox = ix - 0.5;
        } else if (nside == 3) {
            if (ind == 0) {
                oy = -0.5;
            } else {
                ox = (ind - 1.5) * sqrt(3.) * 0.5;
                oy = 0.25;
            }
        } else if (nside == 4) {
            ox = ix - 0.5;
            oy = iy - 0.5;
        }
        sioff[2 * ind] = ox;
        sioff[2 * ind + 1] = oy;
    }
    return 0;
}
