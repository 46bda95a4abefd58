/*
 * Model i0, gx, gy of a Shack-Hartmann powfs from the turbulence-averaged
 * short-exposure OTF: the reference's genseotf -> gensepsf -> gensei chain
 * (lib/gensei.c:42-419, lib/genotf.c:50-300, called from
 * maos/powfs.c:1545-1622), from which MAOS builds its matched filter at
 * setup.  FP64, OpenMP over subapertures.
 *
 * The von Karman covariance is restated from its closed form with this
 * file's own modified Bessel function K_nu (trapezoid rule on
 * K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt, which converges
 * geometrically fast for this integrand).
 */
#include "../../include/maos_setup.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* K_nu(x), x > 0: trapezoid rule with step h on [0, T], T where
 * x cosh T exceeds 745 (the integrand is below the smallest double). */
static double bessel_k(double nu, double x) {
    const double h = 0.05;
    const double tmax = acosh(fmax(745. / x, 1.)) + h;
    double s = 0.5; /* t = 0: exp(-x) cosh 0, halved, scaled by exp(x) below */
    const double ex = exp(-x);
    long n = (long)ceil(tmax / h);
    for (long k = 1; k <= n; k++) {
        const double t = k * h;
        s += exp(-x * (cosh(t) - 1.)) * cosh(nu * t);
    }
    return s * h * ex;
}

/* OPD covariance (m^2) of von Karman turbulence at separation r (turbcov,
 * lib/turbulence.c:400-431): B(r) = c (L0/r0)^(5/3) (2 pi r/L0)^(5/6)
 * K_5/6(2 pi r/L0), c = Gamma(11/6) (24/5 Gamma(6/5))^(5/6)
 * / (2^(5/6) pi^(8/3)) (0.5e-6/(2 pi))^2 (phase at 0.5 um to OPD), and
 * B(0) = c (L0/r0)^(5/3) 2^(-1/6) Gamma(5/6).  L0 = inf: Kolmogorov,
 * B(r) = -3.44 (0.5e-6/(2 pi))^2 r0^(-5/3) r^(5/3) (half the structure
 * function below the variance at rmax = 0). */
int mxs_turbcov(const double *r, long n, double r0, double L0, double *cov) {
    if (!r || !cov || n < 0 || !(r0 > 0) || !(L0 > 0)) return -1;
    const double k05 = 2. * M_PI / 0.5e-6;
    const double tg1 = tgamma(11. / 6.) * pow(24. / 5. * tgamma(6. / 5.), 5. / 6.) * pow(k05, -2) /
                       pow(M_PI, 8. / 3.);
    if (isinf(L0)) {
        const double coeff = 0.5 * 6.88 * pow(k05, -2) * pow(r0, -5. / 3.);
        for (long i = 0; i < n; i++) cov[i] = -coeff * pow(r[i], 5. / 3.);
        return 0;
    }
    const double f0 = 1. / L0, r0f0p = pow(r0 * f0, -5. / 3.);
    const double c = tg1 / pow(2., 5. / 6.), c0 = tg1 * tgamma(5. / 6.) / 2.;
    for (long i = 0; i < n; i++) {
        if (fabs(r[i]) < 1e-15) {
            cov[i] = c0 * r0f0p;
        } else {
            const double x = r[i] * 2. * M_PI * f0;
            cov[i] = c * r0f0p * pow(x, 5. / 6.) * bessel_k(5. / 6., x);
        }
    }
    return 0;
}

/* 3x3 inverse by cofactors; returns 0 when singular */
static int inv3(const double m[9], double o[9]) {
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (!(fabs(det) > 0)) return 0;
    const double id = 1. / det;
    o[0] = c00 * id;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = c01 * id;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = c02 * id;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return 1;
}

/* Remove the amp-weighted piston/tip/tilt fit from every column, then from
 * every row, of the nloc x nloc matrix B (loc_remove_ptt with cov=1,
 * math/loc.c:794-823; the fit weights are amp, loc_mcc_ptt :592-607 and
 * loc_calc_ptt_stride :653-690). */
static int remove_ptt_both(double *B, long nloc, const double *x, const double *y, const double *amp) {
    double m[9] = {0}, im[9];
    for (long i = 0; i < nloc; i++) {
        const double a = amp ? amp[i] : 1.;
        const double v[3] = {1., x[i], y[i]};
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) m[p * 3 + q] += a * v[p] * v[q];
    }
    if (!inv3(m, im)) return -1;
    for (int pass = 0; pass < 2; pass++) {
        /* pass 0: columns (contiguous), pass 1: rows (stride nloc) */
        const long step = pass == 0 ? 1 : nloc, lead = pass == 0 ? nloc : 1;
        for (long c = 0; c < nloc; c++) {
            double *v = B + c * lead;
            double s0 = 0, s1 = 0, s2 = 0;
            for (long i = 0; i < nloc; i++) {
                const double w = v[i * step] * (amp ? amp[i] : 1.);
                s0 += w;
                s1 += w * x[i];
                s2 += w * y[i];
            }
            const double c0 = im[0] * s0 + im[1] * s1 + im[2] * s2;
            const double c1 = im[3] * s0 + im[4] * s1 + im[5] * s2;
            const double c2 = im[6] * s0 + im[7] * s1 + im[8] * s2;
            for (long i = 0; i < nloc; i++) v[i * step] -= c0 + c1 * x[i] + c2 * y[i];
        }
    }
    return 0;
}

/* OTF of one subaperture (genotf_do, lib/genotf.c:50-111): with the
 * covariance B' (p/t/t removed when pttr), E = exp(k^2 B'), d_i =
 * E_ii^(-1/2), OTF(u) = sum over point pairs (i, j = i + u) of
 * d_i d_j E_ji a_i a_j, divided by sum a^2.  u = (im - npsf/2, jm - npsf/2)
 * in grid steps (peak in the centre); the subaperture grid is n x n. */
static int otf_one(cplx *otf, long n, const double *B0, const double *x, const double *y, const double *amp,
                   double wvl, long npsf, int pttr, double *E, double *d) {
    const long nloc = n * n;
    memcpy(E, B0, sizeof(double) * nloc * nloc);
    if (pttr && remove_ptt_both(E, nloc, x, y, amp)) return -1;
    const double k2 = pow(2. * M_PI / wvl, 2);
    for (long i = 0; i < nloc * nloc; i++) E[i] = exp(k2 * E[i]);
    for (long i = 0; i < nloc; i++) d[i] = pow(E[i + i * nloc], -0.5);
    double norm = 0;
    for (long i = 0; i < nloc; i++) norm += amp ? amp[i] * amp[i] : 1.;
    norm = 1. / norm;
    const long h = npsf / 2;
    for (long jm = 0; jm < npsf; jm++) {
        const long uy = jm - h;
        for (long im = 0; im < npsf; im++) {
            const long ux = im - h;
            double s = 0;
            for (long iy = 0; iy < n; iy++) {
                const long jy = iy + uy;
                if (jy < 0 || jy >= n) continue;
                for (long ix = 0; ix < n; ix++) {
                    const long jx = ix + ux;
                    if (jx < 0 || jx >= n) continue;
                    const long i1 = ix + iy * n, i2 = jx + jy * n;
                    double t = d[i1] * d[i2] * E[i2 + i1 * nloc];
                    if (amp) t *= amp[i1] * amp[i2];
                    s += t;
                }
            }
            otf[im + jm * npsf] = s * norm;
        }
    }
    return 0;
}

int mxs_genotf(double *otf_out, long nxsa, double dx, const double *amp, const double *area, long nsa,
               double thres, double wvl, const double *cov, double r0, double L0, long npsf, int pttr) {
    if (!otf_out || nxsa < 1 || nsa < 1 || npsf < 1 || !(dx > 0) || !(wvl > 0)) return -1;
    const long nloc = nxsa * nxsa, np2 = npsf * npsf;
    /* the grid (mksqloc_auto: the p/t/t fit does not depend on its origin) */
    double *x = (double *)malloc(sizeof(double) * nloc), *y = (double *)malloc(sizeof(double) * nloc);
    for (long iy = 0; iy < nxsa; iy++)
        for (long ix = 0; ix < nxsa; ix++) {
            x[ix + iy * nxsa] = (ix - nxsa / 2) * dx;
            y[ix + iy * nxsa] = (iy - nxsa / 2) * dx;
        }
    /* covariance by separation (|dix|, |diy|), then the nloc x nloc matrix */
    double *rs = (double *)malloc(sizeof(double) * nloc), *cs = (double *)malloc(sizeof(double) * nloc);
    for (long iy = 0; iy < nxsa; iy++)
        for (long ix = 0; ix < nxsa; ix++) rs[ix + iy * nxsa] = sqrt((double)(ix * ix + iy * iy)) * dx;
    int rc = cov ? 0 : mxs_turbcov(rs, nloc, r0, L0, cs);
    double *B = (double *)malloc(sizeof(double) * nloc * nloc);
    if (cov)
        memcpy(B, cov, sizeof(double) * nloc * nloc);
    else
        for (long j = 0; j < nloc && !rc; j++)
            for (long i = 0; i < nloc; i++) {
                const long dix = labs(i % nxsa - j % nxsa), diy = labs(i / nxsa - j / nxsa);
                B[i + j * nloc] = cs[dix + diy * nxsa];
            }
    memset(otf_out, 0, sizeof(double) * 2 * np2 * nsa);
    cplx *otf = (cplx *)otf_out;
    /* subapertures with area above thres share the OTF of the largest one
     * (genotf, lib/genotf.c:263-276; as there, only when that is not the
     * first subaperture); those at or below 0.01 stay zero */
    long isafull = -1;
    if (nsa > 1 && area) {
        double maxa = 0;
        for (long isa = 0; isa < nsa; isa++)
            if (area[isa] > maxa) {
                maxa = area[isa];
                isafull = isa;
            }
    }
    cplx *full = NULL;
    if (!rc && isafull > 0) {
        full = (cplx *)malloc(sizeof(cplx) * np2);
        double *E = (double *)malloc(sizeof(double) * nloc * nloc), *d = (double *)malloc(sizeof(double) * nloc);
        rc = otf_one(full, nxsa, B, x, y, amp ? amp + isafull * nloc : NULL, wvl, npsf, pttr, E, d);
        free(E);
        free(d);
    }
    int err = rc;
#pragma omp parallel if (!rc)
    {
        double *E = (double *)malloc(sizeof(double) * nloc * nloc), *d = (double *)malloc(sizeof(double) * nloc);
#pragma omp for schedule(dynamic)
        for (long isa = 0; isa < nsa; isa++) {
            if (rc) continue;
            if (full && (!area || area[isa] > thres)) {
                memcpy(otf + isa * np2, full, sizeof(cplx) * np2);
            } else if (!area || area[isa] > 0.01) {
                if (otf_one(otf + isa * np2, nxsa, B, x, y, amp ? amp + isa * nloc : NULL, wvl, npsf, pttr, E,
                            d)) {
#pragma omp atomic write
                    err = -2;
                }
            }
        }
        free(E);
        free(d);
    }
    free(full);
    free(B);
    free(rs);
    free(cs);
    free(x);
    free(y);
    return err;
}

/* centred embed of an nin x nin array into nout x nout, cropping or
 * zero-padding (embed with theta=0, math/matmath.c:1428-1458) */
static void cembed_centre(cplx *out, long nout, const cplx *in, long nin) {
    memset(out, 0, sizeof(cplx) * nout * nout);
    const long skip = nout / 2 - nin / 2;
    const long s = nin > nout ? -skip : 0, e = nin > nout ? s + nout : nin;
    for (long iy = s; iy < e; iy++)
        for (long ix = s; ix < e; ix++) out[(ix + skip) + (iy + skip) * nout] = in[ix + iy * nin];
}

/* quadrant swap for even sizes (fftshift, math/mat.c:697-726) */
static void shift2(cplx *a, long n, cplx *tmp) {
    const long h = n / 2;
    for (long iy = 0; iy < n; iy++)
        for (long ix = 0; ix < n; ix++) tmp[((ix + h) % n) + ((iy + h) % n) * n] = a[ix + iy * n];
    memcpy(a, tmp, sizeof(cplx) * n * n);
}

int mxs_gensepsf(double *psf, const double *otf, long npsf, long nsa, const double *lotf, long nlotf,
                 const double *saa, long notf) {
    if (!psf || !otf || npsf < 2 || (npsf & 1) || nsa < 1 || notf < 1 || (lotf && nlotf < 1)) return -1;
    /* gensepsf, lib/gensei.c:136-206: the PSF is formed at the OTF's
     * sampling (npsf) and then cropped or zero-padded, centred, to notf */
    const long m2 = npsf * npsf, n2 = notf * notf;
    cplx *lo = NULL;
    if (lotf) {
        lo = (cplx *)malloc(sizeof(cplx) * m2);
        cembed_centre(lo, npsf, (const cplx *)lotf, nlotf);
    }
#pragma omp parallel
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * m2), *tmp = (cplx *)malloc(sizeof(cplx) * m2),
             *o = (cplx *)malloc(sizeof(cplx) * n2);
#pragma omp for schedule(static)
        for (long isa = 0; isa < nsa; isa++) {
            const double norm = (saa ? saa[isa] : 1.) / (double)m2;
            memcpy(a, (const cplx *)otf + isa * m2, sizeof(cplx) * m2);
            if (lo)
                for (long i = 0; i < m2; i++) a[i] *= conj(lo[i]);
            shift2(a, npsf, tmp);
            mxs_fft2((double *)a, npsf, npsf, 1);
            shift2(a, npsf, tmp);
            for (long i = 0; i < m2; i++) a[i] = creal(a[i]) * norm;
            cembed_centre(o, notf, a, npsf);
            for (long i = 0; i < n2; i++) psf[isa * n2 + i] = creal(o[i]);
        }
        free(a);
        free(tmp);
        free(o);
    }
    free(lo);
    return 0;
}

/* thresholded centre of gravity with background removed (dcog,
 * math/matmath.c:717-761, no mask, no flux override) */
static void cog_thres(double g[2], const double *im, long n, double offx, double offy, double thres,
                      double bkgrnd) {
    double s = 0, sx = 0, sy = 0;
    for (long iy = 0; iy < n; iy++)
        for (long ix = 0; ix < n; ix++) {
            const double v = im[ix + iy * n] - bkgrnd;
            if (v > thres) {
                s += v;
                sx += v * ix;
                sy += v * iy;
            }
        }
    if (s > thres) {
        g[0] = sx / s - ((double)(n - 1) * 0.5 + offx);
        g[1] = sy / s - ((double)(n - 1) * 0.5 + offy);
    } else {
        g[0] = g[1] = 0;
    }
}

/* CSC sampling with weight alpha: out[p] += alpha sum_k si[p,k] Re(img[k])
 * (dspmulcreal, math/spmm.c:77-86) */
static void si_apply(double *out, const long *pp, const long *pi, const double *px, long ncol,
                     const cplx *img, double alpha) {
    for (long k = 0; k < ncol; k++) {
        const double v = creal(img[k]) * alpha;
        for (long z = pp[k]; z < pp[k + 1]; z++) out[pi[z]] += px[z] * v;
    }
}

int mxs_gensei(double *i0, double *gx, double *gy, const double *psf, long nsa, long notf,
               const double *etf, const double *nominal, long npix, const long *pp, const long *pi,
               const double *px, double dtheta, double wvlsig, int shift2center, int i0scale,
               const double *saa) {
    if (!psf || nsa < 1 || notf < 2 || (notf & 1) || npix < 1 || !pp || !pi || !px || !(dtheta > 0) ||
        (!etf && !nominal) || (i0scale && (!saa || !i0)))
        return -1;
    const long n2 = notf * notf, h = notf / 2;
    /* derivative factors (mkdtf, lib/mkdtf.c:167-180): -i 2 pi f du with
     * f the signed frequency index, du = 1/(dtheta notf) */
    cplx *U = (cplx *)malloc(sizeof(cplx) * notf);
    const double du = 1. / (dtheta * notf);
    for (long i = 0; i < notf; i++) U[i] = -I * 2. * M_PI * (double)(i < h ? i : i - notf) * du;
#pragma omp parallel
    {
        cplx *k = (cplx *)malloc(sizeof(cplx) * n2), *j = (cplx *)malloc(sizeof(cplx) * n2),
             *tmp = (cplx *)malloc(sizeof(cplx) * n2);
#pragma omp for schedule(static)
        for (long isa = 0; isa < nsa; isa++) {
            /* gensei, lib/gensei.c:334-412 */
            const double *p = psf + isa * n2;
            double pg[2] = {0, 0};
            if (shift2center) {
                double pmax = p[0];
                for (long i = 1; i < n2; i++) pmax = p[i] > pmax ? p[i] : pmax;
                cog_thres(pg, p, notf, 0.5, 0.5, 0.1 * pmax, 0.2 * pmax);
            }
            for (long i = 0; i < n2; i++) k[i] = p[i];
            shift2(k, notf, tmp);
            mxs_fft2((double *)k, notf, notf, -1);
            for (long i = 0; i < n2; i++) k[i] /= (double)n2;
            if (pg[0] != 0 || pg[1] != 0) {
                /* ctilt by (-pg) with the peak in the corner (ctilt2,
                 * math/matcomp.c:287-330): exp(-2 pi i s f / n) per axis */
                for (long iy = 0; iy < notf; iy++) {
                    const double fy = (double)(iy < h ? iy : iy - notf);
                    for (long ix = 0; ix < notf; ix++) {
                        const double fx = (double)(ix < h ? ix : ix - notf);
                        k[ix + iy * notf] *= cexp(-I * 2. * M_PI * (-pg[0] * fx + -pg[1] * fy) / (double)notf);
                    }
                }
            }
            if (etf) {
                const cplx *e = (const cplx *)etf + isa * n2;
                for (long i = 0; i < n2; i++) k[i] *= e[i];
            } else {
                for (long i = 0; i < n2; i++) k[i] *= nominal[i];
            }
            memcpy(j, k, sizeof(cplx) * n2);
            mxs_fft2((double *)k, notf, notf, 1);
            if (i0) si_apply(i0 + isa * npix, pp, pi, px, n2, k, wvlsig);
            if (gx || gy) {
                memcpy(k, j, sizeof(cplx) * n2);
                for (long iy = 0; iy < notf; iy++)
                    for (long ix = 0; ix < notf; ix++) {
                        k[ix + iy * notf] *= U[ix];
                        j[ix + iy * notf] *= U[iy];
                    }
                if (gx) {
                    mxs_fft2((double *)k, notf, notf, 1);
                    si_apply(gx + isa * npix, pp, pi, px, n2, k, wvlsig);
                }
                if (gy) {
                    mxs_fft2((double *)j, notf, notf, 1);
                    si_apply(gy + isa * npix, pp, pi, px, n2, j, wvlsig);
                }
            }
            if (i0scale) {
                double s = 0;
                for (long q = 0; q < npix; q++) s += i0[isa * npix + q];
                const double sc = saa[isa] / s;
                for (long q = 0; q < npix; q++) {
                    i0[isa * npix + q] *= sc;
                    if (gx) gx[isa * npix + q] *= sc;
                    if (gy) gy[isa * npix + q] *= sc;
                }
            }
        }
        free(k);
        free(j);
        free(tmp);
    }
    free(U);
    return 0;
}
