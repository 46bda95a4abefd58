/*
 * The LGS matched-filter loop around the WFS step (host, FP64): the
 * tip/tilt dither's phase-locked loop and amplitude estimation
 * (wfsgrad_dither, maos/wfsgrad.c:576-752; calc_dither_amp,
 * lib/libmisc.c:347-414) and the sodium-profile fit that rebuilds i0, gx,
 * gy from the accumulated images (sodium_fit, maos/powfs_utils.c:146-360).
 * The per-pixel sums these steps consume are made on the GPU
 * (mx_wfsgrad_dither_*); what is here is scalar loop control and a small
 * least-squares problem, which MAOS also runs on the host.
 */
#include "../../include/maos_setup.h"
#include "internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Amplitude of a dither at npoint frames per cycle in a signal sampled
 * every step (nmod x nsim, column per step, every dtrat-th column a WFS
 * frame) by synchronous detection against the dither phase
 * pi/2 + 2 pi k/npoint of frame k.  detrend: a linear drift, measured as the
 * mean change between the first and the last whole cycle, is removed first
 * (a periodic signal cancels in that difference).  combine (nmod = 2, a
 * circular tip/tilt dither): the in-phase and quadrature sums of the two
 * modes together, amplitude sqrt(I^2 + Q^2)/nframe; else per mode
 * 2 sqrt(I^2 + Q^2)/nframe into res (may be NULL if nmod = 1).  Returns the
 * (first) amplitude. */
double mxs_calc_dither_amp(double *res, const double *signal, long nmod, long nsim, long dtrat, long npoint,
                           int detrend, int combine) {
    if (!signal || nmod < 1 || nsim < 1 || dtrat < 1 || npoint < 1 || (combine && nmod != 2)) return -1;
    const long nframe = (nsim - 1) / dtrat + 1;
    const long span = (nframe / npoint - 1) * npoint; /* frames between the first and the last cycle */
    double slope[64] = {0};
    if (nmod > 64) return -1;
    if (detrend && span > 0) {
        for (long m = 0; m < nmod; m++) {
            double s = 0;
            for (long p = 0; p < npoint; p++)
                s += signal[m + (p + span) * dtrat * nmod] - signal[m + p * dtrat * nmod];
            slope[m] = s / (double)(npoint * span);
        }
    }
    const double step = 2. * M_PI / (double)npoint;
    if (combine) {
        double ip = 0, qd = 0;
        for (long f = 0; f < nframe; f++) {
            const double a = 0.5 * M_PI + step * (double)f, c = cos(a), s = sin(a);
            const double x = signal[f * dtrat * 2] - slope[0] * f, y = signal[f * dtrat * 2 + 1] - slope[1] * f;
            ip += x * c + y * s;
            qd += x * s - y * c;
        }
        const double amp = sqrt(ip * ip + qd * qd) / (double)nframe;
        if (res) res[0] = amp;
        return amp;
    }
    double first = 0;
    for (long m = 0; m < nmod; m++) {
        double ip = 0, qd = 0;
        for (long f = 0; f < nframe; f++) {
            const double a = 0.5 * M_PI + step * (double)f;
            const double v = signal[m + f * dtrat * nmod] - slope[m] * f;
            ip += v * cos(a);
            qd += v * sin(a);
        }
        const double amp = 2. * sqrt(ip * ip + qd * qd) / (double)nframe;
        if (res) res[m] = amp;
        if (m == 0) first = amp;
    }
    return first;
}

/* One WFS step of the tip/tilt dither (maos/wfsgrad.c:584-606): from step
 * pllskip on, the measured residual tip/tilt fsmerr projected on the
 * quadrature of the dither phase, err = (-ss e_x + cs e_y)/amp, moves the
 * loop's delay estimate by gpll err / pllrat.  The phase (cs, ss) of the
 * step (dither_position with the last PLL output deltam) is returned in
 * both cases; the caller hands it to mx_wfsgrad_set_dither. */
int mxs_dither_step(mxs_dither_t *d, long isim, const double *fsmerr, double *cs, double *ss) {
    if (!d || !cs || !ss || d->npoint < 1 || d->pllrat < 1 || d->dtrat < 1) return -1;
    mxs_dither_position(cs, ss, d->alfsm, d->dtrat, d->npoint, isim, d->deltam);
    if (isim < d->pllskip) return 0;
    if (!fsmerr || !(d->amp != 0)) return -1;
    const double err = (-(*ss) * fsmerr[0] + (*cs) * fsmerr[1]) / d->amp;
    d->delta += d->gpll * (err / (double)d->pllrat);
    return 0;
}

/* A PLL output (maos/wfsgrad.c:638-653 and 707-731): the phase used from
 * now on, deltam = delta + gdrift deltao; the commanded dither amplitude
 * a2m from the FSM command history and the measured one a2me from the FSM
 * error history (2 x ncol each, ncol = (pllrat - 1) dtrat + 1 steps ending
 * at this one; detrended for NGS, d->detrend); and the scales of the i0 and
 * gx/gy folds, s0 = 1/pllrat and sg = 2/(pllrat a2m) (0 while a2m is 0),
 * which mx_wfsgrad_dither_fold takes. */
int mxs_dither_output(mxs_dither_t *d, const double *fsmcmds, const double *fsmerrs, long ncol, double *s0,
                      double *sg) {
    if (!d || !fsmcmds || !fsmerrs || ncol < 1) return -1;
    d->deltam = d->delta + d->deltao * d->gdrift;
    d->a2m = mxs_calc_dither_amp(NULL, fsmcmds, 2, ncol, d->dtrat, d->npoint, d->detrend, 1);
    d->a2me = mxs_calc_dither_amp(NULL, fsmerrs, 2, ncol, d->dtrat, d->npoint, d->detrend, 1);
    if (d->a2m < 0 || d->a2me < 0) return -1;
    if (s0) *s0 = 1. / (double)d->pllrat;
    if (sg) *sg = d->a2m != 0 ? 2. / ((double)d->pllrat * d->a2m) : 0;
    return 0;
}

/* The dither's own tip/tilt, removed from the FSM error and the gradients
 * so that it does not perturb the loop (maos/wfsgrad.c:734-751):
 * tt = -(cs, ss) a2me. */
void mxs_dither_tt(const mxs_dither_t *d, double cs, double ss, double *tt) {
    tt[0] = -cs * d->a2me;
    tt[1] = -ss * d->a2me;
}

/* Sodium profile from the accumulated images (sodium_fit with one
 * iteration, the matched-filter path, maos/powfs_utils.c:146-360, with no
 * gradient offsets).  The model image of each height bin
 * h_k = 80 km + k dh (up to 105 km) is the image (gensei) through the ETF
 * of a one-sample profile at h_k (mketf's direct sum); the data i0in is a
 * non-negative combination of them.  The profile is the least-squares fit
 * with the normal matrix inverted by its eigen-decomposition, keeping
 * eigenvalues above svdthres times the largest and above 1e-3 times the
 * previous kept one (svd_pow2, math/blas.c:441-485), scaled to sum 1.
 * Then, if i0, gx, gy are given, they are replaced by the model of the
 * fitted profile.  Layouts: i0in, i0, gx, gy [ni0][nsa][npix]; sepsf
 * [nwvl][nsa][notf^2] (mxs_gensepsf); nominal [nwvl][notf^2]; si of each
 * wavelength [nwvl][notf^2 + 1] / [nwvl][si_nzmax] (mxs_si); siglev [ni0].
 * prof receives nh x 2 (heights, then strengths).  Returns nh, or -1 for
 * bad arguments, -2 if maxnh is too small. */
long mxs_sodium_fit(double *prof, long maxnh, double *i0, double *gx, double *gy, const double *i0in, long ni0,
                    long nsa, long notf, long npix, long nwvl, const double *sepsf, const double *nominal,
                    const long *si_pp, const long *si_pi, const double *si_px, long si_nzmax,
                    const double *dtheta, const double *wvlwts, const double *siglev, const double *saa,
                    const double *srot, const double *srsa, double dh, double hs, double htel, double za,
                    double svdthres) {
    const double hmin = 80000., hmax = 105000.;
    if (!prof || !i0in || ni0 < 1 || nsa < 1 || notf < 2 || npix < 1 || nwvl < 1 || !sepsf || !nominal ||
        !si_pp || !si_pi || !si_px || !dtheta || !wvlwts || !siglev || !srot || !srsa || !(dh > 0))
        return -1;
    const long nh = (long)floor((hmax - hmin) / dh) + 1;
    if (nh > maxnh) return -2;
    const long n2 = notf * notf, m = ni0 * nsa * npix;
    double *A = (double *)calloc((size_t)m * nh, sizeof(double)); /* column k: the model of bin k */
    double *etf = (double *)malloc(sizeof(double) * 2 * nsa * n2);
    int rc = 0;
    /* the model of each height bin: a one-sample profile through mketf's direct sum */
    for (long k = 0; k < nh && !rc; k++) {
        const double h1[2] = {hmin + k * dh, 1.};
        for (long iw = 0; iw < nwvl && !rc; iw++) {
            if (mxs_mketf(etf, NULL, nominal + iw * n2, notf, dtheta[iw], h1, 1, srot, srsa, nsa, hs, htel, za, 1))
                rc = -1;
            for (long j = 0; j < ni0 && !rc; j++)
                if (mxs_gensei(A + k * m + j * nsa * npix, NULL, NULL, sepsf + iw * nsa * n2, nsa, notf, etf, NULL,
                               npix, si_pp + iw * (n2 + 1), si_pi + iw * si_nzmax, si_px + iw * si_nzmax, dtheta[iw],
                               wvlwts[iw] * siglev[j], 0, 0, saa))
                    rc = -1;
        }
    }
    if (rc) {
        free(etf);
        free(A);
        return -1;
    }
    /* normal equations A^T A x = A^T b, pseudo-inverse by eigen-decomposition */
    double *ata = (double *)malloc(sizeof(double) * nh * nh), *atb = (double *)malloc(sizeof(double) * nh);
    double *V = (double *)malloc(sizeof(double) * nh * nh), *d = (double *)malloc(sizeof(double) * nh);
    long *ord = (long *)malloc(sizeof(long) * nh);
#pragma omp parallel for schedule(dynamic, 4)
    for (long a = 0; a < nh; a++) {
        for (long b = a; b < nh; b++) {
            double s = 0;
            for (long i = 0; i < m; i++) s += A[a * m + i] * A[b * m + i];
            ata[a + b * nh] = ata[b + a * nh] = s;
        }
        double s = 0;
        for (long i = 0; i < m; i++) s += A[a * m + i] * i0in[i];
        atb[a] = s;
    }
    mxs_sym_eig(ata, (int)nh, d, V);
    for (long i = 0; i < nh; i++) ord[i] = i;
    for (long i = 1; i < nh; i++) /* insertion sort, descending eigenvalues */
        for (long j = i; j > 0 && d[ord[j]] > d[ord[j - 1]]; j--) {
            const long t = ord[j];
            ord[j] = ord[j - 1];
            ord[j - 1] = t;
        }
    const double dmax = d[ord[0]];
    double t1 = fabs(svdthres) * dmax, t2 = 1e-3 * dmax;
    double *x = prof + nh;
    for (long i = 0; i < nh; i++) {
        prof[i] = hmin + i * dh;
        x[i] = 0;
    }
    for (long r = 0; r < nh; r++) {
        const long e = ord[r];
        if (!((!svdthres || d[e] > t1) && d[e] > t2)) break;
        t2 = d[e] * 1e-3;
        double c = 0;
        for (long i = 0; i < nh; i++) c += V[i + e * nh] * atb[i];
        c /= d[e];
        for (long i = 0; i < nh; i++) x[i] += c * V[i + e * nh];
    }
    double sum = 0;
    for (long i = 0; i < nh; i++) sum += x[i];
    if (sum != 0)
        for (long i = 0; i < nh; i++) x[i] /= sum;
    /* i0, gx, gy of the fitted profile */
    if (i0 || gx || gy) {
        for (long i = 0; i < m; i++) {
            if (i0) i0[i] = 0;
            if (gx) gx[i] = 0;
            if (gy) gy[i] = 0;
        }
        for (long iw = 0; iw < nwvl && !rc; iw++) {
            if (mxs_mketf(etf, NULL, nominal + iw * n2, notf, dtheta[iw], prof, nh, srot, srsa, nsa, hs, htel, za, 1))
                rc = -1;
            for (long j = 0; j < ni0 && !rc; j++) {
                const long o = j * nsa * npix;
                if (mxs_gensei(i0 ? i0 + o : NULL, gx ? gx + o : NULL, gy ? gy + o : NULL, sepsf + iw * nsa * n2, nsa,
                               notf, etf, NULL, npix, si_pp + iw * (n2 + 1), si_pi + iw * si_nzmax,
                               si_px + iw * si_nzmax, dtheta[iw], wvlwts[iw] * siglev[j], 0, 0, saa))
                    rc = -1;
            }
        }
    }
    free(ord);
    free(d);
    free(V);
    free(atb);
    free(ata);
    free(etf);
    free(A);
    return rc ? -1 : nh;
}
