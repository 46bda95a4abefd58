/*
 * FP64 CPU oracle for the MAOS Shack-Hartmann physical-optics WFS step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): a restatement of the reference
 * algorithm used as the checker and as the timed CPU baseline.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
#define ORC_EPS 1.e-15 /* sys/common.h:58-63, double build */

/* ---------------------------------------------------------------------
 * FFT: cached radix-2 plans (bit reversal + twiddles); other sizes by a
 * recursive mixed-radix Cooley-Tukey decimation in time (smallest prime
 * factor first, a direct DFT for a prime length).  Sign and scaling follow
 * FFTW as used by cfft2 (math/fft.c:253-281): dir=-1 is exp(-i 2 pi jk/n),
 * no normalisation.
 * ------------------------------------------------------------------- */
typedef struct {
    long n;
    long *rev;
    cplx *tw; /* exp(-i 2 pi k/n), k<n/2 */
} plan_t;
#define NPLAN 24
static plan_t plans[NPLAN];
static int nplans = 0;

static const plan_t *get_plan(long n) {
    for (int i = 0; i < nplans; i++)
        if (plans[i].n == n) return &plans[i];
    const plan_t *out = NULL;
#pragma omp critical(orc_plan)
    {
        for (int i = 0; i < nplans; i++)
            if (plans[i].n == n) out = &plans[i];
        if (!out && nplans < NPLAN) {
            plan_t p;
            p.n = n;
            p.rev = (long *)malloc(sizeof(long) * n);
            p.tw = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
            int lg = 0;
            while ((1L << lg) < n) lg++;
            for (long i = 0; i < n; i++) {
                long r = 0;
                for (int b = 0; b < lg; b++)
                    if (i & (1L << b)) r |= 1L << (lg - 1 - b);
                p.rev[i] = r;
            }
            for (long k = 0; k <= n / 2; k++) {
                double a = -2. * M_PI * (double)k / (double)n;
                p.tw[k] = cos(a) + I * sin(a);
            }
            plans[nplans] = p;
            __atomic_thread_fence(__ATOMIC_SEQ_CST);
            out = &plans[nplans];
            nplans++;
        }
    }
    return out;
}

static void fft_pow2(cplx *x, long n, int dir, const plan_t *p) {
    for (long i = 0; i < n; i++) {
        long j = p->rev[i];
        if (i < j) {
            cplx t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (long len = 2; len <= n; len <<= 1) {
        long h = len >> 1, step = n / len;
        for (long i0 = 0; i0 < n; i0 += len) {
            for (long j = 0; j < h; j++) {
                cplx w = p->tw[j * step];
                if (dir > 0) w = conj(w);
                cplx u = x[i0 + j], v = x[i0 + j + h] * w;
                x[i0 + j] = u + v;
                x[i0 + j + h] = u - v;
            }
        }
    }
}

/* out[k] = sum_j in[j*stride] w^(j k) for the length n = ntop/S sub-transform;
 * w^t = tw[(t*S) mod ntop] with tw[t] = exp(dir i 2 pi t/ntop).  scratch
 * holds n values. */
static void fft_mixed(const cplx *in, long stride, cplx *out, long n, long S, long ntop, const cplx *tw,
                      cplx *scratch) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    long p = n;
    for (long f = 2; f * f <= n; f++)
        if (n % f == 0) {
            p = f;
            break;
        }
    if (p == n) { /* prime length: direct DFT */
        for (long k = 0; k < n; k++) {
            cplx s = 0;
            for (long j = 0; j < n; j++) s += in[j * stride] * tw[((j * k) % n) * S % ntop];
            scratch[k] = s;
        }
        memcpy(out, scratch, sizeof(cplx) * n);
        return;
    }
    const long m = n / p;
    for (long r = 0; r < p; r++) fft_mixed(in + r * stride, stride * p, out + r * m, m, S * p, ntop, tw, scratch);
    /* X[k + q m] = sum_r w^(r (k + q m)) Y_r[k] */
    for (long k = 0; k < m; k++)
        for (long q = 0; q < p; q++) {
            cplx s = 0;
            for (long r = 0; r < p; r++) s += out[r * m + k] * tw[((r * (k + q * m)) % n) * S % ntop];
            scratch[k + q * m] = s;
        }
    memcpy(out, scratch, sizeof(cplx) * n);
}

/* work holds 4 n values */
static void fft_any(cplx *x, long n, int dir, cplx *work) {
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        fft_pow2(x, n, dir, get_plan(n));
        return;
    }
    cplx *tw = work, *out = work + n, *scratch = work + 2 * n;
    for (long t = 0; t < n; t++) {
        double a = dir * 2. * M_PI * (double)t / (double)n;
        tw[t] = cos(a) + I * sin(a);
    }
    fft_mixed(x, 1, out, n, 1, n, tw, scratch);
    memcpy(x, out, sizeof(cplx) * n);
}

void orc_fft2(double *c, long nx, long ny, int dir) {
    cplx *a = (cplx *)c;
    long nmax = nx > ny ? nx : ny;
    cplx stackbuf[5 * 1024];
    cplx *work = nmax <= 1024 ? stackbuf : (cplx *)malloc(sizeof(cplx) * 5 * nmax);
    cplx *col = work + 4 * nmax;
    for (long iy = 0; iy < ny; iy++) fft_any(a + iy * nx, nx, dir, work);
    if (ny > 1) {
        for (long ix = 0; ix < nx; ix++) {
            for (long iy = 0; iy < ny; iy++) col[iy] = a[ix + iy * nx];
            fft_any(col, ny, dir, work);
            for (long iy = 0; iy < ny; iy++) a[ix + iy * nx] = col[iy];
        }
    }
    if (work != stackbuf) free(work);
}

/* ---------------------------------------------------------------------
 * Ray tracing: lib/accphi_grid_do.c
 * ------------------------------------------------------------------- */
static inline long ifloor_(double x) { return (long)floor(x); }
static inline long iceil_(double x) { return (long)ceil(x); }

/* prop_grid_block for wrap=1, TRANSPOSE=0 (lib/accphi_grid_do.c:49-302) */
static long prop_block_wrap(const double *phiin, long nxin, long nyin, double *phiout,
                            long nxout, long nyout, double dxout, double dyout,
                            double oxout, double oyout, double alpha) {
    const long wrapx = nxin - 1, wrapy = nyin - 1;
    if (fabs(dxout - 1.) < ORC_EPS && fabs(dyout - 1.) < ORC_EPS) {
        /* equal sampling: one set of weights for the whole block (:59-146) */
        double dplocx = oxout, dplocy = oyout;
        dplocx -= floor(dplocx / (double)nxin) * nxin;
        long nplocx = ifloor_(dplocx);
        dplocx -= nplocx;
        long rowdiv0 = wrapx - nplocx;
        if (rowdiv0 > nxout) rowdiv0 = nxout;
        if (rowdiv0 < 0) rowdiv0 = 0;
        dplocy -= floor(dplocy / (double)nyin) * nyin;
        long nplocy0 = ifloor_(dplocy);
        dplocy -= nplocy0;
        const double bl = dplocx * dplocy;
        const double br = (1. - dplocx) * dplocy;
        const double tl = dplocx * (1 - dplocy);
        const double tr = (1. - dplocx) * (1 - dplocy);
        const double *pin = phiin + nplocx;
        for (long icol = 0; icol < nyout; icol++) {
            double *phiout2 = phiout + icol * nxout;
            long nplocy = nplocy0 + icol;
            while (nplocy >= nyin) nplocy -= nyin;
            const double *phicol = pin + nplocy * nxin;
            const double *phicol2 = pin + (nplocy == wrapy ? 0 : (nplocy + 1)) * nxin;
            long irow = 0, rowdiv = rowdiv0;
            do {
                for (; irow < rowdiv; irow++)
                    phiout2[irow] += alpha * (bl * phicol2[irow + 1] + br * phicol2[irow] +
                                              tl * phicol[irow + 1] + tr * phicol[irow]);
                if (irow < nxout) {
                    phiout2[irow] += alpha * (bl * phicol2[irow - wrapx] + br * phicol2[irow] +
                                              tl * phicol[irow - wrapx] + tr * phicol[irow]);
                    irow++;
                }
                rowdiv = irow + wrapx;
                phicol2 -= nxin;
                phicol -= nxin;
                if (rowdiv > nxout) rowdiv = nxout;
            } while (irow < nxout);
        }
        return 0;
    }
    /* unequal sampling with cached x interpolation (:147-302) */
    const double dxout1 = 1. / dxout;
    oyout -= floor(oyout / (double)nyin) * nyin;
    oxout -= floor(oxout / (double)nxin) * nxin;
    long rowdiv0 = iceil_((wrapx - oxout) * dxout1);
    if (rowdiv0 <= 0) rowdiv0 = 0;
    else if (rowdiv0 > nxout) rowdiv0 = nxout;
    const double dplocx00 = oxout;
    const double xover = 1;
    double *cd = (double *)malloc(sizeof(double) * nxout);
    long *cn = (long *)malloc(sizeof(long) * nxout * 2);
    long *co = cn + nxout;
    {
        long rowdiv = rowdiv0, irow = 0;
        double dplocx = dplocx00;
        do {
            for (; irow < rowdiv; irow++, dplocx += dxout) {
                long n0 = ifloor_(dplocx);
                cd[irow] = dplocx - n0;
                cn[irow] = n0;
                co[irow] = 1;
            }
            for (; irow < nxout && dplocx < wrapx + xover; irow++, dplocx += dxout) {
                long n0 = ifloor_(dplocx);
                cd[irow] = dplocx - n0;
                cn[irow] = n0;
                co[irow] = -wrapx;
            }
            dplocx -= nxin;
            rowdiv = irow + iceil_((wrapx - dplocx) * dxout1);
            if (rowdiv > nxout) rowdiv = nxout;
        } while (irow < nxout);
    }
    for (long icol = 0; icol < nyout; icol++) {
        double *phiout2 = phiout + icol * nxout;
        double dplocy = fma((double)icol, dyout, oyout);
        while (dplocy >= nyin) {
            dplocy -= nyin;
            oyout -= nyin;
        }
        long nplocy = ifloor_(dplocy);
        double dplocy0 = dplocy - nplocy;
        const double *phicol = phiin + nplocy * nxin;
        const double *phicol2 = phiin + (nplocy == wrapy ? 0 : (nplocy + 1)) * nxin;
        const double dplocy1 = 1. - dplocy0;
        for (long irow = 0; irow < nxout; irow++) {
            double dplocx0 = cd[irow];
            long n0 = cn[irow], off = co[irow];
            phiout2[irow] += alpha * (dplocy1 * ((1. - dplocx0) * phicol[n0] + dplocx0 * phicol[n0 + off]) +
                                      dplocy0 * ((1. - dplocx0) * phicol2[n0] + dplocx0 * phicol2[n0 + off]));
        }
    }
    free(cd);
    free(cn);
    return 0;
}

long orc_prop_grid_pts(const double *phiin, long nxin, long nyin, double dxin,
                       double dyin, double oxin, double oyin, const double *ptsx,
                       const double *ptsy, long nsa, long nxsa, double ptsdx,
                       double *phiout, double alpha, double displacex,
                       double displacey, double scale, int wrap, long start, long end) {
    const double dx_in1 = 1. / dxin, dy_in1 = 1. / dyin;
    if (!wrap) {
        /* the DM path (wrap=0, maos/sim_utils.c:1042): prop_grid_block's
         * unequal-sampling rules (lib/accphi_grid_do.c:147-302, the cached
         * form used for grids of more than 10 rows): a point is used when its
         * row lies in [irows, rowdiv0) (x in [0, nxin-1)) and its column in
         * [icols, coldiv) (y in [0, nyin-1]); others are left alone */
        const double dxout = ptsdx * dx_in1 * scale, dyout = ptsdx * dy_in1 * scale;
        const double dxout1 = 1. / dxout, dyout1 = 1. / dyout;
        const double dispx = (displacex - oxin) * dx_in1, dispy = (displacey - oyin) * dy_in1;
        const long wrapx = nxin - 1, wrapy = nyin - 1;
        long missing = 0;
        if (!end) end = nsa;
        for (long isa = start; isa < end; isa++) {
            const double oxout = ptsx[isa] * dx_in1 * scale + dispx;
            const double oyout = ptsy[isa] * dy_in1 * scale + dispy;
            const long icols = oyout < 0 ? (long)ceil(-oyout * dyout1) : 0;
            long nyout = nxsa;
            const long coldiv = (long)floor((wrapy - oyout) * dyout1) + 1;
            if (coldiv < nyout) nyout = coldiv;
            const long irows = oxout < 0 ? (long)ceil(-oxout * dxout1) : 0;
            long rowdiv0 = (long)ceil((wrapx - oxout) * dxout1);
            if (rowdiv0 > nxsa) rowdiv0 = nxsa;
            double *out = phiout + isa * nxsa * nxsa;
            for (long icol = icols; icol < nyout; icol++) {
                const double y = fma((double)icol, dyout, oyout);
                long ny0 = (long)floor(y); /* clamps: rounding at the range edges */
                ny0 = ny0 < 0 ? 0 : (ny0 > wrapy ? wrapy : ny0);
                const double fy = y - ny0;
                const double *r0 = phiin + ny0 * nxin, *r1 = phiin + (ny0 == wrapy ? 0 : ny0 + 1) * nxin;
                for (long irow = irows; irow < rowdiv0; irow++) {
                    const double x = fma((double)irow, dxout, oxout);
                    long nx0 = (long)floor(x);
                    nx0 = nx0 < 0 ? 0 : (nx0 > wrapx - 1 ? wrapx - 1 : nx0);
                    const double fx = x - nx0;
                    out[icol * nxsa + irow] +=
                        alpha * ((1 - fy) * ((1 - fx) * r0[nx0] + fx * r0[nx0 + 1]) +
                                 fy * ((1 - fx) * r1[nx0] + fx * r1[nx0 + 1]));
                }
            }
            const long used = (nyout > icols ? nyout - icols : 0) * (rowdiv0 > irows ? rowdiv0 - irows : 0);
            missing += nxsa * nxsa - used;
        }
        return missing;
    }
    const double dxout = ptsdx * dx_in1 * scale;
    const double dyout = ptsdx * dy_in1 * scale;
    displacex = (displacex - oxin) * dx_in1;
    displacey = (displacey - oyin) * dy_in1;
    if (!end) end = nsa;
    for (long isa = start; isa < end; isa++) {
        const double oxout = ptsx[isa] * dx_in1 * scale + displacex;
        const double oyout = ptsy[isa] * dy_in1 * scale + displacey;
        prop_block_wrap(phiin, nxin, nyin, phiout + isa * nxsa * nxsa, nxsa, nxsa, dxout, dyout,
                        oxout, oyout, alpha);
    }
    return 0;
}

long orc_prop(const double *phiin, long nxin, long nyin, double dxin, double oxin,
              double oyin, double ht, const double *ptsx, const double *ptsy, long nsa,
              long nxsa, double ptsdx, double *phiout, double alpha, double hs,
              double thetax, double thetay, double misregx, double misregy, double shiftx,
              double shifty, int wrap) {
    /* prop_range, lib/accphi.c:223-244 (hc=0 for pts output, dratio=0) */
    if (!alpha) alpha = 1;
    if (hs && ht > hs) return -1;
    const double scale = hs ? (1. - ht / hs) : 1.;
    const double displacex = ht * thetax + scale * misregx + shiftx;
    const double displacey = ht * thetay + scale * misregy + shifty;
    return orc_prop_grid_pts(phiin, nxin, nyin, dxin, dxin, oxin, oyin, ptsx, ptsy, nsa, nxsa,
                             ptsdx, phiout, alpha, displacex, displacey, scale, wrap, 0, nsa);
}

/* ---------------------------------------------------------------------
 * Science PSF: lib/locfft.c:90-188
 * ------------------------------------------------------------------- */
int orc_locfft_psf(double *psf, const double *opd, const double *amp, long nloc, const long *embed,
                   long nembed, long psfsize, double wvl) {
    const double wvk = 2. * M_PI / wvl;
    if (psfsize == 1) {
        cplx s = 0;
        for (long i = 0; i < nloc; i++) s += amp[i] * cexp(I * wvk * opd[i]);
        psf[0] = creal(s) * creal(s) + cimag(s) * cimag(s);
        return 0;
    }
    if (psfsize < 2 || psfsize > nembed || (psfsize & 1)) return -1;
    cplx *f = (cplx *)calloc((size_t)nembed * nembed, sizeof(cplx));
    for (long i = 0; i < nloc; i++) {
        if (embed[i] < 0 || embed[i] >= nembed * nembed) {
            free(f);
            return -1;
        }
        f[embed[i]] = amp[i] * cexp(I * wvk * opd[i]);
    }
    orc_fft2((double *)f, nembed, nembed, -1);
    const long h = psfsize / 2;
    for (long iy = 0; iy < psfsize; iy++) {
        const long ky = (iy - h + nembed) % nembed;
        for (long ix = 0; ix < psfsize; ix++) {
            const cplx v = f[(ix - h + nembed) % nembed + ky * nembed];
            psf[ix + iy * psfsize] = creal(v) * creal(v) + cimag(v) * cimag(v);
        }
    }
    free(f);
    return 0;
}

int orc_loc_calc_ptt(double *rms, double *coeff, const double *locx, const double *locy, const double *amp,
                     const double *opd, long nloc) {
    double m[3][6] = {{0}}, c[3] = {0, 0, 0}, tot = 0;
    for (long i = 0; i < nloc; i++) {
        const double v[3] = {1., locx[i], locy[i]};
        for (int r = 0; r < 3; r++) {
            for (int k = 0; k < 3; k++) m[r][k] += amp[i] * v[r] * v[k];
            c[r] += amp[i] * opd[i] * v[r];
        }
        tot += amp[i] * opd[i] * opd[i];
    }
    const double pcc = m[0][0];
    for (int r = 0; r < 3; r++) m[r][3 + r] = 1.;
    for (int col = 0; col < 3; col++) { /* Gauss-Jordan with partial pivoting */
        int piv = col;
        for (int r = col + 1; r < 3; r++)
            if (fabs(m[r][col]) > fabs(m[piv][col])) piv = r;
        if (m[piv][col] == 0) return -1;
        for (int k = 0; k < 6; k++) {
            const double t = m[col][k];
            m[col][k] = m[piv][k];
            m[piv][k] = t;
        }
        const double d = m[col][col];
        for (int k = 0; k < 6; k++) m[col][k] /= d;
        for (int r = 0; r < 3; r++) {
            if (r == col) continue;
            const double f = m[r][col];
            for (int k = 0; k < 6; k++) m[r][k] -= f * m[col][k];
        }
    }
    double ic[3];
    for (int r = 0; r < 3; r++) ic[r] = m[r][3] * c[0] + m[r][4] * c[1] + m[r][5] * c[2];
    if (coeff)
        for (int r = 0; r < 3; r++) coeff[r] = ic[r];
    const double pis = c[0] * c[0] / pcc, ptt = c[0] * ic[0] + c[1] * ic[1] + c[2] * ic[2];
    rms[0] = tot - pis;
    rms[1] = ptt - pis;
    rms[2] = tot - ptt;
    return 0;
}

/* ---------------------------------------------------------------------
 * Imaging: math/matcomp.c, maos/wfsints.c
 * ------------------------------------------------------------------- */
void orc_embed_wvf(double *wvf_, long nwvf, const double *opd, const double *amp, long nopd,
                   double wvl) {
    cplx *psf = (cplx *)wvf_;
    const double wvk = 2. * M_PI / wvl;
    memset(psf, 0, sizeof(cplx) * nwvf * nwvf);
    const long skip = (nwvf - nopd) / 2;
    cplx *psf0 = psf + skip * nwvf + skip;
    for (long iy = 0; iy < nopd; iy++) {
        for (long ix = 0; ix < nopd; ix++) {
            double ph = wvk * opd[iy * nopd + ix];
            psf0[iy * nwvf + ix] = amp[iy * nopd + ix] * (cos(ph) + I * sin(ph));
        }
    }
}

/* ccpcorner, math/matcomp.c:176-216 (C_FULL) */
static void ccpcorner(cplx *out, long nx, long ny, const cplx *in, long ninx, long niny);
void orc_ccpcorner(double *out, long nx, long ny, const double *in, long ninx, long niny) {
    ccpcorner((cplx *)out, nx, ny, (const cplx *)in, ninx, niny);
}
static void ccpcorner(cplx *out, long nx, long ny, const cplx *in, long ninx, long niny) {
    memset(out, 0, sizeof(cplx) * nx * ny);
    const long ny2 = (ny < niny) ? ny / 2 : niny / 2;
    const long nx2 = (nx < ninx) ? nx / 2 : ninx / 2;
    for (long i = 0; i < ny2; i++) {
        memcpy(out + (ny - ny2 + i) * nx + (nx - nx2), in + (niny - ny2 + i) * ninx + (ninx - nx2), sizeof(cplx) * nx2);
        memcpy(out + (ny - ny2 + i) * nx, in + (niny - ny2 + i) * ninx, sizeof(cplx) * nx2);
        memcpy(out + i * nx + (nx - nx2), in + i * ninx + (ninx - nx2), sizeof(cplx) * nx2);
        memcpy(out + i * nx, in + i * ninx, sizeof(cplx) * nx2);
    }
}

void orc_lotf(double *lotfc_, const double *lltopd, const double *lltamp, long nlx, long embfac,
              long nwvf, long notf, double wvl) {
    const long nlwvf = nlx * embfac;
    cplx *lwvf = (cplx *)malloc(sizeof(cplx) * nlwvf * nlwvf);
    cplx *lotfc = (cplx *)lotfc_;
    orc_embed_wvf((double *)lwvf, nlwvf, lltopd, lltamp, nlx, wvl);
    orc_fft2((double *)lwvf, nlwvf, nlwvf, -1);
    for (long i = 0; i < nlwvf * nlwvf; i++) {
        double a = creal(lwvf[i]), b = cimag(lwvf[i]);
        lwvf[i] = (a * a + b * b) * (1. / ((double)nlwvf * nlwvf));
    }
    int islotf = 0;
    cplx *lotfc2;
    if (nlwvf != nwvf) {
        islotf = 1;
        orc_fft2((double *)lwvf, nlwvf, nlwvf, -1);
        lotfc2 = (cplx *)malloc(sizeof(cplx) * nwvf * nwvf);
        ccpcorner(lotfc2, nwvf, nwvf, lwvf, nlwvf, nlwvf);
    } else {
        lotfc2 = lwvf;
    }
    if (nwvf != notf) {
        if (islotf) {
            orc_fft2((double *)lotfc2, nwvf, nwvf, 1);
            for (long i = 0; i < nwvf * nwvf; i++) lotfc2[i] *= 1. / ((double)nwvf * nwvf);
            islotf = 0;
        }
        ccpcorner(lotfc, notf, notf, lotfc2, nwvf, nwvf);
    } else {
        memcpy(lotfc, lotfc2, sizeof(cplx) * notf * notf);
    }
    if (!islotf) orc_fft2((double *)lotfc, notf, notf, -1);
    if (lotfc2 != lwvf) free(lotfc2);
    free(lwvf);
}

void orc_wfsints(double *ints, const double *opd, const double *lotfc_, const orc_powfs_t *p,
                 const orc_wfs_t *w, long isa_start, long isa_end) {
    const long nopd = p->nxsa, nwvf = nopd * p->embfac, notf = p->notf;
    const long npix = p->pixpsax * p->pixpsay, notf2 = notf * notf;
    const long saoffset = nopd * nopd;
    const int hasllt = p->llt_nx > 0 && p->etf;
    const double norm_psf = sqrt(p->areascale) / (double)(nopd * nwvf);
    const double norm_ints = w->sigsim * norm_psf * norm_psf / ((double)notf * notf);
    cplx *wvf = (cplx *)malloc(sizeof(cplx) * nwvf * nwvf);
    cplx *psf = nwvf != notf ? (cplx *)malloc(sizeof(cplx) * notf2) : wvf;
    for (long iwvl = 0; iwvl < p->nwvl; iwvl++) {
        const double wvl = p->wvl[iwvl];
        const cplx *lotfc = lotfc_ ? (const cplx *)lotfc_ + iwvl * notf2 : NULL;
        const long ntab = p->si_multi ? p->nsa : 1;
        const double *nominal0 = (!hasllt && p->nominal) ? p->nominal + iwvl * ntab * notf2 : NULL;
        const long *pp0 = p->si_pp + iwvl * ntab * (notf2 + 1);
        const long *pi0 = p->si_pi + iwvl * ntab * p->si_nzmax;
        const double *px0 = p->si_px + iwvl * ntab * p->si_nzmax;
        const cplx *etf1 = p->etf ? (const cplx *)p->etf + iwvl * p->nsa * notf2 : NULL;
        const cplx *etf2 = p->etf2 ? (const cplx *)p->etf2 + iwvl * p->nsa * notf2 : NULL;
        double etf2wt = etf2 ? p->etf2wt : 0, etf1wt = etf2 ? 1. - p->etf2wt : 1;
        const double alpha = w->wvlwts[iwvl] * norm_ints;
        for (long isa = isa_start; isa < isa_end; isa++) {
            /* the subaperture's own tables when multi_nominal */
            const long it = p->si_multi ? isa : 0;
            const double *nominal = nominal0 ? nominal0 + it * notf2 : NULL;
            const long *pp = pp0 + it * (notf2 + 1), *pi = pi0 + it * p->si_nzmax;
            const double *px = px0 + it * p->si_nzmax;
            orc_embed_wvf((double *)wvf, nwvf, opd + isa * saoffset, p->amp + isa * saoffset, nopd, wvl);
            orc_fft2((double *)wvf, nwvf, nwvf, -1);
            if (psf != wvf) ccpcorner(psf, notf, notf, wvf, nwvf, nwvf);
            for (long i = 0; i < notf2; i++) {
                double a = creal(psf[i]), b = cimag(psf[i]);
                psf[i] = a * a + b * b;
            }
            orc_fft2((double *)psf, notf, notf, -1);
            if (lotfc)
                for (long i = 0; i < notf2; i++) psf[i] *= conj(lotfc[i]);
            /* apply_dtf_etf, maos/wfsints.c:38-64 */
            const cplx *B = (hasllt && etf1) ? etf1 + isa * notf2 : NULL;
            const cplx *C = (hasllt && etf2) ? etf2 + isa * notf2 : NULL;
            if (C) {
                for (long i = 0; i < notf2; i++)
                    psf[i] *= (nominal ? nominal[i] : 1.) * (B[i] * etf1wt + C[i] * etf2wt);
            } else if (B) {
                for (long i = 0; i < notf2; i++) psf[i] *= (nominal ? nominal[i] : 1.) * B[i];
            } else if (nominal) {
                for (long i = 0; i < notf2; i++) psf[i] *= nominal[i];
            }
            orc_fft2((double *)psf, notf, notf, 1);
            /* dspmulcreal, math/spmm.c:77-86 */
            double *out = ints + isa * npix;
            for (long icol = 0; icol < notf2; icol++)
                for (long k = pp[icol]; k < pp[icol + 1]; k++)
                    out[pi[k]] += alpha * px[k] * creal(psf[icol]);
        }
    }
    if (psf != wvf) free(psf);
    free(wvf);
}

/* pistatout of wfsints (maos/wfsints.c:262-275): the downlink OTF (no
 * uplink, DTF or ETF), ctilt by -g dtheta1 pixels with the peak in the corner
 * (math/matcomp.c:287-336), the inverse transform (+1), cfftshift, |.|,
 * and norm_pistat = norm_psf^2 / notf^2 times that added to acc. */
void orc_pistat(double *acc, const double *opd, const orc_powfs_t *p, const double *gx, const double *gy,
                long isa_start, long isa_end) {
    const long nopd = p->nxsa, nwvf = nopd * p->embfac, notf = p->notf, notf2 = notf * notf;
    const long saoffset = nopd * nopd;
    const double norm_psf = sqrt(p->areascale) / (double)(nopd * nwvf);
    const double norm_pistat = norm_psf * norm_psf / ((double)notf * notf);
    cplx *wvf = (cplx *)malloc(sizeof(cplx) * nwvf * nwvf);
    cplx *psf = nwvf != notf ? (cplx *)malloc(sizeof(cplx) * notf2) : wvf;
    cplx *ux = (cplx *)malloc(sizeof(cplx) * notf * 2), *uy = ux + notf;
    for (long iwvl = 0; iwvl < p->nwvl; iwvl++) {
        const double wvl = p->wvl[iwvl];
        const double dtheta1 = (double)nwvf * p->dx / wvl;
        for (long isa = isa_start; isa < isa_end; isa++) {
            orc_embed_wvf((double *)wvf, nwvf, opd + isa * saoffset, p->amp + isa * saoffset, nopd, wvl);
            orc_fft2((double *)wvf, nwvf, nwvf, -1);
            if (psf != wvf) ccpcorner(psf, notf, notf, wvf, nwvf, nwvf);
            for (long i = 0; i < notf2; i++) {
                double a = creal(psf[i]), b = cimag(psf[i]);
                psf[i] = a * a + b * b;
            }
            orc_fft2((double *)psf, notf, notf, -1);
            /* ctilt(psf, sx, sy, 0): exp(-2 pi i u s / n) with u the signed
             * frequency of each index (u >= n/2 taken as u - n) */
            const double sx = -gx[isa] * dtheta1, sy = -gy[isa] * dtheta1;
            for (long i = 0; i < notf; i++) {
                const double u = (double)(i < notf / 2 ? i : i - notf);
                ux[i] = cexp(-2. * I * M_PI * u * sx / (double)notf);
                uy[i] = cexp(-2. * I * M_PI * u * sy / (double)notf);
            }
            for (long iy = 0; iy < notf; iy++)
                for (long ix = 0; ix < notf; ix++) psf[ix + iy * notf] *= ux[ix] * uy[iy];
            orc_fft2((double *)psf, notf, notf, 1);
            /* cfftshift, cabstoreal and creal2d: peak moved to the centre */
            double *out = acc + (iwvl * p->nsa + isa) * notf2;
            for (long iy = 0; iy < notf; iy++)
                for (long ix = 0; ix < notf; ix++)
                    out[(ix + notf / 2) % notf + ((iy + notf / 2) % notf) * notf] +=
                        norm_pistat * cabs(psf[ix + iy * notf]);
        }
    }
    free(ux);
    if (psf != wvf) free(psf);
    free(wvf);
}

/* psfout of wfsints (maos/wfsints.c:243-253): the complex PSF (the field's
 * transform, cropped or padded to notf by ccpcorner) scaled by
 * norm_psf / nwvf^2, transformed back (+1), and its centre (notf/2 + 2)^2
 * taken as cembed crops (math/matmath.c:1428-1458: out index i holds in
 * index i + notf/2 - nout/2). */
void orc_psfout(double *out_, const double *opd, const orc_powfs_t *p, long isa_start, long isa_end) {
    const long nopd = p->nxsa, nwvf = nopd * p->embfac, notf = p->notf, notf2 = notf * notf;
    const long nout = notf / 2 + 2, off = notf / 2 - nout / 2;
    const long saoffset = nopd * nopd;
    const double norm_psf = sqrt(p->areascale) / (double)(nopd * nwvf);
    cplx *out = (cplx *)out_;
    cplx *wvf = (cplx *)malloc(sizeof(cplx) * nwvf * nwvf);
    cplx *psf = nwvf != notf ? (cplx *)malloc(sizeof(cplx) * notf2) : wvf;
    for (long iwvl = 0; iwvl < p->nwvl; iwvl++) {
        const double wvl = p->wvl[iwvl];
        for (long isa = isa_start; isa < isa_end; isa++) {
            orc_embed_wvf((double *)wvf, nwvf, opd + isa * saoffset, p->amp + isa * saoffset, nopd, wvl);
            orc_fft2((double *)wvf, nwvf, nwvf, -1);
            if (psf != wvf) ccpcorner(psf, notf, notf, wvf, nwvf, nwvf);
            for (long i = 0; i < notf2; i++) psf[i] *= norm_psf / ((double)nwvf * nwvf);
            orc_fft2((double *)psf, notf, notf, 1);
            cplx *o = out + (iwvl * p->nsa + isa) * nout * nout;
            for (long iy = 0; iy < nout; iy++)
                for (long ix = 0; ix < nout; ix++) o[ix + iy * nout] = psf[(ix + off) + (iy + off) * notf];
        }
    }
    if (psf != wvf) free(psf);
    free(wvf);
}

/* ---------------------------------------------------------------------
 * Pixel processing: math/matmath.c, maos/utils.c
 * ------------------------------------------------------------------- */
void orc_dcog(double *grad, const double *im, long nx, long ny, double offsetx,
              double offsety, double thres, double bkgrnd, double flux) {
    double sum = 0, sumx = 0, sumy = 0;
    for (long iy = 0; iy < ny; iy++) {
        for (long ix = 0; ix < nx; ix++) {
            double iI = im[ix + iy * nx] - bkgrnd;
            if (iI > thres) {
                sum += iI;
                sumx += iI * ix;
                sumy += iI * iy;
            }
        }
    }
    if (flux) sum = flux;
    if (sum > thres) {
        grad[0] = sumx / sum - ((double)(nx - 1) * 0.5 + offsetx);
        grad[1] = sumy / sum - ((double)(ny - 1) * 0.5 + offsety);
    } else {
        grad[0] = 0;
        grad[1] = 0;
    }
}

void orc_shwfs_grad(double *grad, const double *ints, const orc_powfs_t *p, double siglev) {
    (void)siglev;
    const long nsa = p->nsa, npix = p->pixpsax * p->pixpsay;
    if (p->phytype == 1 && (!p->mtche || (p->sigmatch == 1 && !p->i0sum))) return; /* no MF yet */
    double *gx = grad, *gy = grad + nsa;
    double i1sum = 0;
    if (p->sigmatch == 2)
        for (long isa = 0; isa < nsa; isa++)
            for (long i = 0; i < npix; i++) i1sum += ints[isa * npix + i];
    double sigtot = NAN;
    if (p->sigmatch == 0) sigtot = p->siglev_dtrat;
    else if (p->sigmatch == 2) sigtot = i1sum / p->saasum;
    for (long isa = 0; isa < nsa; isa++) {
        double geach[3] = {0, 0, 1};
        if (p->saa[isa] < p->saat) continue;
        const double *im = ints + isa * npix;
        if (p->phytype == 1) {
            double scale = 1.;
            if (p->sigmatch == 1) {
                double s = 0;
                for (long i = 0; i < npix; i++) s += im[i];
                scale = p->i0sum[isa] / s;
                if (scale <= 0 || scale > 10 || !isfinite(scale)) scale = 1;
            } else if (p->sigmatch == 2) {
                scale = p->i0sumsum / i1sum;
            }
            /* dmulvec, math/matmath.c:384-396: y += A*x*alpha, A is 2 x npix */
            const double *m = p->mtche + isa * 2 * npix;
            for (long iy = 0; iy < npix; iy++) {
                geach[0] += (im[iy] * scale) * m[0 + 2 * iy];
                geach[1] += (im[iy] * scale) * m[1 + 2 * iy];
            }
        } else {
            double sumi = 0;
            if (p->sigmatch == 0) sumi = p->i0sum ? p->i0sum[isa] : sigtot * p->saa[isa];
            else if (p->sigmatch == 2) sumi = sigtot * p->saa[isa];
            orc_dcog(geach, im, p->pixpsax, p->pixpsay, 0., 0., p->cogthres, p->cogoff, sumi);
            geach[0] *= p->pixthetax;
            geach[1] *= p->pixthetay;
        }
        gx[isa] = geach[0];
        gy[isa] = geach[1];
    }
}

/* ---------------------------------------------------------------------
 * One WFS step: maos/wfsgrad.c:122-501 (open loop, noise free)
 * ------------------------------------------------------------------- */
void orc_wfsgrad_iwfs(double *ints, double *grad, const orc_powfs_t *p, const orc_wfs_t *w,
                      const orc_atm_t *atm, long isim, double dt, double atmscale, int nthread,
                      double *opd_out, double *lotf_out) {
    const long nsa = p->nsa, nxsa = p->nxsa, npsa = nxsa * nxsa;
    const long npix = p->pixpsax * p->pixpsay;
    const long notf2 = p->notf * p->notf;
#ifdef _OPENMP
    if (nthread <= 0) nthread = omp_get_max_threads();
#else
    nthread = 1;
#endif
    double *opd = opd_out ? opd_out : (double *)malloc(sizeof(double) * nsa * npsa);
    memset(opd, 0, sizeof(double) * nsa * npsa);
    /* atmosphere ray trace, one layer after another, subapertures split
       over threads (maos/wfsgrad.c:175-194, maos/sim_utils.c:1004-1033) */
    for (long ips = 0; ips < atm->nps; ips++) {
        const double ht = atm->ht[ips];
        if (w->hs && ht > w->hs) continue;
        const double scale = w->hs ? (1. - ht / w->hs) : 1.;
        const double shiftx = -atm->vx[ips] * dt * isim;
        const double shifty = -atm->vy[ips] * dt * isim;
        const double dispx = ht * w->thetax + scale * w->misregx + shiftx;
        const double dispy = ht * w->thetay + scale * w->misregy + shifty;
#pragma omp parallel for num_threads(nthread) schedule(static)
        for (long isa = 0; isa < nsa; isa++) {
            orc_prop_grid_pts(atm->screens[ips], atm->nx, atm->ny, atm->dx, atm->dx, atm->ox,
                              atm->oy, p->ptsx, p->ptsy, nsa, nxsa, p->dx, opd, atmscale, dispx,
                              dispy, scale, 1, isa, isa + 1);
        }
    }
    /* LLT uplink (maos/wfsgrad.c:329-386) */
    double *lotfc = NULL;
    if (p->llt_nx > 0) {
        const long nlx = p->llt_nx;
        double *lltopd = (double *)calloc(nlx * nlx, sizeof(double));
        double lx = p->llt_origx, ly = p->llt_origy;
        for (long ips = 0; ips < atm->nps; ips++) {
            const double hl = atm->ht[ips];
            const double scale = 1. - hl / w->hs;
            if (scale < 0) continue;
            const double thetax = w->thetax - p->llt_ox / w->hs;
            const double thetay = w->thetay - p->llt_oy / w->hs;
            orc_prop(atm->screens[ips], atm->nx, atm->ny, atm->dx, atm->ox, atm->oy, hl, &lx, &ly, 1,
                     nlx, p->dx, lltopd, atmscale, w->hs, thetax, thetay, 0, 0,
                     p->llt_ox - atm->vx[ips] * isim * dt, p->llt_oy - atm->vy[ips] * isim * dt, 1);
        }
        lotfc = lotf_out ? lotf_out : (double *)malloc(sizeof(double) * 2 * notf2 * p->nwvl);
        for (long iwvl = 0; iwvl < p->nwvl; iwvl++)
            orc_lotf(lotfc + iwvl * 2 * notf2, lltopd, p->llt_amp, nlx, p->embfac, nxsa * p->embfac,
                     p->notf, p->wvl[iwvl]);
        free(lltopd);
    }
    memset(ints, 0, sizeof(double) * nsa * npix);
    /* wfsints split over subapertures (maos/sim_utils.c:1069-1079) */
    long chunk = (nsa + nthread - 1) / nthread;
#pragma omp parallel for num_threads(nthread) schedule(static, 1)
    for (long it = 0; it < nthread; it++) {
        long s0 = it * chunk, s1 = s0 + chunk;
        if (s1 > nsa) s1 = nsa;
        if (s0 < s1) orc_wfsints(ints, opd, lotfc, p, w, s0, s1);
    }
    if (grad) orc_shwfs_grad(grad, ints, p, 0);
    if (lotfc && lotfc != lotf_out) free(lotfc);
    if (opd != opd_out) free(opd);
}

/* ---------------------------------------------------------------------
 * Pixel noise: lib/libmisc.c:31-57 and math/random.c:186-209, with the
 * uniforms of a Philox4x32-10 stream per (step, unit, pixel).
 * ------------------------------------------------------------------- */
void orc_philox4x32_10(unsigned ctr[4], unsigned k0, unsigned k1) {
    for (int r = 0; r < 10; r++) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * ctr[0];
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * ctr[2];
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        const unsigned n0 = hi1 ^ ctr[1] ^ k0, n2 = hi0 ^ ctr[3] ^ k1;
        ctr[0] = n0;
        ctr[1] = lo1;
        ctr[2] = n2;
        ctr[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

typedef struct {
    unsigned k0, k1, pix, unit, step, blk;
    unsigned buf[4];
    int left;
} orc_stream;

/* uniform in (0, 1] */
static double st_uniform(orc_stream *s) {
    if (s->left == 0) {
        unsigned c[4] = {s->blk++, s->pix, s->unit, s->step};
        orc_philox4x32_10(c, s->k0, s->k1);
        memcpy(s->buf, c, sizeof c);
        s->left = 4;
    }
    return ((double)s->buf[4 - s->left--] + 1.) * 2.3283064365386962890625e-10;
}
/* Box-Muller, one normal per pair of uniforms */
static double st_normal(orc_stream *s) {
    const double u1 = st_uniform(s), u2 = st_uniform(s);
    return sqrt(-2. * log(u1)) * cos(2. * M_PI * u2);
}
/* randp; the multiplication loop stops after 1000 factors */
static double st_poisson(orc_stream *s, double xm) {
    if (xm > 200.) return rint(xm + st_normal(s) * sqrt(xm));
    double x = 0.;
    while (xm > 0.) {
        const double xmu = xm > 12. ? 12. : xm;
        xm -= xmu;
        const double g = exp(-xmu);
        double t = 1.;
        int xu = -1;
        while (t > g && xu < 1000) {
            xu++;
            t *= st_uniform(s);
        }
        x += xu;
    }
    return x;
}

double orc_randp(double xm, unsigned key0, unsigned key1, unsigned pix, unsigned unit, unsigned step,
                 unsigned *blk, int *used) {
    orc_stream s = {key0, key1, pix, unit, step, *blk, {0, 0, 0, 0}, 0};
    /* resume inside the current block of four */
    if (*used) {
        s.blk = *blk - 1;
        unsigned c[4] = {s.blk++, pix, unit, step};
        orc_philox4x32_10(c, key0, key1);
        memcpy(s.buf, c, sizeof c);
        s.left = 4 - *used;
    }
    const double x = st_poisson(&s, xm);
    *blk = s.blk;
    *used = s.left ? 4 - s.left : 0;
    return x;
}

void orc_addnoise(double *ints, long npix, double bkgrnd, double bkgrndc, double rne, unsigned key0,
                  unsigned key1, unsigned unit, unsigned step) {
    for (long p = 0; p < npix; p++) {
        orc_stream s = {key0, key1, (unsigned)p, unit, step, 0, {0, 0, 0, 0}, 0};
        const double tot = ints[p] + bkgrnd;
        const double c = st_poisson(&s, tot);
        ints[p] = c + rne * st_normal(&s) - bkgrndc;
    }
}

/* ---------------------------------------------------------------------
 * Geometric Zernike-tilt gradients: pts_mcc_ptt (math/loc.c:612-650), its
 * inverse (dcellinvspd_each), pts_ztilt (math/loc.c:839-880).
 * ------------------------------------------------------------------- */
void orc_ztilt(double *grad, const double *opd, const double *amp, const double *ptsx,
               const double *ptsy, long nsa, long nxsa, double dx) {
    for (long isa = 0; isa < nsa; isa++) {
        const double *a = amp + isa * nxsa * nxsa, *o = opd + isa * nxsa * nxsa;
        double m[3][3] = {{0}}, c[3] = {0, 0, 0};
        for (long iy = 0; iy < nxsa; iy++)
            for (long ix = 0; ix < nxsa; ix++) {
                const double x = ptsx[isa] + ix * dx, y = ptsy[isa] + iy * dx, w = a[iy * nxsa + ix];
                const double v[3] = {1., x, y};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) m[i][j] += w * v[i] * v[j];
                if (w != 0)
                    for (int i = 0; i < 3; i++) c[i] += w * o[iy * nxsa + ix] * v[i];
            }
        /* Gauss-Jordan on the symmetric 3x3 system m g = c */
        double aug[3][4];
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) aug[i][j] = m[i][j];
            aug[i][3] = c[i];
        }
        int ok = 1;
        for (int k = 0; k < 3 && ok; k++) {
            int piv = k;
            for (int i = k + 1; i < 3; i++)
                if (fabs(aug[i][k]) > fabs(aug[piv][k])) piv = i;
            if (aug[piv][k] == 0) {
                ok = 0;
                break;
            }
            for (int j = 0; j < 4; j++) {
                const double t = aug[k][j];
                aug[k][j] = aug[piv][j];
                aug[piv][j] = t;
            }
            for (int i = 0; i < 3; i++)
                if (i != k) {
                    const double f = aug[i][k] / aug[k][k];
                    for (int j = 0; j < 4; j++) aug[i][j] -= f * aug[k][j];
                }
        }
        grad[isa] = ok ? aug[1][3] / aug[1][1] : 0;
        grad[isa + nsa] = ok ? aug[2][3] / aug[2][2] : 0;
    }
}

/* ---------------------------------------------------------------------
 * Geometric average-gradient (G-tilt) gradients, GS0 of mkg
 * (lib/mkg.c:108-468, maos/powfs.c:603-630, do_partial = 1) applied to an
 * OPD: per subaperture, g = (1/dx) int A dPhi / int A over the cells of the
 * common point grid that meet the subaperture square and whose four corners
 * exist with amplitude > 1e-14; A and Phi are bilinear.  Each cell's part
 * is integrated by 2x2 Gauss-Legendre quadrature (exact: the integrand is
 * at most quadratic in each direction).  A subaperture whose amplitude
 * integral is at most 1% of its area, or with saa < saat (an empty row,
 * applied by dspmm at maos/wfsgrad.c:308), gives 0.
 * ------------------------------------------------------------------- */
void orc_gtilt(double *grad, const double *opd, const double *amp, const double *ptsx, const double *ptsy,
               long nsa, long nxsa, double dx, const double *saa, double saat) {
    const long npsa = nxsa * nxsa;
    /* points keyed by integer grid coordinates relative to subaperture 0 */
    long minx = 0, miny = 0, maxx = 0, maxy = 0;
    long *gx = (long *)malloc(sizeof(long) * nsa), *gy = (long *)malloc(sizeof(long) * nsa);
    for (long isa = 0; isa < nsa; isa++) {
        gx[isa] = (long)floor((ptsx[isa] - ptsx[0]) / dx + 0.5);
        gy[isa] = (long)floor((ptsy[isa] - ptsy[0]) / dx + 0.5);
        if (gx[isa] < minx) minx = gx[isa];
        if (gy[isa] < miny) miny = gy[isa];
        if (gx[isa] > maxx) maxx = gx[isa];
        if (gy[isa] > maxy) maxy = gy[isa];
    }
    const long w = maxx - minx + nxsa + 4, hgt = maxy - miny + nxsa + 4;
    long *idx = (long *)malloc(sizeof(long) * w * hgt);
    for (long i = 0; i < w * hgt; i++) idx[i] = -1;
#define ORC_AT(X, Y) idx[((X) - minx + 2) + ((Y) - miny + 2) * w]
    for (long isa = 0; isa < nsa; isa++)
        for (long iy = 0; iy < nxsa; iy++)
            for (long ix = 0; ix < nxsa; ix++) ORC_AT(gx[isa] + ix, gy[isa] + iy) = isa * npsa + iy * nxsa + ix;
    const double gq = 0.5 / sqrt(3.);
    for (long isa = 0; isa < nsa; isa++) {
        if (saa && saat > 0 && saa[isa] < saat) { /* an empty row of GS0 */
            grad[isa] = grad[isa + nsa] = 0;
            continue;
        }
        double sa = 0, sx = 0, sy = 0;
        const double lo_x = gx[isa] - 0.5, hi_x = gx[isa] + nxsa - 0.5;
        const double lo_y = gy[isa] - 0.5, hi_y = gy[isa] + nxsa - 0.5;
        for (long cy = gy[isa] - 1; cy <= gy[isa] + nxsa - 1; cy++)
            for (long cx = gx[isa] - 1; cx <= gx[isa] + nxsa - 1; cx++) {
                const double u1 = lo_x - cx > 0 ? lo_x - cx : 0, u2 = hi_x - cx < 1 ? hi_x - cx : 1;
                const double v1 = lo_y - cy > 0 ? lo_y - cy : 0, v2 = hi_y - cy < 1 ? hi_y - cy : 1;
                if (u2 <= u1 || v2 <= v1) continue;
                double av[2][2], ph[2][2];
                int ok = 1;
                for (int j = 0; j < 2; j++)
                    for (int i = 0; i < 2; i++) {
                        const long k = ORC_AT(cx + i, cy + j);
                        av[j][i] = k < 0 ? 0 : (amp ? amp[k] : 1.);
                        ph[j][i] = k < 0 ? 0 : opd[k];
                        if (!(fabs(av[j][i]) > 1e-14)) ok = 0;
                    }
                if (!ok) continue;
                const double um = 0.5 * (u1 + u2), ur = 0.5 * (u2 - u1), vm = 0.5 * (v1 + v2), vr = 0.5 * (v2 - v1);
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        const double u = um + (a ? gq : -gq) * 2 * ur, v = vm + (b ? gq : -gq) * 2 * vr;
                        const double wq = ur * vr; /* each of the four nodes weighs a quarter of the area */
                        const double A = av[0][0] * (1 - u) * (1 - v) + av[0][1] * u * (1 - v) +
                                         av[1][0] * (1 - u) * v + av[1][1] * u * v;
                        const double du = (ph[0][1] - ph[0][0]) * (1 - v) + (ph[1][1] - ph[1][0]) * v;
                        const double dv = (ph[1][0] - ph[0][0]) * (1 - u) + (ph[1][1] - ph[0][1]) * u;
                        sa += wq * A;
                        sx += wq * A * du;
                        sy += wq * A * dv;
                    }
            }
        if (sa > 0.01 * (double)npsa) {
            grad[isa] = sx / (sa * dx);
            grad[isa + nsa] = sy / (sa * dx);
        } else {
            grad[isa] = grad[isa + nsa] = 0;
        }
    }
#undef ORC_AT
    free(idx);
    free(gx);
    free(gy);
}

/* ---------------------------------------------------------------------
 * Pyramid WFS: lib/pywfs.c
 * ------------------------------------------------------------------- */
/* bilinear weight of the image at grid coordinate (fx, fy), with mkht's
 * rules (lib/mkh.c:62-150): weights <= EPS and points outside the grid are
 * dropped, the rest are not renormalised */
static double py_sample(const double *img, long n, double fx, double fy) {
    const long ix = (long)floor(fx), iy = (long)floor(fy);
    const double wx[2] = {1. - (fx - ix), fx - ix}, wy[2] = {1. - (fy - iy), fy - iy};
    double s = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const double w = wx[b] * wy[a];
            const long x = ix + b, y = iy + a;
            if (w > 1e-15 && x >= 0 && x < n && y >= 0 && y < n) s += w * img[x + y * n];
        }
    return s;
}

void orc_pywfs_ints(double *ints, const double *opd, const orc_pywfs_t *c, double siglev) {
    const long n = c->npsf, n2 = n / 2, nn = n * n;
    cplx *psf = (cplx *)malloc(sizeof(cplx) * nn * c->nwvl);
    cplx *otf = (cplx *)malloc(sizeof(cplx) * nn);
    double *pup = (double *)calloc(nn, sizeof(double));
    double ampnorm = 0;
    for (long i = 0; i < c->nloc; i++) ampnorm += c->amp[i] * c->amp[i];
    /* locfft_psf with sum2one (lib/locfft.c:110-188): centred complex PSF */
    for (int iw = 0; iw < c->nwvl; iw++) {
        cplx *p = psf + iw * nn;
        memset(p, 0, sizeof(cplx) * nn);
        const double wvk = 2. * M_PI / c->wvl[iw];
        for (long i = 0; i < c->nloc; i++) p[c->embed[i]] = c->amp[i] * cexp(I * (wvk * opd[i]));
        orc_fft2((double *)p, n, n, -1);
        const double norm = 1. / (sqrt(ampnorm) * (double)n);
        /* cfftshift: corner sample (x, y) to ((x + n2) % n, (y + n2) % n) */
        memcpy(otf, p, sizeof(cplx) * nn);
        for (long y = 0; y < n; y++)
            for (long x = 0; x < n; x++) p[(x + n2) % n + ((y + n2) % n) * n] = otf[x + y * n] * norm;
    }
    for (int ir = 0; ir < c->modulring; ir++) {
        const double ri = c->modulate * (ir + 1) / c->modulring;
        const int npos = c->modulpos * (ir + 1) / c->modulring;
        for (int ip = 0; ip < npos; ip++) {
            const double th = 2. * M_PI * ((double)ip / npos);
            const double px = cos(th) * ri, py = sin(th) * ri;
            for (int iw = 0; iw < c->nwvl; iw++) {
                const double dtheta = c->wvl[iw] / (c->dx * (double)n);
                const long ox = (long)round(px / dtheta), oy = (long)round(py / dtheta);
                const cplx *mask = (const cplx *)c->mask + iw * nn;
                const cplx *p = psf + iw * nn;
                for (long y = 0; y < n; y++)
                    for (long x = 0; x < n; x++) {
                        const long sx = x + ox, sy = y + oy;
                        otf[x + y * n] = (sx >= 0 && sx < n && sy >= 0 && sy < n) ? p[sx + sy * n] * mask[x + y * n] : 0;
                    }
                orc_fft2((double *)otf, n, n, 1);
                const double a = c->wvlwts[iw] / ((double)nn * npos * c->modulring);
                for (long i = 0; i < nn; i++) pup[i] += a * (creal(otf[i]) * creal(otf[i]) + cimag(otf[i]) * cimag(otf[i]));
            }
        }
    }
    /* detector transfer function: FFT, times nominal, inverse FFT */
    for (long i = 0; i < nn; i++) otf[i] = pup[i];
    orc_fft2((double *)otf, n, n, -1);
    for (long i = 0; i < nn; i++) otf[i] *= c->nominal[i];
    orc_fft2((double *)otf, n, n, 1);
    for (long i = 0; i < nn; i++) pup[i] = creal(otf[i]);
    /* sampling at the subaperture centres of each pupil image
       (pywfs_mksi, lib/pywfs.c:34-107, on the grid of lib/pywfs.c:279) */
    const double alpha = (double)c->nsa * siglev / (double)nn;
    for (int is = 0; is < c->nside; is++)
        for (long isa = 0; isa < c->nsa; isa++) {
            const double fx = c->salocx[isa] / c->dx + c->sioff[2 * is] * n2 + n2 - 0.5;
            const double fy = c->salocy[isa] / c->dx + c->sioff[2 * is + 1] * n2 + n2 - 0.5;
            ints[isa + is * c->nsa] = alpha * py_sample(pup, n, fx, fy);
        }
    free(pup);
    free(otf);
    free(psf);
}

void orc_pywfs_grad(double *grad, const double *ints, const orc_pywfs_t *c) {
    const long nsa = c->nsa;
    const int raw = c->raw || c->nside < 3;
    const int ng = raw ? c->nside : 2;
    double imean = 0;
    if (c->sigmatch == 2) {
        for (long i = 0; i < nsa * c->nside; i++) imean += ints[i];
        imean /= (double)nsa;
    }
    for (long isa = 0; isa < nsa; isa++) {
        const double *I0 = ints + isa;
        if (raw) {
            for (int s = 0; s < c->nside; s++) grad[isa + nsa * s] = I0[s * nsa] * c->gain / c->siglev;
            continue;
        }
        double isum = 0;
        if (c->sigmatch == 0) isum = c->siglev * c->saa[isa];
        else if (c->sigmatch == 1)
            for (int s = 0; s < c->nside; s++) isum += I0[s * nsa];
        else isum = imean * c->saa[isa];
        const double a = c->gain / isum;
        if (c->nside == 3) {
            grad[isa] = (I0[nsa] - I0[2 * nsa]) * a * sqrt(3.) * 0.5;
            grad[isa + nsa] = (I0[0] - 0.5 * (I0[nsa] + I0[2 * nsa])) * a;
        } else {
            grad[isa] = (I0[0] - I0[nsa] + I0[2 * nsa] - I0[3 * nsa]) * a;
            grad[isa + nsa] = (I0[0] + I0[nsa] - I0[2 * nsa] - I0[3 * nsa]) * a;
        }
    }
    if (c->gradoff)
        for (long i = 0; i < ng * nsa; i++) grad[i] -= c->gradoff[i];
}
