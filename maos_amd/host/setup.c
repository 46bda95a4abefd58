/*
 * Host-side FP64 table builders for the SH physical-optics WFS path.
 *
 * These restate MAOS's setup routines so the GPU path can be fed the same
 * tables MAOS would hand it (see include/maos_setup.h).  Layout follows
 * MAOS: column-major, x fastest; complex arrays interleaved (re, im).
 */
#include "../../include/maos_setup.h"
#include "internal.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define EPS_D 1.e-15 /* EPS of a double build, sys/common.h:58-63 */

/* ------------------------------------------------------------------------
 * FP64 FFT: iterative radix-2 for powers of two, direct DFT otherwise.
 * Unnormalised; dir=-1 forward (exp(-i)), dir=+1 backward.
 * --------------------------------------------------------------------- */
static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft1d(cplx *x, long n, int dir, cplx *work) {
    if (n <= 1) return;
    if (!is_pow2(n)) {
        for (long k = 0; k < n; k++) {
            cplx s = 0;
            for (long j = 0; j < n; j++) {
                double ang = dir * 2. * M_PI * (double)((j * k) % n) / (double)n;
                s += x[j] * (cos(ang) + I * sin(ang));
            }
            work[k] = s;
        }
        memcpy(x, work, sizeof(cplx) * n);
        return;
    }
    /* bit reversal */
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (long len = 2; len <= n; len <<= 1) {
        long h = len >> 1;
        for (long j = 0; j < h; j++) {
            double ang = dir * 2. * M_PI * (double)j / (double)len;
            cplx w = cos(ang) + I * sin(ang);
            for (long i = j; i < n; i += len) {
                cplx u = x[i], v = x[i + h] * w;
                x[i] = u + v;
                x[i + h] = u - v;
            }
        }
    }
}

void mxs_fft2(double *c, long nx, long ny, int dir) {
    cplx *a = (cplx *)c;
    long nmax = nx > ny ? nx : ny;
#pragma omp parallel
    {
        cplx *work = (cplx *)malloc(sizeof(cplx) * nmax * 2);
        cplx *col = work + nmax;
#pragma omp for schedule(static)
        for (long iy = 0; iy < ny; iy++) fft1d(a + iy * nx, nx, dir, work);
        if (ny > 1) {
#pragma omp for schedule(static)
            for (long ix = 0; ix < nx; ix++) {
                for (long iy = 0; iy < ny; iy++) col[iy] = a[ix + iy * nx];
                fft1d(col, ny, dir, work);
                for (long iy = 0; iy < ny; iy++) a[ix + iy * nx] = col[iy];
            }
        }
        free(work);
    }
}

/* quadrant swap for even sizes, as fftshift (math/mat.c:697-726) */
static void cfftshift(cplx *a, long nx, long ny) {
    long nx2 = nx / 2, ny2 = ny / 2;
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * nx * ny);
    for (long iy = 0; iy < ny; iy++)
        for (long ix = 0; ix < nx; ix++)
            tmp[((ix + nx2) % nx) + ((iy + ny2) % ny) * nx] = a[ix + iy * nx];
    memcpy(a, tmp, sizeof(cplx) * nx * ny);
    free(tmp);
}

/* nextfftsize, math/mathmisc.c:109-146 */
long mxs_nextfftsize(long n) {
    static const int radixs[4] = {2, 3, 5, 7};
    if (n == 0) n = 1;
    n = (n + 1) / 2 * 2;
    long n2, n3;
    do {
        n2 = n;
        n3 = 1;
        for (int irad = 0; irad < 4;) {
            int radix = radixs[irad];
            long ratio = n2 / radix;
            if (ratio * radix == n2) {
                n2 = ratio;
                n3 *= radix;
            } else {
                irad++;
            }
        }
        n += 2;
    } while (n2 > 1 || (n3 & 1) == 1);
    return n3;
}

/* Embedding of a point set in a square FFT grid (loc_create_embed,
 * math/loc.c:147-179): the grid side is ceil(max(nx, ny) * oversize), raised
 * by nextfftsize when fftpad, unless *nembed > 0 gives it (below 0.6 of the
 * minimum it is an error); the points are centred with the odd pad on the
 * left. */
long mxs_loc_create_embed(long *embed, long *nembed, long nloc, const double *locx, const double *locy, double dx,
                          double dy, double oversize, int fftpad) {
    if (!embed || !nembed || nloc < 1 || !locx || !locy || !(dx > 0) || !(dy > 0)) return -1;
    double xmin = locx[0], xmax = locx[0], ymin = locy[0], ymax = locy[0];
    for (long i = 1; i < nloc; i++) {
        xmin = fmin(xmin, locx[i]);
        xmax = fmax(xmax, locx[i]);
        ymin = fmin(ymin, locy[i]);
        ymax = fmax(ymax, locy[i]);
    }
    const double idx = 1. / dx, idy = 1. / dy;
    const long nx = lround((xmax - xmin) * idx) + 1, ny = lround((ymax - ymin) * idy) + 1;
    long nxy = (long)ceil((nx > ny ? nx : ny) * oversize);
    if (fftpad) nxy = mxs_nextfftsize(nxy);
    if (*nembed > 0) {
        if (*nembed < (long)(nxy * 0.6)) return -2;
        nxy = *nembed;
    }
    *nembed = nxy;
    xmin -= (nxy - nx + 1) / 2 * dx;
    ymin -= (nxy - ny + 1) / 2 * dy;
    for (long i = 0; i < nloc; i++) {
        const long ix = lround((locx[i] - xmin) * idx), iy = lround((locy[i] - ymin) * idy);
        if (ix < 0 || iy < 0 || ix >= nxy || iy >= nxy) return -3;
        embed[i] = ix + iy * nxy;
    }
    return nxy;
}

/* ------------------------------------------------------------------------
 * Geometry: maos/powfs.c:255-462
 * --------------------------------------------------------------------- */
long mxs_saloc(int order, double dsa, double din, double misregrmax,
               double *locx, double *locy, long maxn) {
    double offset = (order & 1) ? -0.5 : 0;
    int orderpad = (int)(order + (order > 2 ? (1 + misregrmax / dsa) * 2 : 0));
    double r2max = pow(order * 0.5, 2);
    double r2min = dsa < din ? pow(din / dsa / 2, 2) : -1;
    long count = 0;
    for (int j = -orderpad / 2; j <= (orderpad - 1) / 2; j++) {
        for (int i = -orderpad / 2; i <= (orderpad - 1) / 2; i++) {
            double xc = i + offset, yc = j + offset;
            double r1 = pow(xc + 1, 2) + pow(yc + 1, 2);
            double r2 = pow(xc + 1, 2) + pow(yc, 2);
            double r3 = pow(xc, 2) + pow(yc + 1, 2);
            double r4 = pow(xc, 2) + pow(yc, 2);
            if (r1 < r2max || r2 < r2max || r3 < r2max || r4 < r2max || r1 > r2min ||
                r2 > r2min || r3 > r2min || r4 > r2min) {
                if (count < maxn) {
                    locx[count] = xc * dsa;
                    locy[count] = yc * dsa;
                }
                count++;
            }
        }
    }
    return count < maxn ? count : maxn;
}

void mxs_loc_circle_add(double *phi, const double *locx, const double *locy,
                        long nloc, double dx, double dy, double cx, double cy,
                        double r, double rin, double val) {
    const int nres = 10;
    const double cres = (nres - 1.) / 2.;
    const double res = 1. / nres;
    const double dxres = res * dx, dyres = res * dy, res2 = res * res;
    const double r2 = r * r;
    const double r2l = (r - dx * 1.5) * (r - dy * 1.5);
    const double r2u = (r + dx * 1.5) * (r + dy * 1.5);
    for (long iloc = 0; iloc < nloc; iloc++) {
        double x = locx[iloc], y = locy[iloc];
        double r2r = pow(x - cx, 2) + pow(y - cy, 2);
        if (r2r < r2l) {
            phi[iloc] += val;
        } else if (r2r < r2u) {
            long tot = 0;
            for (int jres = 0; jres < nres; jres++) {
                double iiy = y + (jres - cres) * dyres;
                double rr2y = (iiy - cy) * (iiy - cy);
                for (int ires = 0; ires < nres; ires++) {
                    double iix = x + (ires - cres) * dxres;
                    if ((iix - cx) * (iix - cx) + rr2y <= r2) tot++;
                }
            }
            phi[iloc] += (double)tot * res2 * val;
        }
    }
    if (rin > 0) mxs_loc_circle_add(phi, locx, locy, nloc, dx, dy, cx, cy, rin, 0, -val);
}

void mxs_saa(const double *amp, long nsa, long npsa, double areascale, double *saa) {
    for (long isa = 0; isa < nsa; isa++) {
        double area = 0;
        for (long i = 0; i < npsa; i++) area += amp[isa * npsa + i] * amp[isa * npsa + i];
        saa[isa] = area / (double)npsa * areascale;
    }
}

long mxs_sa_reduce(double *salocx, double *salocy, double *ptsx, double *ptsy,
                   double *amp, double *saa, long nsa, long npsa, double dsa,
                   double saat) {
    if (nsa > 4) {
        /* integer grid index of each subaperture, then a lookup map */
        double xmin = INFINITY, ymin = INFINITY, xmax = -INFINITY, ymax = -INFINITY;
        for (long i = 0; i < nsa; i++) {
            if (ptsx[i] < xmin) xmin = ptsx[i];
            if (ptsy[i] < ymin) ymin = ptsy[i];
            if (ptsx[i] > xmax) xmax = ptsx[i];
            if (ptsy[i] > ymax) ymax = ptsy[i];
        }
        long nx = (long)round((xmax - xmin) / dsa) + 3, ny = (long)round((ymax - ymin) / dsa) + 3;
        long *map = (long *)calloc(nx * ny, sizeof(long));
        long *ixs = (long *)malloc(sizeof(long) * nsa), *iys = (long *)malloc(sizeof(long) * nsa);
        for (long i = 0; i < nsa; i++) {
            ixs[i] = (long)round((ptsx[i] - xmin) / dsa) + 1;
            iys[i] = (long)round((ptsy[i] - ymin) / dsa) + 1;
            map[ixs[i] + iys[i] * nx] = i + 1;
        }
        int changed;
        do {
            changed = 0;
            for (long isa = 0; isa < nsa; isa++) {
                if (saa[isa] < saat) continue;
                int nedge = 0, ncorner = 0;
                for (int jx = -1; jx < 2; jx++) {
                    for (int jy = -1; jy < 2; jy++) {
                        if (jx == 0 && jy == 0) continue;
                        long jsa = map[(ixs[isa] + jx) + (iys[isa] + jy) * nx];
                        if (jsa && saa[jsa - 1] >= saat) {
                            if (abs(jx + jy) == 1) nedge++;
                            else ncorner++;
                        }
                    }
                }
                if (ncorner + nedge <= 2) {
                    saa[isa] = 0;
                    changed++;
                }
            }
        } while (changed);
        free(map);
        free(ixs);
        free(iys);
    }
    long count = 0;
    for (long isa = 0; isa < nsa; isa++) {
        if (saa[isa] >= saat) {
            if (count != isa) {
                salocx[count] = salocx[isa];
                salocy[count] = salocy[isa];
                ptsx[count] = ptsx[isa];
                ptsy[count] = ptsy[isa];
                memmove(amp + count * npsa, amp + isa * npsa, sizeof(double) * npsa);
                saa[count] = saa[isa];
            }
            count++;
        }
    }
    return count;
}

/* ------------------------------------------------------------------------
 * Detector transfer function and sampler: lib/mkdtf.c, lib/mkh.c
 * --------------------------------------------------------------------- */
static double sinc(double x) {
    /* MAOS sinc(x)=sin(pi x)/(pi x) */
    if (fabs(x) < 1.e-5) return 1.;
    return sin(M_PI * x) / (M_PI * x);
}

void mxs_dtf_nominal(double *nominal, long notfx, long notfy, double pdthetax,
                     double pdthetay, double pixrot, double pixblur) {
    const long notfx2 = notfx >> 1, notfy2 = notfy >> 1;
    double ct = pixrot != 0 ? cos(pixrot) : 1;
    double st = pixrot != 0 ? sin(pixrot) : 0;
    double duxp = pdthetax / notfx, duyp = pdthetay / notfy;
    pixblur = fabs(pixblur);
    double bxp2 = -2 * pow(M_PI * pixblur * duxp, 2);
    double byp2 = -2 * pow(M_PI * pixblur * duyp, 2);
    for (long iy = 0; iy < notfy; iy++) {
        long jy = iy - notfy2;
        for (long ix = 0; ix < notfx; ix++) {
            long jx = ix - notfx2;
            double ir = ct * jx + st * jy;
            double ia = -st * jx + ct * jy;
            nominal[ix + iy * notfx] =
                (pdthetax ? sinc(ir * duxp) * sinc(ia * duyp) * pdthetax * pdthetay : 1) *
                (pixblur ? exp(bxp2 * (ir * ir) + byp2 * (ia * ia)) : 1);
        }
    }
    /* treat: fftshift, fft(-1), fftshift, fft(+1), 1/n (lib/mkdtf.c:26-32) */
    long n = notfx * notfy;
    cplx *w = (cplx *)malloc(sizeof(cplx) * n);
    for (long i = 0; i < n; i++) w[i] = nominal[i];
    cfftshift(w, notfx, notfy);
    mxs_fft2((double *)w, notfx, notfy, -1);
    cfftshift(w, notfx, notfy);
    mxs_fft2((double *)w, notfx, notfy, 1);
    for (long i = 0; i < n; i++) nominal[i] = creal(w[i]) / (double)n;
    free(w);
}

long mxs_si(long notfx, long notfy, double dtheta, long pixpsax, long pixpsay,
            double pixthetax, double pixthetay, double pixoffx, double pixoffy,
            double pixrot, long *pp, long *pi, double *px) {
    /* loc_psf = mksqloc(notfx, notfy, dtheta, dtheta, -notfx2*dtheta, ...);
       loc_ccd = mksqloc(pixpsax, pixpsay, pixthetax, pixthetay, pxo-dx, pyo-dy)
       H = mkh(loc_psf, loc_ccd, 0, 0, 1, -theta)            lib/mkdtf.c:144-161 */
    const long notfx2 = notfx >> 1, notfy2 = notfy >> 1;
    const double pxo = -(pixpsax * 0.5 - 0.5) * pixthetax;
    const double pyo = -(pixpsay * 0.5 - 0.5) * pixthetay;
    const double ox = pxo - pixoffx * pixthetax, oy = pyo - pixoffy * pixthetay;
    const double angle = -pixrot;
    const double ct = cos(angle), st = sin(angle);
    const long npix = pixpsax * pixpsay;
    const long npsf = notfx * notfy;
    /* mkht builds H^T column by column (one column per pixel) */
    long *tp = (long *)malloc(sizeof(long) * (npix + 1));
    long *ti = (long *)malloc(sizeof(long) * 4 * npix);
    double *tx = (double *)malloc(sizeof(double) * 4 * npix);
    long count = 0;
    const double dx1 = 1. / dtheta;
    /* displacex=(0-map->ox)/dx, map->ox = locin min = -notfx2*dtheta; the map
       padding in loc_create_map is undone by loc_map_get, so index = floor. */
    const double dispx = notfx2, dispy = notfy2;
    for (long ipix = 0; ipix < npix; ipix++) {
        long ix0 = ipix % pixpsax, iy0 = ipix / pixpsax;
        double lx = ix0 * pixthetax + ox, ly = iy0 * pixthetay + oy;
        tp[ipix] = count;
        double x2 = ct * lx + st * ly;
        double y2 = -st * lx + ct * ly;
        double fx1 = x2 * dx1 + dispx, fy1 = y2 * dx1 + dispy;
        long nplocx = (long)floor(fx1), nplocy = (long)floor(fy1);
        double fx[2], fy[2];
        fx[1] = fx1 - nplocx;
        fy[1] = fy1 - nplocy;
        fx[0] = 1. - fx[1];
        fy[0] = 1. - fy[1];
        for (int iy = 0; iy < 2; iy++) {
            for (int ix = 0; ix < 2; ix++) {
                double weight = fx[ix] * fy[iy];
                long jx = nplocx + ix, jy = nplocy + iy;
                if (weight > EPS_D && jx >= 0 && jx < notfx && jy >= 0 && jy < notfy) {
                    ti[count] = jx + jy * notfx;
                    tx[count] = weight;
                    count++;
                }
            }
        }
        /* The renormalisation in mkht (lib/mkh.c:142-146) writes past the
           column and has no effect; edge pixels keep raw partial weights. */
    }
    tp[npix] = count;
    /* dspdroptol(EPS) then transpose to H (npix x npsf), CSC by psf column */
    for (long j = 0; j <= npsf; j++) pp[j] = 0;
    for (long k = 0; k < count; k++)
        if (fabs(tx[k]) > EPS_D) pp[ti[k] + 1]++;
    for (long j = 0; j < npsf; j++) pp[j + 1] += pp[j];
    long *next = (long *)malloc(sizeof(long) * npsf);
    for (long j = 0; j < npsf; j++) next[j] = pp[j];
    for (long ipix = 0; ipix < npix; ipix++) {
        for (long k = tp[ipix]; k < tp[ipix + 1]; k++) {
            if (fabs(tx[k]) > EPS_D) {
                long dst = next[ti[k]]++;
                pi[dst] = ipix;
                px[dst] = tx[k];
            }
        }
    }
    long nnz = pp[npsf];
    free(next);
    free(tp);
    free(ti);
    free(tx);
    return nnz;
}

/* ------------------------------------------------------------------------
 * LGS: srot/srsa, sodium smoothing, ETF (maos/powfs.c, lib/mkdtf.c)
 * --------------------------------------------------------------------- */
void mxs_srot_srsa(const double *salocx, const double *salocy, long nsa,
                   double dsa, double lltox, double lltoy, double *srot,
                   double *srsa) {
    double ox2 = lltox - dsa * 0.5, oy2 = lltoy - dsa * 0.5;
    for (long isa = 0; isa < nsa; isa++) {
        double ddx = salocx[isa] - ox2, ddy = salocy[isa] - oy2;
        srot[isa] = atan2(ddy, ddx);
        srsa[isa] = sqrt(ddx * ddx + ddy * ddy);
    }
}

long mxs_smooth(const double *prof, long nxin, double dxnew, double *out, long maxout) {
    const double x0in = prof[0];
    const double dxin = (prof[nxin - 1] - x0in) / (nxin - 1);
    if (!(dxnew > dxin * 2)) {
        long n = nxin < maxout ? nxin : maxout;
        for (long i = 0; i < n; i++) {
            out[i] = prof[i];
            out[i + n] = prof[i + nxin];
        }
        return n;
    }
    long nxnew = (long)ceil((prof[nxin - 1] - x0in) / dxnew);
    if (nxnew > maxout) nxnew = maxout;
    /* mkht(loc_out, loc_in): each input sample spreads onto the two nearest
       output samples with linear weights; missing neighbours drop out. */
    for (long i = 0; i < nxnew; i++) {
        out[i] = x0in + dxnew * i;
        out[i + nxnew] = 0;
    }
    double insum = 0, outsum = 0;
    for (long i = 0; i < nxin; i++) {
        double f = (prof[i] - x0in) / dxnew;
        long k = (long)floor(f);
        double w1 = f - k, w0 = 1. - w1;
        if (w0 > EPS_D && k >= 0 && k < nxnew) out[k + nxnew] += w0 * prof[i + nxin];
        if (w1 > EPS_D && k + 1 >= 0 && k + 1 < nxnew) out[k + 1 + nxnew] += w1 * prof[i + nxin];
        insum += prof[i + nxin];
    }
    /* dvecnormalize_sumabs: scale so sum(|out|)=sum(input) */
    for (long i = 0; i < nxnew; i++) outsum += fabs(out[i + nxnew]);
    if (outsum > 0)
        for (long i = 0; i < nxnew; i++) out[i + nxnew] *= insum / outsum;
    return nxnew;
}

/* The ETF as the Fourier transform of the sodium column projected on the
 * sky (the direct-sum form of mketf, lib/mkdtf.c:255-278, selected there by
 * llt.na_interp=0 or a one-sample profile).  A sample of strength w at
 * altitude h (range r = (h - htel)/cos(za)) seen from a subaperture at
 * distance rsa from the launch axis sits at the radial angle
 * theta = rsa/r - rsa/hs from the focus; along the radial direction
 * (cos srot, sin srot) it contributes w exp(-2 pi i k_r theta) at spatial
 * frequency k_r = cos(srot) kx + sin(srot) ky, with kx, ky signed multiples
 * of 1/(dtheta notf) in corner (FFT) order.  The phase is separable in kx
 * and ky, so each sample costs two 1-D phasor tables and one outer product.
 * No renormalisation: ETF(0) = sum of the strengths. */
static void mketf_direct(cplx *etf2d_all, const double *nominal, long notf, double dtheta,
                         const double *px, const double *pna, long nhp, const double *srot,
                         const double *srsa, long nsa, double hs, double htel, double cosza) {
    const double du = 1. / (dtheta * notf);
#pragma omp parallel
    {
        cplx *ex = (cplx *)malloc(sizeof(cplx) * 2 * notf);
        cplx *ey = ex + notf;
#pragma omp for schedule(dynamic, 8)
        for (long isa = 0; isa < nsa; isa++) {
            cplx *e = etf2d_all + isa * notf * notf;
            memset(e, 0, sizeof(cplx) * notf * notf);
            const double cx = cos(srot[isa]) * du, cy = sin(srot[isa]) * du;
            for (long ih = 0; ih < nhp; ih++) {
                if (!(pna[ih] > 0)) continue;
                const double theta = srsa[isa] * cosza / (px[ih] - htel) - srsa[isa] / hs;
                for (long i = 0; i < notf; i++) {
                    const double k = (double)(i < notf / 2 ? i : i - notf);
                    ex[i] = cexp(-2. * M_PI * I * (theta * cx * k));
                    ey[i] = pna[ih] * cexp(-2. * M_PI * I * (theta * cy * k));
                }
                for (long iy = 0; iy < notf; iy++)
                    for (long ix = 0; ix < notf; ix++) e[ix + iy * notf] += ex[ix] * ey[iy];
            }
            if (nominal)
                for (long i = 0; i < notf * notf; i++) e[i] *= nominal[i];
        }
        free(ex);
    }
}

int mxs_mketf(double *etf_out, double *etf1d_out, const double *nominal, long notf,
              double dtheta, const double *prof, long nhp, const double *srot,
              const double *srsa, long nsa, double hs, double htel, double za,
              int no_interp) {
    const double *px = prof, *pna = prof + nhp;
    const double cosza = cos(za);
    if (nhp < 1 || notf < 2 || (notf & 1)) return -1;
    if (no_interp || nhp == 1) {
        mketf_direct((cplx *)etf_out, nominal, notf, dtheta, px, pna, nhp, srot, srsa, nsa, hs, htel, cosza);
        return 0;
    }
    double hpmin = (px[0] - htel) / cosza;
    double dhp1 = cosza / (px[1] - px[0]);
    if (fabs((px[nhp - 1] - px[0]) / ((nhp - 1) * (px[1] - px[0])) - 1.) > 1.e-3) return -2;
    double i0scale = 0;
    for (long i = 0; i < nhp; i++) i0scale += pna[i];
    const long notfx = notf, notfy = notf;
    const long notfx2 = notfx >> 1, notfy2 = notfy >> 1;
    const int npad = 2, nover = 2;
    const long netf = notfx * nover * npad;
    const double dtetf = dtheta / nover;
    const double dusc = (netf * dtetf) / (dtheta * notfx);
    const long netf2 = netf >> 1;
    const long etf0 = netf2 - (long)round(notfx2 * (dtheta / dtetf));
    const long etf1 = etf0 + (long)round(notfx * (dtheta / dtetf));
    if (etf0 < 0) return -3;
    cplx *etf2d_all = (cplx *)etf_out;
#pragma omp parallel
    {
        cplx *etf = (cplx *)malloc(sizeof(cplx) * netf);
#pragma omp for schedule(dynamic, 16)
        for (long isa = 0; isa < nsa; isa++) {
            cplx *etf2d = etf2d_all + isa * notfx * notfy;
            memset(etf2d, 0, sizeof(cplx) * notfx * notfy);
            double rsa = srsa[isa];
            double etf2sum = 0;
            memset(etf, 0, sizeof(cplx) * netf);
            if (!rsa) {
                for (long i = 0; i < notfx * notfy; i++) etf2d[i] = 1;
            } else {
                for (long icomp = etf0; icomp < etf1; icomp++) {
                    const double itheta = (icomp - netf2) * dtetf;
                    const double ih = hs * rsa / (rsa + itheta * hs);
                    const double iih = (ih - hpmin) * dhp1;
                    const long iihf = (long)floor(iih);
                    const double iihw = iih - iihf;
                    if (iihf < 0 || iihf > nhp - 2) {
                        etf[icomp] = 0.;
                    } else {
                        double tmp = pna[iihf] * (1. - iihw) + pna[iihf + 1] * iihw;
                        etf[icomp] = tmp;
                        etf2sum += tmp;
                    }
                }
            }
            if (etf2sum) {
                for (long i = 0; i < netf; i++) etf[i] *= i0scale / etf2sum;
                cfftshift(etf, netf, 1);
                mxs_fft2((double *)etf, netf, 1, -1);
                cfftshift(etf, netf, 1);
                if (etf1d_out) memcpy(etf1d_out + isa * netf * 2, etf, sizeof(cplx) * netf);
                const double theta = srot[isa];
                const double ct = cos(theta), st = sin(theta);
                for (long icompy = 0; icompy < notfy; icompy++) {
                    double iy = (double)(icompy - notfy2);
                    for (long icompx = 0; icompx < notfx; icompx++) {
                        double ix = (double)(icompx - notfx2);
                        double ir = (dusc * (ct * ix + st * iy)) + netf2;
                        long iir = (long)floor(ir);
                        ir = ir - iir;
                        if (iir >= 0 && iir < netf - 1) {
                            etf2d[icompx + icompy * notfx] = etf[iir] * (1. - ir) + etf[iir + 1] * ir;
                        }
                    }
                }
                cfftshift(etf2d, notfx, notfy);
            } else if (etf1d_out) {
                memset(etf1d_out + isa * netf * 2, 0, sizeof(cplx) * netf);
            }
            /* fuse nominal (ccwmd, lib/mkdtf.c:389-393) */
            if (nominal)
                for (long i = 0; i < notfx * notfy; i++) etf2d[i] *= nominal[i];
        }
        free(etf);
    }
    return 0;
}

long mxs_llt_amp(double lltd, double widthp, double dx, double dsa, double *amps,
                 long maxn, double *origin) {
    double lltdsa = lltd > dsa ? lltd : dsa;
    long nxsa = (long)round(lltdsa / dx);
    lltdsa = dx * nxsa;
    double o = (dx - lltdsa) * 0.5;
    if (origin) *origin = o;
    if (nxsa * nxsa > maxn) return -nxsa;
    double sumamp2 = 0;
    const double l2max = pow(lltd * 0.5, 2);
    const double r2waist = pow(lltd * 0.5 * widthp, 2);
    for (long i = 0; i < nxsa * nxsa; i++) amps[i] = 0;
    for (long iy = 0; iy < nxsa; iy++) {
        double yy = iy * dx + o;
        for (long ix = 0; ix < nxsa; ix++) {
            double xx = ix * dx + o;
            double r2 = xx * xx + yy * yy;
            double amp = exp(-r2 / r2waist);
            sumamp2 += amp * amp;
            if (r2 <= l2max) amps[iy * nxsa + ix] = amp;
        }
    }
    sumamp2 = 1. / sqrt(sumamp2);
    for (long i = 0; i < nxsa * nxsa; i++) amps[i] *= sumamp2;
    return nxsa;
}

/* ------------------------------------------------------------------------
 * Atmosphere: lib/turbulence.c
 * --------------------------------------------------------------------- */
typedef struct {
    unsigned long long s[4];
} rng_t;
static unsigned long long rotl(unsigned long long x, int k) { return (x << k) | (x >> (64 - k)); }
static unsigned long long rng_next(rng_t *r) {
    unsigned long long *s = r->s;
    unsigned long long result = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}
static void rng_seed(rng_t *r, unsigned long long seed) {
    /* splitmix64 expansion of the seed */
    for (int i = 0; i < 4; i++) {
        unsigned long long z = (seed += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        r->s[i] = z ^ (z >> 31);
    }
}
static double rng_unif(rng_t *r) { return ((rng_next(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

void mxs_genscreen(double *screens, long nx, long ny, double dx, double r0, double L0,
                   const double *wt, long nlayer, unsigned long long seed) {
    const double strength = 0.0229 * pow(r0, -5. / 3.) * pow((0.5e-6) / (2. * M_PI), 2);
    const double slope = -11. / 3. * 0.5 / 2.; /* slope*power/2 with power=0.5 */
    const double dfx = 1. / (nx * dx), dfy = 1. / (ny * dx);
    const double zerofreq2 = L0 == 0 ? 0 : pow(L0, -2);
    const double scrnstr = pow(strength * (dfx * dfy), 0.5);
    const long nx2 = nx / 2, ny2 = ny / 2;
    double *spect = (double *)malloc(sizeof(double) * nx * ny);
    for (long iy = 0; iy < ny; iy++) {
        double r2y = pow((iy < ny2 ? iy : iy - ny) * dfy, 2);
        for (long ix = 0; ix < nx; ix++) {
            double r2 = pow((ix < nx2 ? ix : ix - nx) * dfx, 2) + r2y;
            spect[ix + iy * nx] = pow(r2 + zerofreq2, slope) * scrnstr;
        }
    }
    spect[0] = 0; /* remove piston */
    rng_t rs;
    rng_seed(&rs, seed);
    cplx *c = (cplx *)malloc(sizeof(cplx) * nx * ny);
    for (long il = 0; il < nlayer; il += 2) {
        for (long i = 0; i < nx * ny; i++) {
            /* Box-Muller pair: real and imaginary parts */
            double u1 = rng_unif(&rs), u2 = rng_unif(&rs);
            double rr = sqrt(-2. * log(u1));
            c[i] = (rr * cos(2 * M_PI * u2) + I * rr * sin(2 * M_PI * u2)) * spect[i];
        }
        mxs_fft2((double *)c, nx, ny, -1);
        double s0 = sqrt(wt[il]);
        double *l0 = screens + il * nx * ny;
        for (long i = 0; i < nx * ny; i++) l0[i] = creal(c[i]) * s0;
        if (il + 1 < nlayer) {
            double s1 = sqrt(wt[il + 1]);
            double *l1 = screens + (il + 1) * nx * ny;
            for (long i = 0; i < nx * ny; i++) l1[i] = cimag(c[i]) * s1;
        }
    }
    free(c);
    free(spect);
}

/* ------------------------------------------------------------------------
 * Matched filter: lib/mtch.c
 * --------------------------------------------------------------------- */
void mxs_derive_by_fft(double *out, const double *i0, long nx, long ny, double theta) {
    cplx *otf = (cplx *)malloc(sizeof(cplx) * nx * ny);
    for (long i = 0; i < nx * ny; i++) otf[i] = i0[i];
    mxs_fft2((double *)otf, nx, ny, -1);
    double sx = cos(theta) * 2. * M_PI / nx;
    double sy = sin(theta) * 2. * M_PI / ny;
    long ny2 = ny / 2, nx2 = nx / 2;
    for (long iy = 0; iy < ny; iy++)
        for (long ix = 0; ix < nx; ix++)
            otf[ix + iy * nx] *= -I * ((ix < nx2 ? ix : (ix - nx)) * sx + (iy < ny2 ? iy : (iy - ny)) * sy);
    mxs_fft2((double *)otf, nx, ny, 1);
    for (long i = 0; i < nx * ny; i++) out[i] = creal(otf[i]) / (double)(nx * ny);
    free(otf);
}

/* Symmetric eigen-decomposition by cyclic Jacobi: A = V diag(d) V^T. */
static void jacobi_eig(double *A, int n, double *d, double *V) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i + j * n] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p + q * n] * A[p + q * n];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = A[p + q * n];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p + p * n], aqq = A[q + q * n];
                double tau = (aqq - app) / (2 * apq);
                double t = (tau >= 0 ? 1. : -1.) / (fabs(tau) + sqrt(1 + tau * tau));
                double c = 1. / sqrt(1 + t * t), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = c * akp - s * akq;
                    A[k + q * n] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = c * apk - s * aqk;
                    A[q + k * n] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = c * vkp - s * vkq;
                    V[k + q * n] = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < n; i++) d[i] = A[i + i * n];
}

void mxs_sym_eig(double *A, int n, double *d, double *V) { jacobi_eig(A, n, d, V); }

/* (A^T W A)^-1 A^T W with the SVD thresholds of dpinv (math/blas.c:170-206,
 * svd_pow2 :441-485 with thres1=1e-7, thres2=1e-3).  A is m x n. */
static void dpinv_w(double *out, const double *A, const double *wt, int m, int n) {
    double cc[64], d[8], V[64], inv[64];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < m; k++) s += A[k + i * m] * wt[k] * A[k + j * m];
            cc[i + j * n] = s;
        }
    jacobi_eig(cc, n, d, V);
    /* sort eigenvalues descending (SVD order) */
    int idx[8];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (d[idx[j]] > d[idx[i]]) {
                int t = idx[i];
                idx[i] = idx[j];
                idx[j] = t;
            }
    double s[8];
    double maxeig = d[idx[0]];
    double thres10 = 1e-7 * maxeig, thres20 = 1e-3 * maxeig;
    int cut = 0;
    for (int i = 0; i < n; i++) {
        double ev = d[idx[i]];
        if (!cut && ev > thres10 && ev > thres20) {
            thres20 = ev * 1e-3;
            s[i] = 1. / ev;
        } else {
            cut = 1;
            s[i] = 0;
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double acc = 0;
            for (int k = 0; k < n; k++) acc += V[i + idx[k] * n] * s[k] * V[j + idx[k] * n];
            inv[i + j * n] = acc;
        }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < m; k++) {
            double acc = 0;
            for (int j = 0; j < n; j++) acc += inv[i + j * n] * A[k + j * m];
            out[i + k * n] = acc * wt[k];
        }
}

/* mki0sh (lib/mtch.c:23-33) */
static void mki0sh(double *i0x1, double *i0x2, const double *i0, long nx, long ny,
                   double scale, long sx, long sy) {
    for (long iy = 0; iy < ny - sy; iy++)
        for (long ix = 0; ix < nx - sx; ix++) {
            i0x1[(ix + sx) + (iy + sy) * nx] = i0[ix + iy * nx] * scale;
            i0x2[ix + iy * nx] = i0[(ix + sx) + (iy + sy) * nx] * scale;
        }
}

/* mtch (lib/mtch.c:93-211) for pixrot=0, no qe, no per-pixel background. */
static void mtch_one(double *mtche, double *nea2, const double *i0, const double *gx0,
                     const double *gy0, long nx, long ny, double bkgrnd, double bkgrndc,
                     double rne, double pixthetax, double pixthetay, int cr) {
    const long i0n = nx * ny;
    const double bkgrnd_res = bkgrnd - bkgrndc;
    const double kpx = 1. / pixthetax, kpy = 1. / pixthetay;
    int nmod = 3, mtchcrx = 0, mtchcry = 0;
    if (cr == 1 || cr == 2) {
        mtchcrx = nmod;
        nmod += 2;
    }
    if (cr == 1 || cr == 3) {
        mtchcry = nmod;
        nmod += 2;
    }
    double i0m[2 * 7] = {0};
    double *i0g = (double *)calloc(i0n * nmod, sizeof(double));
    double *wt = (double *)malloc(sizeof(double) * i0n);
    double *gxb = NULL, *gyb = NULL;
    i0m[0 + 2 * 0] = 1;
    i0m[1 + 2 * 1] = 1;
    if (mtchcrx) {
        i0m[0 + 2 * mtchcrx] = 1;
        i0m[1 + 2 * mtchcrx] = 0;
        i0m[0 + 2 * (mtchcrx + 1)] = -1;
        i0m[1 + 2 * (mtchcrx + 1)] = 0;
    }
    if (mtchcry) {
        i0m[0 + 2 * mtchcry] = 0;
        i0m[1 + 2 * mtchcry] = 1;
        i0m[0 + 2 * (mtchcry + 1)] = 0;
        i0m[1 + 2 * (mtchcry + 1)] = -1;
    }
    const double *gx = gx0, *gy = gy0;
    if (!gx) {
        gxb = (double *)malloc(sizeof(double) * i0n);
        gyb = (double *)malloc(sizeof(double) * i0n);
        mxs_derive_by_fft(gxb, i0, nx, ny, 0);
        mxs_derive_by_fft(gyb, i0, nx, ny, M_PI / 2);
        for (long i = 0; i < i0n; i++) {
            gxb[i] *= kpx;
            gyb[i] *= kpy;
        }
        gx = gxb;
        gy = gyb;
    }
    /* addvec with a NULL per-pixel background still adds its offset theta,
       so column 2 receives bkgrnd_res twice and the shifted columns once
       (lib/mtch.c:150-166). */
    for (long i = 0; i < i0n; i++) {
        i0g[i + 0 * i0n] = gx[i];
        i0g[i + 1 * i0n] = gy[i];
        i0g[i + 2 * i0n] = i0[i] * kpx + 2 * bkgrnd_res;
    }
    if (mtchcrx) {
        mki0sh(i0g + mtchcrx * i0n, i0g + (mtchcrx + 1) * i0n, i0, nx, ny, kpx, 1, 0);
        for (long i = 0; i < i0n; i++) {
            i0g[i + mtchcrx * i0n] += bkgrnd_res;
            i0g[i + (mtchcrx + 1) * i0n] += bkgrnd_res;
        }
    }
    if (mtchcry) {
        mki0sh(i0g + mtchcry * i0n, i0g + (mtchcry + 1) * i0n, i0, nx, ny, kpy, 0, 1);
        for (long i = 0; i < i0n; i++) {
            i0g[i + mtchcry * i0n] += bkgrnd_res;
            i0g[i + (mtchcry + 1) * i0n] += bkgrnd_res;
        }
    }
    const double rne2 = rne * rne;
    for (long i = 0; i < i0n; i++) {
        if (i0[i] < 0) wt[i] = 1. / rne2;
        else wt[i] = 1. / (rne2 + bkgrnd + i0[i]);
    }
    double *tmp = (double *)malloc(sizeof(double) * nmod * i0n);
    dpinv_w(tmp, i0g, wt, (int)i0n, nmod);
    /* mtche = i0m * tmp : 2 x i0n */
    for (int r = 0; r < 2; r++)
        for (long k = 0; k < i0n; k++) {
            double acc = 0;
            for (int j = 0; j < nmod; j++) acc += i0m[r + 2 * j] * tmp[j + k * nmod];
            mtche[r + 2 * k] = acc;
        }
    /* nea2 = mtche diag(1/wt) mtche^T (dtmcc) */
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double acc = 0;
            for (long k = 0; k < i0n; k++) acc += mtche[a + 2 * k] * mtche[b + 2 * k] / wt[k];
            nea2[a + 2 * b] = acc;
        }
    free(tmp);
    free(i0g);
    free(wt);
    free(gxb);
    free(gyb);
}

int mxs_mtch(double *mtche, double *sanea, double *i0sum, double *i0sumsum,
             const double *i0, const double *gx, const double *gy, long nsa,
             long pixpsax, long pixpsay, double bkgrnd, double bkgrndc, double rne,
             double pixthetax, double pixthetay, int mtchcr, double sigratio) {
    const long npix = pixpsax * pixpsay;
    if (!sigratio) sigratio = 1;
    double i0max = 0;
    for (long isa = 0; isa < nsa; isa++) {
        double s = 0;
        for (long i = 0; i < npix; i++) s += i0[isa * npix + i];
        i0sum[isa] = s;
        if (isa == 0 || s > i0max) i0max = s;
    }
    double i0thres = i0max * 0.1 > 10 * rne ? i0max * 0.1 : 10 * rne;
    double nea2thres = pixthetax * pixthetay * 100;
    double sumsum = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : sumsum)
    for (long isa = 0; isa < nsa; isa++) {
        double nea2[4];
        double *m = mtche + isa * 2 * npix;
        mtch_one(m, nea2, i0 + isa * npix, gx ? gx + isa * npix : NULL,
                 gy ? gy + isa * npix : NULL, pixpsax, pixpsay, bkgrnd, bkgrndc, rne,
                 pixthetax, pixthetay, mtchcr);
        if (fabs(sigratio - 1) > 1e-5)
            for (int k = 0; k < 4; k++) nea2[k] /= sigratio;
        sumsum += i0sum[isa];
        if (i0sum[isa] < i0thres || nea2[0] > nea2thres || nea2[3] > nea2thres) {
            nea2[0] = nea2[3] = nea2thres;
            nea2[1] = nea2[2] = 0;
            for (long i = 0; i < 2 * npix; i++) m[i] = 0;
        }
        if (sanea) {
            sanea[isa + 0 * nsa] = nea2[0];
            sanea[isa + 1 * nsa] = nea2[3];
            sanea[isa + 2 * nsa] = nea2[1];
        }
    }
    if (i0sumsum) *i0sumsum = sumsum;
    return 0;
}

/* dither_position (maos/utils.c:884-899): the dither phase at the start of
 * the integration, shifted by the FSM delay (alfsm + 1 - dtrat frames),
 * and the next one, blended by the fractional delay beta; the frame counts
 * are integer divisions as there. */
void mxs_dither_position(double *cs, double *ss, int alfsm, int dtrat, int npoint, long isim,
                         double deltam) {
    const long adjust = alfsm + 1 - dtrat, adjust2 = dtrat - 1;
    const double dphi = 2. * M_PI / npoint, phi0 = 0.5 * M_PI;
    const double a1 = phi0 + (double)((isim - adjust - adjust2) / dtrat) * dphi + deltam;
    const double a2 = phi0 + (double)((isim - adjust) / dtrat) * dphi + deltam;
    const double delay = (double)adjust / dtrat;
    const double beta = 1 + delay + floor(-delay);
    const double sc = 1. / (beta * beta + (1 - beta) * (1 - beta));
    *cs = (beta * cos(a1) + (1 - beta) * cos(a2)) * sc;
    *ss = (beta * sin(a1) + (1 - beta) * sin(a2)) * sc;
}
