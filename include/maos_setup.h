/*
 * maos_setup.h -- host-side FP64 table builders for the Shack-Hartmann
 * physical-optics WFS path (libmaos_setup.so).
 *
 * In a MAOS run these tables come from MAOS's own setup (maos/powfs.c,
 * lib/mkdtf.c, lib/mtch.c, lib/turbulence.c) and are handed to the GPU
 * library through maos_mi355x.h.  This library restates those builders so
 * that the GPU path can be run and benchmarked without MAOS.  Every function
 * cites the reference routine it follows.  Arrays are column-major with x
 * fastest, as in MAOS's dmat/cmat (math/type.h:31-41).  Complex arrays are
 * interleaved (re, im) doubles.
 */
#ifndef MAOS_SETUP_H
#define MAOS_SETUP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* 2-D unnormalised FFT in place on nx*ny complex (x fastest). dir=-1 is the
 * forward transform exp(-i...), dir=+1 the backward one.  Follows the FFTW
 * convention used by cfft2 (math/fft.c:253-281). */
void mxs_fft2(double *c, long nx, long ny, int dir);

/* nextfftsize (math/mathmisc.c:109-146). */
long mxs_nextfftsize(long n);

/* loc_create_embed (math/loc.c:147-179): embed[i] = ix + iy * nembed, the
 * index of point i in a square grid of side nembed.  *nembed > 0 gives the
 * side (error -2 below 0.6 of the minimum ceil(max(nx, ny) * oversize));
 * otherwise it is that minimum, raised by nextfftsize when fftpad, and is
 * written back.  Returns the side, or < 0 on error (-3: a point outside). */
long mxs_loc_create_embed(long *embed, long *nembed, long nloc, const double *locx, const double *locy, double dx,
                          double dy, double oversize, int fftpad);

/* Subaperture lower-left corners on the padded grid, before area pruning
 * (maos/powfs.c:302-340).  Returns the count written (<= maxn). */
long mxs_saloc(int order, double dsa, double din, double misregrmax,
               double *locx, double *locy, long maxn);

/* Anti-aliased annulus sampled at each point: loc_circle_add
 * (math/loc.c:942-986) with val=1, added into amp. */
void mxs_loc_circle_add(double *amp, const double *locx, const double *locy,
                        long nloc, double dx, double dy, double cx, double cy,
                        double r, double rin, double val);

/* Normalised subaperture area from the amplitude map: wfsamp2saa
 * (maos/powfs.c:86-99), then scaled by areascale (maos/powfs.c:412-414). */
void mxs_saa(const double *amp, long nsa, long npsa, double areascale,
             double *saa);

/* Isolated-subaperture rule and area pruning (sa_reduce, maos/powfs.c:
 * 141-237).  Compacts salocx/y, ptsx/y, amp and saa in place; returns the
 * new nsa. */
long mxs_sa_reduce(double *salocx, double *salocy, double *ptsx, double *ptsy,
                   double *amp, double *saa, long nsa, long npsa, double dsa,
                   double saat);

/* Detector transfer function (dtf_otf, lib/mkdtf.c:51-80) with treat=1
 * (dotf_treat, lib/mkdtf.c:26-39). nominal is notfx*notfy real. */
void mxs_dtf_nominal(double *nominal, long notfx, long notfy, double pdthetax,
                     double pdthetay, double pixrot, double pixblur);

/* Bilinear PSF->pixel sampler (mkdtf's mkh call, lib/mkdtf.c:147-161 and
 * lib/mkh.c:45-155).  Output is MAOS's dsp layout of H: CSC with nx=npix
 * rows and ny=notfx*notfy columns.  pp has notfx*notfy+1 entries; pi/px
 * need at most 4*npix entries.  Returns nnz. */
long mxs_si(long notfx, long notfy, double dtheta, long pixpsax, long pixpsay,
            double pixthetax, double pixthetay, double pixoffx,
            double pixoffy, double pixrot, long *pp, long *pi, double *px);

/* LGS subaperture polar coordinates relative to the LLT
 * (maos/powfs.c:744-758). */
void mxs_srot_srsa(const double *salocx, const double *salocy, long nsa,
                   double dsa, double lltox, double lltoy, double *srot,
                   double *srsa);

/* Sodium profile resampling by linear binning (smooth, lib/mkdtf.c:
 * 421-452).  prof is nh x 2 (column 0 height, column 1 strength).  Writes
 * up to maxout rows to out (maxout x 2, column-major) and returns the row
 * count; returns nh and copies when no smoothing is needed. */
long mxs_smooth(const double *prof, long nh, double dxnew, double *out,
                long maxout);

/* Elongation transfer function (mketf, lib/mkdtf.c:186-396), fused with
 * the nominal (lib/mkdtf.c:389-393; nominal may be NULL).  prof is nh x 2
 * (column-major: heights, then strengths).  etf is nsa*notf*notf complex,
 * peak in corner.  Two forms, as MAOS's llt.na_interp selects them:
 * no_interp=0 (the default there) interpolates the profile onto a uniform
 * angle grid, FFTs it and rotates the 1-D result onto the 2-D grid
 * (lib/mkdtf.c:281-378); it needs evenly spaced heights (returns -2
 * otherwise, where MAOS stops with an error).  no_interp=1, or a profile of
 * one sample, sums the Fourier phasor of every sample directly
 * (lib/mkdtf.c:255-278): any spacing, no renormalisation.  If etf1d is not
 * NULL the interpolating form also writes the per-subaperture 1-D ETF
 * (nsa * 4*notf complex, peak in centre).  Returns 0, or -1 for bad sizes,
 * -3 if the FoV exceeds the interpolation grid. */
int mxs_mketf(double *etf, double *etf1d, const double *nominal, long notf,
              double dtheta, const double *prof, long nh, const double *srot,
              const double *srsa, long nsa, double hs, double htel,
              double za, int no_interp);

/* LLT pupil (setup_powfs_llt, maos/powfs.c:1182-1242): Gaussian beam of
 * waist widthp*d/2 clipped at d/2, normalised so sum(amp^2)=1 for the
 * unclipped beam.  Returns nx; amp holds nx*nx values, *origin the pts
 * origin (same in x and y). */
long mxs_llt_amp(double lltd, double widthp, double dx, double dsa,
                 double *amp, long maxn, double *origin);

/* Seeded von Karman FFT screens (spect_screen_do, lib/turbulence.c:70-171;
 * spatial_psd :436-480).  Two layers come from one complex FFT.  Output is
 * nlayer*nx*ny; each layer is scaled by sqrt(wt[i]).  The random stream is
 * this library's own (xoshiro256** + Box-Muller), not MAOS's MT19937. */
void mxs_genscreen(double *screens, long nx, long ny, double dx, double r0,
                   double L0, const double *wt, long nlayer,
                   unsigned long long seed);

/* Derivative by FFT (derive_by_fft, lib/mtch.c:39-57). */
void mxs_derive_by_fft(double *out, const double *i0, long nx, long ny,
                       double theta);

/* Matched filter for all subapertures (mtch_cell, lib/mtch.c:223-335, with
 * mtch :93-211).  i0/gx/gy are nsa*npix (gx/gy may be NULL: FFT derivative,
 * as with powfs.mtchfft).  Outputs mtche nsa*(2*npix) (each a 2 x npix
 * column-major matrix), sanea nsa*3 (optional), i0sum nsa, *i0sumsum. */
int mxs_mtch(double *mtche, double *sanea, double *i0sum, double *i0sumsum,
             const double *i0, const double *gx, const double *gy, long nsa,
             long pixpsax, long pixpsay, double bkgrnd, double bkgrndc,
             double rne, double pixthetax, double pixthetay, int mtchcr,
             double sigratio);


/* Average-gradient (G-tilt) operator GS0 (mkg, lib/mkg.c:108-468, as
 * setup_shwfs_grad calls it, maos/powfs.c:603-630): CSR with 2*nsa rows
 * (x gradients, then y) over the nsa*nxsa^2 OPD points (subaperture isa's
 * point (ix, iy) at ptsx[isa] + ix dx, ptsy[isa] + iy dx; the grids form
 * one common grid, so a subaperture's edge cells use its neighbours'
 * points).  Row r: the amplitude-weighted average over the subaperture
 * square of the gradient of the bilinear OPD, over cells whose four corners
 * exist with amplitude > 1e-14; zero if the amplitude integral is at most 1%
 * of the area; empty if saa < saat (saa may be NULL).  rowp has 2*nsa+1
 * entries.  Returns nnz, or -2 if more than nzmax entries are needed
 * ((nxsa+2)^2 per row suffice). */
long mxs_mkg(long *rowp, long *coli, double *val, long nzmax, const double *ptsx, const double *ptsy,
             long nsa, long nxsa, double dx, const double *amp, const double *saa, double saat);

/* Dither phase of a WFS for step isim, averaged over the integration
 * (dither_position, maos/utils.c:884-899): alfsm the FSM delay, dtrat the
 * WFS frame ratio, npoint steps per dither period, deltam the PLL phase. */
void mxs_dither_position(double *cs, double *ss, int alfsm, int dtrat, int npoint, long isim,
                         double deltam);

/* Model i0 of a powfs from the turbulence-averaged short-exposure OTF
 * (genseotf -> gensepsf -> gensei, lib/gensei.c:42-419 and
 * lib/genotf.c:50-300, as maos/powfs.c:1545-1622 builds it for the matched
 * filter).  Complex arrays are interleaved (re, im). */

/* OPD covariance (m^2) at separations r of von Karman turbulence (L0
 * finite) or Kolmogorov (L0 = inf) (turbcov, lib/turbulence.c:400-431). */
int mxs_turbcov(const double *r, long n, double r0, double L0, double *cov);

/* Turbulence OTF of each subaperture of an nxsa x nxsa grid of step dx
 * (genotf, lib/genotf.c:230-300): npsf x npsf per subaperture, peak in the
 * centre, OTF(0) = 1.  amp is nsa*nxsa^2 (NULL: 1); area nsa (NULL: all
 * computed), subapertures above thres share the largest one's OTF, those
 * at or below 0.01 are zero.  cov (nxsa^2 x nxsa^2) replaces the von Karman
 * covariance of r0, L0 when not NULL.  pttr removes the amp-weighted
 * piston/tip/tilt from both sides of the covariance. */
int mxs_genotf(double *otf, long nxsa, double dx, const double *amp, const double *area, long nsa,
               double thres, double wvl, const double *cov, double r0, double L0, long npsf, int pttr);

/* Short-exposure PSF of each subaperture (gensepsf, lib/gensei.c:136-206):
 * otf nsa*npsf^2 (peak in the centre) times conj(lotf) (nlotf^2, centred
 * embed; NULL: none), inverse FFT at npsf, times saa/npsf^2 (saa NULL: 1),
 * then cropped or zero-padded, centred, to psf nsa*notf^2 (real, peak in
 * the centre). */
int mxs_gensepsf(double *psf, const double *otf, long npsf, long nsa, const double *lotf, long nlotf,
                 const double *saa, long notf);

/* Adds one wavelength's model i0, gx, gy (nsa*npix each; any may be NULL)
 * for sepsf psf (gensei, lib/gensei.c:213-419): OTF of the PSF, optionally
 * re-centred by its thresholded CoG (shift2center, powfs.mtchstc), times the
 * ETF (etf nsa*notf^2 complex, nominal fused) or the nominal (notf^2 real),
 * inverse FFT, sampled by si (CSC, as mxs_si) with weight wvlsig; gx, gy
 * use the derivative factors -2 pi i f/(dtheta notf).  i0scale rescales
 * each subaperture's sums to saa. */
int mxs_gensei(double *i0, double *gx, double *gy, const double *psf, long nsa, long notf,
               const double *etf, const double *nominal, long npix, const long *pp, const long *pi,
               const double *px, double dtheta, double wvlsig, int shift2center, int i0scale,
               const double *saa);
/* MAOS .bin files (sys/bin.c:693-800, sys/bin.h:39-84): header = optional
 * comment block, M_SKIP, magic, nx, ny (64-bit), then nx*ny elements, x
 * fastest; ".gz" names are gzip-compressed (zfopen, sys/bin.c:301).  Every
 * call returns 0 on success, 1 at the end of a file (read_header), negative
 * on error. */
#define MXS_M_DBL 0x6402u     /* double */
#define MXS_M_INT64 0x6403u   /* int64 */
#define MXS_M_CMP 0x6404u     /* double complex */
#define MXS_M_FLT 0x6408u     /* float */
#define MXS_MCC_ANY 0x6421u   /* cell of any type */
#define MXS_M_COMMENT 0x6500u /* keyword text */
#define MXS_M_SKIP 0x6600u    /* padding word */
#define MXS_M_EOD 0x64FFu     /* end of data */
typedef struct mxs_bin_s mxs_bin_t;
int mxs_bin_open(mxs_bin_t **out, const char *fn, int write);
int mxs_bin_close(mxs_bin_t *f);
int mxs_bin_write_header(mxs_bin_t *f, uint32_t magic, uint64_t nx, uint64_t ny,
                         const char *keywords);
int mxs_bin_write_data(mxs_bin_t *f, const void *data, uint64_t bytes);
/* keywords (may be NULL) receives the concatenated comment text, up to
 * klen - 1 characters */
int mxs_bin_read_header(mxs_bin_t *f, uint32_t *magic, uint64_t *nx, uint64_t *ny,
                        char *keywords, uint64_t klen);
int mxs_bin_read_data(mxs_bin_t *f, void *data, uint64_t bytes);

/* The LGS matched-filter loop (host, FP64; maos_amd/host/loop.c).
 *
 * mxs_calc_dither_amp: the amplitude of a dither of npoint frames per cycle
 * in a signal (nmod x nsim, a column per step, a WFS frame every dtrat
 * columns) by synchronous detection, optionally detrended; combine=1 for a
 * circular tip/tilt dither (nmod = 2) (calc_dither_amp, lib/libmisc.c:
 * 347-414).  res (nmod, or NULL) receives each mode's amplitude; returns
 * the first, or -1 on bad arguments. */
double mxs_calc_dither_amp(double *res, const double *signal, long nmod, long nsim, long dtrat, long npoint,
                           int detrend, int combine);
/* State and keys of one WFS's tip/tilt dither loop (dither_t, maos/types.h:
 * 470-492; powfs.dither_* keys).  Fill the keys, zero the state. */
typedef struct mxs_dither_t {
    int npoint, pllrat, pllskip, dtrat, alfsm, detrend; /* detrend: 1 for NGS, 0 with an LLT */
    double amp, gpll, gdrift;                           /* dither_amp (rad), dither_gpll, dither_gdrift */
    double delta, deltam, deltao, a2m, a2me;            /* state */
} mxs_dither_t;
/* Per step (maos/wfsgrad.c:584-606): the step's dither phase (cs, ss) from
 * the last PLL output, and, from step pllskip on, the PLL update from the
 * residual tip/tilt fsmerr[2].  Returns 0 or -1. */
int mxs_dither_step(mxs_dither_t *d, long isim, const double *fsmerr, double *cs, double *ss);
/* At a PLL output (maos/wfsgrad.c:638-653, 707-731): deltam, a2m and a2me
 * from the FSM command and error histories (2 x ncol, ncol = (pllrat - 1)
 * dtrat + 1 steps ending now), and the fold scales s0, sg of
 * mx_wfsgrad_dither_fold.  Returns 0 or -1. */
int mxs_dither_output(mxs_dither_t *d, const double *fsmcmds, const double *fsmerrs, long ncol, double *s0,
                      double *sg);
/* The dither tip/tilt tt[2] to add to the FSM error and, through the TT
 * operator, to the gradients (maos/wfsgrad.c:734-751). */
void mxs_dither_tt(const mxs_dither_t *d, double cs, double ss, double *tt);
/* Sodium profile fit of accumulated LGS images (sodium_fit, one iteration,
 * matched-filter path, maos/powfs_utils.c:146-360; no gradient offsets):
 * model images of one-sample profiles every dh from 80 to 105 km (mketf's
 * direct sum, then gensei), least squares with the eigenvalue thresholds of
 * svd_pow2 (svdthres and 1e-3 of the previous), the profile scaled to sum
 * 1 into prof (nh x 2); i0, gx, gy (any may be NULL) replaced by the model
 * of the fitted profile.  i0in, i0, gx, gy: [ni0][nsa][npix]; sepsf
 * [nwvl][nsa][notf^2]; nominal [nwvl][notf^2]; si per wavelength as mxs_si
 * ([nwvl][notf^2 + 1], [nwvl][si_nzmax]); siglev [ni0].  Returns nh, -1 on
 * bad arguments, -2 if maxnh is short. */
long mxs_sodium_fit(double *prof, long maxnh, double *i0, double *gx, double *gy, const double *i0in, long ni0,
                    long nsa, long notf, long npix, long nwvl, const double *sepsf, const double *nominal,
                    const long *si_pp, const long *si_pi, const double *si_px, long si_nzmax,
                    const double *dtheta, const double *wvlwts, const double *siglev, const double *saa,
                    const double *srot, const double *srsa, double dh, double hs, double htel, double za,
                    double svdthres);

/* Pyramid WFS tables (pywfs_new, lib/pywfs.c:110-370).
 *
 * mxs_pywfs_pyramid: the focal-plane phase mask of one wavelength, notf x
 * notf complex (x fastest), apex at (notf/2, notf/2): exp(i pi/2 h) with h
 * the facet height in focal-plane samples, |x| + |y| for nside 4, |x| for a
 * roof (nside 2), the three-facet form of nside 3, a pi dot for nside 1;
 * zero outside the clear aperture of diameter nstop samples (notf without a
 * field stop, ceil(fieldstop/dtheta/2)*2 with one, lib/pywfs.c:176-186);
 * eskip/vskip are the half-widths (samples) of flat edges and vertex
 * (flate/dtheta/2, flatv/dtheta/2).  Returns 0, or -1 for bad sizes. */
int mxs_pywfs_pyramid(double *mask, long notf, int nside, long nstop, double eskip,
                      double vskip);
/* The detector transfer function of the pupil images, notf x notf real in
 * FFT (corner) order: pixel box of side dsa and Gaussian blur of rms
 * pixblur*dsa, for pupil-image sampling dx2 (lib/pywfs.c:236-262). */
void mxs_pywfs_nominal(double *nominal, long notf, double dsa, double dx2, double pixblur);
/* Detector subaperture centres: order x order grid of pitch dsa centred on
 * the axis (lib/pywfs.c:281-296); returns order^2 or -1 if maxn is short. */
long mxs_pywfs_saloc(double *x, double *y, long maxn, long order, double dsa);
/* Centre of each of the nside pupil images relative to the grid centre, in
 * units of notf/2 samples, [nside][2] (pywfs_mksi, lib/pywfs.c:34-107). */
int mxs_pywfs_sioff(double *sioff, int nside);

/* zfarr (math/zfarr.c:33-110): a cell file of nstep x 1 elements written one
 * per step, as MAOS saves telemetry (wfs%d_gradcl_%d.bin and the like,
 * maos/sim_utils.c:955-998).  Steps not pushed are written as empty
 * elements; push takes steps in increasing order. */
typedef struct mxs_zfarr_s mxs_zfarr_t;
int mxs_zfarr_open(mxs_zfarr_t **out, const char *fn, long nstep,
                   const char *keywords);
int mxs_zfarr_push(mxs_zfarr_t *z, long istep, uint32_t magic, uint64_t nx,
                   uint64_t ny, const void *data, uint64_t elsize);
int mxs_zfarr_close(mxs_zfarr_t *z);

#ifdef __cplusplus
}
#endif
#endif
