/*
 * FP64 CPU oracle for the MAOS Shack-Hartmann physical-optics WFS step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/ to check the GPU path, by smoke() as the checker
 * and by bench.py as the timed CPU baseline.  The product library never
 * links, loads or calls it.
 *
 * Each function cites the reference routine it follows.  Arrays are
 * column-major with x fastest (math/type.h:31-41); complex arrays are
 * interleaved (re, im) doubles.
 */
#ifndef MAOS_ORACLE_H
#define MAOS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* 2-D unnormalised FFT, dir=-1 forward, +1 backward (cfft2, math/fft.c:253-281) */
void orc_fft2(double *c, long nx, long ny, int dir);

/* prop_grid_pts + prop_grid_block (lib/accphi_grid_do.c:49-352, 408-446),
 * non-transposed, with both the equal-sampling and the unequal-sampling
 * code paths for wrap=1 (screens), and the unequal-sampling rules for
 * wrap=0 (DM maps: points outside [0, nxin-1) x [0, nyin-1] are left
 * alone).  phiout (nsa*nxsa*nxsa) is accumulated.  Returns the number of
 * points that fell outside the map (0 with wrap). */
long orc_prop_grid_pts(const double *phiin, long nxin, long nyin, double dxin,
                       double dyin, double oxin, double oyin, const double *ptsx,
                       const double *ptsy, long nsa, long nxsa, double ptsdx,
                       double *phiout, double alpha, double displacex,
                       double displacey, double scale, int wrap, long start,
                       long end);

/* prop_range (lib/accphi.c:223-244) for a screen (ht, dratio=0) onto pts:
 * derives scale, displacement and alpha from the ray parameters, then calls
 * orc_prop_grid_pts.  Returns -1 if the layer is above the source. */
long orc_prop(const double *phiin, long nxin, long nyin, double dxin, double oxin,
              double oyin, double ht, const double *ptsx, const double *ptsy,
              long nsa, long nxsa, double ptsdx, double *phiout, double alpha,
              double hs, double thetax, double thetay, double misregx,
              double misregy, double shiftx, double shifty, int wrap);

/* Static per-powfs description (powfs_t / parms->powfs fields the step uses). */
typedef struct orc_powfs_t {
    long nsa, nxsa, embfac, notf, pixpsax, pixpsay, nwvl;
    double dx, areascale, saat;
    const double *ptsx, *ptsy; /* pts->origx/origy [nsa] */
    const double *amp;         /* [nsa*nxsa*nxsa] */
    const double *saa;         /* [nsa] */
    const double *wvl;         /* [nwvl] */
    const double *nominal;     /* [nwvl][notf*notf] real, or NULL */
    const long *si_pp;         /* [nwvl][notf*notf+1] CSC of H (npix x notf^2) */
    const long *si_pi;         /* [nwvl][si_nzmax] */
    const double *si_px;       /* [nwvl][si_nzmax] */
    long si_nzmax;
    const double *etf;         /* [nwvl][nsa][notf*notf] complex, or NULL */
    const double *etf2;        /* second ETF for interpolation, or NULL */
    double etf2wt;             /* weight of etf2; etf gets 1-etf2wt */
    /* LLT (uplink); llt_nx==0 disables it */
    long llt_nx;
    double llt_origx, llt_origy, llt_ox, llt_oy;
    const double *llt_amp;     /* [llt_nx*llt_nx] */
    /* pixel processing */
    int phytype;               /* 1: matched filter, 2: CoG */
    int sigmatch;
    double cogthres, cogoff, pixthetax, pixthetay;
    double siglev_dtrat;       /* siglev*dtrat (sigmatch 0 CoG) */
    double saasum;             /* sum(saa) (sigmatch 2) */
    const double *mtche;       /* [nsa][2*npix] for this WFS */
    const double *i0sum;       /* [nsa] */
    double i0sumsum;
    /* multi_nominal (maos/wfsints.c:200-233): nonzero when nominal and si
     * hold one table per subaperture: nominal [nwvl][nsa][notf^2], si_pp
     * [nwvl][nsa][notf^2+1], si_pi / si_px [nwvl][nsa][si_nzmax] */
    long si_multi;
} orc_powfs_t;

typedef struct orc_wfs_t {
    double thetax, thetay, hs, sigsim;
    double misregx, misregy;
    const double *wvlwts; /* [nwvl] */
} orc_wfs_t;

typedef struct orc_atm_t {
    long nps, nx, ny;
    double dx, ox, oy;
    const double *const *screens; /* [nps] each nx*ny */
    const double *ht, *vx, *vy;   /* [nps] */
} orc_atm_t;

/* cembed_wvf without rotation (math/matcomp.c:56-82) */
void orc_embed_wvf(double *wvf, long nwvf, const double *opd, const double *amp,
                   long nopd, double wvl);

/* The uplink OTF lotfc of wfsints (maos/wfsints.c:149-199). */
void orc_lotf(double *lotfc, const double *lltopd, const double *lltamp, long nlx,
              long embfac, long nwvf, long notf, double wvl);

/* wfsints over [isa_start, isa_end) (maos/wfsints.c:79-302), accumulating
 * into ints [nsa][npix].  lotfc is [nwvl][notf*notf] complex or NULL. */
void orc_wfsints(double *ints, const double *opd, const double *lotfc,
                 const orc_powfs_t *p, const orc_wfs_t *w, long isa_start,
                 long isa_end);

/* pistatout of wfsints (maos/wfsints.c:262-275) over [isa_start,
 * isa_end): the downlink OTF of each subaperture (no uplink, DTF or ETF)
 * tilted by ctilt (math/matcomp.c:287-336) by -g dtheta1 pixels, g =
 * (gx[isa], gy[isa]) the step's geometric gradient in radians and dtheta1 =
 * nwvf dx / wvl, back to a PSF, fftshifted (peak at the centre), |.| times
 * norm_pistat = norm_psf^2 / notf^2 added to acc [nwvl][nsa][notf*notf]
 * (x fastest). */
void orc_pistat(double *acc, const double *opd, const orc_powfs_t *p, const double *gx, const double *gy,
                long isa_start, long isa_end);

/* psfout of wfsints (maos/wfsints.c:243-253) over [isa_start, isa_end):
 * the subaperture's complex PSF scaled by norm_psf / nwvf^2, transformed
 * back, centre (notf/2 + 2)^2 as cembed crops it; out [nwvl][nsa][nout^2]
 * complex (x fastest). */
void orc_psfout(double *out, const double *opd, const orc_powfs_t *p, long isa_start, long isa_end);

/* dcog (math/matmath.c:717-761) on an nx*ny image */
void orc_dcog(double *grad, const double *im, long nx, long ny, double offsetx,
              double offsety, double thres, double bkgrnd, double flux);

/* shwfs_grad for MF and CoG (maos/utils.c:983-1122). grad is 2*nsa
 * (x block then y block); entries of vignetted subapertures are left as is. */
void orc_shwfs_grad(double *grad, const double *ints, const orc_powfs_t *p,
                    double siglev);

/* One physical-optics WFS step of wfsgrad_iwfs (maos/wfsgrad.c:122-501,
 * noise-free, open loop): ray trace every layer (FP64), LLT uplink, wfsints
 * over all subapertures with OpenMP (nthread<=0: all), then shwfs_grad.
 * ints [nsa*npix] is zeroed first (dtrat=1).  opd_out/lotf_out are optional
 * debug outputs. */
void orc_wfsgrad_iwfs(double *ints, double *grad, const orc_powfs_t *p,
                      const orc_wfs_t *w, const orc_atm_t *atm, long isim,
                      double dt, double atmscale, int nthread, double *opd_out,
                      double *lotf_out);

/* ccpcorner (math/matcomp.c:176-216): the four corner quadrants of the
 * ninx x niny complex array in (FFT order) copied into the corners of the
 * nx x ny array out, zero elsewhere (padding or cropping). */
void orc_ccpcorner(double *out, long nx, long ny, const double *in, long ninx, long niny);

/* Geometric Zernike-tilt gradients (pts_ztilt, math/loc.c:839-880, with
 * the moments of pts_mcc_ptt, math/loc.c:612-650): the tip and tilt of the
 * amplitude-weighted least-squares fit of [1, x, y] to each subaperture's
 * OPD.  grad is 2*nsa (x block then y block), overwritten. */
void orc_ztilt(double *grad, const double *opd, const double *amp, const double *ptsx,
               const double *ptsy, long nsa, long nxsa, double dx);

/* Geometric average-gradient (G-tilt) gradients: GS0 of mkg
 * (lib/mkg.c:108-468, maos/powfs.c:603-630) applied to opd (nsa*nxsa^2,
 * the subaperture grids forming one common grid): the amplitude-weighted
 * average of the bilinear OPD's gradient over each subaperture square, by
 * Gauss-Legendre quadrature per cell.  grad is 2*nsa; subapertures with
 * saa < saat get 0 (an empty row). */
void orc_gtilt(double *grad, const double *opd, const double *amp, const double *ptsx, const double *ptsy,
               long nsa, long nxsa, double dx, const double *saa, double saat);

/* locfft_psf (lib/locfft.c:110-188) for one OPD and wavelength, unscaled
 * and squared as psfcomp_r accumulates it (cuda/sim/perfevl.cu:255-290):
 * the field amp exp(i 2 pi opd / wvl) written at embed[iloc] of a zero
 * nembed^2 grid in point order (a later point at the same index replaces
 * the earlier one), forward FFT, and |.|^2 of the centre psfsize^2 window
 * (ccpcorner2center / corner2center_abs2_do, cuda/math/kernel.cu:292-307):
 * psf[ix + iy psfsize] = |F[(ix - psfsize/2) mod nembed, (iy - psfsize/2)
 * mod nembed]|^2.  psfsize 1: |sum amp exp(i 2 pi opd / wvl)|^2
 * (strehlcomp, lib/locfft.c:90-100).  Returns 0, or -1 on a bad size. */
int orc_locfft_psf(double *psf, const double *opd, const double *amp, long nloc, const long *embed,
                   long nembed, long psfsize, double wvl);

/* loc_calc_ptt (math/loc.c:653-709) with mcc from loc_mcc_ptt
 * (math/loc.c:612-650) inverted by Gauss-Jordan elimination: rms = {piston
 * removed, tip/tilt, piston/tip/tilt removed} variances, coeff = the fitted
 * piston, tip and tilt.  Returns 0, or -1 if mcc is singular. */
int orc_loc_calc_ptt(double *rms, double *coeff, const double *locx, const double *locy, const double *amp,
                     const double *opd, long nloc);

/* Pyramid WFS (lib/pywfs.c): the per-powfs tables and configuration. */
typedef struct orc_pywfs_t {
    long nloc;               /* pupil points */
    const double *amp;       /* [nloc] */
    const long *embed;       /* [nloc] index ix + iy npsf of each point in the FFT grid */
    long npsf;               /* FFT grid side (= notf) */
    double dx;               /* pupil sampling (m) */
    int nwvl;
    const double *wvl, *wvlwts;
    const double *mask;      /* [nwvl][npsf^2] complex pyramid phase mask (apex at npsf/2) */
    const double *nominal;   /* [npsf^2] detector transfer function, corner order */
    long nsa;
    const double *salocx, *salocy; /* [nsa] detector subaperture centres (m) */
    int nside;
    const double *sioff;     /* [nside][2] pupil-image centres in units of npsf/2 */
    double modulate;         /* modulation radius (rad) */
    int modulpos, modulring;
    /* gradients (pywfs_grad) */
    int sigmatch, raw;
    double siglev, gain;     /* cfg siglev, gain (inverse optical gain) */
    const double *saa;       /* [nsa], mean 1 */
    const double *gradoff;   /* [ng nsa] or NULL */
} orc_pywfs_t;

/* pywfs_ints (lib/pywfs.c:370-452): ints [nside][nsa] (overwritten).  The
 * complex PSF of each wavelength, centred and scaled so that |psf|^2 sums
 * to 1, is shifted to each modulation position (rounded to a sample),
 * multiplied by the pyramid mask and inverse transformed; |.|^2 summed over
 * positions and wavelengths with weight wvlwts / (npsf^2 npos nring) is
 * the pupil intensity, which is convolved with the detector transfer
 * function and sampled bilinearly at the subaperture centres of each pupil
 * image, times nsa siglev / npsf^2.  The shifted window is the full valid
 * overlap (as cuda/sim/pywfs.cu:213-222 copies it). */
void orc_pywfs_ints(double *ints, const double *opd, const orc_pywfs_t *c, double siglev);
/* pywfs_grad (lib/pywfs.c:457-522): grad [ng nsa] (ng = 2 for three or
 * four sides unless raw; else nside) minus gradoff. */
void orc_pywfs_grad(double *grad, const double *ints, const orc_pywfs_t *c);

/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC 2011): ctr is replaced
 * by the four output words. */
void orc_philox4x32_10(unsigned ctr[4], unsigned key0, unsigned key1);

/* Poisson sample of mean xm by randp (math/random.c:186-209) drawing from the
 * stream (key, pixel pix, unit, step); blk and used carry the stream position
 * between calls (start at 0). */
double orc_randp(double xm, unsigned key0, unsigned key1, unsigned pix, unsigned unit,
                 unsigned step, unsigned *blk, int *used);

/* addnoise (lib/libmisc.c:31-57) on the npix pixels of subaperture unit
 * (= WFS * nsa + subaperture) at step: ints = randp(ints + bkgrnd) +
 * rne N(0,1) - bkgrndc, one Philox stream per pixel, normals by Box-Muller. */
void orc_addnoise(double *ints, long npix, double bkgrnd, double bkgrndc, double rne,
                  unsigned key0, unsigned key1, unsigned unit, unsigned step);

#ifdef __cplusplus
}
#endif
#endif
