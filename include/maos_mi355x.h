/*
 * maos_mi355x.h -- C ABI of the MI355X (gfx950) Shack-Hartmann
 * physical-optics WFS library (libmaos_mi355x.so).
 *
 * The library is a drop-in for the WFS step that MAOS runs through
 * wfsgrad_iwfs -> prop -> wfsints -> shwfs_grad (maos/wfsgrad.c:122-501,
 * maos/wfsints.c:79-302, maos/utils.c:983-1122), with the lifecycle of the
 * reference's GPU plugin (cuda/sim/gpu_sim.h:25-55): init, upload screens,
 * queue a step, sync the gradients, update ETF / matched filter.
 *
 * Conventions (as MAOS): arrays are column-major with x fastest; complex
 * arrays are interleaved (re, im) doubles; sparse matrices are MAOS dsp
 * (CSC) arrays.  All host buffers are owned by the caller and only read or
 * written during the call.  Every entry point returns 0 on success or a
 * negative MX_E* code; mx_last_error() describes the last failure.
 * Compute runs in FP32 on the GPU (the MFMA imaging form: BF16 hi/lo
 * operand pairs, three products, FP32 accumulation) except the
 * screen-coordinate arithmetic and the long-exposure i0 accumulation, which
 * are FP64.
 */
#ifndef MAOS_MI355X_H
#define MAOS_MI355X_H
#ifdef __cplusplus
extern "C" {
#endif

#define MX_OK 0
#define MX_EINVAL -1   /* bad argument or unsupported configuration */
#define MX_EHIP -2     /* HIP runtime error */
#define MX_ESTATE -3   /* call out of order (e.g. no screens uploaded) */

#define MX_VARIANT_AUTO -1      /* MFMA form where it applies, else FP32 FFT */
#define MX_VARIANT_FFT32 0      /* FP32 radix-2 FFTs on the vector ALUs */
#define MX_VARIANT_MFMA_BF16 1  /* DFTs as split-BF16 matrix products on MFMA */
#define MX_VARIANT_MFMA_FP16 2  /* the same with split-FP16 operands (comparison form) */

#define MX_PTYPE_MF 1  /* matched filter (PTYPE_MF) */
#define MX_PTYPE_COG 2 /* thresholded centre of gravity (PTYPE_COG) */

/* One turbulence layer: the fields of map_t that ray tracing reads
 * (lib/accphi.h:31-49, math/type.h map_t). */
typedef struct mx_atm_layer_t {
    const double *p;     /* nx*ny screen values (m) */
    long nx, ny;
    double dx, dy;       /* sampling (m) */
    double ox, oy;       /* coordinate of p[0] (m) */
    double ht;           /* conjugate height (m) */
    double vx, vy;       /* frozen-flow wind (m/s) */
} mx_atm_layer_t;

/* Static description of one powfs: the parms->powfs[ipowfs] keys and the
 * powfs_t geometry that wfsints/shwfs_grad read (maos/types.h:76-134). */
typedef struct mx_powfs_t {
    long nsa;            /* powfs.saloc->nloc */
    long nxsa;           /* powfs.pts->nxsa: OPD points per subaperture side */
    long embfac;         /* parms->powfs.embfac (nwvf = nxsa*embfac) */
    long notf;           /* powfs.notfx (== notfy) */
    long pixpsax, pixpsay;
    long nwvl;
    double dx;           /* powfs.pts->dx */
    const double *ptsx;  /* powfs.pts->origx [nsa] */
    const double *ptsy;  /* powfs.pts->origy [nsa] */
    const double *amp;   /* powfs.amp[0] [nsa*nxsa*nxsa] */
    const double *saa;   /* powfs.saa[0] [nsa] */
    double saat;         /* parms->powfs.saat */
    double areascale;    /* powfs.areascale */
    const double *wvl;   /* parms->powfs.wvl [nwvl] */
    /* pixel processing */
    int phytype;         /* MX_PTYPE_MF or MX_PTYPE_COG (phytype_sim) */
    int sigmatch;        /* parms->powfs.sigmatch: 0, 1 or 2 */
    double cogthres, cogoff;
    double pixthetax;    /* radpixtheta */
    double pixthetay;    /* pixtheta */
    double siglev_dtrat; /* parms->powfs.siglev*dtrat (CoG sigmatch 0) */
    double saasum;       /* powfs.saasum (CoG sigmatch 2) */
    /* LLT uplink (powfs.llt); llt_nx == 0 means no LLT */
    long llt_nx;         /* llt->pts->nxsa */
    double llt_origx, llt_origy; /* llt->pts->origx/origy[0] */
    double llt_ox, llt_oy;       /* parms->powfs.llt->ox/oy of this LLT */
    const double *llt_amp;       /* llt->amp [llt_nx*llt_nx] */
} mx_powfs_t;

/* Per-WFS parameters (parms->wfs[iwfs]). */
typedef struct mx_wfs_t {
    double thetax, thetay; /* direction (rad) */
    double hs;             /* source range (m), INFINITY for NGS */
    double sigsim;         /* photons per subaperture per frame */
    double misregx, misregy;
    const double *wvlwts;  /* [nwvl] */
} mx_wfs_t;

typedef struct mx_wfsgrad_s mx_wfsgrad_t;

/* Select the GPU (gpu_init, maos/parms.c:3796). */
int mx_init(int device);
const char *mx_last_error(void);

/* Create the WFS engine for one powfs and its nwfs WFS
 * (gpu_wfsgrad_init, cuda/sim/wfsinit.cu:159; called at maos/setup.c:131).
 * Copies geometry; tables are supplied with the calls below. */
int mx_wfsgrad_init(mx_wfsgrad_t **out, const mx_powfs_t *powfs,
                    const mx_wfs_t *wfs, int nwfs);
/* The same engine for nreal independent realisations advanced together
 * (1 <= nreal <= 32): one MAOS run per realisation, each with its own
 * screens (mx_atm2gpu_real), gradients, intensities, i0 and matched filter,
 * sharing the powfs tables.  Every other call addresses WFS ireal*nwfs + iwfs.
 * No counterpart in the reference, which runs one realisation per process. */
int mx_wfsgrad_init_batch(mx_wfsgrad_t **out, const mx_powfs_t *powfs,
                          const mx_wfs_t *wfs, int nwfs, int nreal);
int mx_wfsgrad_free(mx_wfsgrad_t *h);

/* Detector transfer tables of one wavelength: powfs.dtf[iwvl].nominal (real
 * notf*notf, treated; ignored when the powfs has an LLT, whose ETF carries
 * the nominal) and powfs.dtf[iwvl].si (dsp: si_nx = npix rows, si_ny =
 * notf*notf columns, CSC pp/pi/px) (lib/mkdtf.c:89-184). */
int mx_wfsgrad_set_dtf(mx_wfsgrad_t *h, int iwvl, const double *nominal,
                       long si_nx, long si_ny, const long *si_pp,
                       const long *si_pi, const double *si_px);
/* The same tables per subaperture (multi_nominal, maos/wfsints.c:200-233:
 * powfs.dtf[iwvl].nominal and .si with nsa entries, which mkdtf makes for
 * radial-pixel detectors or per-subaperture pixel offsets, lib/mkdtf.c:
 * 140-161).  Subaperture isa's nominal is nominal + isa*notf*notf (NULL
 * with an LLT), its si CSC si_pp + isa*(si_ny + 1), si_pi and si_px +
 * isa*si_nzmax.  The MFMA and FP32 radix-8 forms need the subapertures'
 * sampled image rows and columns to number at most 16 together; otherwise
 * the FP32 register form runs. */
int mx_wfsgrad_set_dtf_sa(mx_wfsgrad_t *h, int iwvl, const double *nominal,
                          long si_nx, long si_ny, long si_nzmax, const long *si_pp,
                          const long *si_pi, const double *si_px);

/* Sodium elongation transfer function of one wavelength
 * (gpu_wfsgrad_update_etf, cuda/sim/wfsinit.cu:32; maos/sim_utils.c:388):
 * etf is powfs.etfsim[iwvl].etf, nsa*notf*notf complex with the nominal
 * fused; etf2 (or NULL) is etfsim2 for profile interpolation with weight
 * etf2wt = (isim%dtrat)/dtrat (maos/wfsints.c:215-223).  Waits for the
 * engine's queued steps before replacing the tables. */
int mx_wfsgrad_update_etf(mx_wfsgrad_t *h, int iwvl, const double *etf,
                          const double *etf2, double etf2wt);
/* Per-step weight of the second sodium profile's ETF, w2 = (isim %
 * coldtrat)/coldtrat (maos/wfsints.c:211-220; the first gets 1 - w2),
 * without re-uploading the tables; needs a second ETF from
 * mx_wfsgrad_update_etf.  The MFMA forms blend the two half-plane tables
 * in a kernel of the step. */
int mx_wfsgrad_set_etf_weight(mx_wfsgrad_t *h, double etf2wt);

/* Matched filter and reference flux of WFS iwfs (gpu_wfsgrad_update_ref,
 * cuda/sim/wfsinit.cu:107; maos/wfsgrad.c:1207): mtche is nsa matrices of
 * 2 x npix (intstat->mtche column of this WFS), i0sum is nsa values
 * (intstat->i0sum).  iwfs = -1 loads the same tables into every WFS.
 * Either pointer may be NULL to leave that table unchanged.  Waits for the
 * engine's queued steps before replacing the tables. */
int mx_wfsgrad_update_ref(mx_wfsgrad_t *h, int iwfs, const double *mtche,
                          const double *i0sum);

/* Upload the turbulence screens (gpu_atm2gpu, cuda/sim/atm.cu:106; called
 * at maos/sim.c:79).  Values are stored as FP32; geometry stays FP64.
 * atmscale multiplies every layer (simu->atmscale).  The call first waits
 * for all queued GPU work (every engine and evaluation handle), so a step
 * or trace queued before it reads the previous screens. */
int mx_atm2gpu(const mx_atm_layer_t *layers, int nps, double atmscale);
/* The screens of realisation ireal (0 <= ireal < 32) for batched engines;
 * mx_atm2gpu uploads realisation 0.  All realisations of an engine must have
 * the same number of layers. */
int mx_atm2gpu_real(int ireal, const mx_atm_layer_t *layers, int nps,
                    double atmscale);

/* The deformable-mirror surfaces seen by every WFS of later steps
 * (maos/wfsgrad.c:236-242: prop from sim.cachedm or dmreal with alpha=-1
 * and no wrap, maos/sim_utils.c:1035-1052): each map (p, nx, ny, dx, dy, ox,
 * oy and the conjugate height ht; vx, vy unused) is traced onto the
 * subaperture grids after the atmosphere and subtracted; points beyond a
 * map's last row or column are left alone.  Replaces the previous set
 * (ndm=0 removes them); waits for queued steps first.  At most 8 DMs;
 * mx_dm2gpu_real sets realisation ireal's.  No counterpart call in the
 * reference's GPU plugin, which keeps the DM on the GPU with its own
 * reconstructor. */
int mx_dm2gpu(const mx_atm_layer_t *dms, int ndm);
int mx_dm2gpu_real(int ireal, const mx_atm_layer_t *dms, int ndm);

/* Enqueue step isim for every WFS of the engine (gpu_wfsgrad_queue,
 * cuda/sim/wfsgrad.cu:483): ray trace, LLT uplink, imaging, gradients and
 * (when enabled) FP64 i0 accumulation.  Asynchronous. */
int mx_wfsgrad_queue(mx_wfsgrad_t *h, long isim, double dt);

/* Wait for the queued step and copy WFS iwfs's gradients (2*nsa, x block
 * then y block; entries of subapertures with saa<saat keep their previous
 * value) and, if ints is not NULL, its pixel intensities (nsa*npix)
 * (gpu_wfsgrad_sync, cuda/sim/wfsgrad.cu:747). */
int mx_wfsgrad_sync(mx_wfsgrad_t *h, int iwfs, double *grad, double *ints);

/* Long-exposure FP64 accumulation of ints (maos/wfsgrad.c:449-467).
 * enable=1 accumulates every later step; reset=1 clears the sums. */
int mx_wfsgrad_i0_control(mx_wfsgrad_t *h, int enable, int reset);
/* Copy WFS iwfs's accumulated sum (nsa*npix) and the number of frames. */
int mx_wfsgrad_i0_get(mx_wfsgrad_t *h, int iwfs, double *i0sum_out,
                      long *nframe);

/* Tip/tilt-removed PSF statistics (powfs.pistatout: maos/wfsints.c:262-275,
 * GPU form cuda/sim/wfsints.cu:458-472, accumulated from pistatstart by
 * maos/wfsgrad.c:155-156 and saved by gpu_save_pistat, cuda/sim/
 * wfsgrad.cu:777-800).  With enable=1 every later step adds, per WFS,
 * wavelength and subaperture, the downlink PSF (no uplink, DTF or ETF)
 * shifted by the step's geometric gradient of that subaperture (gtype 1:
 * Zernike tilt, pts_ztilt; 2: G-tilt, GS0 from mx_wfsgrad_set_gs0), peak at
 * the centre, scaled by norm_psf^2 as MAOS scales it.  The nxsa 32, embfac
 * 2, notf 64 geometry only; not with geometric gradients.  reset=1 clears
 * the sums.  The sums are FP32 on the GPU, as the reference GPU code keeps
 * them. */
int mx_wfsgrad_pistat_control(mx_wfsgrad_t *h, int enable, int gtype, int reset);
/* Copy WFS iwfs's sums, [nwvl][nsa][notf*notf] (x fastest; MAOS's dcell
 * nsa x nwvl of notf x notf), and the number of steps they hold
 * (save_pistat, maos/save.c:29-58, divides by it). */
int mx_wfsgrad_pistat_get(mx_wfsgrad_t *h, int iwfs, double *out, long *nstep);

/* The complex field of every subaperture of WFS iwfs at the last queued
 * step (powfs.psfout: maos/wfsints.c:243-253, GPU form cuda/sim/
 * wfsints.cu:443-450, pushed to wvfout_*.bin each step by maos/wfsgrad.c:
 * 403): the subaperture's complex PSF scaled by norm_psf / nwvf^2,
 * transformed back, centre nout^2 with nout = notf/2 + 2.  out is
 * [nwvl][nsa][nout*nout] complex interleaved (x fastest; MAOS's ccell nsa x
 * nwvl).  Needs nxsa*embfac == notf; waits for the queued steps. */
int mx_wfsgrad_psfout(mx_wfsgrad_t *h, int iwfs, double *out);

/* TT-dither statistics for the matched-filter update (maos/wfsgrad.c:
 * 449-466).  With enable=1 every later step adds, per WFS and pixel, ints,
 * cs*ints and ss*ints (FP64) to the sums imb, imx, imy; (cs, ss) is the
 * WFS's dither phase for the step (dither_position, maos/utils.c:884-899;
 * mxs_dither_position computes it), set by mx_wfsgrad_set_dither before
 * the step is queued.  The first call allocates the sums and the model
 * i0/gx/gy (zero); reset=1 clears the sums and the folded sums. */
int mx_wfsgrad_dither_control(mx_wfsgrad_t *h, int enable, int reset);
int mx_wfsgrad_set_dither(mx_wfsgrad_t *h, int iwfs, double cs, double ss);
/* At a PLL output (maos/wfsgrad.c:709-731): folded i0 += s0*imb, folded
 * gx += sg*imx, folded gy += sg*imy (MAOS: s0 = 1/pllrat, sg =
 * 2/(pllrat*a2m)), then imb, imx, imy are cleared.  iwfs=-1: all WFS.
 * Queued on the engine's stream. */
int mx_wfsgrad_dither_fold(mx_wfsgrad_t *h, int iwfs, double s0, double sg);
/* Model i0, gx, gy (nsa*npix each, NULL: zero) that the update blends
 * into, e.g. the setup-time model of mxs_gensei (intstat->i0/gx/gy). */
int mx_wfsgrad_set_i0model(mx_wfsgrad_t *h, int iwfs, const double *i0, const double *gx,
                           const double *gy);
/* Matched-filter update on the GPU (maos/wfsgrad.c:1036-1047 then
 * mtch_cell, lib/mtch.c:223-335): model = keep*model + acc*folded (MAOS:
 * keep = 1-g2, acc = scale1*g2) for i0, gx, gy, the folded sums cleared,
 * then mtche and i0sum rebuilt in FP64 for the MF stage, with bkgrnd,
 * bkgrndc, rne, mtchcr and sigratio as mxs_mtch takes them.  No host round
 * trip; queued on the engine's stream.  iwfs=-1: all WFS.  This entry
 * point has no single counterpart in the reference: it replaces the host
 * update and the gpu_wfsgrad_update_ref upload (cuda/sim/wfsgrad.cu). */
int mx_wfsgrad_mtch_update(mx_wfsgrad_t *h, int iwfs, double keep, double acc, double bkgrnd,
                           double bkgrndc, double rne, int mtchcr, double sigratio);
/* Copies of the per-step dither sums (nsa*npix each, any may be NULL) and
 * the number of frames since the last reset. */
int mx_wfsgrad_dither_get(mx_wfsgrad_t *h, int iwfs, double *imb, double *imx, double *imy,
                          long *nframe);
/* The MF stage's filter (nsa*2*npix, as mx_wfsgrad_update_ref takes it)
 * and, after an update, its i0sum (nsa), NEA^2 (nsa*3: xx, yy, xy, as
 * mxs_mtch's sanea) and the model i0, gx, gy; any may be NULL. */
int mx_wfsgrad_mtch_get(mx_wfsgrad_t *h, int iwfs, double *mtche, double *i0sum, double *sanea,
                        double *i0, double *gx, double *gy);

/* Debug/parity outputs of the last step: the WFS OPD (nsa*nxsa*nxsa, m)
 * and the uplink OTF lotfc of wavelength iwvl (notf*notf complex, as
 * wfsints.c:149-199 leaves it before conjugation). */
int mx_wfsgrad_debug_opd(mx_wfsgrad_t *h, int iwfs, double *opd);
int mx_wfsgrad_debug_lotf(mx_wfsgrad_t *h, int iwfs, int iwvl, double *lotf);

/* Time the kernels of subsequent steps with HIP events on the engine's
 * stream (each timed step waits for its kernels).  enable=1 starts
 * recording and clears the sums; mx_wfsgrad_timing reports the number of
 * timed steps and the summed milliseconds of the uplink-OTF kernel, the
 * ray-trace kernel and the imaging kernel. */
int mx_wfsgrad_timing_enable(mx_wfsgrad_t *h, int enable);
int mx_wfsgrad_timing(mx_wfsgrad_t *h, long *nlaunch, double *ms_llt,
                      double *ms_opd, double *ms_ints);

/* Select how the per-subaperture transforms are computed (no counterpart in
 * the reference, whose GPU code always uses cuFFT).  MX_VARIANT_MFMA_BF16
 * and MX_VARIANT_MFMA_FP16 need notf=64, nxsa=32 and at most 16 sampled
 * image rows and columns per wavelength (a second ETF is blended in a kernel
 * of the step); AUTO (the default) uses the BF16 form when those hold.  Other sizes run the FP32 FFT forms.
 * get_variant reports the form the next step will use. */
int mx_wfsgrad_set_variant(mx_wfsgrad_t *h, int variant);
int mx_wfsgrad_get_variant(mx_wfsgrad_t *h, int *variant);

/* Closed-loop inputs of the WFS OPD, added after the ray trace of every
 * later step (not to the LLT): a static surface OPD per WFS
 * (powfs.opdadd, maos/wfsgrad.c:228-230; nsa*nxsa*nxsa values in the OPD
 * layout, iwfs=-1 for every WFS, opd=NULL for zero) and a piston / tip /
 * tilt per WFS (loc_add_ptt, maos/wfsgrad.c:243-256: opd += ptt[0] +
 * ptt[1]*x + ptt[2]*y with the TTM/FSM commands; the values are copied at
 * the call and used from the next queued step on). */
int mx_wfsgrad_set_opdadd(mx_wfsgrad_t *h, int iwfs, const double *opd);
int mx_wfsgrad_set_ptt(mx_wfsgrad_t *h, int iwfs, const double *ptt);

/* Gradient type: mode 0 (the default) physical optics; mode 1 geometric
 * Zernike-tilt gradients from the OPD (pts_ztilt, math/loc.c:839-880, as
 * maos/wfsgrad.c:303-306 computes them for gtype_sim = GTYPE_Z); mode 2
 * geometric average-gradient (G-tilt) gradients GS0 * opd (gtype_sim =
 * GTYPE_G, maos/wfsgrad.c:307-308), with GS0 from mx_wfsgrad_set_gs0.  The
 * geometric steps run the ray trace and one reduction per subaperture (or
 * row), leave ints untouched, and need no DTF, ETF or matched filter. */
int mx_wfsgrad_set_geom(mx_wfsgrad_t *h, int mode);
/* The G-tilt operator GS0 (powfs.GS0 of setup_shwfs_grad, maos/powfs.c:
 * 603-630; mxs_mkg builds it): CSR with 2*nsa rows (x gradients, then y)
 * over the nsa*nxsa^2 OPD points, shared by all WFS.  The structure is
 * checked (rowp monotone from 0 to nnz, columns in range) before upload. */
int mx_wfsgrad_set_gs0(mx_wfsgrad_t *h, const long *rowp, const long *coli, const double *val, long nnz);

/* Photon and read noise on the pixel intensities of every later step
 * (addnoise, lib/libmisc.c:31-57, as applied at maos/wfsgrad.c:423-442;
 * seeding replaces gpu_wfsgrad_seeding, cuda/sim/wfsinit.cu:532):
 * ints = randp(ints + bkgrnd) + rne N(0,1) - bkgrndc with MAOS's randp
 * (math/random.c:186-209: a normal approximation above 200, else the
 * multiplication method in chunks of 12).  bkgrnd is powfs.bkgrnd*dtrat and
 * bkgrndc bkgrnd*powfs.bkgrndc in photons per pixel per frame.  The uniforms
 * come from Philox4x32-10 keyed by seed, one stream per (step, WFS,
 * subaperture, pixel), so the result does not depend on launch order; the
 * streams differ from MAOS's MT19937 and the reference GPU code's curand by
 * design.  Gradients and the i0 sums use the noisy intensities.  Give each
 * engine its own seed.  enable=0 (the default) is noise-free. */
int mx_wfsgrad_set_noise(mx_wfsgrad_t *h, int enable, double rne, double bkgrnd,
                         double bkgrndc, unsigned long long seed);

/* Exposure of dtrat steps (powfs.dtrat; maos/wfsgrad.c:164-167 and the
 * gradout branch :411-495): the frames of steps isim with isim % dtrat == 0
 * up to (isim + 1) % dtrat == 0 add up on the GPU; only the last step of an
 * exposure adds noise (give mx_wfsgrad_set_noise bkgrnd*dtrat), writes the
 * intensities, adds to the i0 and dither sums (which count exposures) and
 * forms the gradients; the other steps leave them as they were.  Physical
 * optics only (geometric gradients are per step).  dtrat 1 (the default)
 * is one frame per exposure.  Waits for queued steps. */
int mx_wfsgrad_set_dtrat(mx_wfsgrad_t *h, int dtrat);

/* Launch each step's kernels as one HIP graph (enable=1) or one by one
 * (enable=0, the default).  No counterpart in the reference, whose GPU code
 * launches kernels directly; the graph is built on the next queued step and
 * rebuilt when the kernel set changes (imaging form, uplink present); every
 * step updates its nodes' arguments.  Steps with timing enabled are launched
 * one by one. */
int mx_wfsgrad_set_graph(mx_wfsgrad_t *h, int enable);

/* A group of engines (e.g. the powfs of one simulation: LGS, low-order NGS,
 * tip-tilt NGS) stepped together by one HIP graph launch per step, each
 * engine a branch of the graph that runs concurrently with the others.  No
 * counterpart in the reference, which queues each WFS on its own stream
 * (gpu_wfsgrad_queue, cuda/sim/wfsgrad.cu:483).  mx_wfsgroup_queue plans
 * every engine's step as mx_wfsgrad_queue would (same checks, same
 * results), starts after each engine's earlier work, and makes each
 * engine's later calls (sync, wait, queue) wait for it.  Group steps are
 * not timed.  Engines must stay alive while the group exists. */
typedef struct mx_wfsgroup_s mx_wfsgroup_t;
int mx_wfsgroup_create(mx_wfsgroup_t **out, mx_wfsgrad_t *const *engines, int n);
int mx_wfsgroup_queue(mx_wfsgroup_t *g, long isim, double dt);
int mx_wfsgroup_free(mx_wfsgroup_t *g);

/* Pick the fastest way to run this engine's step on this GPU (no
 * counterpart in the reference, which always uses cuFFT): times nrep steps
 * from isim (kernel time, HIP events) for each candidate -- imaging form
 * (MFMA split-BF16 where it applies, FP32 FFT; the large-field FFT for
 * other sizes), ray-trace tile side (32 or 16 points; an LLT grid is one
 * tile), threads per large-field block (128, 256, 512) -- and keeps the
 * fastest.  Every candidate gives results within the parity bar of the
 * FP64 reference.  The i0 sums and frame count are left unchanged; ints and
 * gradients hold the last tuning step.  report (may be NULL) receives one
 * line per candidate, the chosen one marked with '*'. */
int mx_wfsgrad_autotune(mx_wfsgrad_t *h, long isim, double dt, int nrep,
                        char *report, long report_len);

/* Wait for all queued work of the engine. */
int mx_wfsgrad_wait(mx_wfsgrad_t *h);

/* ---------------------------------------------------------------------
 * Science-field evaluation: OPD on the aperture points and long-exposure
 * PSF per (direction, wavelength).
 * ------------------------------------------------------------------- */
typedef struct mx_evl_s mx_evl_t;

/* Replaces gpu_perfevl_init_sim's PSF tables (cuda/sim/perfinit.cu:28-90)
 * built from locfft_init (lib/locfft.c:32-68).  nloc aperture points with
 * amplitude amp (not normalised here: the mean PSF is divided by
 * (sum amp)^2).  locx/locy (may be NULL when only the PSF of a given OPD is
 * wanted) are the point coordinates for mx_evl_trace.  ndir (1..64)
 * directions thetax/thetay (rad; NULL: 0) at height hs (m; 0 or NULL: at
 * infinity).  For each of nwvl (1..4) wavelengths wvl (m): the FFT side
 * nembed (a power of two from 32 to 4096; mxs_loc_create_embed computes
 * it), psfsize (1 for the Strehl-only form, else even and at most nembed),
 * and embed[iwvl * nloc + iloc], the index ix + iy * nembed of each point
 * in the FFT grid (loc_create_embed, math/loc.c:147-179).  Two points at
 * one index: the later one is used. */
int mx_evl_init(mx_evl_t **out, long nloc, const double *locx, const double *locy,
                const double *amp, int ndir, const double *thetax, const double *thetay,
                const double *hs, int nwvl, const double *wvl, const long *nembed,
                const long *psfsize, const long *embed);
int mx_evl_free(mx_evl_t *h);

/* Replaces psfcomp_r (cuda/sim/perfevl.cu:255-290) and the psfmean branch
 * of perfevl_psfcl (maos/perfevl.c:64-78): adds, for every direction and
 * wavelength, |FFT(amp exp(2 pi i opd / wvl))|^2 over the centred
 * psfsize^2 window (psfsize 1: |sum amp exp(2 pi i opd / wvl)|^2) to the
 * accumulators.  opd is [ndir][nloc] in metres (host, FP64; converted to
 * FP32).  Queued on the handle's stream. */
int mx_evl_psf_add(mx_evl_t *h, const double *opd);
/* The same from an FP32 OPD [ndir][nloc] already in device memory, which
 * must stay unchanged until the handle's work is done (mx_evl_wait). */
int mx_evl_psf_add_device(mx_evl_t *h, const float *d_opd);

/* Replaces the atmosphere and DM ray trace onto the evaluation points
 * (maos/perfevl.c:275-300, gpu_atm2loc in cuda/sim/accphi.cu): the screens
 * of realisation 0 (mx_atm2gpu) at step isim, frozen flow v dt isim,
 * bilinear with periodic wrap, times atmscale; minus the DM surfaces
 * (mx_dm2gpu; no wrap, points outside a surface take nothing).  add_psf=1
 * then adds the PSF of that OPD (mx_evl_psf_add).  Queued. */
int mx_evl_trace(mx_evl_t *h, long isim, double dt, int add_psf);
/* The realisation (0 <= ireal < 32, of mx_atm2gpu_real / mx_dm2gpu_real)
 * whose screens and DM surfaces mx_evl_trace reads; 0 by default.  No
 * counterpart in the reference, which runs one realisation per process. */
int mx_evl_set_real(mx_evl_t *h, int ireal);
/* The last OPD traced or added from the host, [ndir][nloc] (waits). */
int mx_evl_opd_get(mx_evl_t *h, double *opd);

/* Wavefront error of the last OPD traced or added from the host, per
 * direction (calc_ptt + calc_ptt_post, cuda/sim/perfevl.cu:60-127, with the
 * FP64 sums of loc_calc_ptt, math/loc.c:653-704): from
 * c = sum amp opd [1, x, y], tot = sum amp opd^2 and
 * mcc = sum amp [1, x, y]^T [1, x, y] (loc_mcc_ptt, math/loc.c:612-650),
 * rms[3 d + 0..2] = {tot - c0^2 / mcc00 (piston removed),
 * c^T mcc^-1 c - c0^2 / mcc00 (tip/tilt), tot - c^T mcc^-1 c (piston, tip
 * and tilt removed)} and coeff[3 d + 0..2] = mcc^-1 c (piston, tip, tilt).
 * Variances in m^2 when amp sums to 1.  Either output may be NULL; needs the
 * point coordinates.  Waits. */
int mx_evl_wfe(mx_evl_t *h, double *rms, double *coeff);
/* mx_evl_wfe of an FP32 OPD [ndir][nloc] already in device memory
 * (d_opd), allocating nothing: no hipMalloc, hipFree or allocating library
 * call.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  The
 * partial sums go to a static device array of the library (130 KiB per GPU,
 * held until the process ends; room for 64 directions, the most a handle
 * has), whose address is looked up at mx_init (or at the first call on a
 * GPU).  Every call on one GPU shares that array: calls that overlap on one
 * GPU (from several threads or streams) race.  Same results as mx_evl_wfe
 * (the same sums in the same order).  Waits for its kernels. */
int mx_evl_wfe_noalloc(mx_evl_t *h, const float *d_opd, double *rms, double *coeff);

/* The mean PSF of direction idir and wavelength iwvl, psfsize^2 values with
 * x fastest, centred as corner2center (cuda/math/kernel.cu:292-307):
 * accumulated sum / (nacc (sum amp)^2), so a diffraction-limited peak is
 * the Strehl ratio 1 (gpu_perfevl_save, cuda/sim/perfevl.cu:618-660).
 * nacc (may be NULL) receives the number of PSFs added.  Waits. */
int mx_evl_psf_get(mx_evl_t *h, int idir, int iwvl, double *psf, long *nacc);
/* Clear the accumulators (queued). */
int mx_evl_psf_reset(mx_evl_t *h);
int mx_evl_wait(mx_evl_t *h);
/* Kernel time, summed over timed calls, of the trace and of the PSF's row
 * and column FFT passes (all wavelengths), from HIP events on the handle's
 * stream; every timed call waits for its kernels.  n counts traces. */
int mx_evl_timing_enable(mx_evl_t *h, int enable);
int mx_evl_timing(mx_evl_t *h, long *n, double *ms_trace, double *ms_rows, double *ms_cols);

/* ---------------------------------------------------------------------
 * Pyramid WFS (lib/pywfs.c; the reference's GPU form cuda/sim/pywfs.cu).
 * ------------------------------------------------------------------- */
typedef struct mx_pywfs_s mx_pywfs_t;

/* Static description of one pyramid WFS: the pywfs_t tables pywfs_new
 * builds (lib/pywfs.c:110-370; mxs_pywfs_* in maos_setup.h make them) and
 * the pywfs_cfg_t keys the step reads. */
typedef struct mx_pywfs_cfg_t {
    long nloc;                   /* pupil points (powfs.loc) */
    const double *locx, *locy;   /* [nloc] (m), NULL if only mx_pywfs_ints_opd is used */
    const double *amp;           /* [nloc] */
    const long *embed;           /* [nloc] ix + iy nembed in the FFT grid (locfft->embed) */
    long nembed;                 /* FFT grid side (locfft->nembed), even, factors 2, 3, 5, 7, <= 4096 */
    double dx;                   /* pupil sampling (loc->dx) */
    int nwvl;
    const double *wvl, *wvlwts;  /* [nwvl] */
    const double *mask;          /* [nwvl][nembed^2] complex pyramid mask (pywfs->pyramid) */
    const double *nominal;       /* [nembed^2] detector transfer function (pywfs->nominal, real) */
    long nsa;
    const double *salocx, *salocy; /* [nsa] detector subaperture centres (pywfs->saloc) */
    int nside;                   /* pupil images: 4 (pyramid), 3, or 2 (roof) */
    const double *sioff;         /* [nside][2] pupil-image centres in units of nembed/2 (pywfs->sioff) */
    double modulate;             /* modulation radius (rad) */
    int modulpos, modulring;
    int sigmatch, raw;           /* gradient normalisation (pywfs_cfg_t) */
    double siglev;               /* pywfs_cfg_t siglev (sigmatch 0 and raw) */
    double thetax, thetay, hs;   /* direction (rad) and range (m, INFINITY: NGS) for the trace */
} mx_pywfs_cfg_t;

/* Create the pyramid WFS (the GPU tables of cuda/sim/pywfs.cu, set up at
 * gpu_wfsgrad_init).  The gain is 1, saa 1 and no gradient offset until
 * mx_pywfs_set_calib. */
int mx_pywfs_init(mx_pywfs_t **out, const mx_pywfs_cfg_t *cfg);
int mx_pywfs_free(mx_pywfs_t *h);
/* The calibration pywfs_new derives from the imaging itself
 * (lib/pywfs.c:298-368): the inverse optical gain, the normalised
 * subaperture areas saa [nsa] and the gradients of a flat wavefront
 * gradoff [ng nsa] (NULL: none).  Waits for queued steps. */
int mx_pywfs_set_calib(mx_pywfs_t *h, double gain, const double *saa, const double *gradoff);
/* pywfs_ints + pywfs_grad (lib/pywfs.c:370-522) of a given pupil OPD
 * [nloc] (m): the intensities at siglev photons per subaperture.  Queued. */
int mx_pywfs_ints_opd(mx_pywfs_t *h, const double *opd, double siglev);
/* One step (the pyramid branch of wfsgrad_iwfs, maos/wfsgrad.c:400-410, as
 * gpu_wfsgrad_queue runs it): the screens of mx_atm2gpu (and DM surfaces of
 * mx_dm2gpu) traced onto the pupil points at step isim in the WFS
 * direction, then the intensities and gradients.  Queued. */
int mx_pywfs_queue(mx_pywfs_t *h, long isim, double dt, double siglev);
/* Wait and copy the gradients [ng nsa] (x block then y block; ng = 2 for
 * three or four sides, else nside) and the intensities [nside][nsa]; either
 * may be NULL. */
int mx_pywfs_sync(mx_pywfs_t *h, double *grad, double *ints);
/* The last traced or given OPD [nloc] (waits). */
int mx_pywfs_opd_get(mx_pywfs_t *h, double *opd);
/* The realisation whose screens mx_pywfs_queue reads (0 by default). */
int mx_pywfs_set_real(mx_pywfs_t *h, int ireal);
/* Time later steps with HIP events (each timed step waits): the summed
 * milliseconds from the first kernel of the imaging to the gradients. */
int mx_pywfs_timing_enable(mx_pywfs_t *h, int enable);
int mx_pywfs_timing(mx_pywfs_t *h, long *nstep, double *ms);
int mx_pywfs_wait(mx_pywfs_t *h);

#ifdef __cplusplus
}
#endif
#endif
