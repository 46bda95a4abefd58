// MI355X (gfx950) Shack-Hartmann physical-optics WFS step.
//
// One 64-lane wave images 64/N subapertures of an N x N (N = notf) grid:
// every lane owns one line (row or column) of its subaperture's field and
// keeps it in registers, where a fully unrolled radix-2 FFT runs with the
// twiddles folded into instruction literals.  Rows and columns are exchanged
// through a 32 KiB XOR-swizzled LDS tile (N = 64), so the chain
//   ray trace (FP64 coordinates) -> embed -> FFT -> |E|^2 -> FFT
//   -> x conj(uplink OTF) x ETF/nominal -> inverse FFT -> pixel sampling
//   -> matched filter / tCoG -> FP64 i0 accumulation
// runs in one launch with no intermediate array in HBM
// (maos/wfsgrad.c:122-501, maos/wfsints.c:79-302, maos/utils.c:983-1122).
#include <hip/hip_runtime.h>
#include <type_traits>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/maos_mi355x.h"
#include "twiddles.h"

#define MX_MAXLAYER 16
#define MX_MAXREAL 32 // realisations one engine can advance together
#define MX_MAXWVL 4

namespace {

// ---------------------------------------------------------------- errors
thread_local char g_err[512] = "";

int set_err(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return set_err(MX_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr int bitrev(int x, int lg) {
    int r = 0;
    for (int i = 0; i < lg; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// a * exp(-/+ 2 pi i k / MX_TWN); k is a compile-time constant after unrolling
template <bool INV>
__device__ __forceinline__ float2 twmul(float2 a, int k) {
    if (k == 0) return a;
    if (k == MX_TWN / 4) return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
    const float c = kTwCos[k], s = kTwSin[k];
    if (!INV) return make_float2(a.x * c + a.y * s, a.y * c - a.x * s);
    return make_float2(a.x * c - a.y * s, a.y * c + a.x * s);
}

// Unnormalised N-point FFTs of a register array.  INV=false is exp(-i...)
// (FFTW_FORWARD, cfft2(.,-1)), INV=true is exp(+i...) (cfft2(.,+1)).
// fft_dif: natural-order input, bit-reversed output (v[k] = X[bitrev(k)]).
// fft_dit: bit-reversed input, natural-order output.  Alternating the two
// keeps every permutation in compile-time indices.
template <int N, bool INV>
__device__ __forceinline__ void fft_dif(float2 (&v)[N]) {
    constexpr int LG = ilog2c(N);
    static_assert((1 << LG) == N && N <= MX_TWN, "power of two up to MX_TWN");
#pragma unroll
    for (int s = LG; s >= 1; s--) {
        const int len = 1 << s, h = len >> 1;
#pragma unroll
        for (int i0 = 0; i0 < N; i0 += len) {
#pragma unroll
            for (int j = 0; j < h; j++) {
                const float2 u = v[i0 + j], w = v[i0 + j + h];
                v[i0 + j] = cadd(u, w);
                v[i0 + j + h] = twmul<INV>(csub(u, w), j * (MX_TWN / len));
            }
        }
    }
}

template <int N, bool INV>
__device__ __forceinline__ void fft_dit(float2 (&v)[N]) {
    constexpr int LG = ilog2c(N);
    static_assert((1 << LG) == N && N <= MX_TWN, "power of two up to MX_TWN");
#pragma unroll
    for (int s = 1; s <= LG; s++) {
        const int len = 1 << s, h = len >> 1;
#pragma unroll
        for (int i0 = 0; i0 < N; i0 += len) {
#pragma unroll
            for (int j = 0; j < h; j++) {
                const float2 u = v[i0 + j];
                const float2 w = twmul<INV>(v[i0 + j + h], j * (MX_TWN / len));
                v[i0 + j] = cadd(u, w);
                v[i0 + j + h] = csub(u, w);
            }
        }
    }
}

// ---------------------------------------------------------------- params
struct DevLayer {
    const float *p;
    int nx, ny;
    double dx1, dy1;   // 1/dx, 1/dy
    double ox, oy, ht, vx, vy;
    double inx, iny;   // 1/nx, 1/ny
};

struct DevWfs {
    double thetax, thetay, hs, misregx, misregy;
    float alpha[MX_MAXWVL]; // wvlwts[iwvl]*norm_ints
    int ireal;              // realisation (atmosphere) of this WFS
};

// Per-wavelength tables; transfer functions are stored transposed
// ([ux][uy]) so that the lanes of a row pass read them coalesced.
struct DevWvl {
    float inv_wvl;
    const float *nominal;
    const float2 *etf, *etf2;
    const int4 *si_idx;
    const float4 *si_w;
    // per-subaperture tables (multi_nominal): element strides between the
    // subapertures' si entries and nominals, 0 when one table serves all
    int si_sa, nom_sa;
};

// ------------------------------------------------------------ pixel noise
// addnoise (lib/libmisc.c:31-57) as applied per frame at
// maos/wfsgrad.c:423-442: ints = randp(ints + bkgrnd) + rne N(0,1) - bkgrndc,
// with randp (math/random.c:186-209): a normal approximation above 200
// photons, else the multiplication method in chunks of at most 12.  The
// uniforms come from Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011) with counter (block, pixel,
// WFS*nsa + subaperture, step) and the 64-bit seed as key: one stream per
// pixel and step, independent of launch order.  Normals are Box-Muller (one
// per pair of uniforms).  Poisson arithmetic is FP64.
// Dither statistics of the TT-dithered matched-filter update
// (maos/wfsgrad.c:449-466): each step adds ints, cs ints and ss ints (FP64)
// to imb, imx, imy, with (cs, ss) of the WFS's dither phase for the step
// (dither_position, maos/utils.c:884-899, on the host).
struct DitherAcc {
    double *imb, *imx, *imy; // [nwfs][nsa][npix], or imb null when off
    const double *cs;        // [nwfs][2]: cs, ss
};
__device__ __forceinline__ void dither_add(const DitherAcc &d, int iwfs, long i, float v) {
    if (!d.imb) return;
    const double x = (double)v;
    d.imb[i] += x;
    d.imx[i] += d.cs[2 * iwfs] * x;
    d.imy[i] += d.cs[2 * iwfs + 1] * x;
}

// A WFS exposure of dtrat steps (maos/wfsgrad.c:164-167, 411-495): the
// frames of one exposure add up in acc [nwfs][nsa][npix]; only the last
// frame (MX_DT_OUT) goes on to the noise, intensities, i0 and dither sums
// and gradients; the first (MX_DT_FIRST) starts the sum.  dtrat 1: both.
constexpr int MX_DT_FIRST = 1, MX_DT_OUT = 2;
struct DtArgs {
    float *acc;
    int phase;
};
// frame value pv of pixel i: accumulated (returns false: not the last
// frame), or the whole exposure (returns true)
__device__ __forceinline__ bool dt_frame(const DtArgs &d, long i, float &pv) {
    if (d.phase == (MX_DT_FIRST | MX_DT_OUT)) return true;
    if (!(d.phase & MX_DT_OUT)) {
        d.acc[i] = (d.phase & MX_DT_FIRST) ? pv : d.acc[i] + pv;
        return false;
    }
    if (!(d.phase & MX_DT_FIRST)) pv += d.acc[i];
    return true;
}

struct NoiseArgs {
    int on;
    float bkgrnd, bkgrndc, rne;
    unsigned key0, key1, isim;
};
__device__ __forceinline__ void philox4x32_10(unsigned (&c)[4], unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const unsigned lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const unsigned lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
struct NoiseStream {
    // the four words of the current Philox block are consumed in order;
    // they are shifted down rather than indexed, so the stream stays in
    // registers (a run-time array index would place it in private memory)
    unsigned k0, k1, pix, unit, step, blk = 0, b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    int left = 0;
    // uniform in (0, 1]
    __device__ double uniform() {
        if (left == 0) {
            unsigned c[4] = {blk++, pix, unit, step};
            philox4x32_10(c, k0, k1);
            b0 = c[0];
            b1 = c[1];
            b2 = c[2];
            b3 = c[3];
            left = 4;
        }
        const unsigned u = b0;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        left--;
        return ((double)u + 1.) * 2.3283064365386962890625e-10;
    }
    __device__ double normal() {
        const double u1 = uniform(), u2 = uniform();
        return sqrt(-2. * log(u1)) * cospi(2. * u2);
    }
    // randp: a multiplication loop ends after at most 1000 factors (its
    // chance of running that long for a mean <= 12 is below 1e-300)
    __device__ double poisson(double xm) {
        if (xm > 200.) return rint(xm + normal() * sqrt(xm));
        double x = 0.;
        while (xm > 0.) {
            const double xmu = xm > 12. ? 12. : xm;
            xm -= xmu;
            const double g = exp(-xmu);
            double t = 1.;
            int xu = -1;
            while (t > g && xu < 1000) {
                xu++;
                t *= uniform();
            }
            x += xu;
        }
        return x;
    }
};
// noisy value of pixel pix of (WFS iwfs, subaperture isa)
__device__ __forceinline__ float add_noise(const NoiseArgs &n, float v, int pix, int iwfs, int nsa, int isa) {
    NoiseStream s;
    s.k0 = n.key0;
    s.k1 = n.key1;
    s.pix = (unsigned)pix;
    s.unit = (unsigned)(iwfs * nsa + isa);
    s.step = n.isim;
    const double tot = (double)v + (double)n.bkgrnd;
    const double c = s.poisson(tot);
    return (float)(c + (double)n.rne * s.normal() - (double)n.bkgrndc);
}

struct KArgs {
    int nsa, npix, pixpsax, pixpsay, nwvl, nwfs;
    const float *amp;
    const DevWvl *wv;   // [nwvl]
    const DevWfs *wfs;
    float etf1wt, etf2wt;
    const float2 *lotf; // [nwfs][nwvl][N*N] conj(lotfc), or null
    // pixel processing
    int phytype, sigmatch;
    float cogthres, cogoff, pixthetax, pixthetay, sigtot;
    const float *mtche; // [nwfs][nsa][2*npix]
    const float *i0sum; // [nwfs][nsa]
    const float *saa;
    float saat;
    const float *opd; // [nwfs][nsa][nxsa^2] from k_wfs_opd
    // outputs
    float *ints;     // [nwfs][nsa][npix]
    float *grad;     // [nwfs][2*nsa]
    double *i0acc;   // [nwfs][nsa][npix] or null
    DitherAcc dith;
    NoiseArgs noise;
    DtArgs dt;
};

struct OpdArgs {
    int nsa, nwfs, nps;
    double ptsdx;
    const double *ptsx, *ptsy;
    const DevLayer *layers;  // [MX_MAXREAL][MX_MAXLAYER]
    const double *atmscale;  // [MX_MAXREAL]
    const DevWfs *wfs;
    long isim;
    double dt;
    float *opd; // [nwfs][nsa][nxsa^2]
    // LLT uplink grids (nllt = nwfs llt_tiles^2 blocks, or 0 without LLT):
    // origin and launch offset; WFS w's grid is llt_tiles x llt_tiles blocks
    // of the kernel's NXSA, stored with side llt_tiles NXSA (its first llt_nx
    // rows and columns are the LLT's points)
    int nllt, llt_tiles;
    double llt_origx, llt_origy, llt_ox, llt_oy;
    float *llt_opd; // [nwfs][side^2]
    int tiles;      // a subaperture grid is tiles x tiles blocks of the kernel's NXSA
    int tworow;     // interpolate every layer from two tile rows a point (verification)
};

// Transfer-function tables of the MFMA imaging kernel (N = 64) cover the half
// plane uy < 32 in the C-tile order of its stage 2b: element (ux, uy) sits in
// 16x16 tile t = 2 (2 e + hf) + uy/16 (e = ux%2, hf = (ux/2)/16), element
// v = uy%4, lane 16 ((uy%16)/4) + (ux/2)%16.
constexpr int MX_HALF = 2048;
__host__ __device__ inline int half_slot(int ux, int uy) {
    const int e = ux & 1, c = ux >> 1, hf = c >> 4, col = c & 15;
    const int mt = uy >> 4, r = uy & 15;
    const int t = (2 * e + hf) * 2 + mt;
    return (t * 4 + (r & 3)) * 64 + 16 * (r >> 2) + col;
}

struct LltArgs {
    int nwvl, nwfs;
    const float *amp; // nlx*nlx
    const float *opd; // [nwfs][nlx*nlx] uplink OPD from k_wfs_opd
    const DevWvl *wv;
    float norm;       // 1/nlwvf^2
    float2 *lotf;     // [nwfs][nwvl][N*N] conj, transposed
    float2 *lotf_t;   // [nwfs][nwvl][MX_HALF] conj, half plane in half_slot order (N = 64), or null
};

// XOR-swizzled element (x, y) of an N x N complex tile
template <int N>
__device__ __forceinline__ int swz(int x, int y) { return y * N + (x ^ (y & (N - 1))); }

template <int N>
__device__ __forceinline__ void seg_sum(float &v) {
#pragma unroll
    for (int o = N / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
}
// sum over the 64 lanes of a wave (every lane gets it): DPP steps within
// rows of 16 lanes (lane xor 1, xor 2, mirror in 8, mirror in 16), then the
// four row sums by readlane; no LDS round trips
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum64(float v) {
    v += dpp_f<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    v += dpp_f<0x4E>(v);  // quad_perm [2, 3, 0, 1]
    v += dpp_f<0x141>(v); // row_half_mirror
    v += dpp_f<0x140>(v); // row_mirror
    const int b = __float_as_int(v);
    return (__int_as_float(__builtin_amdgcn_readlane(b, 0)) + __int_as_float(__builtin_amdgcn_readlane(b, 16))) +
           (__int_as_float(__builtin_amdgcn_readlane(b, 32)) + __int_as_float(__builtin_amdgcn_readlane(b, 48)));
}

// Ray trace every layer onto one NXSA x NXSA point grid whose first point is
// at (origx, origy), writing the sum over layers to opd (LDS, x fastest).
// Follows prop_range + prop_grid_pts with wrap=1 (lib/accphi.c:223-244,
// lib/accphi_grid_do.c:147-302): the point coordinate on each screen, its
// periodic wrap, integer part and fraction are FP64; the four weights and
// the screen values are FP32.  Coordinates of the NXSA columns and NXSA
// rows are computed once per layer by 2*NXSA lanes and shared through LDS.
// Screen column (isx) or row index and fraction of OPD grid point i on layer
// L, FP64 (prop_grid_pts, lib/accphi_grid_do.c:408-446, with the frozen-flow
// displacement of maos/wfsgrad.c:175-194): disp = ht theta + scale misreg +
// shift0 - v dt isim, scale = 1 - ht/hs (cone), periodic wrap.
// ihs = 1/hs (0 for an NGS).  The wrap count floor(X/n) comes from X * (1/n)
// except within 1e-9 of an integer quotient, where the division decides (the
// wrapped coordinate then matches the division's rounding exactly).
// The map is split into its grid-uniform part, point i of the axis at
// screen pixel X = i dout + oout (axis_map), and the per-point wrap
// (screen_wrap).
__device__ __forceinline__ void axis_map(const DevLayer &L, bool isx, double ihs, double theta, double misreg,
                                         double shift0, long isim, double dt, double ptsdx, double orig,
                                         double &dout, double &oout) {
    const double scale = 1. - L.ht * ihs;
    const double disp = L.ht * theta + scale * misreg + shift0 - (isx ? L.vx : L.vy) * dt * (double)isim;
    const double d1 = isx ? L.dx1 : L.dy1;
    const double o0 = isx ? L.ox : L.oy;
    dout = ptsdx * d1 * scale;
    oout = orig * d1 * scale + (disp - o0) * d1;
}
__device__ __forceinline__ void screen_wrap(const DevLayer &L, bool isx, int i, double dout, double oout, int &ii,
                                            float &ff) {
    const double nn = (double)(isx ? L.nx : L.ny);
    double X = fma((double)i, dout, oout);
    const double q = X * (isx ? L.inx : L.iny);
    double fq = floor(q);
    if (q - fq < 1e-9 || q - fq > 1. - 1e-9) fq = floor(X / nn);
    X -= fq * nn;
    const double fl = floor(X);
    ii = (int)fl;
    double fr = X - fl;
    const int n = isx ? L.nx : L.ny;
    if (ii >= n) ii -= n;
    if (ii < 0) {
        ii += n;
        fr = 0.;
    }
    ff = (float)fr;
}
__device__ __forceinline__ void screen_coord(const DevLayer &L, bool isx, int i, double ihs, double theta,
                                             double misreg, double shift0, long isim, double dt,
                                             double ptsdx, double orig, int &ii, float &ff) {
    double dout, oout;
    axis_map(L, isx, ihs, theta, misreg, shift0, isim, dt, ptsdx, orig, dout, oout);
    screen_wrap(L, isx, i, dout, oout, ii, ff);
}

// four floats at 4-byte (a screen row) and 16-byte (an LDS tile row)
// alignment
typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float float4_a16 __attribute__((ext_vector_type(4), aligned(16)));

// row stride (floats) of the staged screen tile of an NXSA x NXSA grid: its
// NXSA + 2 columns rounded up to 16 bytes
constexpr int rt_tile_stride(int nxsa) { return (nxsa + 2 + 3) / 4 * 4; }
// rows of that tile: its NXSA + 2 rows rounded up to whole loads of
// 64 / (stride / 4) rows
constexpr int rt_tile_rows(int nxsa) {
    return (nxsa + 2 + 64 / (rt_tile_stride(nxsa) / 4) - 1) / (64 / (rt_tile_stride(nxsa) / 4)) *
           (64 / (rt_tile_stride(nxsa) / 4));
}

// OPD of one NXSA x NXSA grid (origin origx, origy, step ptsdx) summed over
// the layers, by the 64 lanes of one wave (one block).  Per layer, lanes
// [0, 2 NXSA) find the column and row coordinates in FP64; the (NXSA+2)^2
// screen tile that the grid touches (its step on the screen is <= 1 pixel)
// is staged in LDS with coalesced loads, and the bilinear interpolation runs
// from LDS.  The layers are software-pipelined: the tile of the next layer
// is fetched into registers while the current one is interpolated, and the
// coordinates of the next layer are computed while the current tile is
// still in flight.  The tile origin (column of the first point, row of the
// first row) is wave-uniform, so the screen base address is scalar and each
// lane addresses the tile with a 32-bit offset; the coordinate lanes store
// tile-relative offsets (column, row * tile width) with the periodic wrap
// already applied.  A tile that crosses the screen edge wraps periodically;
// the common case that does not (a uniform test) skips the wrap arithmetic.
// DIRECT (a grid coarser than some screen, whose footprint exceeds the
// tile) reads the four neighbours of each point from the screen itself.
// cidx, cfrac hold 4 NXSA entries (two layers' coordinates).
template <int N, int NXSA, bool DIRECT>
__device__ __forceinline__ void ray_trace(const DevLayer *layers, int nps, double hs,
                                          double thetax, double thetay, double misregx,
                                          double misregy, double shift0x, double shift0y,
                                          long isim, double dt, double atmscale, double ptsdx,
                                          double origx, double origy, int line, float *opd,
                                          int *cidx, float *cfrac, float *tile, double *ctab,
                                          bool tworow, int ostride = NXSA) {
    static_assert(N == 64 && 2 * NXSA <= N, "one wave holds the coordinate lanes");
    constexpr int PPT = NXSA * NXSA / N;
    constexpr int TW = NXSA + 2, TS = rt_tile_stride(NXSA);
    static_assert(PPT * N == NXSA * NXSA, "points per lane");
    const double ihs = hs != 0. ? 1. / hs : 0.;
    float acc[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) acc[j] = 0.f;
    // layers above the source are skipped (uniform)
    auto next_layer = [&](int l) {
        for (l++; l < nps; l++)
            if (!(hs != 0. && layers[l].ht > hs)) break;
        return l;
    };
    const bool isx = line < NXSA;
    const int ci = isx ? line : line - NXSA;
    // this lane's points: column ix of the PPT consecutive rows from jb
    const int ix = line % NXSA, jb = line / NXSA * PPT;
    // FP64 coordinates of layer L, as (ii, ff) on the lanes [0, 2 NXSA)
    auto coords = [&](const DevLayer &L, int &ii, float &ff) {
        ii = 0;
        ff = 0.f;
        if (line < 2 * NXSA)
            screen_coord(L, isx, ci, ihs, isx ? thetax : thetay, isx ? misregx : misregy,
                         isx ? shift0x : shift0y, isim, dt, ptsdx, isx ? origx : origy, ii, ff);
    };
    if constexpr (DIRECT) {
        for (int l = next_layer(-1); l < nps; l = next_layer(l)) {
            const DevLayer L = layers[l];
            __syncthreads(); // the previous layer's coordinates fully read
            int ii;
            float ff;
            coords(L, ii, ff);
            if (line < 2 * NXSA) {
                cidx[line] = ii;
                cfrac[line] = ff;
            }
            __syncthreads();
            const int nx = L.nx, ny = L.ny;
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                const int iy = jb + j;
                const int c0 = cidx[ix], r0i = cidx[NXSA + iy];
                const int c1 = c0 + 1 == nx ? 0 : c0 + 1, r1i = r0i + 1 == ny ? 0 : r0i + 1;
                const float *p0 = L.p + (long)r0i * nx, *p1 = L.p + (long)r1i * nx;
                const float fx = cfrac[ix], fy = cfrac[NXSA + iy];
                const float r0 = fmaf(fx, p0[c1] - p0[c0], p0[c0]);
                const float r1 = fmaf(fx, p1[c1] - p1[c0], p1[c0]);
                acc[j] += fmaf(fy, r1 - r0, r0);
            }
        }
    } else {
        // The tile is TH = KM RPI >= TW rows of TS = TW rounded up to 4
        // floats (the padding rows and columns are loaded, never read).
        // Lane line loads 16 bytes: row k RPI + line / LPR, columns
        // 4 (line % LPR) .. +3, LPR lanes a row; the row step lives in the
        // 32-bit offset and the LDS offset immediates.  Lanes past RPI LPR
        // repeat lane 0's slot (the same value to the same address), so no
        // load is masked.  The screen rows are 4-byte aligned only: the
        // 16-byte loads are unaligned, which this target performs (the
        // compiler splits them where it cannot).
        constexpr int LPR = TS / 4, RPI = N / LPR, KM = (TW + RPI - 1) / RPI, TH = KM * RPI;
        static_assert(TH <= rt_tile_rows(NXSA), "tile rows");
        const bool lane_on = line < RPI * LPR;
        const int mrow = lane_on ? line / LPR : 0, mcol = lane_on ? 4 * (line % LPR) : 0;
        float4_a4 v[KM];
        // Screen and tile origin of a staged layer.  A tile inside the screen
        // is prefetched into v (its loads stay in flight until the tile is
        // stored); one that crosses the screen edge (rare) is loaded with
        // the periodic wrap when it is stored, so that v has one producer
        // and the compiler need not wait for the prefetch to merge values.
        struct Staged {
            const float *p;
            int nx, ny, tx0, ty0;
            bool inside;
        };
        auto fetch = [&](const DevLayer &L, int tx0, int ty0) {
            Staged t{L.p, L.nx, L.ny, tx0, ty0, tx0 <= L.nx - TS && ty0 <= L.ny - TH};
            if (t.inside) {
                const char *base = (const char *)(L.p + (long)ty0 * L.nx + tx0);
                const unsigned moff = (unsigned)(mrow * L.nx + mcol) * 4u, rstep = (unsigned)(RPI * L.nx) * 4u;
#pragma unroll
                for (int k = 0; k < KM; k++)
                    v[k] = *(const float4_a4 *)(base + (size_t)(moff + (unsigned)k * rstep));
            }
            return t;
        };
        auto store = [&](const Staged &t) {
            if (t.inside) {
#pragma unroll
                for (int k = 0; k < KM; k++) *(float4_a16 *)(tile + (k * RPI + mrow) * TS + mcol) = v[k];
            } else {
#pragma unroll
                for (int k = 0; k < KM; k++)
                    if (k * RPI + mrow < TW) {
                        int gy = t.ty0 + k * RPI + mrow;
                        gy -= gy >= t.ny ? t.ny : 0;
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            if (mcol + q < TW) {
                                int gx = t.tx0 + mcol + q;
                                gx -= gx >= t.nx ? t.nx : 0;
                                tile[(k * RPI + mrow) * TS + mcol + q] = t.p[(long)gy * t.nx + gx];
                            }
                    }
            }
        };
        // tile-relative coordinates of layer L into slot s; returns its
        // tile origin through tx0, ty0
        // jump: some row step of the grid is not 0 or 1 tile row (a step
        // of one pixel rounding over a row): that layer interpolates each
        // point from two tile rows instead of carrying the upper one
        auto publish = [&](const DevLayer &L, int lidx, int s, int &tx0, int &ty0, bool &jump) {
            int ii = 0;
            float ff = 0.f;
            if (line < 2 * NXSA) {
                const double *m = ctab + 2 * (2 * lidx + (isx ? 0 : 1));
                screen_wrap(L, isx, ci, m[0], m[1], ii, ff);
            }
            tx0 = __builtin_amdgcn_readlane(ii, 0);
            ty0 = __builtin_amdgcn_readlane(ii, NXSA);
            int rel = ii - (isx ? tx0 : ty0);
            rel += rel < 0 ? (isx ? L.nx : L.ny) : 0;
            const int prev = __shfl_up(rel, 1);
            jump = tworow || __any(line > NXSA && line < 2 * NXSA && (rel - prev > 1 || rel < prev));
            if (line < 2 * NXSA) {
                cidx[s * 2 * NXSA + line] = isx ? rel : rel * TS;
                cfrac[s * 2 * NXSA + line] = ff;
            }
        };
        // grid-uniform axis maps of every layer, lane 2 l + (y axis), once
        if (line < 2 * nps) {
            const bool ax = (line & 1) == 0;
            double dout, oout;
            axis_map(layers[line >> 1], ax, ihs, ax ? thetax : thetay, ax ? misregx : misregy,
                     ax ? shift0x : shift0y, isim, dt, ptsdx, ax ? origx : origy, dout, oout);
            ctab[2 * line] = dout;
            ctab[2 * line + 1] = oout;
        }
        __syncthreads();
        int l = next_layer(-1), s = 0;
        Staged cur{};
        bool jcur = false;
        if (l < nps) {
            int tx0, ty0;
            const DevLayer L0 = layers[l];
            publish(L0, l, 0, tx0, ty0, jcur);
            cur = fetch(L0, tx0, ty0);
        }
        // (also keeps the layer table loads in the loop scalar: the compiler
        // proves them unclobbered across a barrier, not across the stores
        // above)
        __syncthreads();
        while (l < nps) {
            const int ln = next_layer(l);
            const DevLayer L1 = layers[ln < nps ? ln : l];
            int tx1 = 0, ty1 = 0;
            bool jnext = false;
            if (ln < nps) publish(L1, ln, s ^ 1, tx1, ty1, jnext); // overlaps the loads of layer l
            __syncthreads(); // layer l-1's interpolation done with the tile
            store(cur);
            __syncthreads();
            if (ln < nps) cur = fetch(L1, tx1, ty1); // in flight during the interpolation
            // The lane's rows jb .. jb + PPT - 1 are consecutive and their
            // tile rows advance by 0 or 1 (a step <= 1 pixel; else jump):
            // the x-interpolated upper row of one point is the lower row of
            // the next one that advances, so each point reads one tile row.
            // Scheduling barriers every 4 points bound the registers that
            // preloaded coordinates take.
            const int *cx = cidx + s * 2 * NXSA;
            const float *cf = cfrac + s * 2 * NXSA;
            const float *tc = tile + cx[ix];
            const float fx = cf[ix];
            auto xlerp = [&](int roff) {
                const float *t = tc + roff;
                return fmaf(fx, t[1] - t[0], t[0]);
            };
            if (!jcur) {
                int rp = cx[NXSA + jb];
                float lo = xlerp(rp), up = xlerp(rp + TS);
                acc[0] += fmaf(cf[NXSA + jb], up - lo, lo);
#pragma unroll
                for (int j = 1; j < PPT; j++) {
                    if (j % 4 == 0) __builtin_amdgcn_sched_barrier(0);
                    const int r = cx[NXSA + jb + j];
                    const float nu = xlerp(r + TS);
                    if (r != rp) {
                        lo = up;
                        up = nu;
                    }
                    rp = r;
                    acc[j] += fmaf(cf[NXSA + jb + j], up - lo, lo);
                }
            } else {
#pragma unroll
                for (int j = 0; j < PPT; j++) {
                    if (j % 4 == 0) __builtin_amdgcn_sched_barrier(0);
                    const int r = cx[NXSA + jb + j];
                    const float lo = xlerp(r), up = xlerp(r + TS);
                    acc[j] += fmaf(cf[NXSA + jb + j], up - lo, lo);
                }
            }
            l = ln;
            s ^= 1;
            jcur = jnext;
        }
    }
#pragma unroll
    for (int j = 0; j < PPT; j++) opd[(jb + j) * ostride + ix] = (float)atmscale * acc[j];
}

// XCD-aware unit index: consecutive units (the WFS of one subaperture
// group, which share the ETF) land on the same XCD and L2.
__device__ __forceinline__ int xcd_unit(int b, int total) {
    const int nx = 8;
    const int xcd = b % nx, k = b / nx;
    const int q = total / nx, r = total % nx;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// ------------------------------------------------------------ main kernel
template <int N, int NXSA, int PPL>
__global__ __launch_bounds__(64) void k_shwfs_step(KArgs a) {
    constexpr int SPW = 64 / N;        // subapertures per wave
    constexpr int SKIP = (N - NXSA) / 2;
    constexpr int LG = ilog2c(N);
    __shared__ float2 tile[SPW][N * N];
    const int lane = threadIdx.x;
    const int sub = lane / N, line = lane % N;
    const int ngroup = (a.nsa + SPW - 1) / SPW;
    const int unit = xcd_unit(blockIdx.x, ngroup * a.nwfs);
    const int iwfs = unit % a.nwfs;
    const int isa_raw = (unit / a.nwfs) * SPW + sub;
    const bool valid = isa_raw < a.nsa;
    const int isa = valid ? isa_raw : a.nsa - 1;
    const DevWfs &W = a.wfs[iwfs];
    float2 *T = tile[sub];
    float *opd = reinterpret_cast<float *>(T);
    float *amp = opd + NXSA * NXSA;

    float pv[PPL];
#pragma unroll
    for (int j = 0; j < PPL; j++) pv[j] = 0.f;
    float2 v[N];
    for (int iwvl = 0; iwvl < a.nwvl; iwvl++) {
        // stage: OPD (from k_wfs_opd) and amplitude into LDS
        {
            const float *og = a.opd + ((long)iwfs * a.nsa + isa) * NXSA * NXSA;
            const float *ag = a.amp + (long)isa * NXSA * NXSA;
            for (int pt = line; pt < NXSA * NXSA; pt += N) {
                opd[pt] = og[pt];
                amp[pt] = ag[pt];
            }
        }
        __syncthreads();
        // piston (unweighted mean) is removed before exp(i k opd); |E|^2 is
        // invariant to it and the FP32 phase stays small
        float psum = 0.f;
        for (int pt = line; pt < NXSA * NXSA; pt += N) psum += opd[pt];
        seg_sum<N>(psum);
        const float mean = psum * (1.f / (NXSA * NXSA));
        const DevWvl WV = a.wv[iwvl];
        const float inv_wvl = WV.inv_wvl;
        // pass 0 (rows y): embed amp*exp(+i 2 pi opd/wvl) (cembed_wvf,
        // math/matcomp.c:56-82) and FFT along x
        const int y = line;
        const bool inband = y >= SKIP && y < SKIP + NXSA;
        if (inband) {
            const int iy = y - SKIP;
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = make_float2(0.f, 0.f);
#pragma unroll
            for (int ix = 0; ix < NXSA; ix++) {
                const float o = opd[iy * NXSA + ix] - mean;
                const float am = amp[iy * NXSA + ix];
                float sn, cs;
                sincospif(2.f * o * inv_wvl, &sn, &cs);
                v[SKIP + ix] = make_float2(am * cs, am * sn);
            }
            fft_dif<N, false>(v);
        }
        __syncthreads(); // opd/amp reads done before the tile is overwritten
#pragma unroll
        for (int k = 0; k < N; k++) T[swz<N>(bitrev(k, LG), y)] = inband ? v[k] : make_float2(0.f, 0.f);
        __syncthreads();
        // pass 1 (columns x): FFT along y, |E|^2, first half of PSF->OTF FFT
        {
            const int x = line;
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = T[swz<N>(x, k)];
            fft_dif<N, false>(v);
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = make_float2(v[k].x * v[k].x + v[k].y * v[k].y, 0.f);
            fft_dit<N, false>(v);
#pragma unroll
            for (int k = 0; k < N; k++) T[swz<N>(x, k)] = v[k];
        }
        __syncthreads();
        // pass 2 (rows uy): finish OTF, x conj(lotfc) (ccwmc), x ETF or
        // nominal (apply_dtf_etf, maos/wfsints.c:38-64), inverse FFT along x
        {
            const int uy = line;
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = T[swz<N>(k, uy)];
            fft_dif<N, false>(v); // v[k] = OTF(ux = bitrev(k), uy)
            if (a.lotf) {
                const float2 *lt = a.lotf + ((long)iwfs * a.nwvl + iwvl) * N * N + uy;
#pragma unroll
                for (int k = 0; k < N; k++) v[k] = cmul(v[k], lt[bitrev(k, LG) * N]);
            }
            if (WV.etf) {
                const float2 *e1 = WV.etf + (long)isa * N * N + uy;
                if (WV.etf2) {
                    const float2 *e2 = WV.etf2 + (long)isa * N * N + uy;
                    const float w1 = a.etf1wt, w2 = a.etf2wt;
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        const float2 e = e1[bitrev(k, LG) * N], f = e2[bitrev(k, LG) * N];
                        v[k] = cmul(v[k], make_float2(e.x * w1 + f.x * w2, e.y * w1 + f.y * w2));
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < N; k++) v[k] = cmul(v[k], e1[bitrev(k, LG) * N]);
                }
            } else if (WV.nominal) {
                const float *nm = WV.nominal + (long)isa * WV.nom_sa + uy;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const float s = nm[bitrev(k, LG) * N];
                    v[k] = make_float2(v[k].x * s, v[k].y * s);
                }
            }
            fft_dit<N, true>(v);
#pragma unroll
            for (int k = 0; k < N; k++) T[swz<N>(k, uy)] = v[k];
        }
        __syncthreads();
        // pass 3 (columns x): inverse FFT along y -> image, real part to LDS
        {
            const int x = line;
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = T[swz<N>(x, k)];
            fft_dif<N, true>(v);
            __syncthreads();
            float *img = reinterpret_cast<float *>(T);
#pragma unroll
            for (int k = 0; k < N; k++) img[bitrev(k, LG) * N + x] = v[k].x;
        }
        __syncthreads();
        // bilinear pixel sampler (dspmulcreal with si, math/spmm.c:77-86)
        {
            const float *img = reinterpret_cast<const float *>(T);
            const int4 *sidx = WV.si_idx + (long)isa * WV.si_sa;
            const float4 *sw = WV.si_w + (long)isa * WV.si_sa;
            const float alpha = W.alpha[iwvl];
#pragma unroll
            for (int j = 0; j < PPL; j++) {
                const int p = line + j * N;
                if (p < a.npix) {
                    const int4 id = sidx[p];
                    const float4 w = sw[p];
                    pv[j] += alpha * (w.x * img[id.x] + w.y * img[id.y] + w.z * img[id.z] + w.w * img[id.w]);
                }
            }
        }
        __syncthreads(); // image reads done before the next wavelength
    }

    // outputs and pixel processing (shwfs_grad, maos/utils.c:983-1122)
    const long ibase = ((long)iwfs * a.nsa + isa) * a.npix;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int j = 0; j < PPL; j++) {
        const int p = line + j * N;
        if (p < a.npix && valid) dt_frame(a.dt, ibase + p, pv[j]);
    }
    if (!(a.dt.phase & MX_DT_OUT)) return; // uniform: a frame inside a longer exposure
    if (a.noise.on)
#pragma unroll
        for (int j = 0; j < PPL; j++) {
            const int p = line + j * N;
            if (p < a.npix && valid) pv[j] = add_noise(a.noise, pv[j], p, iwfs, a.nsa, isa);
        }
#pragma unroll
    for (int j = 0; j < PPL; j++) {
        const int p = line + j * N;
        if (p < a.npix) {
            if (valid) {
                a.ints[ibase + p] = pv[j];
                if (a.i0acc) a.i0acc[ibase + p] += (double)pv[j];
                dither_add(a.dith, iwfs, ibase + p, pv[j]);
            }
            if (a.phytype == MX_PTYPE_MF) {
                const float *m = a.mtche + ibase * 2 + 2 * p;
                s0 += m[0] * pv[j];
                s1 += m[1] * pv[j];
                s2 += pv[j];
            } else {
                const float iI = pv[j] - a.cogoff;
                if (iI > a.cogthres) {
                    s0 += iI * (float)(p % a.pixpsax);
                    s1 += iI * (float)(p / a.pixpsax);
                    s2 += iI;
                }
            }
        }
    }
    seg_sum<N>(s0);
    seg_sum<N>(s1);
    seg_sum<N>(s2);
    if (line == 0 && valid && !(a.saa[isa] < a.saat)) {
        float gx = 0.f, gy = 0.f;
        if (a.sigmatch == 2) { // raw moments; k_sigmatch2 forms the gradients
            gx = s0;
            gy = s1;
        } else if (a.phytype == MX_PTYPE_MF) {
            float scale = 1.f;
            if (a.sigmatch == 1) {
                scale = a.i0sum[(long)iwfs * a.nsa + isa] / s2;
                if (scale <= 0.f || scale > 10.f || !isfinite(scale)) scale = 1.f;
            }
            gx = s0 * scale;
            gy = s1 * scale;
        } else {
            float sum = s2;
            if (a.sigmatch == 0)
                sum = a.i0sum ? a.i0sum[(long)iwfs * a.nsa + isa] : a.sigtot * a.saa[isa];
            if (sum > a.cogthres) {
                gx = (s0 / sum - (float)(a.pixpsax - 1) * 0.5f) * a.pixthetax;
                gy = (s1 / sum - (float)(a.pixpsay - 1) * 0.5f) * a.pixthetay;
            }
        }
        a.grad[(long)iwfs * 2 * a.nsa + isa] = gx;
        a.grad[(long)iwfs * 2 * a.nsa + a.nsa + isa] = gy;
    }
}

// ------------------------------------------------------------ ray trace
// OPD of every (WFS, subaperture) at step isim (the atmosphere loop of
// wfsgrad_iwfs, maos/wfsgrad.c:175-194).  Small register and LDS footprint
// (6 waves per SIMD for the one-tile 32 x 32 form), and the next layer's
// tile in flight during each interpolation, hide the screen-fetch latency;
// the FP32 OPD goes to HBM for the imaging kernel.
// threads of a ray-trace block (one grid per block, one wave)
constexpr int MX_OPD_THREADS = 64;
// smallest screen side that the staged tile (<= 36 pixels wide) wraps at
// most once
constexpr int MX_TILE_MINSCREEN = 40;
#ifndef MX_OPD_WAVES
#define MX_OPD_WAVES 6
#endif
// TILED: a subaperture grid is tiles x tiles blocks (else one block, with
// the register budget of the one-tile case untouched).
template <int NXSA, bool DIRECT, bool TILED>
__global__ __launch_bounds__(MX_OPD_THREADS) __attribute__((amdgpu_waves_per_eu(DIRECT || TILED ? MX_OPD_WAVES - 1 : MX_OPD_WAVES))) void k_wfs_opd(OpdArgs a) {
    __shared__ int cidx[4 * NXSA];
    __shared__ float cfrac[4 * NXSA];
    __shared__ double ctab[4 * MX_MAXLAYER];
    __shared__ __align__(16) float tile[rt_tile_rows(NXSA) * rt_tile_stride(NXSA)];
    const int b = (int)blockIdx.x;
    if (b < a.nllt) {
        // LLT uplink grid of WFS b / llt_tiles^2 (maos/wfsgrad.c:329-386):
        // theta - o/hs, shift o - v dt isim.  These blocks come first so that
        // they start early.
        const int lt = a.llt_tiles, lt2 = lt * lt, iw = b / lt2, t = b % lt2, tx = t % lt, ty = t / lt;
        const int side = NXSA * lt;
        const DevWfs &W = a.wfs[iw];
        ray_trace<MX_OPD_THREADS, NXSA, DIRECT>(a.layers + W.ireal * MX_MAXLAYER, a.nps, W.hs, W.thetax - a.llt_ox / W.hs,
                                        W.thetay - a.llt_oy / W.hs, 0., 0., a.llt_ox, a.llt_oy, a.isim, a.dt,
                                        a.atmscale[W.ireal], a.ptsdx, a.llt_origx + a.ptsdx * (double)(NXSA * tx),
                                        a.llt_origy + a.ptsdx * (double)(NXSA * ty), threadIdx.x,
                                        a.llt_opd + (long)iw * side * side + (long)ty * NXSA * side + tx * NXSA, cidx,
                                        cfrac, tile, ctab, a.tworow, side);
        return;
    }
    if constexpr (!TILED) {
        const int unit = xcd_unit(b - a.nllt, a.nsa * a.nwfs);
        const int iwfs = unit % a.nwfs, isa = unit / a.nwfs;
        const DevWfs &W = a.wfs[iwfs];
        ray_trace<MX_OPD_THREADS, NXSA, DIRECT>(a.layers + W.ireal * MX_MAXLAYER, a.nps, W.hs, W.thetax, W.thetay,
                                                W.misregx, W.misregy, 0., 0., a.isim, a.dt, a.atmscale[W.ireal],
                                                a.ptsdx, a.ptsx[isa], a.ptsy[isa], threadIdx.x,
                                                a.opd + ((long)iwfs * a.nsa + isa) * NXSA * NXSA, cidx, cfrac, tile, ctab, a.tworow);
        return;
    }
    // unit: (WFS, subaperture, tile); a subaperture grid of side NXSA * tiles
    // is traced as tiles x tiles blocks of NXSA x NXSA points
    const int T2 = a.tiles * a.tiles;
    const int unit = xcd_unit(b - a.nllt, a.nsa * a.nwfs * T2);
    const int iwfs = unit % a.nwfs, rest = unit / a.nwfs;
    const int isa = rest / T2, it = rest % T2, tx = it % a.tiles, ty = it / a.tiles;
    const DevWfs &W = a.wfs[iwfs];
    const int side = NXSA * a.tiles;
    ray_trace<MX_OPD_THREADS, NXSA, DIRECT>(a.layers + W.ireal * MX_MAXLAYER, a.nps, W.hs, W.thetax, W.thetay, W.misregx,
                                    W.misregy, 0., 0., a.isim, a.dt, a.atmscale[W.ireal], a.ptsdx,
                                    a.ptsx[isa] + a.ptsdx * (double)(NXSA * tx), a.ptsy[isa] + a.ptsdx * (double)(NXSA * ty),
                                    threadIdx.x,
                                    a.opd + ((long)iwfs * a.nsa + isa) * side * side + (long)ty * NXSA * side + tx * NXSA,
                                    cidx, cfrac, tile, ctab, a.tworow, side);
}

// ------------------------------------------------------------ OPD inputs
// Static surface OPD (powfs.opdadd, maos/wfsgrad.c:228-230) and per-step
// piston/tip/tilt (loc_add_ptt, maos/wfsgrad.c:243-256: TTM/FSM commands)
// added to the traced OPD of every subaperture point:
// opd += opdadd + ptt0 + ptt1 x + ptt2 y, x = origx + ix dx.
struct OpdAddArgs {
    int nsa, nxsa, nwfs;
    double dx;
    const double *ptsx, *ptsy;
    const float *opdadd; // [nwfs][nsa][nxsa^2] or null
    const double *ptt;   // [nwfs][3] or null
    float *opd;
};
__global__ __launch_bounds__(256) void k_opd_add(OpdAddArgs a) {
    const long np = (long)a.nxsa * a.nxsa, total = np * a.nsa * a.nwfs;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int pt = (int)(i % np);
        const long sw = i / np;
        const int isa = (int)(sw % a.nsa), iwfs = (int)(sw / a.nsa);
        float v = a.opd[i];
        if (a.opdadd) v += a.opdadd[i];
        if (a.ptt) {
            const double *t = a.ptt + 3 * iwfs;
            const double x = a.ptsx[isa] + a.dx * (double)(pt % a.nxsa), y = a.ptsy[isa] + a.dx * (double)(pt / a.nxsa);
            v += (float)(t[0] + t[1] * x + t[2] * y);
        }
        a.opd[i] = v;
    }
}

// ------------------------------------------------------------ dither fold
// At each PLL output of the TT dither (maos/wfsgrad.c:709-731): the folded
// sums take fi0 += s0 imb, fgx += sg imx, fgy += sg imy, and the per-step
// sums are cleared.
struct FoldArgs {
    long n;
    double s0, sg;
    double *imb, *imx, *imy, *fi0, *fgx, *fgy;
};
__global__ __launch_bounds__(256) void k_dither_fold(FoldArgs a) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x) {
        a.fi0[i] += a.s0 * a.imb[i];
        a.fgx[i] += a.sg * a.imx[i];
        a.fgy[i] += a.sg * a.imy[i];
        a.imb[i] = 0.;
        a.imx[i] = 0.;
        a.imy[i] = 0.;
    }
}

// ------------------------------------------------------------ MF update
// Matched filter rebuilt on the GPU (FP64) from the model i0, gx, gy after
// the low-pass blend of maos/wfsgrad.c:1036-1047, i0 = keep i0 + acc fi0
// (gx, gy likewise; the folded sums are cleared), following mtch
// (lib/mtch.c:93-211): columns [gx, gy, i0/pixtheta + 2 bkgrnd_res] and,
// with the constraints, i0 shifted by one pixel either way in x and/or y
// (+ bkgrnd_res); weights 1/(rne^2 + bkgrnd + i0) (1/rne^2 where i0 < 0);
// mtche = i0m pinv(A, W) with the eigenvalue thresholds of dpinv (1e-7 of
// the largest, and 1e-3 of the previous kept); NEA^2 = mtche W^-1 mtche^T.
// One block per (WFS, subaperture); the per-WFS thresholds of mtch_cell
// (lib/mtch.c:300-318) follow in k_mtch_fin.
constexpr int MX_MTCH_THREADS = 256;
struct MtchArgs {
    int nsa, npix, pixpsax, pixpsay, iwfs0;
    int crx, cry; // column of the x, y constraint pairs, or 0
    int nmod;
    double keep, acc;
    double bkres, bkgrnd, rne2, kpx, kpy, sigratio, nea2thres, rne;
    double *i0, *gx, *gy;    // model [nwfs][nsa][npix], updated in place
    double *fi0, *fgx, *fgy; // folded sums, cleared
    double *nea;             // [nwfs][nsa][4]
    double *i0sum;           // [nwfs][nsa]
    float *mtche;            // [nwfs][nsa][npix][2]
    float *i0sumf;           // [nwfs][nsa]
};

// column k of the mtch design matrix at pixel p (mki0sh, lib/mtch.c:23-33)
__device__ __forceinline__ double mtch_col(const MtchArgs &a, const double *i0, const double *gx,
                                          const double *gy, int p, int k) {
    if (k == 0) return gx[p];
    if (k == 1) return gy[p];
    if (k == 2) return i0[p] * a.kpx + 2. * a.bkres;
    const int ix = p % a.pixpsax, iy = p / a.pixpsax;
    if (a.crx && (k == a.crx || k == a.crx + 1)) {
        // i0x1 = i0 shifted +1 in x, i0x2 = shifted -1 (zero where outside)
        if (k == a.crx) return (ix >= 1 ? i0[p - 1] * a.kpx : 0.) + a.bkres;
        return (ix + 1 < a.pixpsax ? i0[p + 1] * a.kpx : 0.) + a.bkres;
    }
    if (k == a.cry) return (iy >= 1 ? i0[p - a.pixpsax] * a.kpy : 0.) + a.bkres;
    return (iy + 1 < a.pixpsay ? i0[p + a.pixpsax] * a.kpy : 0.) + a.bkres;
}

// symmetric eigen-decomposition by cyclic Jacobi rotations (n <= 7)
__device__ void jacobi_eig_dev(double *A, int n, double *d, double *V) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i + j * n] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p + q * n] * A[p + q * n];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p + q * n];
                if (fabs(apq) < 1e-300) continue;
                const double app = A[p + p * n], aqq = A[q + q * n];
                const double tau = (aqq - app) / (2 * apq);
                const double t = (tau >= 0 ? 1. : -1.) / (fabs(tau) + sqrt(1 + tau * tau));
                const double c = 1. / sqrt(1 + t * t), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = c * akp - s * akq;
                    A[k + q * n] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = c * apk - s * aqk;
                    A[q + k * n] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = c * vkp - s * vkq;
                    V[k + q * n] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) d[i] = A[i + i * n];
}

// block sum of v (all threads get it)
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x / 64;
    __syncthreads(); // red free
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < NT / 64; i++) s += red[i];
    return s;
}

__global__ __launch_bounds__(MX_MTCH_THREADS) void k_mtch(MtchArgs a) {
    __shared__ double red[MX_MTCH_THREADS / 64];
    __shared__ double inv[49];
    const int isa = (int)(blockIdx.x % (unsigned)a.nsa), iwfs = a.iwfs0 + (int)(blockIdx.x / (unsigned)a.nsa);
    const long base = ((long)iwfs * a.nsa + isa) * a.npix;
    double *i0 = a.i0 + base, *gx = a.gx + base, *gy = a.gy + base;
    // the low-pass blend of the folded sums into the model
    for (int p = threadIdx.x; p < a.npix; p += MX_MTCH_THREADS) {
        i0[p] = a.keep * i0[p] + a.acc * a.fi0[base + p];
        gx[p] = a.keep * gx[p] + a.acc * a.fgx[base + p];
        gy[p] = a.keep * gy[p] + a.acc * a.fgy[base + p];
        a.fi0[base + p] = 0.;
        a.fgx[base + p] = 0.;
        a.fgy[base + p] = 0.;
    }
    __syncthreads();
    const int nm = a.nmod;
    // normal matrix A^T W A (upper triangle) and the i0 sum
    double cc[28];
    for (int k = 0; k < 28; k++) cc[k] = 0.;
    double s0 = 0.;
    for (int p = threadIdx.x; p < a.npix; p += MX_MTCH_THREADS) {
        const double w = i0[p] < 0 ? 1. / a.rne2 : 1. / (a.rne2 + a.bkgrnd + i0[p]);
        double col[7];
        for (int k = 0; k < nm; k++) col[k] = mtch_col(a, i0, gx, gy, p, k);
        int t = 0;
        for (int i = 0; i < nm; i++)
            for (int j = i; j < nm; j++) cc[t++] += col[i] * w * col[j];
        s0 += i0[p];
    }
    const int ncc = nm * (nm + 1) / 2;
    for (int k = 0; k < ncc; k++) cc[k] = block_sum<MX_MTCH_THREADS>(cc[k], red);
    s0 = block_sum<MX_MTCH_THREADS>(s0, red);
    if (threadIdx.x == 0) {
        double A[49], d[7], V[49];
        int t = 0;
        for (int i = 0; i < nm; i++)
            for (int j = i; j < nm; j++) {
                A[i + j * nm] = cc[t];
                A[j + i * nm] = cc[t];
                t++;
            }
        jacobi_eig_dev(A, nm, d, V);
        int idx[7];
        for (int i = 0; i < nm; i++) idx[i] = i;
        for (int i = 0; i < nm; i++)
            for (int j = i + 1; j < nm; j++)
                if (d[idx[j]] > d[idx[i]]) {
                    const int tt = idx[i];
                    idx[i] = idx[j];
                    idx[j] = tt;
                }
        double sv[7];
        const double maxeig = d[idx[0]];
        const double thres10 = 1e-7 * maxeig;
        double thres20 = 1e-3 * maxeig;
        bool cut = false;
        for (int i = 0; i < nm; i++) {
            const double ev = d[idx[i]];
            if (!cut && ev > thres10 && ev > thres20) {
                thres20 = ev * 1e-3;
                sv[i] = 1. / ev;
            } else {
                cut = true;
                sv[i] = 0.;
            }
        }
        for (int i = 0; i < nm; i++)
            for (int j = 0; j < nm; j++) {
                double acc = 0;
                for (int k = 0; k < nm; k++) acc += V[i + idx[k] * nm] * sv[k] * V[j + idx[k] * nm];
                inv[i + j * nm] = acc;
            }
    }
    __syncthreads();
    // mtche = i0m inv A^T W; NEA^2 = mtche W^-1 mtche^T
    double n00 = 0, n01 = 0, n11 = 0;
    for (int p = threadIdx.x; p < a.npix; p += MX_MTCH_THREADS) {
        const double w = i0[p] < 0 ? 1. / a.rne2 : 1. / (a.rne2 + a.bkgrnd + i0[p]);
        double col[7];
        for (int k = 0; k < nm; k++) col[k] = mtch_col(a, i0, gx, gy, p, k);
        double m[2] = {0., 0.};
        for (int i = 0; i < nm; i++) {
            double ti = 0;
            for (int j = 0; j < nm; j++) ti += inv[i + j * nm] * col[j];
            ti *= w;
            // i0m: rows (gx | +x, -x), (gy | +y, -y)
            if (i == 0 || (a.crx && i == a.crx)) m[0] += ti;
            if (a.crx && i == a.crx + 1) m[0] -= ti;
            if (i == 1 || (a.cry && i == a.cry)) m[1] += ti;
            if (a.cry && i == a.cry + 1) m[1] -= ti;
        }
        a.mtche[(base + p) * 2] = (float)m[0];
        a.mtche[(base + p) * 2 + 1] = (float)m[1];
        n00 += m[0] * m[0] / w;
        n01 += m[0] * m[1] / w;
        n11 += m[1] * m[1] / w;
    }
    n00 = block_sum<MX_MTCH_THREADS>(n00, red);
    n01 = block_sum<MX_MTCH_THREADS>(n01, red);
    n11 = block_sum<MX_MTCH_THREADS>(n11, red);
    if (threadIdx.x == 0) {
        const long u = (long)iwfs * a.nsa + isa;
        const double r = fabs(a.sigratio - 1.) > 1e-5 ? 1. / a.sigratio : 1.;
        a.nea[4 * u] = n00 * r;
        a.nea[4 * u + 1] = n01 * r;
        a.nea[4 * u + 2] = n01 * r;
        a.nea[4 * u + 3] = n11 * r;
        a.i0sum[u] = s0;
        a.i0sumf[u] = (float)s0;
    }
}

// per WFS: subapertures with too little signal (i0sum below max(0.1 of the
// largest, 10 rne)) or too large an NEA (above 100 pixtheta^2) get a zero
// filter and NEA^2 = nea2thres (mtch_cell, lib/mtch.c:300-318)
__global__ __launch_bounds__(MX_MTCH_THREADS) void k_mtch_fin(MtchArgs a) {
    __shared__ double red[MX_MTCH_THREADS / 64];
    const int iwfs = a.iwfs0 + (int)blockIdx.x;
    const long u0 = (long)iwfs * a.nsa;
    double mx = -INFINITY;
    for (int isa = threadIdx.x; isa < a.nsa; isa += MX_MTCH_THREADS) mx = fmax(mx, a.i0sum[u0 + isa]);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = mx;
    __syncthreads();
    mx = red[0];
    for (int i = 1; i < MX_MTCH_THREADS / 64; i++) mx = fmax(mx, red[i]);
    const double i0thres = mx * 0.1 > 10 * a.rne ? mx * 0.1 : 10 * a.rne;
    for (int isa = 0; isa < a.nsa; isa++) {
        const long u = u0 + isa;
        const bool zero = a.i0sum[u] < i0thres || a.nea[4 * u] > a.nea2thres || a.nea[4 * u + 3] > a.nea2thres;
        if (!zero) continue; // uniform: every thread reads the same values
        for (int p = threadIdx.x; p < 2 * a.npix; p += MX_MTCH_THREADS) a.mtche[u * 2 * a.npix + p] = 0.f;
        __syncthreads(); // all threads have read nea before thread 0 changes it
        if (threadIdx.x == 0) {
            a.nea[4 * u] = a.nea[4 * u + 3] = a.nea2thres;
            a.nea[4 * u + 1] = a.nea[4 * u + 2] = 0.;
        }
    }
}

// ------------------------------------------------------------ DM surfaces
// The deformable mirrors seen by each WFS (maos/wfsgrad.c:236-242 with the
// propdata of maos/sim_utils.c:1035-1052): every DM surface map is traced
// onto the subaperture grids like a screen at the DM's conjugate height
// (cone scale 1 - ht/hs, displacement ht theta + scale misreg, no wind) with
// alpha = -1 and without wrapping: a point is used when its row lies in
// [irows, rowdiv0) and its column in [icols, coldiv), the ranges of
// prop_grid_block's unequal-sampling form (lib/accphi_grid_do.c:147-302),
// so points beyond the map's last column or row are left alone.
constexpr int MX_MAXDM = 8;
struct DmArgs {
    int nsa, nxsa, nwfs, ndm;
    double ptsdx;
    const double *ptsx, *ptsy;
    const DevLayer *dms; // [MX_MAXREAL][MX_MAXDM]
    const DevWfs *wfs;
    float *opd;
};
__global__ __launch_bounds__(64) void k_dm_opd(DmArgs a) {
    const int unit = xcd_unit((int)blockIdx.x, a.nsa * a.nwfs);
    const int iwfs = unit % a.nwfs, isa = unit / a.nwfs;
    const DevWfs &W = a.wfs[iwfs];
    const int n = a.nxsa;
    float *op = a.opd + ((long)iwfs * a.nsa + isa) * n * n;
    const double ihs = W.hs != 0. ? 1. / W.hs : 0.;
    for (int d = 0; d < a.ndm; d++) {
        const DevLayer L = a.dms[W.ireal * MX_MAXDM + d];
        if (W.hs != 0. && L.ht > W.hs) continue;
        const double scale = 1. - L.ht * ihs;
        const double dxout = a.ptsdx * L.dx1 * scale, dyout = a.ptsdx * L.dy1 * scale;
        const double oxout = a.ptsx[isa] * L.dx1 * scale + (L.ht * W.thetax + scale * W.misregx - L.ox) * L.dx1;
        const double oyout = a.ptsy[isa] * L.dy1 * scale + (L.ht * W.thetay + scale * W.misregy - L.oy) * L.dy1;
        const double dxout1 = 1. / dxout, dyout1 = 1. / dyout;
        const int wrapx = L.nx - 1, wrapy = L.ny - 1;
        const long icols = oyout < 0 ? (long)ceil(-oyout * dyout1) : 0;
        const long coldiv = (long)floor((wrapy - oyout) * dyout1) + 1;
        const long irows = oxout < 0 ? (long)ceil(-oxout * dxout1) : 0;
        const long rowdiv0 = (long)ceil((wrapx - oxout) * dxout1);
        for (int pt = threadIdx.x; pt < n * n; pt += 64) {
            const int ix = pt % n, iy = pt / n;
            if (iy < icols || iy >= coldiv || ix < irows || ix >= rowdiv0) continue;
            const double x = fma((double)ix, dxout, oxout), y = fma((double)iy, dyout, oyout);
            // the ranges keep x in [0, nx-1) and y in [0, ny-1]; the clamps
            // only catch rounding at those edges
            const int nx0 = min(max((int)floor(x), 0), wrapx - 1), ny0 = min(max((int)floor(y), 0), wrapy);
            const float fx = (float)(x - nx0), fy = (float)(y - ny0);
            const float *r0 = L.p + (long)ny0 * L.nx, *r1 = L.p + (long)(ny0 == wrapy ? 0 : ny0 + 1) * L.nx;
            const float v0 = fmaf(fx, r0[nx0 + 1] - r0[nx0], r0[nx0]);
            const float v1 = fmaf(fx, r1[nx0 + 1] - r1[nx0], r1[nx0]);
            op[pt] -= fmaf(fy, v1 - v0, v0);
        }
    }
}

// ------------------------------------------------------------ geometric
// Zernike-tilt gradients of each (WFS, subaperture) from its OPD (pts_ztilt,
// math/loc.c:839-880, used at maos/wfsgrad.c:303-306): moments
// a = sum amp opd [1, x, y] in FP64, times the inverse of
// M = sum amp [1, x, y]^T [1, x, y] (pts_mcc_ptt, math/loc.c:612-650,
// inverted on the host); the gradients are the x and y terms.
// Geometric average-gradient (G-tilt) gradients: grad = GS0 opd
// (dspmm at maos/wfsgrad.c:308) with GS0 from mkg (lib/mkg.c:108-468), a
// CSR of 2 nsa rows (x, then y) over the nsa nxsa^2 OPD points, shared by
// all WFS.  One wave per row computes the row for every WFS (the row's
// entries are read once), FP64 sums; an empty row gives 0.
constexpr int MX_GT_WFS = 8; // WFS per pass of a row
struct GtiltArgs {
    int nsa, nwfs;
    long npts; // OPD points per WFS
    const int *rowp, *coli;
    const double *val;
    const float *opd; // [nwfs][npts]
    float *grad;      // [nwfs][2*nsa]
};
__global__ __launch_bounds__(64) void k_gtilt(GtiltArgs a) {
    const int r = (int)blockIdx.x, lane = (int)threadIdx.x;
    const int s = a.rowp[r], e = a.rowp[r + 1];
    for (int w0 = 0; w0 < a.nwfs; w0 += MX_GT_WFS) {
        double acc[MX_GT_WFS];
#pragma unroll
        for (int j = 0; j < MX_GT_WFS; j++) acc[j] = 0.;
        for (int k = s + lane; k < e; k += 64) {
            const double v = a.val[k];
            const long c = a.coli[k];
#pragma unroll
            for (int j = 0; j < MX_GT_WFS; j++)
                if (w0 + j < a.nwfs) acc[j] += v * (double)a.opd[(long)(w0 + j) * a.npts + c];
        }
#pragma unroll
        for (int j = 0; j < MX_GT_WFS; j++) {
            double t = acc[j];
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
            if (lane == 0 && w0 + j < a.nwfs) a.grad[(long)(w0 + j) * 2 * a.nsa + r] = (float)t;
        }
    }
}

struct ZtiltArgs {
    int nsa, nxsa, nwfs;
    double dx;
    const double *ptsx, *ptsy;
    const float *amp;   // [nsa][nxsa^2]
    const double *imcc; // [nsa][9] inverse moment matrices
    const float *opd;
    const float *saa;
    float saat;
    float *grad;        // [nwfs][2*nsa]
    int all;            // every subaperture (else those with saa >= saat)
};
__global__ __launch_bounds__(64) void k_ztilt(ZtiltArgs a) {
    const int unit = xcd_unit((int)blockIdx.x, a.nsa * a.nwfs);
    const int iwfs = unit % a.nwfs, isa = unit / a.nwfs;
    const long np = (long)a.nxsa * a.nxsa;
    const float *op = a.opd + ((long)iwfs * a.nsa + isa) * np, *am = a.amp + (long)isa * np;
    double a0 = 0., a1 = 0., a2 = 0.;
    for (int pt = threadIdx.x; pt < np; pt += 64) {
        const float w = am[pt];
        if (w != 0.f) {
            const double t = (double)w * (double)op[pt];
            a0 += t;
            a1 += t * (a.ptsx[isa] + a.dx * (double)(pt % a.nxsa));
            a2 += t * (a.ptsy[isa] + a.dx * (double)(pt / a.nxsa));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, 64);
        a1 += __shfl_xor(a1, o, 64);
        a2 += __shfl_xor(a2, o, 64);
    }
    if (threadIdx.x == 0 && (a.all || !(a.saa[isa] < a.saat))) {
        const double *m = a.imcc + 9 * (long)isa; // column-major 3x3
        a.grad[(long)iwfs * 2 * a.nsa + isa] = (float)(m[1] * a0 + m[4] * a1 + m[7] * a2);
        a.grad[(long)iwfs * 2 * a.nsa + a.nsa + isa] = (float)(m[2] * a0 + m[5] * a1 + m[8] * a2);
    }
}

// ------------------------------------------------------------ uplink OTF
// lotfc of wfsints (maos/wfsints.c:149-199) for nlwvf == nwvf == notf: from
// the LLT OPD (traced by k_wfs_opd), embed, FFT, |.|^2/nlwvf^2, FFT.
// Stored conjugated and transposed for the imaging kernels.
template <int N, int NLX>
__global__ __launch_bounds__(64) void k_llt_otf(LltArgs a) {
    constexpr int SKIP = (N - NLX) / 2;
    constexpr int LG = ilog2c(N);
    __shared__ float2 T[N * N];
    const int lane = threadIdx.x;
    const int line = lane % N;
    const bool active = lane < N;
    const int iwfs = blockIdx.x / a.nwvl, iwvl = blockIdx.x % a.nwvl;
    float *opd = reinterpret_cast<float *>(T);
    float *amp = opd + NLX * NLX;
    if (active)
        for (int pt = line; pt < NLX * NLX; pt += N) {
            amp[pt] = a.amp[pt];
            opd[pt] = a.opd[(long)iwfs * NLX * NLX + pt];
        }
    __syncthreads();
    float psum = 0.f;
    if (active)
        for (int pt = line; pt < NLX * NLX; pt += N) psum += opd[pt];
    seg_sum<N>(psum);
    const float mean = psum * (1.f / (NLX * NLX));
    float2 v[N];
    const float inv_wvl = a.wv[iwvl].inv_wvl;
    const int y = line;
    const bool inband = active && y >= SKIP && y < SKIP + NLX;
    if (inband) {
        const int iy = y - SKIP;
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = make_float2(0.f, 0.f);
#pragma unroll
        for (int ix = 0; ix < NLX; ix++) {
            const float o = opd[iy * NLX + ix] - mean;
            const float am = amp[iy * NLX + ix];
            float sn, cs;
            sincospif(2.f * o * inv_wvl, &sn, &cs);
            v[SKIP + ix] = make_float2(am * cs, am * sn);
        }
        fft_dif<N, false>(v);
    }
    __syncthreads();
    if (active)
#pragma unroll
        for (int k = 0; k < N; k++) T[swz<N>(bitrev(k, LG), y)] = inband ? v[k] : make_float2(0.f, 0.f);
    __syncthreads();
    if (active) {
        const int x = line;
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = T[swz<N>(x, k)];
        fft_dif<N, false>(v);
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = make_float2((v[k].x * v[k].x + v[k].y * v[k].y) * a.norm, 0.f);
        fft_dit<N, false>(v);
#pragma unroll
        for (int k = 0; k < N; k++) T[swz<N>(x, k)] = v[k];
    }
    __syncthreads();
    if (active) {
        const int uy = line;
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = T[swz<N>(k, uy)];
        fft_dif<N, false>(v);
        float2 *out = a.lotf + ((long)iwfs * a.nwvl + iwvl) * N * N + uy;
#pragma unroll
        for (int k = 0; k < N; k++) out[bitrev(k, LG) * N] = make_float2(v[k].x, -v[k].y);
        if (N == 64 && a.lotf_t && uy < 32) {
            // the half plane uy < 32 in the C-tile order of the MFMA imaging
            // kernel's stage 2b
            float2 *ot = a.lotf_t + ((long)iwfs * a.nwvl + iwvl) * MX_HALF;
#pragma unroll
            for (int k = 0; k < N; k++) ot[half_slot(bitrev(k, LG), uy)] = make_float2(v[k].x, -v[k].y);
        }
    }
}


// ------------------------------------------------------------ MFMA form
// The same imaging chain with every 1-D DFT stage computed as a complex
// matrix product on the matrix cores (v_mfma_f32_16x16x32_bf16, or _f16).
// Operands are split into hi + lo pairs of the element type E (BF16, or FP16
// for comparison) and each real product is hi*hi + hi*lo + lo*hi,
// accumulated in FP32.  One workgroup of four waves images one subaperture
// (N = 64, nxsa = 32).  FP16 has 11 significant bits against BF16's 8 but
// the range of +-65504, and a lo part below 2^-14 is subnormal: every stored
// operand is scaled by a power of two so that its bound is 2^14 (field |E| <=
// amp <= 1; Z <= 32; P+- <= 2^21; Q+- <= 2^23; X <= 2^22 times the uplink OTF
// at zero, the sum of the LLT amplitude squared, <= 1024; R <= 64 max X), the
// DFT tables by 2^13, and the image by the product of the scales removed at
// the end.
//
// The field is the A operand (LDS rows contiguous along the summed index) and
// the symmetric DFT matrix F[k][n] = exp(-2 pi i k n / 64) the B operand.  A
// stage writes its output transposed, split into four BF16 planes (re hi, re
// lo, im hi, im lo), contiguous along the index the next stage sums over.
//   1a Z[ky][x]  = sum_y W[x][y] F[y][ky]       (32x32 band, K = 32)
//   1b Y[kx][ky] = sum_x Z[ky][x] F[x][kx]      -> P = |Y|^2
//   2a Q[uy][kx] = sum_ky P[kx][ky] F[ky][uy]   (uy < 32)
//   2b O[ux][uy] = sum_kx Q[uy][kx] F[kx][ux]   -> X = O conj(lotf) ETF
//   3a R[ys][ux] = sum_uy X[ux][uy] w(uy) G[uy][ys]   (sampled image rows)
//   3b I[ys][xs] = Re sum_ux R[ys][ux] G[ux][xs]      (sampled image columns)
// (G = conj F at <= 16 sampled coordinates), then the si sampler, pixel
// processing and FP64 i0 as in k_shwfs_step.
//
// Two properties halve the inverse transform:
// - P is real, so O and X = O conj(lotf) ETF are Hermitian (every factor is
//   the transform of a real function); the field fills 32 of 64 samples per
//   axis, so O (its autocorrelation) is zero at lag 32.  The half plane
//   uy < 32 carries all of X, with weight w = 1 at uy = 0 and 2 elsewhere.
// - Radix-2 split of the 64-point sums in 2a and 2b: F[k + 32][n] = (-1)^n
//   F[k][n], so even outputs sum (x[k] + x[k+32]) and odd outputs (x[k] -
//   x[k+32]) over k < 32.  The producing stage forms both in FP32 before the
//   split (its C tile holds k and k + 32 in the same lane) and stores them side
//   by side, columns [0, 32) and [32, 64).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
template <typename E>
using v8_t = E __attribute__((ext_vector_type(8)));
template <typename E>
constexpr bool is_f16 = std::is_same<E, _Float16>::value;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS plane layout: element (row, k) of a plane sits at row * MLDB + (k ^
// swizzle(row)).  The swizzle permutes whole 8-element (16-byte) chunks, so
// the 16-byte operand reads and the 8-byte stores stay contiguous.
#ifndef MX_LDS_SWZ
#define MX_LDS_SWZ 2
#endif
constexpr int MLDB = MX_LDS_SWZ ? 64 : 72; // bf16 row stride of the LDS planes
constexpr int MPLANE = 64 * MLDB;
__device__ __forceinline__ int lds_at(int row, int k) {
    if constexpr (MX_LDS_SWZ == 1) return row * MLDB + (k ^ ((row & 7) << 3));
    else if constexpr (MX_LDS_SWZ == 2) return row * MLDB + (k ^ (((row >> 1) & 7) << 3));
    else return row * MLDB + k;
}
constexpr int MX_MFMA_NS = 16;  // sampled image rows/columns kept by stage 3

// Table entries: [.. entry ..][part][lane] 8 elements (BF16 or FP16 tables),
// parts re hi, re lo, im hi,
// im lo, -im hi, -im lo; a lane holds 8 consecutive k of one n (B-operand
// layout).  The negated imaginary part lets every complex product accumulate
// with additions only.
constexpr int MX_TPARTS = 6;
// DFT table entries (each MX_TPARTS x 64 bf16x8), B operands of
// v_mfma_f32_16x16x32_bf16
constexpr int MX_F1A = 0;  // + nt: n = 16 nt + l%16, k = 16 + 8 (l/16) + j (the band)
constexpr int MX_F2 = 4;   // + 2 e + hf: n = 2 (16 hf + l%16) + e, k = 8 (l/16) + j
constexpr int MX_FENT = 8;
struct MfmaWvl {
    const float2 *etf_t;     // [nsa][MX_HALF] ETF (nominal fused), half_slot order, or null
    const float *nominal_t;  // [MX_HALF] nominal when there is no ETF, or null
    const void *gx;          // [2 ks][6][64] conj(F) at the sampled columns
    const void *gy;          // [6][64] w(uy) conj(F) at the sampled rows, uy < 32
    const int4 *si_c;        // [npix] compact indices (iy*16 + ix) into the 16x16 image
    const float4 *si_w;      // [npix]
    int si_sa, nom_sa;       // per-subaperture strides of si_c / si_w and nominal_t (0: shared)
    float inv_wvl;
    // FP32 radix-8 form: w(uy) exp(+2 pi i uy ys / 64) [uy < 32][16 sampled
    // rows ys], and the 16 sampled columns xs (the list padded with its first)
    const float2 *fy;
    const int *fxs;
};

struct MfmaArgs {
    int nsa, npix, pixpsax, pixpsay, nwvl, nwfs;
    const float *amp;
    const float *opd;
    const void *F;            // [MX_FENT][6][64], element type of the form
    const MfmaWvl *wv;
    const float2 *lotf_t;     // [nwfs][nwvl][MX_HALF] conj(lotfc), or null
    const DevWfs *wfs;
    int phytype, sigmatch;
    float cogthres, cogoff, pixthetax, pixthetay, sigtot;
    const float *mtche, *i0sum, *saa;
    float saat;
    float *ints, *grad;
    double *i0acc;
    DitherAcc dith;
    NoiseArgs noise;
    DtArgs dt;
    // uplink mode (one workgroup per WFS and wavelength): LLT OPD
    // [nwfs][32*32] and amplitude, 1/nlwvf^2, and the two outputs conj(lotfc)
    // [nwfs][nwvl][64*64] ([ux][uy]) and its half plane [nwfs][nwvl][MX_HALF]
    const float *llt_opd, *llt_amp;
    float llt_norm;
    float2 *lotf_full, *lotf_half;
    // FUSED instances trace their own grid first (the screens and geometry
    // of k_wfs_opd in tr) into LDS, and store the traced OPD to opd_out
    // ([nwfs][nsa][32*32]) or llt_opd_out ([nwfs][32*32]) for later readers
    OpdArgs tr;
    float *opd_out, *llt_opd_out;
};
// per-wave LDS scratch of the fused ray trace (cidx, cfrac, the layer table,
// the staged 16 x 16 tile), carved from the imaging planes
constexpr int MX_FUSE_SCR = 3072;

template <typename E>
__device__ __forceinline__ f32x4 mf16(v8_t<E> a, v8_t<E> b, f32x4 c) {
    if constexpr (is_f16<E>) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// c += A * B for one K step with split operands: hi*hi + hi*lo + lo*hi
template <typename E>
__device__ __forceinline__ void mac3(f32x4 &c, v8_t<E> ah, v8_t<E> al, v8_t<E> bh, v8_t<E> bl) {
    c = mf16<E>(ah, bh, c);
    c = mf16<E>(ah, bl, c);
    c = mf16<E>(al, bh, c);
}

// (cr, ci) += D * B with D the field (A operand, from LDS) and B = b[0] + i
// b[1]; b[2] = -b[1]
template <typename E, bool DREAL>
__device__ __forceinline__ void cmacd(f32x4 &cr, f32x4 &ci, const E *const (&L)[4], int row, int k,
                                      const v8_t<E> (&b)[3][2], int poff = 0) {
    // poff: element offset of the real data's hi/lo planes (DREAL)
    const v8_t<E> dh = *reinterpret_cast<const v8_t<E> *>(L[0] + poff + lds_at(row, k));
    const v8_t<E> dl = *reinterpret_cast<const v8_t<E> *>(L[1] + poff + lds_at(row, k));
    mac3<E>(cr, dh, dl, b[0][0], b[0][1]);
    mac3<E>(ci, dh, dl, b[1][0], b[1][1]);
    if (!DREAL) {
        const v8_t<E> eh = *reinterpret_cast<const v8_t<E> *>(L[2] + lds_at(row, k));
        const v8_t<E> el = *reinterpret_cast<const v8_t<E> *>(L[3] + lds_at(row, k));
        mac3<E>(cr, eh, el, b[2][0], b[2][1]);
        mac3<E>(ci, eh, el, b[0][0], b[0][1]);
    }
}

template <typename E>
__device__ __forceinline__ void loadT(const v8_t<E> *T, int lane, v8_t<E> (&b)[3][2]) {
#pragma unroll
    for (int p = 0; p < 3; p++) {
        b[p][0] = T[(2 * p) * 64 + lane];
        b[p][1] = T[(2 * p + 1) * 64 + lane];
    }
}

// hi = E(a) (round to nearest even), lo = E(a - hi) for a pair, packed two
// per dword: one packed convert, two unpacks, two subtractions, one packed
// convert
template <typename E>
__device__ __forceinline__ void split2(float a0, float a1, unsigned &h, unsigned &l) {
    if constexpr (is_f16<E>) {
        const f16x2 hb = __builtin_convertvector((float2v){a0, a1}, f16x2);
        h = __builtin_bit_cast(unsigned, hb);
        const f16x2 lb = __builtin_convertvector((float2v){a0 - (float)hb[0], a1 - (float)hb[1]}, f16x2);
        l = __builtin_bit_cast(unsigned, lb);
    } else {
        const bf16x2 hb = __builtin_convertvector((float2v){a0, a1}, bf16x2);
        h = __builtin_bit_cast(unsigned, hb);
        const float h0 = __uint_as_float(h << 16), h1 = __uint_as_float(h & 0xffff0000u);
        const bf16x2 lb = __builtin_convertvector((float2v){a0 - h0, a1 - h1}, bf16x2);
        l = __builtin_bit_cast(unsigned, lb);
    }
}

// split 4 values into hi/lo and store them at plane row, columns k..k+3
template <typename E>
__device__ __forceinline__ void st4(E *ph, E *pl, int row, int k, float a0, float a1, float a2, float a3) {
    uint2 h, l;
    split2<E>(a0, a1, h.x, l.x);
    split2<E>(a2, a3, h.y, l.y);
    *reinterpret_cast<uint2 *>(ph + lds_at(row, k)) = h;
    *reinterpret_cast<uint2 *>(pl + lds_at(row, k)) = l;
}

// amp * exp(+i 2 pi x) for x in revolutions: the phase is reduced to
// [-1/2, 1/2] exactly, then the hardware sine and cosine (argument in
// revolutions) are used
__device__ __forceinline__ void cis_rev(float x, float am, float &re, float &im) {
    const float r = x - rintf(x);
    re = am * __builtin_amdgcn_cosf(r);
    im = am * __builtin_amdgcn_sinf(r);
}

// Outputs and pixel processing of one subaperture (shwfs_grad,
// maos/utils.c:983-1122) by one wave: lane p < npix holds pixel p's
// intensity pv (npix <= 64).  Noise, the intensities, the FP64 i0 and dither
// sums, then the matched-filter or CoG gradient (raw moments for sigmatch 2,
// finished by k_sigmatch2).
// Inputs of the pixel epilogue that do not depend on the image, loaded at
// the start of a subaperture so that their latency is not on the tail of
// the workgroup: the pixel's matched-filter row and the subaperture's area
// and model flux.
struct PixPre {
    float m0, m1, saa, i0s;
};
__device__ __forceinline__ PixPre pix_pre(const MfmaArgs &a, int iwfs, int isa, int p) {
    PixPre q{0.f, 0.f, a.saa[isa], a.i0sum ? a.i0sum[(long)iwfs * a.nsa + isa] : 0.f};
    if (a.phytype == MX_PTYPE_MF && p < a.npix) {
        const float2 m = *reinterpret_cast<const float2 *>(a.mtche + (((long)iwfs * a.nsa + isa) * a.npix + p) * 2);
        q.m0 = m.x;
        q.m1 = m.y;
    }
    return q;
}
__device__ __forceinline__ void shwfs_pixels(const MfmaArgs &a, int iwfs, int isa, int lane, float pv,
                                             const PixPre &q) {
    const long ibase = ((long)iwfs * a.nsa + isa) * a.npix;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    const int p = lane;
    if (!(a.dt.phase & MX_DT_OUT)) { // uniform: a frame inside a longer exposure
        if (p < a.npix) dt_frame(a.dt, ibase + p, pv);
        return;
    }
    if (p < a.npix) {
        dt_frame(a.dt, ibase + p, pv);
        if (a.noise.on) pv = add_noise(a.noise, pv, p, iwfs, a.nsa, isa);
        a.ints[ibase + p] = pv;
        if (a.i0acc) a.i0acc[ibase + p] += (double)pv;
        dither_add(a.dith, iwfs, ibase + p, pv);
        if (a.phytype == MX_PTYPE_MF) {
            s0 = q.m0 * pv;
            s1 = q.m1 * pv;
            s2 = pv;
        } else {
            const float iI = pv - a.cogoff;
            if (iI > a.cogthres) {
                s0 = iI * (float)(p % a.pixpsax);
                s1 = iI * (float)(p / a.pixpsax);
                s2 = iI;
            }
        }
    }
    s0 = wave_sum64(s0);
    s1 = wave_sum64(s1);
    s2 = wave_sum64(s2);
    if (lane == 0 && !(q.saa < a.saat)) {
        float gx = 0.f, gy = 0.f;
        if (a.sigmatch == 2) { // raw moments; k_sigmatch2 forms the gradients
            gx = s0;
            gy = s1;
        } else if (a.phytype == MX_PTYPE_MF) {
            float scale = 1.f;
            if (a.sigmatch == 1) {
                scale = q.i0s / s2;
                if (scale <= 0.f || scale > 10.f || !isfinite(scale)) scale = 1.f;
            }
            gx = s0 * scale;
            gy = s1 * scale;
        } else {
            float sum = s2;
            if (a.sigmatch == 0)
                sum = a.i0sum ? q.i0s : a.sigtot * q.saa;
            if (sum > a.cogthres) {
                gx = (s0 / sum - (float)(a.pixpsax - 1) * 0.5f) * a.pixthetax;
                gy = (s1 / sum - (float)(a.pixpsay - 1) * 0.5f) * a.pixthetay;
            }
        }
        a.grad[(long)iwfs * 2 * a.nsa + isa] = gx;
        a.grad[(long)iwfs * 2 * a.nsa + a.nsa + isa] = gy;
    }
}

// waves per SIMD the register budget is sized for
#ifndef MX_MFMA_WAVES
#define MX_MFMA_WAVES 4
#endif
// UPLINK: the same stages through 2b on the LLT grid give the uplink OTF
// lotfc = FFT(|FFT(embed)|^2) / nlwvf^2 (maos/wfsints.c:149-199); one
// workgroup per (WFS, wavelength) writes conj(lotfc) and stops there.
// PP: the stages alternate between two halves of the planes, logical
// columns 16..47 (region A) and 0..15 + 48..63 (region B), each stage
// reading one and writing the other, so a stage's writes need no barrier
// after its reads (one barrier per stage instead of two).  Arrays wider
// than 32 columns fold their second half into rows 32..63.
template <bool UPLINK, typename E, bool PP = true, bool FUSED = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MX_MFMA_WAVES))) void k_shwfs_mfma(MfmaArgs a) {
    using V8 = v8_t<E>;
    // range scales of the FP16 form (powers of two; tables carry 2^13): the
    // field, the stage outputs 1a, 1b (|Y|^2), 2a, 2b, 3a, the image and the
    // uplink OTF
    constexpr bool H = is_f16<E>;
    constexpr float SE = H ? 0x1p14f : 1.f, F1 = H ? 0x1p-18f : 1.f, F2 = H ? 0x1p-51f : 1.f,
                    F3 = H ? 0x1p-15f : 1.f, F4 = H ? 0x1p-22f : 1.f, F5 = H ? 0x1p-19f : 1.f,
                    FI = H ? 0x1p11f : 1.f, FO = H ? 0x1p-4f : 1.f;
    __shared__ __attribute__((aligned(16))) E Lb[4 * MPLANE];
    __shared__ float img[MX_MFMA_NS * MX_MFMA_NS];
    __shared__ float red[4];
    E *const L[4] = {Lb, Lb + MPLANE, Lb + 2 * MPLANE, Lb + 3 * MPLANE};
    const E *const Lc[4] = {Lb, Lb + MPLANE, Lb + 2 * MPLANE, Lb + 3 * MPLANE};
    // logical column of region-A / region-B column c < 32
    auto colA = [](int c) { return 16 + c; };
    auto colB = [](int c) { return c < 16 ? c : c + 32; };
    const V8 *const F = static_cast<const V8 *>(a.F);
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    int iwfs, isa, iwvl0, iwvl1;
    if constexpr (UPLINK) {
        iwfs = (int)blockIdx.x / a.nwvl;
        isa = 0;
        iwvl0 = (int)blockIdx.x % a.nwvl;
        iwvl1 = iwvl0 + 1;
    } else {
        const int unit = xcd_unit(blockIdx.x, a.nsa * a.nwfs);
        iwfs = unit % a.nwfs;
        isa = unit / a.nwfs;
        iwvl0 = 0;
        iwvl1 = a.nwvl;
    }
    const DevWfs &W = a.wfs[iwfs];
    constexpr int NX = 32, B0 = 16; // nonzero block [16, 48)
    // role of the wave; it rotates with the block so that the SIMDs of a CU
    // share the unequal parts (stage 3b) evenly
    const int lw = (wave + (int)blockIdx.x) & 3;
    const int mb = lw % 2, nb = lw / 2;
    const int h32 = lane / 32, l32 = lane % 32, h16 = lane / 16, l16 = lane % 16;
    // OPD and amplitude: thread (ix = tid%32, iy = 4 (tid/32) + j)
    const int ix = tid % 32, iy0 = 4 * (tid / 32);
    const float *og = UPLINK ? a.llt_opd + (long)iwfs * NX * NX : a.opd + ((long)iwfs * a.nsa + isa) * NX * NX;
    const float *ag = UPLINK ? a.llt_amp : a.amp + (long)isa * NX * NX;
    __shared__ float opdl[FUSED ? NX * NX : 1];
    if constexpr (FUSED) {
        // the grid's OPD (k_wfs_opd's ray trace, maos/wfsgrad.c:175-194 and
        // the LLT's :329-386): wave w traces the 16 x 16 quadrant (w%2, w/2)
        // into opdl with its scratch in the planes, which the embed
        // overwrites only after the barrier below
        const OpdArgs &t = a.tr;
        char *scr = reinterpret_cast<char *>(Lb) + wave * MX_FUSE_SCR;
        int *cidx = reinterpret_cast<int *>(scr);
        float *cfrac = reinterpret_cast<float *>(scr + 256);
        double *ctab = reinterpret_cast<double *>(scr + 512);
        float *tile = reinterpret_cast<float *>(scr + 1024);
        const int qx = wave & 1, qy = wave >> 1;
        const double o16 = t.ptsdx * 16.;
        if constexpr (UPLINK)
            ray_trace<64, 16, false>(t.layers + W.ireal * MX_MAXLAYER, t.nps, W.hs, W.thetax - t.llt_ox / W.hs,
                                     W.thetay - t.llt_oy / W.hs, 0., 0., t.llt_ox, t.llt_oy, t.isim, t.dt,
                                     t.atmscale[W.ireal], t.ptsdx, t.llt_origx + o16 * qx, t.llt_origy + o16 * qy,
                                     lane, opdl + 16 * NX * qy + 16 * qx, cidx, cfrac, tile, ctab, t.tworow, NX);
        else
            ray_trace<64, 16, false>(t.layers + W.ireal * MX_MAXLAYER, t.nps, W.hs, W.thetax, W.thetay, W.misregx,
                                     W.misregy, 0., 0., t.isim, t.dt, t.atmscale[W.ireal], t.ptsdx,
                                     t.ptsx[isa] + o16 * qx, t.ptsy[isa] + o16 * qy, lane,
                                     opdl + 16 * NX * qy + 16 * qx, cidx, cfrac, tile, ctab, t.tworow, NX);
        __syncthreads();
        float *go = UPLINK ? a.llt_opd_out + (long)iwfs * NX * NX : a.opd_out + ((long)iwfs * a.nsa + isa) * NX * NX;
#pragma unroll
        for (int j = 0; j < 4; j++) go[tid + 256 * j] = opdl[tid + 256 * j];
        og = opdl;
    }
    float mean;
    {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) s += og[(iy0 + j) * NX + ix];
        s = wave_sum64(s);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        mean = (red[0] + red[1] + red[2] + red[3]) * (1.f / (NX * NX));
    }
    float pv = 0.f; // pixel tid (< npix <= 64, wave 0)
    PixPre pq{};
    if (!UPLINK && wave == 0) pq = pix_pre(a, iwfs, isa, lane);
    for (int iwvl = iwvl0; iwvl < iwvl1; iwvl++) {
        const MfmaWvl &WV = a.wv[iwvl];
        // embed amp*exp(+i 2 pi opd/wvl) (cembed_wvf) as W[x][y], y contiguous;
        // OPD and amplitude are read again per wavelength (cache hits) rather
        // than held in registers
        {
            float re[4], im[4];
#pragma unroll
            for (int j = 0; j < 4; j++)
                cis_rev((og[(iy0 + j) * NX + ix] - mean) * WV.inv_wvl, ag[(iy0 + j) * NX + ix] * SE, re[j], im[j]);
            st4(L[0], L[1], B0 + ix, B0 + iy0, re[0], re[1], re[2], re[3]);
            st4(L[2], L[3], B0 + ix, B0 + iy0, im[0], im[1], im[2], im[3]);
        }
        // B operand of stages 1a and 1b: F over the band [16, 48) at the 16
        // columns 16 lw + l%16, fetched after the embed (which then waits for
        // the OPD and amplitude only; the scheduling barrier keeps the loads
        // from being hoisted above it) and in flight over the barrier
        __builtin_amdgcn_sched_barrier(0);
        V8 b1[3][2];
        loadT<E>(F + (MX_F1A + lw) * MX_TPARTS * 64, lane, b1);
        __syncthreads();
        // 1a: C[x][ky] = sum_{y in band} W[x][y] F[y][ky]; wave: ky in [16 lw, 16 lw + 16),
        // both 16-row blocks of x; stored as Z row ky, columns x
        {
            f32x4 zr[2] = {}, zi[2] = {};
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
                cmacd<E, false>(zr[mt], zi[mt], Lc, B0 + 16 * mt + l16, B0 + 8 * h16, b1);
            if (!PP) __syncthreads();
            else __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                const int row = 16 * lw + l16, col = PP ? colB(16 * mt + 4 * h16) : B0 + 16 * mt + 4 * h16;
                st4(L[0], L[1], row, col, zr[mt][0] * F1, zr[mt][1] * F1, zr[mt][2] * F1, zr[mt][3] * F1);
                st4(L[2], L[3], row, col, zi[mt][0] * F1, zi[mt][1] * F1, zi[mt][2] * F1, zi[mt][3] * F1);
            }
        }
        __syncthreads();
        // 1b: C[ky][kx] = sum_{x in band} Z[ky][x] F[x][kx]; P = |C|^2.  Wave:
        // kx in [16 lw, 16 lw + 16) (the same B operand as 1a, F being
        // symmetric), the four 16-row blocks t of ky; blocks t and t + 2 hold ky
        // and ky + 32 in the same lane: P+ = P[ky] + P[ky+32] and P- = P[ky] -
        // P[ky+32] (ky < 32) go to row kx, columns ky and 32 + ky
        V8 b2[3][2];
        {
            f32x4 yr[4] = {}, yi[4] = {};
#pragma unroll
            for (int t = 0; t < 4; t++)
                cmacd<E, false>(yr[t], yi[t], Lc, 16 * t + l16, PP ? colB(8 * h16) : B0 + 8 * h16, b1);
            loadT<E>(F + (MX_F2 + 2 * nb) * MX_TPARTS * 64, lane, b2); // stage 2a's operand
            if (!PP) __syncthreads();
            else __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int v = 0; v < 4; v++) yr[t][v] = (yr[t][v] * yr[t][v] + yi[t][v] * yi[t][v]) * F2;
            const int kx = 16 * lw + l16;
#pragma unroll
            for (int t = 0; t < 2; t++) {
                // PP: P+ in planes 0, 1 and P- in planes 2, 3 of region A
                const int col = 16 * t + 4 * h16;
                st4(L[0], L[1], kx, PP ? colA(col) : col, yr[t][0] + yr[t + 2][0], yr[t][1] + yr[t + 2][1],
                    yr[t][2] + yr[t + 2][2], yr[t][3] + yr[t + 2][3]);
                st4(L[PP ? 2 : 0], L[PP ? 3 : 1], kx, PP ? colA(col) : 32 + col, yr[t][0] - yr[t + 2][0],
                    yr[t][1] - yr[t + 2][1], yr[t][2] - yr[t + 2][2], yr[t][3] - yr[t + 2][3]);
            }
        }
        __syncthreads();
        // 2a: Q[uy][kx] = sum_{k<32} P+-[kx][k] F[k][uy] for uy < 32 of parity e
        // (P+ for even uy, P- for odd); wave: e = lw/2, kx in [16 mb, 16 mb + 16)
        // and the same + 32.  Q+ = Q[kx] + Q[kx+32], Q- = Q[kx] - Q[kx+32]
        // (kx < 32) go to row uy, columns kx and 32 + kx
        {
            const int e = nb;
            f32x4 qr[2] = {}, qi[2] = {};
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
                cmacd<E, true>(qr[mt], qi[mt], Lc, 16 * mb + 32 * mt + l16, PP ? colA(8 * h16) : 32 * e + 8 * h16,
                               b2, PP ? 2 * e * MPLANE : 0);
            loadT<E>(F + (MX_F2 + 2 * nb + mb) * MX_TPARTS * 64, lane, b1); // stage 2b's operand
            if (!PP) __syncthreads();
            else __builtin_amdgcn_sched_barrier(0);
            const int uy = 2 * l16 + e, col = PP ? colB(16 * mb + 4 * h16) : 16 * mb + 4 * h16;
            // PP: Q+ (columns < 32) in rows uy, Q- in rows 32 + uy, region B
            const int rq = PP ? uy + 32 : uy, cq = PP ? col : 32 + col;
            if constexpr (H)
#pragma unroll
                for (int mt = 0; mt < 2; mt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        qr[mt][v] *= F3;
                        qi[mt][v] *= F3;
                    }
            st4(L[0], L[1], uy, col, qr[0][0] + qr[1][0], qr[0][1] + qr[1][1], qr[0][2] + qr[1][2],
                qr[0][3] + qr[1][3]);
            st4(L[2], L[3], uy, col, qi[0][0] + qi[1][0], qi[0][1] + qi[1][1], qi[0][2] + qi[1][2],
                qi[0][3] + qi[1][3]);
            st4(L[0], L[1], rq, cq, qr[0][0] - qr[1][0], qr[0][1] - qr[1][1], qr[0][2] - qr[1][2],
                qr[0][3] - qr[1][3]);
            st4(L[2], L[3], rq, cq, qi[0][0] - qi[1][0], qi[0][1] - qi[1][1], qi[0][2] - qi[1][2],
                qi[0][3] - qi[1][3]);
        }
        __syncthreads();
        // 2b: O[uy][ux] = sum_{k<32} Q+-[uy][k] F[k][ux], uy < 32, for ux of parity
        // e (Q+ for even ux, Q- for odd); wave: e = lw/2, ux = 2 (16 hf + l%16) + e
        // with hf = lw%2, both 16-row blocks of uy.  X = O conj(lotf) ETF goes
        // to row ux, columns uy
        {
            const int e = nb, hf = mb;
            f32x4 xr[2] = {}, xi[2] = {};
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
                cmacd<E, false>(xr[mt], xi[mt], Lc, PP ? 16 * mt + l16 + 32 * e : 16 * mt + l16,
                                PP ? colB(8 * h16) : 32 * e + 8 * h16, b1);
            // element (mt, v) of this lane is table slot ((2 lw + mt) 4 + v) 64 + lane
            const long tbase = (long)(2 * lw) * 4 * 64 + lane;
            const int ux = 2 * (16 * hf + l16) + e;
            if constexpr (UPLINK) {
                // conj(lotfc): the half plane in slot order, and the full plane
                // [ux][uy] with uy > 32 from conj(lotfc)(-u) = lotfc(u) and the
                // lag-32 row zero (the LLT field fills 32 of 64 samples)
                float2 *hp = a.lotf_half + ((long)iwfs * a.nwvl + iwvl) * MX_HALF + tbase;
                float2 *fp = a.lotf_full + ((long)iwfs * a.nwvl + iwvl) * 64 * 64;
#pragma unroll
                for (int mt = 0; mt < 2; mt++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int uy = 16 * mt + 4 * h16 + v;
                        const float re = xr[mt][v] * (a.llt_norm * FO), im = xi[mt][v] * (a.llt_norm * FO);
                        hp[(mt * 4 + v) * 64] = make_float2(re, -im);
                        fp[ux * 64 + uy] = make_float2(re, -im);
                        if (uy) fp[((64 - ux) & 63) * 64 + (64 - uy)] = make_float2(re, im);
                    }
                if (h16 == 0) fp[ux * 64 + 32] = make_float2(0.f, 0.f);
                continue;
            }
            // transfer functions: conj(lotf) and the ETF (LGS), or the nominal
            // (NGS); one uniform branch, every load issued before the first use
            if (a.lotf_t && WV.etf_t) {
                const float2 *lt = a.lotf_t + ((long)iwfs * a.nwvl + iwvl) * MX_HALF + tbase;
                const float2 *et = WV.etf_t + (long)isa * MX_HALF + tbase;
#pragma unroll
                for (int mt = 0; mt < 2; mt++) {
                    float2 tl[4], te[4];
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        tl[v] = lt[(mt * 4 + v) * 64];
                        te[v] = et[(mt * 4 + v) * 64];
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const float2 t = make_float2(tl[v].x * te[v].x - tl[v].y * te[v].y,
                                                     tl[v].x * te[v].y + tl[v].y * te[v].x);
                        const float re = xr[mt][v], im = xi[mt][v];
                        xr[mt][v] = re * t.x - im * t.y;
                        xi[mt][v] = re * t.y + im * t.x;
                    }
                }
            } else if (WV.nominal_t) {
                const float *nm = WV.nominal_t + (long)isa * WV.nom_sa + tbase;
                float tn[8];
#pragma unroll
                for (int i = 0; i < 8; i++) tn[i] = nm[i * 64];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    xr[i / 4][i % 4] *= tn[i];
                    xi[i / 4][i % 4] *= tn[i];
                }
            }
            loadT<E>(static_cast<const V8 *>(WV.gy), lane, b2); // stage 3a's operand
            if (!PP) __syncthreads();
            else __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                const int col = PP ? colA(16 * mt + 4 * h16) : 16 * mt + 4 * h16;
                st4(L[0], L[1], ux, col, xr[mt][0] * F4, xr[mt][1] * F4, xr[mt][2] * F4, xr[mt][3] * F4);
                st4(L[2], L[3], ux, col, xi[mt][0] * F4, xi[mt][1] * F4, xi[mt][2] * F4, xi[mt][3] * F4);
            }
        }
        __syncthreads();
        // 3a: R[ys][ux] = sum_{uy < 32} X[ux][uy] w(uy) Gy[uy][ys], w = 1 at uy = 0
        // and 2 elsewhere (folded into the table); role r: ux in [16r, 16r+16)
        {
            f32x4 rr = {}, ri = {};
            cmacd<E, false>(rr, ri, Lc, 16 * lw + l16, PP ? colA(8 * h16) : 8 * h16, b2);
            if (lw == 3) loadT<E>(static_cast<const V8 *>(WV.gx), lane, b1); // stage 3b's operand, first K step
            if (!PP) __syncthreads();
            else __builtin_amdgcn_sched_barrier(0);
            // element v: ux = 16 lw + 4 (l/16) + v, ys = l%16 -> row ys, columns ux
            // (PP: ux >= 32 in rows 32 + ys, region B)
            const int rr3 = PP ? l16 + 32 * (lw >> 1) : l16;
            const int cr3 = PP ? colB(16 * (lw & 1) + 4 * h16) : 16 * lw + 4 * h16;
            st4(L[0], L[1], rr3, cr3, rr[0] * F5, rr[1] * F5, rr[2] * F5, rr[3] * F5);
            st4(L[2], L[3], rr3, cr3, ri[0] * F5, ri[1] * F5, ri[2] * F5, ri[3] * F5);
        }
        __syncthreads();
        // 3b: I[ys][xs] = Re sum_ux R[ys][ux] Gx[ux][xs]  (role 3)
        if (lw == 3) {
            f32x4 ir = {};
            loadT<E>(static_cast<const V8 *>(WV.gx) + MX_TPARTS * 64, lane, b2);
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const V8(&g)[3][2] = ks ? b2 : b1;
                const int k = PP ? colB(8 * h16) : 32 * ks + 8 * h16, r3 = PP ? l16 + 32 * ks : l16;
                const V8 dh = *reinterpret_cast<const V8 *>(Lc[0] + lds_at(r3, k));
                const V8 dl = *reinterpret_cast<const V8 *>(Lc[1] + lds_at(r3, k));
                const V8 eh = *reinterpret_cast<const V8 *>(Lc[2] + lds_at(r3, k));
                const V8 el = *reinterpret_cast<const V8 *>(Lc[3] + lds_at(r3, k));
                mac3<E>(ir, dh, dl, g[0][0], g[0][1]); // Re R * Re G
                mac3<E>(ir, eh, el, g[2][0], g[2][1]); // Im R * (-Im G)
            }
#pragma unroll
            for (int v = 0; v < 4; v++) img[(4 * (lane / 16) + v) * MX_MFMA_NS + lane % 16] = ir[v] * FI;
        }
        __syncthreads();
        if (tid < a.npix) {
            const int4 id = WV.si_c[isa * WV.si_sa + tid];
            const float4 w = WV.si_w[isa * WV.si_sa + tid];
            pv += W.alpha[iwvl] * (w.x * img[id.x] + w.y * img[id.y] + w.z * img[id.z] + w.w * img[id.w]);
        }
        // (PP: the next wavelength writes img only after further barriers)
        if (!PP) __syncthreads();
    }
    if constexpr (UPLINK) return;
    // outputs and pixel processing, wave 0
    if (wave == 0) shwfs_pixels(a, iwfs, isa, lane, pv, pq);
}

#include "shwfs_fft.hip"

// psfout of wfsints (maos/wfsints.c:243-253; GPU form cuda/sim/wfsints.cu:
// 443-450): the complex PSF scaled by norm_psf / nwvf^2 and transformed
// back, centre nout^2 (nout = notf/2 + 2).  With nwvf == notf the two
// transforms cancel and leave norm_psf amp exp(i 2 pi opd / wvl) at the
// embedded position, zero around it, which is what this writes: one block
// per (subaperture, wavelength) of one WFS.
struct PsfoutArgs {
    int nsa, nxsa, nout, b; // b: output index of the field's first sample
    float norm;             // norm_psf
    float inv_wvl[MX_MAXWVL];
    const float *opd, *amp; // this WFS's [nsa][nxsa^2], [nsa][nxsa^2]
    float2 *out;            // [nwvl][nsa][nout^2]
};
__global__ __launch_bounds__(256) void k_psfout(PsfoutArgs a) {
    const int isa = (int)blockIdx.x, iwvl = (int)blockIdx.y, n2 = a.nout * a.nout;
    float2 *o = a.out + ((long)iwvl * a.nsa + isa) * n2;
    for (int e = (int)threadIdx.x; e < n2; e += 256) {
        const int ix = e % a.nout - a.b, iy = e / a.nout - a.b;
        float2 v = make_float2(0.f, 0.f);
        if (ix >= 0 && ix < a.nxsa && iy >= 0 && iy < a.nxsa) {
            const long pt = (long)isa * a.nxsa * a.nxsa + (long)iy * a.nxsa + ix;
            cis_rev(a.opd[pt] * a.inv_wvl[iwvl], a.amp[pt] * a.norm, v.x, v.y);
        }
        o[e] = v;
    }
}

// The uplink OTF of an LLT grid wider than the subaperture grid
// (maos/wfsints.c:160-185: nlwvf > nwvf, the OTF cropped by ccpcorner to
// nwvf = notf).  FFT(|FFT(E)|^2 / nlwvf^2) is the autocorrelation
// lotfc(u) = sum_x E(x) conj(E(x + u)) of the LLT field E (its support,
// nlx <= nlwvf / 2, leaves no wrap-around), and the crop keeps the lags
// -notf/2 <= u < notf/2, so each is summed directly.  One block per (WFS,
// wavelength, row uy); conj(lotfc) goes to the full plane [ux][uy] and, for
// the half-plane forms, to the half plane in half_slot order.
struct LltAcorrArgs {
    int nlx, side, nwvl, N; // N = notf = nwvf <= 64, nlx <= 64
    float inv_wvl[MX_MAXWVL];
    const float *opd;       // [nwfs][side^2], the LLT points in the first nlx rows and columns
    const float *amp;       // [nlx^2]
    float2 *lotf_full;      // [nwfs][nwvl][N*N]
    float2 *lotf_half;      // [nwfs][nwvl][MX_HALF] or null
};
__global__ __launch_bounds__(256) void k_llt_acorr(LltAcorrArgs a) {
    __shared__ float2 E[64 * 64];
    __shared__ float2 part[256];
    __shared__ float red[4];
    const int N = a.N, nlx = a.nlx, tid = (int)threadIdx.x;
    const int uy = (int)blockIdx.x % N, r = (int)blockIdx.x / N, iwvl = r % a.nwvl, iwfs = r / a.nwvl;
    const float *op = a.opd + (long)iwfs * a.side * a.side;
    // the field, piston removed (it cancels in E(x) conj(E(x + u)))
    float s = 0.f;
    for (int e = tid; e < nlx * nlx; e += 256) s += op[(e / nlx) * a.side + e % nlx];
    seg_sum<64>(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    const float mean = (red[0] + red[1] + red[2] + red[3]) / (float)(nlx * nlx);
    for (int e = tid; e < nlx * nlx; e += 256) {
        float re, im;
        cis_rev((op[(e / nlx) * a.side + e % nlx] - mean) * a.inv_wvl[iwvl], a.amp[e], re, im);
        E[e] = make_float2(re, im);
    }
    __syncthreads();
    const int np = 256 / N, ux = tid % N, pp = tid / N;
    const int sx = ux < N / 2 ? ux : ux - N, sy = uy < N / 2 ? uy : uy - N;
    const int x0 = max(0, -sx), x1 = min(nlx, nlx - sx), y0 = max(0, -sy), y1 = min(nlx, nlx - sy);
    float2 acc = make_float2(0.f, 0.f);
    for (int y = y0 + pp; y < y1; y += np) {
        const float2 *e0 = E + y * nlx, *e1 = E + (y + sy) * nlx + sx;
        for (int x = x0; x < x1; x++) {
            const float2 p = e0[x], q = e1[x]; // p conj(q)
            acc.x = fmaf(p.x, q.x, fmaf(p.y, q.y, acc.x));
            acc.y = fmaf(p.y, q.x, fmaf(-p.x, q.y, acc.y));
        }
    }
    part[tid] = acc;
    __syncthreads();
    if (tid < N) {
        float2 t = make_float2(0.f, 0.f);
        for (int k = 0; k < np; k++) t = cadd(t, part[k * N + tid]);
        const long w = (long)iwfs * a.nwvl + iwvl;
        a.lotf_full[w * N * N + (long)tid * N + uy] = make_float2(t.x, -t.y);
        if (a.lotf_half && uy < N / 2) a.lotf_half[w * MX_HALF + half_slot(tid, uy)] = make_float2(t.x, -t.y);
    }
}

// ------------------------------------------------------------ large fields
// Fields larger than the one-workgroup subaperture forms hold: nwvf and notf
// up to 1024 with prime factors 2, 3, 5 and 7 (the 2x2 and 1x1 NGS of 15 and
// 30 m, a 128^2 padding of an LGS subaperture).  Each (WFS, subaperture,
// wavelength) chain c = (iwfs nsa + isa) nwvl + iwvl runs five batched
// line-transform passes; a block transforms L lines held in LDS (rows, or L
// adjacent columns read and written as L-wide row segments):
//   A  embed amp exp(+i 2 pi opd/wvl) (cembed_wvf, math/matcomp.c:56-82) on
//      the nxsa rows of the field band, FFT along x  -> B1 [nxsa][nwvf]
//   B  FFT the columns of B1 (rows outside the band are zero), |E|^2
//                                                     -> PSF [nwvf][nwvf]
//   C  FFT the PSF rows padded to notf by the corner copy (ccpcorner,
//      math/matcomp.c:176-216: x < nwvf/2 stays, the rest moves up by
//      notf - nwvf, the middle is zero), two real rows per complex FFT,
//      columns <= notf/2 kept                       -> B2 [nwvf][notf/2+1]
//   D  FFT those columns of B2 (rows padded the same way), x nominal DTF,
//      inverse FFT, keep the rows the detector samples
//                                                     -> B3 [ny][notf/2+1]
//   E  inverse FFT of those rows (the other columns are the conjugate
//      mirror), keep the sampled columns (real part) -> image [ny][nx]
// then k_lf_sample applies si (maos/wfsints.c:286) and the pixel processing.
constexpr int MX_LF_MAXN = 1024;
constexpr int MX_LF_MAXRAD = 12;
constexpr int MX_LF_THREADS = 512; // most threads of a block (128, 256 or 512 run)
constexpr int MX_LF_LDS = 4352;    // complex elements of one LDS line buffer
constexpr int MX_LF_MAXS = 128;    // most sampled rows / columns kept (else the full image)

// LDS position of element e of a line: one pad element after every 32, so
// that the power-of-two write strides of the early Stockham stages spread
// over the LDS banks
__host__ __device__ constexpr int lf_pos(int e) { return e + (e >> 5); }
// LDS stride of lines of length n: the padded length, made odd
__host__ __device__ constexpr int lf_stride(int n) { return lf_pos(n) | 1; }
// lines of length n per block: a power of two <= 32 that one buffer holds;
// for lines of 768 and more at most 2048 / n of them, so that a one-chain
// plane (the 960^2 and 1024^2 tip-tilt sizes: 2 lines, 480-512 blocks)
// spreads over more blocks than CUs (53 -> 46 us per step for the 1024^2
// tip-tilt imaging; the 512^2 planes, with many chains, keep 8 lines)
__host__ __device__ constexpr int lf_lines(int n) {
    int L = 32;
    while (L > 1 && (L * lf_stride(n) > MX_LF_LDS || (n >= 768 && L * n > 2048))) L /= 2;
    return L;
}
// radix of stage s in the plan of n (radix 8 first, then 4, 2, 3, 5, 7); 0
// past the last stage or if n has another prime factor
__host__ __device__ constexpr int lf_radix(int n, int s) {
    const int rad[6] = {8, 4, 2, 3, 5, 7};
    int r = n, c = 0;
    for (int k = 0; k < 6; k++)
        while (r > 1 && r % rad[k] == 0) {
            if (c == s) return rad[k];
            c++;
            r /= rad[k];
        }
    return 0;
}

struct LfPlan {
    int n, nrad;
    int rad[MX_LF_MAXRAD]; // lf_radix(n, s)
    int off[MX_LF_MAXRAD]; // twiddle offset of each stage
    const float2 *tw;      // stage s: exp(-2 pi i r k / (Ns R)) at off[s] + k (R-1) + r-1
};
struct LfWvl {
    float inv_wvl;
    const float *nominal; // [ux][uy] (transposed), or null
    const float2 *etf, *etf2; // [nsa][ux][uy] ETF with the nominal fused (LLT), second profile or null
    int ny, nx;           // sampled image rows / columns
    const int *rows, *cols;
    const int4 *si_c;     // si in compact indices iy * nx + ix
    const float4 *si_w;
    int si_sa, nom_sa;    // per-subaperture strides of si_c / si_w and nominal (0: shared)
};
struct LfArgs {
    LfPlan pn, pm;        // lengths nwvf and notf
    int nwvf, notf, nxsa, nsa, nwfs, nwvl, nchain;
    int L, lgL;           // lines per block in this pass (lf_lines of its length)
    int maxny, maxnx;     // B3 rows and image capacity per chain
    const float *opd, *amp;
    const LfWvl *wv;
    float2 *b1, *b2, *b3;
    float *psf, *img;
    float psf_scale;      // pass C multiplies the PSF by this (1/nlwvf^2 for the uplink)
    float2 *lotf;         // [nwfs][nwvl][ux][uy] conj(lotfc): written by the uplink, read by imaging; or null
    float etf1wt, etf2wt; // ETF profile weights (maos/wfsints.c:215-223)
};

// a / b for 0 <= a < 2^22 from a rounded reciprocal rb of b, corrected
__device__ __forceinline__ int qdiv(int a, int b, float rb) {
    int q = (int)((float)a * rb);
    const int r = a - q * b;
    q += r >= b ? 1 : (r < 0 ? -1 : 0);
    return q;
}
// a / b: by the compiler when b is a constant (CT), else by qdiv
template <bool CT>
__device__ __forceinline__ int ldiv(int a, int b, float rb) {
    if constexpr (CT) return a / b;
    else return qdiv(a, b, rb);
}

__device__ __forceinline__ float2 mul_ti(float2 z, float t) { return make_float2(-t * z.y, t * z.x); }
// v[q] <- sum_r v[r] exp(t 2 pi i r q / R), t = +-1
template <int R>
__device__ __forceinline__ void small_dft(float2 (&v)[R], float t) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = mul_ti(csub(v[1], v[3]), t);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 8) {
        // even and odd halves by 4-point transforms, then w8^q = exp(t 2 pi i q / 8)
        // on the odd half: w8 = (1 + t i) / sqrt 2, w8^2 = t i, w8^3 = (-1 + t i) / sqrt 2
        const float h = 0.70710678118654752f;
        float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        small_dft<4>(e, t);
        small_dft<4>(o, t);
        const float2 o1 = mul_ti(o[1], t), o3 = mul_ti(o[3], t);
        o[1] = make_float2(h * (o[1].x + o1.x), h * (o[1].y + o1.y));
        o[2] = mul_ti(o[2], t);
        o[3] = make_float2(h * (o3.x - o[3].x), h * (o3.y - o[3].y));
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = cadd(e[q], o[q]);
            v[q + 4] = csub(e[q], o[q]);
        }
    } else if constexpr (R == 3) {
        const float h = 0.86602540378443865f; // sqrt(3)/2
        const float2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
        const float2 r = mul_ti(make_float2(h * d.x, h * d.y), t);
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    } else if constexpr (R == 7) {
        // pairs r, 7 - r: v[k] = v0 + sum_r cos(2 pi r k / 7) (v[r] + v[7-r])
        //                      + t i sum_r sin(2 pi r k / 7) (v[r] - v[7-r])
        constexpr float c[4] = {1.f, 0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f};
        constexpr float sn[4] = {0.f, 0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f};
        float2 a[4], b[4];
#pragma unroll
        for (int r = 1; r <= 3; r++) {
            a[r] = cadd(v[r], v[7 - r]);
            b[r] = csub(v[r], v[7 - r]);
        }
        const float2 v0 = v[0];
#pragma unroll
        for (int k = 1; k <= 3; k++) {
            float2 p = v0, q = make_float2(0.f, 0.f);
#pragma unroll
            for (int r = 1; r <= 3; r++) {
                const int m = (r * k) % 7; // cos(2 pi m / 7), sin of it with the sign of m <= 3
                const float cm = c[m <= 3 ? m : 7 - m], sm = m <= 3 ? sn[m] : -sn[7 - m];
                p = make_float2(p.x + cm * a[r].x, p.y + cm * a[r].y);
                q = make_float2(q.x + sm * b[r].x, q.y + sm * b[r].y);
            }
            q = mul_ti(q, t);
            v[k] = cadd(p, q);
            v[7 - k] = csub(p, q);
        }
        v[0] = cadd(v0, cadd(a[1], cadd(a[2], a[3])));
    } else { // 5: cos and sin of 2 pi/5 and 4 pi/5
        const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
        const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
        const float2 a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
        const float2 b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
        const float2 p1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
        const float2 p2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
        const float2 q1 = mul_ti(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), t);
        const float2 q2 = mul_ti(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), t);
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
}

// One Stockham stage of radix R and span Ns over L lines (line l at l S,
// element e at lf_pos(e)) by a block of T threads: item (l, j), j < n/R,
// takes x[j + r n/R] times w^(r k) (w = exp(-+ 2 pi i / (Ns R)), k = j % Ns),
// transforms and writes x'[(j/Ns) Ns R + k + r Ns].  CT: n, Ns and L are
// compile-time constants.  Forward is FFTW_FORWARD (exp(-i ...)).
template <int R, bool INV, bool CT, int T>
__device__ __forceinline__ void lds_stage(const float2 *src, float2 *dst, int n, int Ns, int L, int S,
                                          const float2 *tw) {
    const int nb = n / R, total = L * nb;
    const float rnb = 1.f / (float)nb, rns = 1.f / (float)Ns;
    for (int it = (int)threadIdx.x; it < total; it += T) {
        const int l = ldiv<CT>(it, nb, rnb), j = it - l * nb;
        const int g = ldiv<CT>(j, Ns, rns), k = j - g * Ns;
        const float2 *x = src + l * S;
        const float2 *w = tw + k * (R - 1);
        float2 v[R];
        v[0] = x[lf_pos(j)];
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 wr = w[r - 1];
            if (INV) wr.y = -wr.y;
            v[r] = cmul(x[lf_pos(j + r * nb)], wr);
        }
        small_dft<R>(v, INV ? 1.f : -1.f);
        float2 *y = dst + l * S;
        const int d = g * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) y[lf_pos(d + r * Ns)] = v[r];
    }
}

// Stockham autosort FFT of L lines of length P.n in src; returns the buffer
// that holds the result (src or dst).  Stages from the plan at run time.
template <bool INV, int T>
__device__ float2 *lds_fft_rt(float2 *src, float2 *dst, const LfPlan &P, int L, int S) {
    int Ns = 1;
    for (int s = 0; s < P.nrad; s++) {
        const int R = P.rad[s];
        const float2 *tw = P.tw + P.off[s];
        if (R == 8) lds_stage<8, INV, false, T>(src, dst, P.n, Ns, L, S, tw);
        else if (R == 4) lds_stage<4, INV, false, T>(src, dst, P.n, Ns, L, S, tw);
        else if (R == 2) lds_stage<2, INV, false, T>(src, dst, P.n, Ns, L, S, tw);
        else if (R == 3) lds_stage<3, INV, false, T>(src, dst, P.n, Ns, L, S, tw);
        else if (R == 5) lds_stage<5, INV, false, T>(src, dst, P.n, Ns, L, S, tw);
        else lds_stage<7, INV, false, T>(src, dst, P.n, Ns, L, S, tw);
        __syncthreads();
        float2 *t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}
// the same for a compile-time length N, from stage SI of span NS on
template <bool INV, int T, int N, int SI, int NS>
__device__ __forceinline__ float2 *lds_fft_ct(float2 *src, float2 *dst, const float2 *tw) {
    constexpr int R = lf_radix(N, SI);
    if constexpr (R == 0) {
        return src;
    } else {
        constexpr int L = lf_lines(N), S = lf_stride(N);
        lds_stage<R, INV, true, T>(src, dst, N, NS, L, S, tw);
        __syncthreads();
        return lds_fft_ct<INV, T, N, SI + 1, NS * R>(dst, src, tw + NS * (R - 1));
    }
}
// FFT of the block's lines: compile-time plan for N > 0, else P's
template <bool INV, int T, int N>
__device__ __forceinline__ float2 *lf_fft(float2 *src, float2 *dst, const LfPlan &P, int L, int S) {
    if constexpr (N > 0) return lds_fft_ct<INV, T, N, 0, 1>(src, dst, P.tw);
    else return lds_fft_rt<INV, T>(src, dst, P, L, S);
}

// block -> (line group g of ngroup, chain c); consecutive groups of a chain
// share an XCD and its L2
__device__ __forceinline__ void lf_unit(const LfArgs &a, int ngroup, int &g, int &c) {
    const int u = xcd_unit((int)blockIdx.x, ngroup * a.nchain);
    g = u % ngroup;
    c = u / ngroup;
}
__device__ __forceinline__ void lf_chain(const LfArgs &a, int c, int &iwfs, int &isa, int &iwvl) {
    iwvl = c % a.nwvl;
    isa = (c / a.nwvl) % a.nsa;
    iwfs = c / (a.nwvl * a.nsa);
}
// padded position -> source index of the corner copy (-1 in the zero middle)
__device__ __forceinline__ int corner_src(int x2, int m, int n) {
    const int h = n / 2;
    return x2 < h ? x2 : (x2 >= m - h ? x2 - (m - n) : -1);
}

// Each pass is a template on the block size T and the length N of its lines
// (0: taken from the arguments at run time).
#define MX_LF_GEOM(LEN, ARGL)                                                   \
    constexpr int LC = N > 0 ? lf_lines(N) : 1, SC = N > 0 ? lf_stride(N) : 1; \
    const int n_ = N > 0 ? N : (LEN), L = N > 0 ? LC : (ARGL), S = N > 0 ? SC : lf_stride(n_);

// pass A: embed the field-band rows and FFT them along x
template <int T, int N>
__global__ __launch_bounds__(T) void k_lf_a(LfArgs a) {
    __shared__ float2 X[MX_LF_LDS], Y[MX_LF_LDS];
    MX_LF_GEOM(a.nwvf, a.L)
    const int n = n_, nxsa = a.nxsa, skip = (n - nxsa) / 2;
    int g, c, iwfs, isa, iwvl;
    lf_unit(a, (nxsa + L - 1) / L, g, c);
    lf_chain(a, c, iwfs, isa, iwvl);
    const long np = (long)nxsa * nxsa;
    const float *op = a.opd + ((long)iwfs * a.nsa + isa) * np;
    const float *am = a.amp + (long)isa * np;
    const float inv_wvl = a.wv[iwvl].inv_wvl;
    const float rn = 1.f / (float)n;
    for (int e = threadIdx.x; e < L * n; e += T) {
        const int l = ldiv<(N > 0)>(e, n, rn), x = e - l * n, iy = g * L + l, ix = x - skip;
        float2 v = make_float2(0.f, 0.f);
        if (iy < nxsa && ix >= 0 && ix < nxsa) {
            float sn, cs;
            sincospif(2.f * op[(long)iy * nxsa + ix] * inv_wvl, &sn, &cs);
            const float A = am[(long)iy * nxsa + ix];
            v = make_float2(A * cs, A * sn);
        }
        X[l * S + lf_pos(x)] = v;
    }
    __syncthreads();
    const float2 *r = lf_fft<false, T, N>(X, Y, a.pn, L, S);
    float2 *out = a.b1 + (long)c * nxsa * n;
    for (int e = threadIdx.x; e < L * n; e += T) {
        const int l = ldiv<(N > 0)>(e, n, rn), x = e - l * n, iy = g * L + l;
        if (iy < nxsa) out[(long)iy * n + x] = r[l * S + lf_pos(x)];
    }
}

// pass B: FFT the columns of B1 and take |E|^2
template <int T, int N>
__global__ __launch_bounds__(T) void k_lf_b(LfArgs a) {
    __shared__ float2 X[MX_LF_LDS], Y[MX_LF_LDS];
    MX_LF_GEOM(a.nwvf, a.L)
    const int n = n_, nxsa = a.nxsa, skip = (n - nxsa) / 2, lgL = N > 0 ? ilog2c(LC) : a.lgL;
    int g, c;
    lf_unit(a, (n + L - 1) / L, g, c);
    const int x0 = g * L;
    const float2 *in = a.b1 + (long)c * nxsa * n;
    for (int e = threadIdx.x; e < L * n; e += T) {
        const int l = e & (L - 1), y = e >> lgL, x = x0 + l, iy = y - skip;
        float2 v = make_float2(0.f, 0.f);
        if (x < n && iy >= 0 && iy < nxsa) v = in[(long)iy * n + x];
        X[l * S + lf_pos(y)] = v;
    }
    __syncthreads();
    const float2 *r = lf_fft<false, T, N>(X, Y, a.pn, L, S);
    float *out = a.psf + (long)c * n * n;
    for (int e = threadIdx.x; e < L * n; e += T) {
        const int l = e & (L - 1), y = e >> lgL, x = x0 + l;
        if (x < n) {
            const float2 v = r[l * S + lf_pos(y)];
            out[(long)y * n + x] = v.x * v.x + v.y * v.y;
        }
    }
}

// pass C: FFT the PSF rows, padded to notf by the corner copy.  The PSF is
// real, so rows 2p and 2p+1 go through one complex FFT as z = a + i b, and
// A[k] = (Z[k] + conj Z[m-k]) / 2, B[k] = (Z[k] - conj Z[m-k]) / 2i are kept
// for k <= m/2 only (the rest is their conjugate mirror): B2 is [nwvf][m/2+1].
template <int T, int N>
__global__ __launch_bounds__(T) void k_lf_c(LfArgs a) {
    __shared__ float2 X[MX_LF_LDS], Y[MX_LF_LDS];
    MX_LF_GEOM(a.notf, a.L)
    const int n = a.nwvf, m = n_, mh = m / 2 + 1, np = n / 2;
    int g, c;
    lf_unit(a, (np + L - 1) / L, g, c);
    const float *in = a.psf + (long)c * n * n;
    const float rm = 1.f / (float)m, rmh = 1.f / (float)mh, s = a.psf_scale;
    for (int e = threadIdx.x; e < L * m; e += T) {
        const int l = ldiv<(N > 0)>(e, m, rm), x2 = e - l * m, p = g * L + l, x = corner_src(x2, m, n);
        X[l * S + lf_pos(x2)] = p < np && x >= 0
                                    ? make_float2(in[(long)(2 * p) * n + x] * s, in[(long)(2 * p + 1) * n + x] * s)
                                    : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2 *r = lf_fft<false, T, N>(X, Y, a.pm, L, S);
    float2 *out = a.b2 + (long)c * n * mh;
    for (int e = threadIdx.x; e < L * mh; e += T) {
        const int l = ldiv<(N > 0)>(e, mh, rmh), k = e - l * mh, p = g * L + l;
        if (p < np) {
            const float2 z = r[l * S + lf_pos(k)], w = r[l * S + lf_pos(k == 0 ? 0 : m - k)];
            out[(long)(2 * p) * mh + k] = make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
            out[(long)(2 * p + 1) * mh + k] = make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x));
        }
    }
}

// pass D: FFT the columns ux <= m/2 of B2 (rows padded by the corner copy),
// x conj(lotf) x ETF (LLT) or x nominal, inverse FFT, keep the sampled rows:
// B3 is [ny][m/2+1].  Every factor is the transform of a real function, so
// the product at -u is the conjugate of that at u and the other columns are
// not needed.  UPLINK: the chains are (WFS, wavelength) LLT fields and the
// column FFT is conj-stored as the uplink OTF lotfc over ux <= m/2
// (maos/wfsints.c:149-199).
template <int T, int N, bool UPLINK>
__global__ __launch_bounds__(T) void k_lf_d(LfArgs a) {
    __shared__ float2 X[MX_LF_LDS], Y[MX_LF_LDS];
    MX_LF_GEOM(a.notf, a.L)
    const int n = a.nwvf, m = n_, mh = m / 2 + 1, lgL = N > 0 ? ilog2c(LC) : a.lgL;
    int g, c, iwfs, isa, iwvl;
    lf_unit(a, (mh + L - 1) / L, g, c);
    lf_chain(a, c, iwfs, isa, iwvl);
    const int x0 = g * L;
    const float2 *in = a.b2 + (long)c * n * mh;
    for (int e = threadIdx.x; e < L * m; e += T) {
        const int l = e & (L - 1), y2 = e >> lgL, x2 = x0 + l, y = corner_src(y2, m, n);
        X[l * S + lf_pos(y2)] = x2 < mh && y >= 0 ? in[(long)y * mh + x2] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2 *r = lf_fft<false, T, N>(X, Y, a.pm, L, S);
    if constexpr (UPLINK) {
        float2 *lo = a.lotf + (long)c * m * m;
        for (int e = threadIdx.x; e < L * m; e += T) {
            const int l = e & (L - 1), y2 = e >> lgL, x2 = x0 + l;
            if (x2 < mh) {
                const float2 z = r[l * S + lf_pos(y2)];
                lo[(long)x2 * m + y2] = make_float2(z.x, -z.y);
            }
        }
        return;
    }
    const LfWvl &W = a.wv[iwvl];
    if (a.lotf) { // x conj(lotfc) x ETF (apply_dtf_etf, maos/wfsints.c:38-64)
        const float2 *lo = a.lotf + ((long)iwfs * a.nwvl + iwvl) * m * m;
        const float2 *e1 = W.etf + (long)isa * m * m, *e2 = W.etf2 ? W.etf2 + (long)isa * m * m : nullptr;
        const float rm = 1.f / (float)m;
        for (int e = threadIdx.x; e < L * m; e += T) {
            const int l = ldiv<(N > 0)>(e, m, rm), y2 = e - l * m, x2 = x0 + l;
            if (x2 < mh) {
                const long i = (long)x2 * m + y2;
                float2 et = e1[i];
                if (e2) {
                    const float2 t2 = e2[i];
                    et = make_float2(et.x * a.etf1wt + t2.x * a.etf2wt, et.y * a.etf1wt + t2.y * a.etf2wt);
                }
                float2 &z = r[l * S + lf_pos(y2)];
                z = cmul(cmul(z, lo[i]), et);
            }
        }
        __syncthreads();
    } else if (W.nominal) {
        const float rm = 1.f / (float)m;
        for (int e = threadIdx.x; e < L * m; e += T) {
            const int l = ldiv<(N > 0)>(e, m, rm), y2 = e - l * m, x2 = x0 + l;
            if (x2 < mh) {
                const float f = W.nominal[(long)isa * W.nom_sa + (long)x2 * m + y2];
                float2 &z = r[l * S + lf_pos(y2)];
                z = make_float2(z.x * f, z.y * f);
            }
        }
        __syncthreads();
    }
    const float2 *q = lf_fft<true, T, N>(r, r == X ? Y : X, a.pm, L, S);
    float2 *out = a.b3 + (long)c * a.maxny * mh;
    for (int e = threadIdx.x; e < L * W.ny; e += T) {
        const int l = e & (L - 1), iyy = e >> lgL, x2 = x0 + l;
        if (x2 < mh) out[(long)iyy * mh + x2] = q[l * S + lf_pos(W.rows[iyy])];
    }
}

// pass E: inverse FFT of the sampled rows, real part at the sampled columns;
// a row's columns past m/2 are the conjugates of columns m - ux
template <int T, int N>
__global__ __launch_bounds__(T) void k_lf_e(LfArgs a) {
    __shared__ float2 X[MX_LF_LDS], Y[MX_LF_LDS];
    MX_LF_GEOM(a.notf, a.L)
    const int m = n_;
    int g, c, iwfs, isa, iwvl;
    lf_unit(a, (a.maxny + L - 1) / L, g, c);
    lf_chain(a, c, iwfs, isa, iwvl);
    const LfWvl &W = a.wv[iwvl];
    if (g * L >= W.ny) return; // uniform: no sampled row in this group
    const int mh = m / 2 + 1;
    const float2 *in = a.b3 + (long)c * a.maxny * mh;
    const float rm = 1.f / (float)m;
    for (int e = threadIdx.x; e < L * m; e += T) {
        const int l = ldiv<(N > 0)>(e, m, rm), x = e - l * m, iyy = g * L + l;
        float2 v = make_float2(0.f, 0.f);
        if (iyy < W.ny) {
            if (x < mh) {
                v = in[(long)iyy * mh + x];
            } else {
                v = in[(long)iyy * mh + m - x];
                v.y = -v.y;
            }
        }
        X[l * S + lf_pos(x)] = v;
    }
    __syncthreads();
    const float2 *r = lf_fft<true, T, N>(X, Y, a.pm, L, S);
    float *out = a.img + (long)c * a.maxny * a.maxnx;
    const float rnx = 1.f / (float)W.nx;
    for (int e = threadIdx.x; e < L * W.nx; e += T) {
        const int l = qdiv(e, W.nx, rnx), ixx = e - l * W.nx, iyy = g * L + l;
        if (iyy < W.ny) out[(long)iyy * W.nx + ixx] = r[l * S + lf_pos(W.cols[ixx])].x;
    }
}
#undef MX_LF_GEOM

// si sampling over the wavelengths and the pixel processing of one (WFS,
// subaperture) (dspmulcreal, maos/wfsints.c:286; shwfs_grad,
// maos/utils.c:983-1122)
struct LfSampleArgs {
    int nsa, nwfs, nwvl, npix, pixpsax, pixpsay, imgstride;
    const float *img;
    const LfWvl *wv;
    const DevWfs *wfs;
    int phytype, sigmatch;
    float cogthres, cogoff, pixthetax, pixthetay, sigtot;
    const float *mtche, *i0sum, *saa;
    float saat;
    float *ints, *grad;
    double *i0acc;
    DitherAcc dith;
    NoiseArgs noise;
    DtArgs dt;
};
__global__ __launch_bounds__(MX_LF_THREADS) void k_lf_sample(LfSampleArgs a) {
    __shared__ float red[3][MX_LF_THREADS / 64];
    const int isa = blockIdx.x % a.nsa, iwfs = blockIdx.x / a.nsa;
    const DevWfs &W = a.wfs[iwfs];
    const long ibase = ((long)iwfs * a.nsa + isa) * a.npix;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int p = threadIdx.x; p < a.npix; p += blockDim.x) {
        float pv = 0.f;
        for (int iwvl = 0; iwvl < a.nwvl; iwvl++) {
            const float *im = a.img + (((long)iwfs * a.nsa + isa) * a.nwvl + iwvl) * a.imgstride;
            const LfWvl &SW = a.wv[iwvl];
            const int4 id = SW.si_c[isa * SW.si_sa + p];
            const float4 w = SW.si_w[isa * SW.si_sa + p];
            pv += W.alpha[iwvl] * (w.x * im[id.x] + w.y * im[id.y] + w.z * im[id.z] + w.w * im[id.w]);
        }
        if (!dt_frame(a.dt, ibase + p, pv)) continue; // uniform: not the exposure's last frame
        if (a.noise.on) pv = add_noise(a.noise, pv, p, iwfs, a.nsa, isa);
        a.ints[ibase + p] = pv;
        if (a.i0acc) a.i0acc[ibase + p] += (double)pv;
        dither_add(a.dith, iwfs, ibase + p, pv);
        if (a.phytype == MX_PTYPE_MF) {
            const float *mt = a.mtche + ibase * 2 + 2 * p;
            s0 += mt[0] * pv;
            s1 += mt[1] * pv;
            s2 += pv;
        } else {
            const float iI = pv - a.cogoff;
            if (iI > a.cogthres) {
                s0 += iI * (float)(p % a.pixpsax);
                s1 += iI * (float)(p / a.pixpsax);
                s2 += iI;
            }
        }
    }
    if (!(a.dt.phase & MX_DT_OUT)) return; // uniform
    seg_sum<64>(s0);
    seg_sum<64>(s1);
    seg_sum<64>(s2);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[0][wave] = s0;
        red[1][wave] = s1;
        red[2][wave] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s0 = s1 = s2 = 0.f;
        for (int w = 0; w < (int)blockDim.x / 64; w++) {
            s0 += red[0][w];
            s1 += red[1][w];
            s2 += red[2][w];
        }
        if (!(a.saa[isa] < a.saat)) {
            float gx = 0.f, gy = 0.f;
            if (a.sigmatch == 2) { // raw moments; k_sigmatch2 forms the gradients
                gx = s0;
                gy = s1;
            } else if (a.phytype == MX_PTYPE_MF) {
                float scale = 1.f;
                if (a.sigmatch == 1) {
                    scale = a.i0sum[(long)iwfs * a.nsa + isa] / s2;
                    if (scale <= 0.f || scale > 10.f || !isfinite(scale)) scale = 1.f;
                }
                gx = s0 * scale;
                gy = s1 * scale;
            } else {
                float sum = s2;
                if (a.sigmatch == 0) sum = a.i0sum ? a.i0sum[(long)iwfs * a.nsa + isa] : a.sigtot * a.saa[isa];
                if (sum > a.cogthres) {
                    gx = (s0 / sum - (float)(a.pixpsax - 1) * 0.5f) * a.pixthetax;
                    gy = (s1 / sum - (float)(a.pixpsay - 1) * 0.5f) * a.pixthetay;
                }
            }
            a.grad[(long)iwfs * 2 * a.nsa + isa] = gx;
            a.grad[(long)iwfs * 2 * a.nsa + a.nsa + isa] = gy;
        }
    }
}
// ------------------------------------------------------------ ETF blend
// The two sodium-profile ETFs of the MFMA forms blended for the step,
// w1 etf + w2 etf2 with w2 = (isim % coldtrat)/coldtrat (maos/wfsints.c:
// 211-220), over the half-plane tables of every wavelength.
struct EtfBlendArgs {
    long n; // elements per wavelength (nsa * half plane)
    int nwvl;
    float w1, w2;
    const float2 *e1[MX_MAXWVL], *e2[MX_MAXWVL];
    float2 *out[MX_MAXWVL];
};
__global__ __launch_bounds__(256) void k_etf_blend(EtfBlendArgs a) {
    const long tot = a.n * a.nwvl;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long)gridDim.x * blockDim.x) {
        const int w = (int)(i / a.n);
        const long k = i - (long)w * a.n;
        const float2 x = a.e1[w][k], y = a.e2[w][k];
        a.out[w][k] = make_float2(x.x * a.w1 + y.x * a.w2, x.y * a.w1 + y.y * a.w2);
    }
}

// ------------------------------------------------------------ sigmatch 2
// Global flux normalisation (shwfs_grad with powfs.sigmatch = 2,
// maos/utils.c:1019-1080): i1sum = sum of every subaperture's pixels of the
// WFS; the matched filter's gradients are scaled by i0sumsum/i1sum (i0sumsum
// the sum of i0sum), and the CoG divides its moments by the flux
// i1sum/saasum * saa[isa] (dcog's flux override).  The imaging epilogues
// left the raw moments in grad; subapertures with saa < saat are skipped
// there and here.  One block per WFS.
struct Sig2Args {
    int nsa, npix, phytype, pixpsax, pixpsay;
    float cogthres, pixthetax, pixthetay, saat;
    double saasum;
    const float *ints, *i0sum, *saa;
    float *grad;
    int phase; // DtArgs phase: the gradients are formed at an exposure's last frame only
};
__global__ __launch_bounds__(256) void k_sigmatch2(Sig2Args a) {
    __shared__ double red[2][4];
    if (!(a.phase & MX_DT_OUT)) return;
    const int iwfs = (int)blockIdx.x;
    const long n = (long)a.nsa * a.npix;
    const float *im = a.ints + (long)iwfs * n;
    double s1 = 0., s0 = 0.;
    for (long i = threadIdx.x; i < n; i += 256) s1 += (double)im[i];
    if (a.phytype == MX_PTYPE_MF)
        for (int isa = threadIdx.x; isa < a.nsa; isa += 256) s0 += (double)a.i0sum[(long)iwfs * a.nsa + isa];
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s0 += __shfl_xor(s0, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x / 64] = s1;
        red[1][threadIdx.x / 64] = s0;
    }
    __syncthreads();
    const double i1sum = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double i0sumg = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    float *gx = a.grad + (long)iwfs * 2 * a.nsa, *gy = gx + a.nsa;
    for (int isa = threadIdx.x; isa < a.nsa; isa += 256) {
        if (a.saa[isa] < a.saat) continue;
        if (a.phytype == MX_PTYPE_MF) {
            const float sc = (float)(i0sumg / i1sum);
            gx[isa] *= sc;
            gy[isa] *= sc;
        } else {
            const float flux = (float)(i1sum / a.saasum * a.saa[isa]);
            if (flux > a.cogthres) {
                gx[isa] = (gx[isa] / flux - (float)(a.pixpsax - 1) * 0.5f) * a.pixthetax;
                gy[isa] = (gy[isa] / flux - (float)(a.pixpsay - 1) * 0.5f) * a.pixthetay;
            } else {
                gx[isa] = gy[isa] = 0.f;
            }
        }
    }
}
} // namespace

// ======================================================================
// C ABI
// ======================================================================
constexpr int MX_MAXLAUNCH = 16; // kernels of one step

// The kernels of one queued step in stream order; arg points into the
// handle's argument blocks (a graph's kernel nodes copy them).
struct StepLaunch {
    const void *f;
    dim3 grid, block;
    void *arg;
    int dep = -2; // the launch it waits for: -2 the previous one, -1 none (a branch's first)
};
struct StepPlan {
    StepLaunch L[MX_MAXLAUNCH];
    int nl = 0;
    int nmain = 0;     // launches of the step itself; the rest compute the next step's uplink OTF
    bool pre_set = false; // the step computes the uplink OTF of isim + 1 (pre_* below)
    long pre_isim = 0;
    double pre_dt = 0;
    int pre_form = 0, pre_buf = 0;
    unsigned pre_gen = 0;
    int key = 0;       // the kernel set: a graph built for one key fits any step of it
    int slot_opd_last = 0; // last kernel of the OPD stage (ray trace, OPD inputs)
    int slot_llt = -1; // index of the first uplink-OTF kernel, or -1 (they run up to slot_img)
    int slot_img = 0;  // index of the first imaging kernel
    bool llt = false;
};

struct mx_wfsgrad_s {
    int device = 0;
    int nreal = 1; // realisations; WFS r * (nwfs / nreal) + i is realisation r's WFS i
    hipStream_t stream = nullptr;
    // the argument blocks of the last queued step, and the step's graph
    OpdArgs arg_opd{};
    KArgs arg_step{};
    MfmaArgs arg_mfma{};
    LltArgs arg_llt{};
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraphNode_t gnode[MX_MAXLAUNCH] = {};
    int gkey = 0;
    long nsa = 0, nxsa = 0, embfac = 0, nwvf = 0, notf = 0, npix = 0, pixpsax = 0, pixpsay = 0,
         nwvl = 0;
    int nwfs = 0;
    mx_powfs_t cfg{};
    std::vector<double> wvl;
    std::vector<mx_wfs_t> wfs_cfg;
    std::vector<double> wvlwts; // nwfs*nwvl
    // device
    double *d_ptsx = nullptr, *d_ptsy = nullptr;
    float *d_amp = nullptr, *d_saa = nullptr;
    DevWfs *d_wfs = nullptr;
    float *d_nominal[MX_MAXWVL] = {};
    float2 *d_etf[MX_MAXWVL] = {}, *d_etf2[MX_MAXWVL] = {};
    float etf2wt = 0;
    int4 *d_si_idx[MX_MAXWVL] = {};
    float4 *d_si_w[MX_MAXWVL] = {};
    bool dtf_set[MX_MAXWVL] = {};
    long dtf_ntab[MX_MAXWVL] = {};  // si / nominal tables: 1, or nsa (one per subaperture)
    float *d_llt_amp = nullptr;
    float *d_llt_opd = nullptr;        // [nwfs][side^2] uplink OPD (side: llt_nx in whole ray-trace tiles)
    LltAcorrArgs arg_lltac{};         // uplink of an LLT grid wider than the subaperture grid
    // The uplink OTF of step isim + 1, computed beside step isim's imaging
    // (it needs only the screens): two lotf buffer pairs, the LLT OPD of the
    // next step, and what the prefetched buffer holds.
    bool uplink_pre = false;          // MX_UPLINK_PREFETCH=1 in the environment turns it on
    float2 *d_lotf2 = nullptr, *d_lotf_t2 = nullptr;
    float *d_llt_opd2 = nullptr;
    int lotf_last = 0;                // buffer the last queued step's imaging read
    bool pre_valid = false;
    long pre_isim = 0;
    double pre_dt = 0;
    int pre_form = 0, pre_buf = 0;
    unsigned pre_gen = 0;
    // The branch runs on a second stream (default priority) beside the step
    // (ev_main: the work queued before the step's chain; ev_branch: the last
    // branch)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_main = nullptr, ev_branch = nullptr;
    bool branch_pending = false;
    OpdArgs arg_opd_pre{};
    MfmaArgs arg_m_pre{};
    LltArgs arg_la_pre{};
    LltAcorrArgs arg_lltac_pre{};
    float2 *d_lotf = nullptr;
    DevWvl *d_wv = nullptr;
    DevWvl h_wv[MX_MAXWVL] = {};
    // MFMA form
    int variant = MX_VARIANT_AUTO;     // requested form
    bool mfma_size = false;            // N = 64, nxsa = 32
    double amp_max = 0;                // largest subaperture or LLT amplitude (the FP16 form needs <= 1)
    double llt_sum2 = 0;               // sum of the LLT amplitudes squared, lotf(0) (the FP16 form needs <= 1024)
    bool mfma_wvl_ok[MX_MAXWVL] = {};  // sampled rows/columns fit in 16
    uint16_t *d_F = nullptr, *d_F16 = nullptr; // DFT tables [MX_FENT][6][64] x 8, BF16 and FP16
    uint16_t *d_gx[MX_MAXWVL] = {}, *d_gy[MX_MAXWVL] = {};     // BF16
    uint16_t *d_gx16[MX_MAXWVL] = {}, *d_gy16[MX_MAXWVL] = {}; // FP16
    int4 *d_sic[MX_MAXWVL] = {};
    float2 *d_fy[MX_MAXWVL] = {};  // FP32 radix-8 form: pass F table
    int *d_fxs[MX_MAXWVL] = {};    // and the sampled columns
    float2 *d_etf_t[MX_MAXWVL] = {};
    // second ETF profile of the MFMA forms (half-plane order) and the blend
    // w1 etf + w2 etf2 that the imaging kernel reads
    float2 *d_etf2_t[MX_MAXWVL] = {}, *d_etfb_t[MX_MAXWVL] = {};
    float *d_nominal_t[MX_MAXWVL] = {};
    float2 *d_lotf_t = nullptr;
    MfmaWvl *d_mwv = nullptr;
    MfmaWvl h_mwv[MX_MAXWVL] = {};
    // large-field form: line-FFT plans of nwvf and notf, per-chain buffers
    bool lf = false;
    int lf_threads = 256;              // threads per large-field block
    int opd_tile = 0;                  // ray-trace tile side (16 or 32), 0: the default
    LfPlan lf_pn{}, lf_pm{};           // twiddles in d_lf_tw
    float2 *d_lf_tw = nullptr;
    float2 *d_b1 = nullptr, *d_b2 = nullptr, *d_b3 = nullptr;
    float *d_psf = nullptr, *d_img = nullptr;
    int lf_ny[MX_MAXWVL] = {}, lf_nx[MX_MAXWVL] = {}; // sampled rows / columns per wavelength
    int *d_lf_rows[MX_MAXWVL] = {}, *d_lf_cols[MX_MAXWVL] = {};
    int4 *d_lf_sic[MX_MAXWVL] = {};
    int lf_maxny = 0, lf_maxnx = 0;    // capacity of B3 and the image per chain
    LfWvl *d_lfw = nullptr;
    LfWvl h_lfw[MX_MAXWVL] = {};
    LfArgs arg_lf_base{}, arg_lf[5] = {}, arg_lf_up[4] = {};
    LfSampleArgs arg_lfs{};
    float *d_mtche = nullptr, *d_i0sum = nullptr;
    bool have_mtche = false, have_i0sum = false;
    float *d_ints = nullptr, *d_grad = nullptr, *d_opd = nullptr;
    double *d_i0acc = nullptr;
    bool i0_enable = false;
    long i0_nframe = 0;
    NoiseArgs noise{}; // isim set per step
    // closed-loop OPD inputs and geometric gradients
    int geom = 0;                     // 1: Zernike-tilt gradients from the OPD, no imaging
    double *d_imcc = nullptr;         // [nsa][9] inverse tip/tilt moment matrices
    int *d_gs_rowp = nullptr, *d_gs_coli = nullptr; // G-tilt operator GS0 (CSR, 2 nsa rows)
    double *d_gs_val = nullptr;
    long gs_nnz = -1;
    GtiltArgs arg_gtilt{};
    float *d_opdadd = nullptr;        // [nwfs][nsa][nxsa^2] static surface OPD
    bool opdadd_set = false;
    double *d_ptt = nullptr;          // [nwfs][3] piston, tip, tilt per step
    std::vector<double> h_ptt;
    bool ptt_set = false, ptt_dirty = false;
    OpdAddArgs arg_opdadd{};
    ZtiltArgs arg_ztilt{};
    DmArgs arg_dm{};
    // TT-dither statistics and the on-GPU matched-filter update
    double *d_imb = nullptr, *d_imx = nullptr, *d_imy = nullptr; // per-step sums
    double *d_fi0 = nullptr, *d_fgx = nullptr, *d_fgy = nullptr; // folded at PLL outputs
    double *d_mi0 = nullptr, *d_mgx = nullptr, *d_mgy = nullptr; // model i0, gx, gy
    double *d_nea = nullptr, *d_i0s64 = nullptr, *d_dcs = nullptr;
    std::vector<double> h_dcs;
    bool dith_on = false, dith_dirty = false;
    long dith_nframe = 0;
    // exposure of dtrat steps (mx_wfsgrad_set_dtrat): frame sums, and whether
    // the step planned last ends an exposure
    int dtrat = 1;
    float *d_dtacc = nullptr;
    bool last_out = true;
    // tip/tilt-removed PSF statistics (mx_wfsgrad_pistat_control)
    bool pistat = false;
    int pistat_gtype = 1;             // shift by 1: Zernike tilt, 2: G-tilt
    long pistat_nstep = 0;
    float *d_pistat = nullptr;        // [nwfs][nwvl][nsa][notf^2]
    float *d_gpis = nullptr;          // [nwfs][2*nsa] the step's geometric gradients
    ZtiltArgs arg_pis_z{};
    GtiltArgs arg_pis_g{};
    PistatArgs arg_pis{};
    float2 *d_psfout = nullptr;       // [nwvl][nsa][nout^2] of one WFS (mx_wfsgrad_psfout)
    FoldArgs arg_fold{};
    MtchArgs arg_mtch{};
    Sig2Args arg_sig2{};
    EtfBlendArgs arg_etfb{};
    bool queued = false;
    // timing
    bool timing = false;
    long t_n = 0;
    double t_llt = 0, t_opd = 0, t_ints = 0;
    std::vector<hipEvent_t> ev;
    bool pending_timing = false;
    // staging
    std::vector<float> hbuf;
};

namespace {
// screens of one realisation
struct AtmState {
    int nps = -1; // -1: not uploaded
    double atmscale = 1;
    DevLayer h_layers[MX_MAXLAYER];
    float *d_data = nullptr;
    size_t bytes = 0;
};
AtmState g_atm[MX_MAXREAL];
unsigned g_atm_gen = 0; // counts screen uploads: an uplink OTF computed ahead is valid for one generation
DevLayer *g_dlayers = nullptr; // device [MX_MAXREAL][MX_MAXLAYER]
// DM surfaces of one realisation
struct DmState {
    int ndm = 0;
    DevLayer h_layers[MX_MAXDM];
    float *d_data = nullptr;
    size_t bytes = 0;
};
DmState g_dm[MX_MAXREAL];
DevLayer *g_ddms = nullptr; // device [MX_MAXREAL][MX_MAXDM]
double *g_datmscale = nullptr; // device [MX_MAXREAL]
int g_device = 0;

// Working memory of the entry points that allocate nothing
// (mx_evl_wfe_noalloc): a fixed array in the code object, one copy per GPU,
// held until the process ends.  On the host, `scratch` names a host-side
// stand-in, not device memory: kernels and copies take the device address
// that hipGetSymbolAddress returns for the current GPU (g_scratch_dev).
constexpr int MX_SCRATCH_DOUBLES = 16640; // 130 KiB: 64 directions x (64 partial + 1 total) x 4 sums
constexpr int MX_SCRATCH_MAXDEV = 64;
static __device__ __align__(256) double scratch[MX_SCRATCH_DOUBLES];
double *g_scratch_dev[MX_SCRATCH_MAXDEV] = {};
std::mutex g_scratch_mu;

// device address of `scratch` on GPU `device` (which must be current),
// looked up once per GPU; MX_OK or an error code with the message set
int scratch_on(int device, double **out) {
    *out = nullptr;
    if (device < 0 || device >= MX_SCRATCH_MAXDEV)
        return set_err(MX_EINVAL, "device %d beyond the %d GPUs the scratch table covers", device, MX_SCRATCH_MAXDEV);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(MX_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                           hipGetErrorString(e));
        g_scratch_dev[device] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[device];
    return MX_OK;
}

template <typename T>
int dev_alloc(T **p, size_t n) {
    if (*p) return 0;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T) + 16);
    if (e != hipSuccess) return set_err(MX_EHIP, "hipMalloc(%zu) failed: %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

// transposed FP32 copy of an N x N table: dst[ux*N+uy] = src[ux+uy*N]
void transpose_c(float2 *dst, const double *src, long N, bool conj) {
    for (long uy = 0; uy < N; uy++)
        for (long ux = 0; ux < N; ux++) {
            const double *s = src + 2 * (ux + uy * N);
            dst[ux * N + uy] = make_float2((float)s[0], (float)(conj ? -s[1] : s[1]));
        }
}

bool supported(long notf, long nxsa, long embfac) {
    return (notf == 64 && nxsa == 32 && embfac == 2) || (notf == 32 && nxsa == 16 && embfac == 2);
}

// Stockham plan of a line length n <= MX_LF_MAXN whose prime factors are 2,
// 3, 5 and 7: radix 8 first, then 4, 2, 3, 5, 7.  False if n does not factor so.
bool lf_plan(long n, LfPlan &P) {
    P = LfPlan{};
    if (n < 2 || n > MX_LF_MAXN) return false;
    P.n = (int)n;
    long prod = 1;
    while (P.nrad < MX_LF_MAXRAD && lf_radix((int)n, P.nrad)) {
        P.rad[P.nrad] = lf_radix((int)n, P.nrad);
        prod *= P.rad[P.nrad++];
    }
    return prod == n;
}

// twiddles exp(-2 pi i r k / (Ns R)) of a plan's stages (k < Ns, 1 <= r <
// R) appended to tw; sets P.off relative to the plan's first entry
void lf_twiddles(LfPlan &P, std::vector<float2> &tw) {
    const size_t base = tw.size();
    int Ns = 1;
    for (int s = 0; s < P.nrad; s++) {
        const int R = P.rad[s];
        P.off[s] = (int)(tw.size() - base);
        for (int k = 0; k < Ns; k++)
            for (int r = 1; r < R; r++) {
                const double ang = -2. * M_PI * (double)(r * k) / (double)(Ns * R);
                tw.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
            }
        Ns *= R;
    }
}

// pass q (0..4 = A..E) of the large-field form for blocks of T threads and
// lines of length n: a kernel compiled for that length where there is one
// (the field and OTF sizes of the configured NGS and padded-LGS powfs)
template <int T, int N>
const void *lf_kernel_ab(int q) {
    return q == 0 ? reinterpret_cast<const void *>(&k_lf_a<T, N>) : reinterpret_cast<const void *>(&k_lf_b<T, N>);
}
template <int T, int N>
const void *lf_kernel_cde(int q) {
    return q == 2 ? reinterpret_cast<const void *>(&k_lf_c<T, N>)
                  : q == 3 ? reinterpret_cast<const void *>(&k_lf_d<T, N, false>)
                  : q == 5 ? reinterpret_cast<const void *>(&k_lf_d<T, N, true>)
                           : reinterpret_cast<const void *>(&k_lf_e<T, N>);
}
template <int T>
const void *lf_kernel_t(int q, int n) {
    if constexpr (T == 512) {
        if (q < 2) {
            if (n == 96) return lf_kernel_ab<T, 96>(q);
            if (n == 128) return lf_kernel_ab<T, 128>(q);
            if (n == 480) return lf_kernel_ab<T, 480>(q);
            if (n == 960) return lf_kernel_ab<T, 960>(q);
        } else {
            if (n == 120) return lf_kernel_cde<T, 120>(q);
            if (n == 128) return lf_kernel_cde<T, 128>(q);
            if (n == 512) return lf_kernel_cde<T, 512>(q);
            if (n == 1024) return lf_kernel_cde<T, 1024>(q);
        }
    }
    return q < 2 ? lf_kernel_ab<T, 0>(q) : lf_kernel_cde<T, 0>(q);
}
// q: 0..4 passes A..E, 5 the uplink's column pass
const void *lf_kernel(int q, int T, int n) {
    return T == 128 ? lf_kernel_t<128>(q, n) : T == 256 ? lf_kernel_t<256>(q, n) : lf_kernel_t<512>(q, n);
}

// the large-field form applies: OPD grid a multiple of 16 points, even
// field and OTF sizes that plan, notf >= nwvf (ccpcorner pads)
bool lf_supported(long notf, long nxsa, long embfac) {
    LfPlan a, b;
    const long nwvf = nxsa * embfac;
    return nxsa >= 16 && nxsa % 16 == 0 && embfac >= 1 && nwvf % 2 == 0 && notf % 2 == 0 && notf >= nwvf &&
           lf_plan(nwvf, a) && lf_plan(notf, b);
}

// FP32 -> BF16 with round-to-nearest-even (finite inputs)
uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// hi/lo BF16 pair of an FP64 value: hi = bf16(x), lo = bf16(x - hi)
void split_bf(double x, uint16_t &hi, uint16_t &lo) {
    hi = f2bf((float)x);
    lo = f2bf((float)(x - (double)bf2f(hi)));
}
// hi/lo FP16 pair (round to nearest even, subnormals kept)
void split_h(double x, uint16_t &hi, uint16_t &lo) {
    const _Float16 h = (_Float16)(float)x, l = (_Float16)(float)(x - (double)(float)h);
    std::memcpy(&hi, &h, 2);
    std::memcpy(&lo, &l, 2);
}
// element type of the table entries being built (BF16 unless set); FP16
// entries carry a scale of 2^13 (the kernel's range plan)
bool g_tab_f16 = false;
// one table entry (6 parts x 64 lanes x 8) of complex values: re, im, -im
void put_entry(uint16_t *dst, const double (*re)[8], const double (*im)[8]) {
    for (int lane = 0; lane < 64; lane++)
        for (int j = 0; j < 8; j++) {
            const double v[3] = {re[lane][j], im[lane][j], -im[lane][j]};
            for (int p = 0; p < 3; p++) {
                uint16_t h, l;
                if (g_tab_f16) split_h(v[p] * 8192., h, l);
                else split_bf(v[p], h, l);
                dst[((2 * p) * 64 + lane) * 8 + j] = h;
                dst[((2 * p + 1) * 64 + lane) * 8 + j] = l;
            }
        }
}
// one B-operand entry of the forward DFT matrix F[k][n] = exp(-2 pi i k n /
// 64): lane l, j -> column ncol(l % w), k = kbase + 8 (l / w) + j, with w = 32
// for v_mfma_f32_32x32x16 and 16 for v_mfma_f32_16x16x32
template <typename NCOL>
void put_F(uint16_t *dst, int w, int kbase, NCOL ncol) {
    double re[64][8], im[64][8];
    for (int lane = 0; lane < 64; lane++)
        for (int j = 0; j < 8; j++) {
            const int n = ncol(lane % w), k = kbase + 8 * (lane / w) + j;
            const double ang = -2. * M_PI * (double)((k * n) % 64) / 64.;
            re[lane][j] = std::cos(ang);
            im[lane][j] = std::sin(ang);
        }
    put_entry(dst, re, im);
}
// the DFT tables of k_shwfs_mfma (entries MX_F1A, MX_F2)
std::vector<uint16_t> build_F64() {
    const size_t ent = (size_t)MX_TPARTS * 64 * 8;
    std::vector<uint16_t> t(MX_FENT * ent);
    for (int nt = 0; nt < 4; nt++)
        put_F(t.data() + (MX_F1A + nt) * ent, 16, 16, [nt](int c) { return 16 * nt + c; });
    for (int e = 0; e < 2; e++)
        for (int hf = 0; hf < 2; hf++)
            put_F(t.data() + (MX_F2 + 2 * e + hf) * ent, 16, 0, [e, hf](int c) { return 2 * (16 * hf + c) + e; });
    return t;
}
// the inverse DFT conj(F) at the sampled coordinates rows[0..n) as B
// operands of v_mfma_f32_16x16x32_bf16: entry ks, lane l, j -> column l%16
// (sampled coordinate rows[l%16]), k = 32 ks + 8 (l/16) + j; columns beyond
// n are zero
std::vector<uint16_t> build_G(const std::vector<int> &rows) {
    std::vector<uint16_t> t(2 * MX_TPARTS * 64 * 8);
    double re[64][8], im[64][8];
    for (int ks = 0; ks < 2; ks++) {
        for (int lane = 0; lane < 64; lane++)
            for (int j = 0; j < 8; j++) {
                const int r = lane % 16, k = 32 * ks + 8 * (lane / 16) + j;
                if (r < (int)rows.size()) {
                    const double ang = 2. * M_PI * (double)((rows[r] * k) % 64) / 64.;
                    re[lane][j] = std::cos(ang);
                    im[lane][j] = std::sin(ang);
                } else {
                    re[lane][j] = im[lane][j] = 0.;
                }
            }
        put_entry(t.data() + (size_t)ks * MX_TPARTS * 64 * 8, re, im);
    }
    return t;
}
// the inverse DFT over the half plane uy in [0, 32) with the Hermitian weight
// w(uy) = 1 at uy = 0 and 2 elsewhere, at the sampled rows rows[0..n), as the B
// operand of one v_mfma_f32_16x16x32_bf16 (entry lane l, j -> column l%16,
// k = 8 (l/16) + j)
std::vector<uint16_t> build_Gh(const std::vector<int> &rows) {
    std::vector<uint16_t> t(MX_TPARTS * 64 * 8);
    double re[64][8], im[64][8];
    for (int lane = 0; lane < 64; lane++)
        for (int j = 0; j < 8; j++) {
            const int r = lane % 16, k = 8 * (lane / 16) + j;
            const double w = k == 0 ? 1. : 2.;
            if (r < (int)rows.size()) {
                const double ang = 2. * M_PI * (double)((rows[r] * k) % 64) / 64.;
                re[lane][j] = w * std::cos(ang);
                im[lane][j] = w * std::sin(ang);
            } else {
                re[lane][j] = im[lane][j] = 0.;
            }
        }
    put_entry(t.data(), re, im);
    return t;
}
// form used by the next step: the requested one, or for AUTO the MFMA form
// whenever its size and sampling limits hold; -1 if the request cannot run
// the half-plane tables (N = 64, nxsa = 32, <= 16 sampled rows and columns
// per wavelength, a second ETF in half-plane form) that the MFMA forms and
// the FP32 radix-8 form read; their pixel epilogue is one lane per pixel of
// one wave, so at most 64 pixels (more run on the register-FFT form)
bool half_tables_ok(const mx_wfsgrad_t *h) {
    bool ok = h->mfma_size && h->npix <= 64;
    for (int w = 0; w < h->nwvl; w++) ok = ok && h->mfma_wvl_ok[w] && (!h->d_etf2[w] || h->d_etf2_t[w]);
    return ok;
}
int effective_variant(const mx_wfsgrad_t *h) {
    const bool ok = half_tables_ok(h);
    if (h->variant == MX_VARIANT_FFT32) return MX_VARIANT_FFT32;
    if (h->variant == MX_VARIANT_MFMA_BF16 || h->variant == MX_VARIANT_MFMA_FP16) return ok ? h->variant : -1;
    return ok ? MX_VARIANT_MFMA_BF16 : MX_VARIANT_FFT32;
}
} // namespace

extern "C" {

const char *mx_last_error(void) { return g_err; }

int mx_init(int device) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(MX_EINVAL, "device %d out of range (%d GPUs)", device, n);
    HIPCHK(hipSetDevice(device));
    g_device = device;
    double *sp = nullptr;
    if (const int rc = scratch_on(device, &sp)) return rc; // the address of this GPU's copy of `scratch`
    return MX_OK;
}

int mx_wfsgrad_init(mx_wfsgrad_t **out, const mx_powfs_t *p, const mx_wfs_t *wfs, int nwfs) {
    return mx_wfsgrad_init_batch(out, p, wfs, nwfs, 1);
}

int mx_wfsgrad_init_batch(mx_wfsgrad_t **out, const mx_powfs_t *p, const mx_wfs_t *wfs0, int nwfs0, int nreal) {
    if (!out || !p || !wfs0 || nwfs0 <= 0) return set_err(MX_EINVAL, "mx_wfsgrad_init: null argument");
    *out = nullptr;
    if (nreal < 1 || nreal > MX_MAXREAL) return set_err(MX_EINVAL, "nreal=%d out of range (1..%d)", nreal, MX_MAXREAL);
    // realisation r's copy of WFS i is WFS r * nwfs0 + i of the engine
    const int nwfs = nwfs0 * nreal;
    std::vector<mx_wfs_t> wfs_all(nwfs);
    for (int i = 0; i < nwfs; i++) wfs_all[i] = wfs0[i % nwfs0];
    const mx_wfs_t *wfs = wfs_all.data();
    const bool lf = !supported(p->notf, p->nxsa, p->embfac);
    if (lf && !lf_supported(p->notf, p->nxsa, p->embfac))
        return set_err(MX_EINVAL,
                       "unsupported size: notf=%ld nxsa=%ld embfac=%ld (supported: 64/32/2, 32/16/2, or nxsa a "
                       "multiple of 16 with nxsa*embfac <= notf <= %d, both even with prime factors 2, 3, 5 and 7)",
                       p->notf, p->nxsa, p->embfac, MX_LF_MAXN);
    if (lf && p->llt_nx > 0 &&
        !(p->llt_nx == p->nxsa && (p->nxsa == 16 || p->nxsa == 32) && p->llt_nx * p->embfac == p->notf))
        return set_err(MX_EINVAL, "with an LLT the large-field form needs llt_nx == nxsa (16 or 32) and "
                       "llt_nx*embfac == notf (got llt_nx=%ld nxsa=%ld notf=%ld)", p->llt_nx, p->nxsa, p->notf);
    if (p->nwvl < 1 || p->nwvl > MX_MAXWVL) return set_err(MX_EINVAL, "nwvl=%ld out of range", p->nwvl);
    const long npix = p->pixpsax * p->pixpsay;
    const long ppl = (npix + p->notf - 1) / p->notf;
    if (npix < 1 || (!lf && ppl > 2) || npix > p->notf * p->notf)
        return set_err(MX_EINVAL, "npix=%ld too large for notf=%ld", npix, p->notf);
    if (p->phytype != MX_PTYPE_MF && p->phytype != MX_PTYPE_COG)
        return set_err(MX_EINVAL, "phytype %d unsupported", p->phytype);
    if (p->sigmatch < 0 || p->sigmatch > 2)
        return set_err(MX_EINVAL, "sigmatch %d unsupported (0, 1 or 2)", p->sigmatch);
    if (p->sigmatch == 2 && !(p->saasum > 0))
        return set_err(MX_EINVAL, "sigmatch 2 needs saasum > 0 (saasum=%g)", p->saasum);
    // an LLT grid wider than the subaperture grid (MAOS makes it max(llt.d,
    // dsa) wide) takes the direct-autocorrelation uplink: nwvf == notf <= 64
    // and at most 64 LLT points a side
    if (p->llt_nx > 0 && p->llt_nx != p->nxsa &&
        !(!lf && p->llt_nx > p->nxsa && p->llt_nx <= 64 && p->nxsa * p->embfac == p->notf && p->notf <= 64))
        return set_err(MX_EINVAL, "LLT grid %ld: needs llt_nx == nxsa, or nxsa < llt_nx <= 64 with "
                       "nxsa*embfac == notf <= 64 (got nxsa=%ld notf=%ld)", p->llt_nx, p->nxsa, p->notf);
    if (p->nsa < 1) return set_err(MX_EINVAL, "nsa=%ld", p->nsa);
    HIPCHK(hipSetDevice(g_device));
    mx_wfsgrad_t *h = new mx_wfsgrad_t();
    h->device = g_device;
    h->cfg = *p;
    h->nsa = p->nsa;
    h->nxsa = p->nxsa;
    h->embfac = p->embfac;
    h->nwvf = p->nxsa * p->embfac;
    h->notf = p->notf;
    h->pixpsax = p->pixpsax;
    h->pixpsay = p->pixpsay;
    h->npix = npix;
    h->nwvl = p->nwvl;
    h->nwfs = nwfs;
    h->nreal = nreal;
    h->wvl.assign(p->wvl, p->wvl + p->nwvl);
    h->wfs_cfg.assign(wfs, wfs + nwfs);
    int rc = 0;
#define ALLOC(ptr, n)                                   \
    if ((rc = dev_alloc(&(ptr), (n))) != 0) {           \
        mx_wfsgrad_free(h);                             \
        return rc;                                      \
    }
    const long nsa = h->nsa, N = h->notf, npsa = h->nxsa * h->nxsa;
    ALLOC(h->d_ptsx, nsa);
    ALLOC(h->d_ptsy, nsa);
    ALLOC(h->d_amp, nsa * npsa);
    ALLOC(h->d_saa, nsa);
    ALLOC(h->d_wfs, nwfs);
    ALLOC(h->d_wv, MX_MAXWVL);
    ALLOC(h->d_ints, nwfs * nsa * npix);
    ALLOC(h->d_grad, nwfs * 2 * nsa);
    ALLOC(h->d_opd, nwfs * nsa * npsa);
    ALLOC(h->d_i0acc, nwfs * nsa * npix);
    ALLOC(h->d_mtche, nwfs * nsa * 2 * npix);
    ALLOC(h->d_i0sum, nwfs * nsa);
    h->mfma_size = (N == 64 && h->nxsa == 32);
    h->lf = lf;
    std::vector<float2> lf_tw;
    if (lf) {
        lf_plan(h->nwvf, h->lf_pn);
        lf_plan(N, h->lf_pm);
        lf_twiddles(h->lf_pn, lf_tw);
        const size_t nt_n = lf_tw.size();
        lf_twiddles(h->lf_pm, lf_tw);
        const long nchain = (long)nwfs * nsa * h->nwvl;
        ALLOC(h->d_lf_tw, lf_tw.size());
        h->lf_pn.tw = h->d_lf_tw;
        h->lf_pm.tw = h->d_lf_tw + nt_n;
        ALLOC(h->d_b1, nchain * h->nxsa * h->nwvf);
        ALLOC(h->d_psf, nchain * h->nwvf * h->nwvf);
        ALLOC(h->d_b2, nchain * h->nwvf * N);
        ALLOC(h->d_lfw, MX_MAXWVL);
    }
    if (p->llt_nx > 0) {
        ALLOC(h->d_llt_amp, p->llt_nx * p->llt_nx);
        // traced in tiles of 16 or 32 points: the grid is stored with its side
        // rounded up to a multiple of 32
        const long side = (p->llt_nx + 31) / 32 * 32;
        ALLOC(h->d_llt_opd, nwfs * side * side);
        ALLOC(h->d_lotf, nwfs * h->nwvl * N * N);
        if (h->mfma_size) ALLOC(h->d_lotf_t, nwfs * h->nwvl * MX_HALF);
        // The next step's uplink computed beside a step's chain, on a second
        // stream (MX_UPLINK_PREFETCH=1): off by default, since the two event
        // waits it adds per step cost more than the uplink (0.2494 against
        // 0.2436 ms per lgs30m step on one box).
        const char *e = getenv("MX_UPLINK_PREFETCH");
        h->uplink_pre = e && atoi(e) != 0 && !lf;
        if (h->uplink_pre) {
            ALLOC(h->d_llt_opd2, nwfs * side * side);
            ALLOC(h->d_lotf2, nwfs * h->nwvl * N * N);
            if (h->mfma_size) ALLOC(h->d_lotf_t2, nwfs * h->nwvl * MX_HALF);
        }
    }
    if (h->d_lotf2 && h->uplink_pre &&
        (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&h->ev_main, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&h->ev_branch, hipEventDisableTiming) != hipSuccess)) {
        mx_wfsgrad_free(h);
        return set_err(MX_EHIP, "second stream or its events could not be created");
    }
    if (h->mfma_size) {
        ALLOC(h->d_F, MX_FENT * MX_TPARTS * 64 * 8);
        ALLOC(h->d_F16, MX_FENT * MX_TPARTS * 64 * 8);
        ALLOC(h->d_mwv, MX_MAXWVL);
    }
#undef ALLOC
    auto fail = [&](int code) {
        mx_wfsgrad_free(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(MX_EHIP, "hipStreamCreate failed"));
    std::vector<float> tmp(nsa * npsa);
    for (long i = 0; i < nsa * npsa; i++) {
        tmp[i] = (float)p->amp[i];
        h->amp_max = std::max(h->amp_max, std::fabs(p->amp[i]));
    }
    for (long i = 0; i < p->llt_nx * p->llt_nx; i++) {
        h->amp_max = std::max(h->amp_max, std::fabs(p->llt_amp[i]));
        h->llt_sum2 += p->llt_amp[i] * p->llt_amp[i];
    }
    if (hipMemcpy(h->d_amp, tmp.data(), sizeof(float) * nsa * npsa, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_ptsx, p->ptsx, sizeof(double) * nsa, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_ptsy, p->ptsy, sizeof(double) * nsa, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(MX_EHIP, "upload of geometry failed"));
    for (long i = 0; i < nsa; i++) tmp[i] = (float)p->saa[i];
    if (hipMemcpy(h->d_saa, tmp.data(), sizeof(float) * nsa, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(MX_EHIP, "upload of saa failed"));
    if (p->llt_nx > 0) {
        std::vector<float> la(p->llt_nx * p->llt_nx);
        for (size_t i = 0; i < la.size(); i++) la[i] = (float)p->llt_amp[i];
        if (hipMemcpy(h->d_llt_amp, la.data(), sizeof(float) * la.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(MX_EHIP, "upload of LLT amplitude failed"));
    }
    // per-WFS constants; norm_ints of wfsints.c:111-115
    const double norm_psf = std::sqrt(p->areascale) / (double)(h->nxsa * h->nwvf);
    std::vector<DevWfs> dw(nwfs);
    for (int i = 0; i < nwfs; i++) {
        dw[i].thetax = wfs[i].thetax;
        dw[i].thetay = wfs[i].thetay;
        dw[i].hs = std::isinf(wfs[i].hs) ? 0. : wfs[i].hs;
        dw[i].misregx = wfs[i].misregx;
        dw[i].misregy = wfs[i].misregy;
        dw[i].ireal = i / nwfs0;
        const double norm_ints = wfs[i].sigsim * norm_psf * norm_psf / ((double)N * N);
        for (int w = 0; w < MX_MAXWVL; w++)
            dw[i].alpha[w] = w < h->nwvl ? (float)(wfs[i].wvlwts[w] * norm_ints) : 0.f;
    }
    if (hipMemcpy(h->d_wfs, dw.data(), sizeof(DevWfs) * nwfs, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(MX_EHIP, "upload of WFS parameters failed"));
    if (lf && hipMemcpy(h->d_lf_tw, lf_tw.data(), sizeof(float2) * lf_tw.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(MX_EHIP, "upload of the FFT twiddles failed"));
    if (h->mfma_size) {
        g_tab_f16 = true;
        std::vector<uint16_t> F16 = build_F64();
        g_tab_f16 = false;
        std::vector<uint16_t> F = build_F64();
        if (hipMemcpy(h->d_F, F.data(), F.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->d_F16, F16.data(), F16.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(MX_EHIP, "upload of the DFT table failed"));
    }
    {
        // inverse tip/tilt moment matrix of each subaperture (pts_mcc_ptt,
        // math/loc.c:612-650, inverted as dcellinvspd_each does)
        std::vector<double> imcc(9 * nsa, 0.);
        for (long isa = 0; isa < nsa; isa++) {
            double m00 = 0, m01 = 0, m02 = 0, m11 = 0, m12 = 0, m22 = 0;
            for (long iy = 0; iy < h->nxsa; iy++)
                for (long ix = 0; ix < h->nxsa; ix++) {
                    const double x = p->ptsx[isa] + ix * p->dx, y = p->ptsy[isa] + iy * p->dx;
                    const double w = p->amp[isa * npsa + iy * h->nxsa + ix];
                    m00 += w;
                    m01 += w * x;
                    m02 += w * y;
                    m11 += w * x * x;
                    m12 += w * x * y;
                    m22 += w * y * y;
                }
            const double c00 = m11 * m22 - m12 * m12, c01 = m02 * m12 - m01 * m22, c02 = m01 * m12 - m02 * m11;
            const double c11 = m00 * m22 - m02 * m02, c12 = m01 * m02 - m00 * m12, c22 = m00 * m11 - m01 * m01;
            const double det = m00 * c00 + m01 * c01 + m02 * c02;
            if (det > 0) {
                const double r[9] = {c00, c01, c02, c01, c11, c12, c02, c12, c22};
                for (int k = 0; k < 9; k++) imcc[9 * isa + k] = r[k] / det;
            }
        }
        if (dev_alloc(&h->d_imcc, imcc.size()) ||
            hipMemcpy(h->d_imcc, imcc.data(), sizeof(double) * imcc.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(MX_EHIP, "upload of the tip/tilt moments failed"));
    }
    if (hipMemset(h->d_grad, 0, sizeof(float) * nwfs * 2 * nsa) != hipSuccess ||
        hipMemset(h->d_i0acc, 0, sizeof(double) * nwfs * nsa * npix) != hipSuccess ||
        hipMemset(h->d_mtche, 0, sizeof(float) * nwfs * nsa * 2 * npix) != hipSuccess ||
        hipMemset(h->d_i0sum, 0, sizeof(float) * nwfs * nsa) != hipSuccess)
        return fail(set_err(MX_EHIP, "clearing buffers failed"));
    *out = h;
    return MX_OK;
}

// release the step graph of a handle (the stream holds no launch of it)
static void drop_graph(mx_wfsgrad_t *h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->gkey = 0;
}

int mx_wfsgrad_free(mx_wfsgrad_t *h) {
    if (!h) return MX_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->stream2) hipStreamSynchronize(h->stream2);
    drop_graph(h);
    void *ptrs[] = {h->d_ptsx, h->d_ptsy, h->d_amp, h->d_saa, h->d_wfs, h->d_wv, h->d_llt_amp, h->d_llt_opd, h->d_lotf,
                    h->d_lotf_t, h->d_F, h->d_F16, h->d_mwv, h->d_imcc, h->d_opdadd, h->d_ptt, h->d_lf_tw, h->d_b1, h->d_b2, h->d_b3, h->d_psf, h->d_img,
                    h->d_lfw,
                    h->d_mtche, h->d_i0sum, h->d_ints, h->d_grad, h->d_opd, h->d_i0acc,
                    h->d_imb, h->d_imx, h->d_imy, h->d_fi0, h->d_fgx, h->d_fgy, h->d_mi0, h->d_mgx, h->d_mgy,
                    h->d_nea, h->d_i0s64, h->d_dcs, h->d_gs_rowp, h->d_gs_coli, h->d_gs_val};
    for (void *q : ptrs)
        if (q) hipFree(q);
    if (h->d_dtacc) hipFree(h->d_dtacc);
    if (h->d_pistat) hipFree(h->d_pistat);
    if (h->d_gpis) hipFree(h->d_gpis);
    if (h->d_psfout) hipFree(h->d_psfout);
    for (void *q : {(void *)h->d_lotf2, (void *)h->d_lotf_t2, (void *)h->d_llt_opd2})
        if (q) hipFree(q);
    for (hipEvent_t e : {h->ev_main, h->ev_branch})
        if (e) hipEventDestroy(e);
    if (h->stream2) hipStreamDestroy(h->stream2);
    for (int w = 0; w < MX_MAXWVL; w++) {
        if (h->d_nominal[w]) hipFree(h->d_nominal[w]);
        if (h->d_gx[w]) hipFree(h->d_gx[w]);
        if (h->d_gy[w]) hipFree(h->d_gy[w]);
        if (h->d_gx16[w]) hipFree(h->d_gx16[w]);
        if (h->d_gy16[w]) hipFree(h->d_gy16[w]);
        if (h->d_sic[w]) hipFree(h->d_sic[w]);
        if (h->d_fy[w]) hipFree(h->d_fy[w]);
        if (h->d_fxs[w]) hipFree(h->d_fxs[w]);
        if (h->d_etf_t[w]) hipFree(h->d_etf_t[w]);
        if (h->d_etf2_t[w]) hipFree(h->d_etf2_t[w]);
        if (h->d_etfb_t[w]) hipFree(h->d_etfb_t[w]);
        if (h->d_nominal_t[w]) hipFree(h->d_nominal_t[w]);
        if (h->d_etf[w]) hipFree(h->d_etf[w]);
        if (h->d_etf2[w]) hipFree(h->d_etf2[w]);
        if (h->d_si_idx[w]) hipFree(h->d_si_idx[w]);
        if (h->d_si_w[w]) hipFree(h->d_si_w[w]);
        if (h->d_lf_rows[w]) hipFree(h->d_lf_rows[w]);
        if (h->d_lf_cols[w]) hipFree(h->d_lf_cols[w]);
        if (h->d_lf_sic[w]) hipFree(h->d_lf_sic[w]);
    }
    for (hipEvent_t e : h->ev) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return MX_OK;
}

namespace {
// The detector tables of one wavelength: ntab = 1 (one si and nominal for
// every subaperture) or nsa (multi_nominal, maos/wfsints.c:200-233).  Table
// t is si_pp + t (si_ny + 1), si_pi / si_px + t si_nzmax and nominal +
// t notf^2.  Device arrays are [ntab][...] with the per-subaperture stride
// in the kernels' argument blocks (0 when ntab = 1).
int set_dtf_tables(mx_wfsgrad_t *h, int iwvl, long ntab, const double *nominal, long si_nx, long si_ny,
                   long si_nzmax, const long *si_pp, const long *si_pi, const double *si_px) {
    const long N = h->notf, npix = h->npix;
    if (si_nx != npix || si_ny != N * N || !si_pp || !si_pi || !si_px)
        return set_err(MX_EINVAL, "si must be %ld x %ld (got %ld x %ld)", npix, N * N, si_nx, si_ny);
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream)); // a queued step may still read the tables replaced here
    // per-pixel list of up to 4 (index, weight) pairs of every table
    std::vector<int4> idx(ntab * npix, make_int4(0, 0, 0, 0));
    std::vector<float4> wt(ntab * npix, make_float4(0, 0, 0, 0));
    for (long t = 0; t < ntab; t++) {
        const long *pp = si_pp + t * (si_ny + 1), *pi = si_pi + t * si_nzmax;
        const double *px = si_px + t * si_nzmax;
        std::vector<int> cnt(npix, 0);
        if (pp[0] < 0 || pp[N * N] > si_nzmax) return set_err(MX_EINVAL, "si table %ld: bad column pointers", t);
        for (long col = 0; col < N * N; col++) {
            for (long k = pp[col]; k < pp[col + 1]; k++) {
                const long r = pi[k];
                if (r < 0 || r >= npix) return set_err(MX_EINVAL, "si row index %ld out of range", r);
                const int c = cnt[r]++;
                if (c >= 4) return set_err(MX_EINVAL, "si pixel %ld has more than 4 entries", r);
                (&idx[t * npix + r].x)[c] = (int)col;
                (&wt[t * npix + r].x)[c] = (float)px[k];
            }
        }
    }
    // a change of table count frees the buffers sized for the old one
    if (h->dtf_ntab[iwvl] != ntab) {
        void **bufs[] = {(void **)&h->d_si_idx[iwvl], (void **)&h->d_si_w[iwvl], (void **)&h->d_nominal[iwvl],
                         (void **)&h->d_sic[iwvl], (void **)&h->d_nominal_t[iwvl], (void **)&h->d_lf_sic[iwvl]};
        for (void **q : bufs)
            if (*q) {
                HIPCHK(hipFree(*q));
                *q = nullptr;
            }
        h->dtf_ntab[iwvl] = ntab;
    }
    int rc;
    if ((rc = dev_alloc(&h->d_si_idx[iwvl], ntab * npix)) || (rc = dev_alloc(&h->d_si_w[iwvl], ntab * npix)))
        return rc;
    HIPCHK(hipMemcpy(h->d_si_idx[iwvl], idx.data(), sizeof(int4) * ntab * npix, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_si_w[iwvl], wt.data(), sizeof(float4) * ntab * npix, hipMemcpyHostToDevice));
    if (nominal && h->cfg.llt_nx == 0) {
        if ((rc = dev_alloc(&h->d_nominal[iwvl], ntab * N * N))) return rc;
        // The large-field form transforms only the columns ux <= N/2 and takes
        // the rest as conjugates, which gives MAOS's real part of the inverse
        // transform exactly when the nominal is even: (n(u) + n(-u)) / 2 is
        // stored for it (the same real part for any nominal; rotated pixels
        // make the nominal uneven on the Nyquist row and column).
        std::vector<float> nt(ntab * N * N);
        for (long t = 0; t < ntab; t++)
            for (long uy = 0; uy < N; uy++)
                for (long ux = 0; ux < N; ux++) {
                    const double *nm = nominal + t * N * N;
                    const double v = nm[ux + uy * N];
                    nt[t * N * N + ux * N + uy] =
                        h->lf ? (float)(0.5 * (v + nm[(N - ux) % N + (N - uy) % N * N])) : (float)v;
                }
        HIPCHK(hipMemcpy(h->d_nominal[iwvl], nt.data(), sizeof(float) * ntab * N * N, hipMemcpyHostToDevice));
    }
    // image rows and columns the samplers read: the union over the tables
    std::vector<int> xs, ys;
    for (long i = 0; i < ntab * npix; i++)
        for (int c = 0; c < 4; c++)
            if ((&wt[i].x)[c] != 0.f) {
                xs.push_back((&idx[i].x)[c] % (int)N);
                ys.push_back((&idx[i].x)[c] / (int)N);
            }
    for (std::vector<int> *v : {&xs, &ys}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    if (h->mfma_size) {
        h->mfma_wvl_ok[iwvl] = xs.size() <= MX_MFMA_NS && ys.size() <= MX_MFMA_NS;
        if (h->mfma_wvl_ok[iwvl]) {
            // si in compact 16x16 indices
            std::vector<int4> sic(ntab * npix);
            for (long i = 0; i < ntab * npix; i++) {
                int *op = &sic[i].x;
                const int *ip = &idx[i].x;
                for (int c = 0; c < 4; c++) {
                    const int cx = (int)(std::lower_bound(xs.begin(), xs.end(), ip[c] % (int)N) - xs.begin());
                    const int cy = (int)(std::lower_bound(ys.begin(), ys.end(), ip[c] / (int)N) - ys.begin());
                    op[c] = (&wt[i].x)[c] != 0.f ? cy * MX_MFMA_NS + cx : 0;
                }
            }
            g_tab_f16 = true;
            std::vector<uint16_t> gx16 = build_G(xs), gy16 = build_Gh(ys);
            g_tab_f16 = false;
            std::vector<uint16_t> gx = build_G(xs), gy = build_Gh(ys);
            if ((rc = dev_alloc(&h->d_gx16[iwvl], gx16.size())) || (rc = dev_alloc(&h->d_gy16[iwvl], gy16.size())))
                return rc;
            HIPCHK(hipMemcpy(h->d_gx16[iwvl], gx16.data(), gx16.size() * 2, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(h->d_gy16[iwvl], gy16.data(), gy16.size() * 2, hipMemcpyHostToDevice));
            if ((rc = dev_alloc(&h->d_gx[iwvl], gx.size())) || (rc = dev_alloc(&h->d_gy[iwvl], gy.size())) ||
                (rc = dev_alloc(&h->d_sic[iwvl], ntab * npix)))
                return rc;
            HIPCHK(hipMemcpy(h->d_gx[iwvl], gx.data(), gx.size() * 2, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(h->d_gy[iwvl], gy.data(), gy.size() * 2, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(h->d_sic[iwvl], sic.data(), sizeof(int4) * ntab * npix, hipMemcpyHostToDevice));
            // the FP32 radix-8 form's last pass: w(uy) exp(+2 pi i uy ys / 64)
            // for uy < 32 at the sampled rows (FP64 angles reduced mod 64),
            // and the sampled columns
            std::vector<float2> fy(32 * MX_MFMA_NS, make_float2(0.f, 0.f));
            for (int uy = 0; uy < 32; uy++)
                for (size_t r = 0; r < ys.size(); r++) {
                    const double w = uy == 0 ? 1. : 2., ang = 2. * M_PI * (double)((uy * ys[r]) % 64) / 64.;
                    fy[uy * MX_MFMA_NS + r] = make_float2((float)(w * std::cos(ang)), (float)(w * std::sin(ang)));
                }
            std::vector<int> fxs(MX_MFMA_NS, xs.empty() ? 0 : xs[0]);
            std::copy(xs.begin(), xs.end(), fxs.begin());
            if ((rc = dev_alloc(&h->d_fy[iwvl], fy.size())) || (rc = dev_alloc(&h->d_fxs[iwvl], fxs.size())))
                return rc;
            HIPCHK(hipMemcpy(h->d_fy[iwvl], fy.data(), sizeof(float2) * fy.size(), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(h->d_fxs[iwvl], fxs.data(), sizeof(int) * fxs.size(), hipMemcpyHostToDevice));
            if (nominal && h->cfg.llt_nx == 0) {
                std::vector<float> nt(ntab * MX_HALF);
                for (long t = 0; t < ntab; t++)
                    for (long uy = 0; uy < N / 2; uy++)
                        for (long ux = 0; ux < N; ux++)
                            nt[t * MX_HALF + half_slot((int)ux, (int)uy)] = (float)nominal[t * N * N + ux + uy * N];
                if ((rc = dev_alloc(&h->d_nominal_t[iwvl], ntab * MX_HALF))) return rc;
                HIPCHK(hipMemcpy(h->d_nominal_t[iwvl], nt.data(), sizeof(float) * ntab * MX_HALF,
                                 hipMemcpyHostToDevice));
            }
        }
    }
    if (h->lf) {
        // all rows / columns past MX_LF_MAXS
        for (std::vector<int> *v : {&xs, &ys})
            if (v->size() > (size_t)MX_LF_MAXS || v->empty()) {
                v->resize(N);
                for (long k = 0; k < N; k++) (*v)[k] = (int)k;
            }
        const int nx = (int)xs.size(), ny = (int)ys.size();
        std::vector<int4> sic(ntab * npix);
        for (long i = 0; i < ntab * npix; i++)
            for (int c = 0; c < 4; c++) {
                const int id = (&idx[i].x)[c];
                const int cx = (int)(std::lower_bound(xs.begin(), xs.end(), id % (int)N) - xs.begin());
                const int cy = (int)(std::lower_bound(ys.begin(), ys.end(), id / (int)N) - ys.begin());
                (&sic[i].x)[c] = (&wt[i].x)[c] != 0.f ? cy * nx + cx : 0;
            }
        for (int **pp : {&h->d_lf_rows[iwvl], &h->d_lf_cols[iwvl]})
            if (*pp) {
                HIPCHK(hipFree(*pp));
                *pp = nullptr;
            }
        if ((rc = dev_alloc(&h->d_lf_rows[iwvl], (size_t)ny)) || (rc = dev_alloc(&h->d_lf_cols[iwvl], (size_t)nx)) ||
            (rc = dev_alloc(&h->d_lf_sic[iwvl], ntab * npix)))
            return rc;
        HIPCHK(hipMemcpy(h->d_lf_rows[iwvl], ys.data(), sizeof(int) * ny, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_lf_cols[iwvl], xs.data(), sizeof(int) * nx, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_lf_sic[iwvl], sic.data(), sizeof(int4) * ntab * npix, hipMemcpyHostToDevice));
        h->lf_ny[iwvl] = ny;
        h->lf_nx[iwvl] = nx;
        // B3 and the image hold the largest sampled set of any wavelength
        int maxny = 0, maxnx = 0;
        for (int w = 0; w < h->nwvl; w++) {
            maxny = std::max(maxny, h->lf_ny[w]);
            maxnx = std::max(maxnx, h->lf_nx[w]);
        }
        if (maxny > h->lf_maxny || maxnx > h->lf_maxnx) {
            const long nchain = (long)h->nwfs * h->nsa * h->nwvl;
            if (h->d_b3) HIPCHK(hipFree(h->d_b3));
            if (h->d_img) HIPCHK(hipFree(h->d_img));
            h->d_b3 = nullptr;
            h->d_img = nullptr;
            h->lf_maxny = h->lf_maxnx = 0;
            if ((rc = dev_alloc(&h->d_b3, (size_t)nchain * maxny * N)) ||
                (rc = dev_alloc(&h->d_img, (size_t)nchain * maxny * maxnx)))
                return rc;
            h->lf_maxny = maxny;
            h->lf_maxnx = maxnx;
        }
    }
    h->dtf_set[iwvl] = true;
    return MX_OK;
}
} // namespace

int mx_wfsgrad_set_dtf(mx_wfsgrad_t *h, int iwvl, const double *nominal, long si_nx, long si_ny,
                       const long *si_pp, const long *si_pi, const double *si_px) {
    if (!h || iwvl < 0 || iwvl >= h->nwvl) return set_err(MX_EINVAL, "mx_wfsgrad_set_dtf: bad handle or iwvl");
    if (!si_pp) return set_err(MX_EINVAL, "mx_wfsgrad_set_dtf: no si");
    return set_dtf_tables(h, iwvl, 1, nominal, si_nx, si_ny, si_pp[h->notf * h->notf], si_pp, si_pi, si_px);
}

int mx_wfsgrad_set_dtf_sa(mx_wfsgrad_t *h, int iwvl, const double *nominal, long si_nx, long si_ny,
                          long si_nzmax, const long *si_pp, const long *si_pi, const double *si_px) {
    if (!h || iwvl < 0 || iwvl >= h->nwvl) return set_err(MX_EINVAL, "mx_wfsgrad_set_dtf_sa: bad handle or iwvl");
    if (si_nzmax < 0) return set_err(MX_EINVAL, "mx_wfsgrad_set_dtf_sa: si_nzmax %ld", si_nzmax);
    return set_dtf_tables(h, iwvl, h->nsa, nominal, si_nx, si_ny, si_nzmax, si_pp, si_pi, si_px);
}

int mx_wfsgrad_update_etf(mx_wfsgrad_t *h, int iwvl, const double *etf, const double *etf2, double etf2wt) {
    if (!h || iwvl < 0 || iwvl >= h->nwvl || !etf) return set_err(MX_EINVAL, "mx_wfsgrad_update_etf: bad argument");
    if (h->cfg.llt_nx == 0) return set_err(MX_EINVAL, "ETF given for a powfs without LLT");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream)); // a queued step may still read the tables replaced here
    const long N = h->notf, nsa = h->nsa;
    std::vector<float2> t(nsa * N * N);
    int rc;
    for (int which = 0; which < 2; which++) {
        const double *src = which ? etf2 : etf;
        if (!src) {
            if (which && h->d_etf2[iwvl]) {
                hipFree(h->d_etf2[iwvl]);
                h->d_etf2[iwvl] = nullptr;
                if (h->d_etf2_t[iwvl]) hipFree(h->d_etf2_t[iwvl]);
                if (h->d_etfb_t[iwvl]) hipFree(h->d_etfb_t[iwvl]);
                h->d_etf2_t[iwvl] = h->d_etfb_t[iwvl] = nullptr;
            }
            continue;
        }
        for (long isa = 0; isa < nsa; isa++) transpose_c(t.data() + isa * N * N, src + 2 * isa * N * N, N, false);
        float2 **dst = which ? &h->d_etf2[iwvl] : &h->d_etf[iwvl];
        if ((rc = dev_alloc(dst, nsa * N * N))) return rc;
        HIPCHK(hipMemcpy(*dst, t.data(), sizeof(float2) * nsa * N * N, hipMemcpyHostToDevice));
        if (h->mfma_size) {
            for (long isa = 0; isa < nsa; isa++)
                for (long uy = 0; uy < N / 2; uy++)
                    for (long ux = 0; ux < N; ux++) {
                        const double *e = src + 2 * (isa * N * N + ux + uy * N);
                        t[isa * MX_HALF + half_slot((int)ux, (int)uy)] = make_float2((float)e[0], (float)e[1]);
                    }
            float2 **dt = which ? &h->d_etf2_t[iwvl] : &h->d_etf_t[iwvl];
            if ((rc = dev_alloc(dt, nsa * MX_HALF))) return rc;
            HIPCHK(hipMemcpy(*dt, t.data(), sizeof(float2) * nsa * MX_HALF, hipMemcpyHostToDevice));
            if (which && (rc = dev_alloc(&h->d_etfb_t[iwvl], nsa * MX_HALF))) return rc;
        }
    }
    h->etf2wt = etf2 ? (float)etf2wt : 0.f;
    return MX_OK;
}

int mx_wfsgrad_set_etf_weight(mx_wfsgrad_t *h, double etf2wt) {
    if (!h || !(etf2wt >= 0.) || !(etf2wt <= 1.)) return set_err(MX_EINVAL, "mx_wfsgrad_set_etf_weight: bad argument");
    bool two = false;
    for (int w = 0; w < h->nwvl; w++) two = two || h->d_etf2[w];
    if (!two) return set_err(MX_ESTATE, "no second ETF (mx_wfsgrad_update_etf with etf2)");
    h->etf2wt = (float)etf2wt;
    return MX_OK;
}

int mx_wfsgrad_update_ref(mx_wfsgrad_t *h, int iwfs, const double *mtche, const double *i0sum) {
    if (!h || iwfs < -1 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_update_ref: bad handle or iwfs");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream)); // a queued step may still read the filter replaced here
    const long n1 = h->nsa * 2 * h->npix;
    const int w0 = iwfs < 0 ? 0 : iwfs, w1 = iwfs < 0 ? h->nwfs : iwfs + 1;
    std::vector<float> t(n1);
    for (int w = w0; w < w1; w++) {
        if (mtche) {
            for (long i = 0; i < n1; i++) t[i] = (float)mtche[i];
            HIPCHK(hipMemcpy(h->d_mtche + w * n1, t.data(), sizeof(float) * n1, hipMemcpyHostToDevice));
        }
        if (i0sum) {
            for (long i = 0; i < h->nsa; i++) t[i] = (float)i0sum[i];
            HIPCHK(hipMemcpy(h->d_i0sum + w * h->nsa, t.data(), sizeof(float) * h->nsa, hipMemcpyHostToDevice));
        }
    }
    if (mtche) h->have_mtche = true;
    if (i0sum) h->have_i0sum = true;
    return MX_OK;
}

int mx_atm2gpu_real(int ireal, const mx_atm_layer_t *layers, int nps, double atmscale) {
    if (ireal < 0 || ireal >= MX_MAXREAL) return set_err(MX_EINVAL, "mx_atm2gpu_real: ireal=%d", ireal);
    if (nps < 0 || nps > MX_MAXLAYER || (nps && !layers)) return set_err(MX_EINVAL, "mx_atm2gpu: nps=%d", nps);
    HIPCHK(hipSetDevice(g_device));
    g_atm_gen++; // uplink OTFs computed ahead from the old screens are stale
    AtmState &A = g_atm[ireal];
    size_t total = 0;
    for (int i = 0; i < nps; i++) {
        if (layers[i].nx < 2 || layers[i].ny < 2 || !layers[i].p || layers[i].dx <= 0 || layers[i].dy <= 0)
            return set_err(MX_EINVAL, "layer %d is invalid", i);
        total += (size_t)layers[i].nx * layers[i].ny;
    }
    // The screens and the layer table are overwritten in place with
    // null-stream copies, which do not wait for the engines' and evaluation
    // handles' non-blocking streams: wait for every queued kernel first.
    HIPCHK(hipDeviceSynchronize());
    if (A.d_data && A.bytes < total * sizeof(float)) {
        HIPCHK(hipFree(A.d_data));
        A.d_data = nullptr;
    }
    if (!A.d_data && total) {
        HIPCHK(hipMalloc(&A.d_data, total * sizeof(float)));
        A.bytes = total * sizeof(float);
    }
    if (!g_dlayers) {
        HIPCHK(hipMalloc(&g_dlayers, sizeof(DevLayer) * MX_MAXREAL * MX_MAXLAYER));
        HIPCHK(hipMalloc(&g_datmscale, sizeof(double) * MX_MAXREAL));
    }
    size_t off = 0;
    std::vector<float> tmp;
    for (int i = 0; i < nps; i++) {
        const mx_atm_layer_t &L = layers[i];
        size_t n = (size_t)L.nx * L.ny;
        tmp.resize(n);
        for (size_t k = 0; k < n; k++) tmp[k] = (float)L.p[k];
        HIPCHK(hipMemcpy(A.d_data + off, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
        DevLayer &D = A.h_layers[i];
        D.p = A.d_data + off;
        D.nx = (int)L.nx;
        D.ny = (int)L.ny;
        D.dx1 = 1. / L.dx;
        D.dy1 = 1. / L.dy;
        D.inx = 1. / (double)L.nx;
        D.iny = 1. / (double)L.ny;
        D.ox = L.ox;
        D.oy = L.oy;
        D.ht = L.ht;
        D.vx = L.vx;
        D.vy = L.vy;
        off += n;
    }
    HIPCHK(hipMemcpy(g_dlayers + (size_t)ireal * MX_MAXLAYER, A.h_layers, sizeof(DevLayer) * MX_MAXLAYER,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g_datmscale + ireal, &atmscale, sizeof(double), hipMemcpyHostToDevice));
    A.nps = nps;
    A.atmscale = atmscale;
    return MX_OK;
}

int mx_atm2gpu(const mx_atm_layer_t *layers, int nps, double atmscale) {
    return mx_atm2gpu_real(0, layers, nps, atmscale);
}

int mx_dm2gpu_real(int ireal, const mx_atm_layer_t *dms, int ndm) {
    if (ireal < 0 || ireal >= MX_MAXREAL) return set_err(MX_EINVAL, "mx_dm2gpu_real: ireal=%d", ireal);
    if (ndm < 0 || ndm > MX_MAXDM || (ndm && !dms)) return set_err(MX_EINVAL, "mx_dm2gpu: ndm=%d (at most %d)", ndm,
                                                                   MX_MAXDM);
    HIPCHK(hipSetDevice(g_device));
    // no queued step may read the surfaces replaced here
    HIPCHK(hipDeviceSynchronize());
    DmState &D = g_dm[ireal];
    size_t total = 0;
    for (int i = 0; i < ndm; i++) {
        if (dms[i].nx < 2 || dms[i].ny < 2 || !dms[i].p || dms[i].dx <= 0 || dms[i].dy <= 0)
            return set_err(MX_EINVAL, "DM %d is invalid", i);
        total += (size_t)dms[i].nx * dms[i].ny;
    }
    if (D.d_data && D.bytes < total * sizeof(float)) {
        HIPCHK(hipFree(D.d_data));
        D.d_data = nullptr;
    }
    if (!D.d_data && total) {
        HIPCHK(hipMalloc(&D.d_data, total * sizeof(float)));
        D.bytes = total * sizeof(float);
    }
    if (!g_ddms) HIPCHK(hipMalloc(&g_ddms, sizeof(DevLayer) * MX_MAXREAL * MX_MAXDM));
    size_t off = 0;
    std::vector<float> tmp;
    for (int i = 0; i < ndm; i++) {
        const mx_atm_layer_t &L = dms[i];
        const size_t n = (size_t)L.nx * L.ny;
        tmp.resize(n);
        for (size_t k = 0; k < n; k++) tmp[k] = (float)L.p[k];
        HIPCHK(hipMemcpy(D.d_data + off, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
        DevLayer &d = D.h_layers[i];
        d = DevLayer{};
        d.p = D.d_data + off;
        d.nx = (int)L.nx;
        d.ny = (int)L.ny;
        d.dx1 = 1. / L.dx;
        d.dy1 = 1. / L.dy;
        d.ox = L.ox;
        d.oy = L.oy;
        d.ht = L.ht;
        off += n;
    }
    HIPCHK(hipMemcpy(g_ddms + (size_t)ireal * MX_MAXDM, D.h_layers, sizeof(DevLayer) * MX_MAXDM,
                     hipMemcpyHostToDevice));
    D.ndm = ndm;
    return MX_OK;
}

int mx_dm2gpu(const mx_atm_layer_t *dms, int ndm) { return mx_dm2gpu_real(0, dms, ndm); }

} // extern "C"

namespace {
// Check the engine's state and fill its argument blocks and kernel list for
// step isim (the body of mx_wfsgrad_queue before any launch).
int plan_step(mx_wfsgrad_t *h, long isim, double dt, StepPlan &sp) {
    if (!h) return set_err(MX_EINVAL, "mx_wfsgrad_queue: null handle");
    for (int r = 0; r < h->nreal; r++) {
        if (g_atm[r].nps < 0) return set_err(MX_ESTATE, "no screens uploaded for realisation %d (mx_atm2gpu)", r);
        if (g_atm[r].nps != g_atm[0].nps)
            return set_err(MX_EINVAL, "realisation %d has %d layers, realisation 0 %d", r, g_atm[r].nps, g_atm[0].nps);
    }
    for (int w = 0; w < h->nwvl && !h->geom; w++)
        if (!h->dtf_set[w]) return set_err(MX_ESTATE, "dtf of wavelength %d not set", w);
    if (h->cfg.llt_nx > 0 && !h->geom)
        for (int w = 0; w < h->nwvl; w++)
            if (!h->d_etf[w]) return set_err(MX_ESTATE, "etf of wavelength %d not set", w);
    // the staged screen tile holds a grid whose step is <= 1 screen pixel
    // and wraps its tile once at most (a screen of >= MX_TILE_MINSCREEN
    // pixels); a coarser grid or a smaller screen gathers from the screens
    // directly
    bool direct = false;
    for (int r = 0; r < h->nreal; r++)
        for (int l = 0; l < g_atm[r].nps; l++) {
            const DevLayer &hl = g_atm[r].h_layers[l];
            if (h->cfg.dx * hl.dx1 > 1. + 1e-9 || h->cfg.dx * hl.dy1 > 1. + 1e-9 || hl.nx < MX_TILE_MINSCREEN ||
                hl.ny < MX_TILE_MINSCREEN)
                direct = true;
        }
    if (h->cfg.phytype == MX_PTYPE_MF && !h->have_mtche && !h->geom)
        return set_err(MX_ESTATE, "matched filter not set");
    if (h->cfg.phytype == MX_PTYPE_MF && h->cfg.sigmatch >= 1 && !h->have_i0sum && !h->geom)
        return set_err(MX_ESTATE, "i0sum not set (matched filter with sigmatch=%d)", h->cfg.sigmatch);
    const int variant = effective_variant(h);
    if (variant < 0 && !h->geom)
        return set_err(MX_EINVAL, "the MFMA form needs notf=64, nxsa=32, <= 64 pixels, <= %d sampled image "
                       "rows and columns and no second ETF", MX_MFMA_NS);
    HIPCHK(hipSetDevice(h->device));
    const long N = h->notf;
    {
        DevWvl hw[MX_MAXWVL] = {};
        for (int w = 0; w < h->nwvl; w++) {
            hw[w].inv_wvl = (float)(1. / h->wvl[w]);
            hw[w].nominal = h->d_nominal[w];
            hw[w].etf = h->d_etf[w];
            hw[w].etf2 = h->d_etf2[w];
            hw[w].si_idx = h->d_si_idx[w];
            hw[w].si_w = h->d_si_w[w];
            hw[w].si_sa = h->dtf_ntab[w] > 1 ? (int)h->npix : 0;
            hw[w].nom_sa = h->dtf_ntab[w] > 1 ? (int)(N * N) : 0;
        }
        if (std::memcmp(hw, h->h_wv, sizeof(hw)) != 0) {
            HIPCHK(hipMemcpyAsync(h->d_wv, hw, sizeof(hw), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            std::memcpy(h->h_wv, hw, sizeof(hw));
        }
    }
    // Argument blocks live in the handle (a graph's kernel nodes copy them).
    const bool llt = h->cfg.llt_nx > 0 && !h->geom; // geometric gradients need no uplink
    OpdArgs &o = h->arg_opd;
    o = OpdArgs{};
    o.nsa = (int)h->nsa;
    o.nwfs = h->nwfs;
    o.nps = g_atm[0].nps;
    o.ptsdx = h->cfg.dx;
    o.ptsx = h->d_ptsx;
    o.ptsy = h->d_ptsy;
    o.layers = g_dlayers;
    o.atmscale = g_datmscale;
    o.wfs = h->d_wfs;
    o.isim = isim;
    o.dt = dt;
    o.opd = h->d_opd;
    o.llt_origx = h->cfg.llt_origx;
    o.llt_origy = h->cfg.llt_origy;
    o.llt_ox = h->cfg.llt_ox;
    o.llt_oy = h->cfg.llt_oy;
    o.llt_opd = h->d_llt_opd;
    // MX_RT_TWOROW=1 in the environment forces the two-row interpolation that
    // layers with a row step over one tile row take (tests compare the forms)
    {
        const char *e = getenv("MX_RT_TWOROW");
        o.tworow = e && atoi(e) != 0;
    }
    // the ray trace covers a subaperture grid with tiles of 32 or 16 points
    // a side (an LLT grid is one tile of nxsa)
    const int opd_nx = h->opd_tile ? h->opd_tile : ((h->lf && h->nxsa % 32 == 0) || h->nxsa == 32 ? 32 : 16);
    o.tiles = (int)(h->nxsa / opd_nx);
    o.llt_tiles = (int)((h->cfg.llt_nx + opd_nx - 1) / opd_nx);
    o.nllt = llt ? h->nwfs * o.llt_tiles * o.llt_tiles : 0;
    KArgs &a = h->arg_step;
    a = KArgs{};
    a.nsa = (int)h->nsa;
    a.npix = (int)h->npix;
    a.pixpsax = (int)h->pixpsax;
    a.pixpsay = (int)h->pixpsay;
    a.nwvl = (int)h->nwvl;
    a.nwfs = h->nwfs;
    a.amp = h->d_amp;
    a.wv = h->d_wv;
    a.etf2wt = h->etf2wt;
    a.etf1wt = 1.f - h->etf2wt;
    a.wfs = h->d_wfs;
    a.lotf = h->cfg.llt_nx > 0 ? h->d_lotf : nullptr;
    a.phytype = h->cfg.phytype;
    a.sigmatch = h->cfg.sigmatch;
    a.cogthres = (float)h->cfg.cogthres;
    a.cogoff = (float)h->cfg.cogoff;
    a.pixthetax = (float)h->cfg.pixthetax;
    a.pixthetay = (float)h->cfg.pixthetay;
    a.sigtot = (float)h->cfg.siglev_dtrat;
    a.mtche = h->d_mtche;
    a.i0sum = h->have_i0sum ? h->d_i0sum : nullptr;
    a.saa = h->d_saa;
    a.saat = (float)h->cfg.saat;
    a.ints = h->d_ints;
    a.grad = h->d_grad;
    a.i0acc = h->i0_enable ? h->d_i0acc : nullptr;
    a.dith = h->dith_on ? DitherAcc{h->d_imb, h->d_imx, h->d_imy, h->d_dcs} : DitherAcc{};
    if (h->dith_on && h->dith_dirty) {
        HIPCHK(hipMemcpyAsync(h->d_dcs, h->h_dcs.data(), sizeof(double) * h->h_dcs.size(), hipMemcpyHostToDevice,
                              h->stream));
        h->dith_dirty = false;
    }
    a.noise = h->noise;
    a.noise.isim = (unsigned)isim;
    // exposure of dtrat steps: isim % dtrat == 0 starts one, (isim + 1) %
    // dtrat == 0 ends it (maos/wfsgrad.c:164, gradout of wfsflags)
    a.dt.acc = h->d_dtacc;
    a.dt.phase = (isim % h->dtrat == 0 ? MX_DT_FIRST : 0) | ((isim + 1) % h->dtrat == 0 ? MX_DT_OUT : 0);
    h->last_out = (a.dt.phase & MX_DT_OUT) != 0;
    a.opd = h->d_opd;
    MfmaArgs &m = h->arg_mfma;
    m = MfmaArgs{};
    const bool f16 = variant == MX_VARIANT_MFMA_FP16;
    // the FP32 form runs the radix-8 kernel on the half-plane tables where
    // they exist, else the register-FFT kernel on the full tables
    // (MX_FFT_REG=1 in the environment keeps the register-FFT kernel, so
    // that tests and the bench can run both FP32 kernels)
    bool fftform = variant == MX_VARIANT_FFT32 && half_tables_ok(h);
    {
        const char *e = getenv("MX_FFT_REG");
        if (e && atoi(e) != 0) fftform = false;
    }
    // (MX_MFMA_PP=0 in the environment: the MFMA imaging kernel with in-place
    // stages, two barriers each, for comparison)
    bool pp = true;
    {
        const char *e = getenv("MX_MFMA_PP");
        if (e && atoi(e) == 0) pp = false;
    }
    // (MX_UPLINK_DIRECT=1 in the environment: the direct-autocorrelation
    // uplink of the wide LLT grid also for an LLT grid of the subaperture's
    // size, for comparison with the transform forms)
    bool uplink_direct = h->cfg.llt_nx != h->nxsa;
    {
        const char *e = getenv("MX_UPLINK_DIRECT");
        if (e && atoi(e) != 0 && !h->lf && h->nwvf == N && N <= 64) uplink_direct = true;
    }
    // The uplink OTF of this step was computed beside the previous step's
    // imaging if that step planned this isim with the same screens, dt and
    // uplink kernel; then the chain skips the LLT trace and the uplink.  The
    // imaging reads buffer pair cur, the next step's uplink goes to 1 - cur.
    const int upform = (variant == MX_VARIANT_MFMA_BF16 ? 1 : 0) | (f16 ? 2 : 0) | (fftform ? 4 : 0) |
                       (uplink_direct ? 8 : 0);
    const bool pre_on = llt && !h->lf && h->uplink_pre && h->d_lotf2 != nullptr;
    const bool pre_hit = pre_on && h->pre_valid && h->pre_isim == isim && h->pre_dt == dt &&
                         h->pre_gen == g_atm_gen && h->pre_form == upform;
    const int cur = pre_hit ? h->pre_buf : 0;
    float2 *const lotf_cur = cur ? h->d_lotf2 : h->d_lotf, *const lotf_t_cur = cur ? h->d_lotf_t2 : h->d_lotf_t;
    float2 *const lotf_nxt = cur ? h->d_lotf : h->d_lotf2, *const lotf_t_nxt = cur ? h->d_lotf_t : h->d_lotf_t2;
    if (pre_hit) o.nllt = 0;
    a.lotf = h->cfg.llt_nx > 0 ? lotf_cur : nullptr;
    if (variant == MX_VARIANT_MFMA_BF16 || f16 || fftform) {
        MfmaWvl hw[MX_MAXWVL] = {};
        for (int w = 0; w < h->nwvl; w++) {
            hw[w].etf_t = h->d_etf2_t[w] ? h->d_etfb_t[w] : h->d_etf_t[w];
            hw[w].nominal_t = h->d_nominal_t[w];
            hw[w].gx = f16 ? h->d_gx16[w] : h->d_gx[w];
            hw[w].gy = f16 ? h->d_gy16[w] : h->d_gy[w];
            hw[w].si_c = h->d_sic[w];
            hw[w].si_w = h->d_si_w[w];
            hw[w].si_sa = h->dtf_ntab[w] > 1 ? (int)h->npix : 0;
            hw[w].nom_sa = h->dtf_ntab[w] > 1 ? MX_HALF : 0;
            hw[w].inv_wvl = (float)(1. / h->wvl[w]);
            hw[w].fy = h->d_fy[w];
            hw[w].fxs = h->d_fxs[w];
        }
        if (std::memcmp(hw, h->h_mwv, sizeof(hw)) != 0) {
            HIPCHK(hipMemcpyAsync(h->d_mwv, hw, sizeof(hw), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            std::memcpy(h->h_mwv, hw, sizeof(hw));
        }
        m.nsa = a.nsa;
        m.npix = a.npix;
        m.pixpsax = a.pixpsax;
        m.pixpsay = a.pixpsay;
        m.nwvl = a.nwvl;
        m.nwfs = a.nwfs;
        m.amp = a.amp;
        m.opd = a.opd;
        m.F = f16 ? h->d_F16 : h->d_F;
        m.wv = h->d_mwv;
        m.lotf_t = h->cfg.llt_nx > 0 ? lotf_t_cur : nullptr;
        m.wfs = a.wfs;
        m.phytype = a.phytype;
        m.sigmatch = a.sigmatch;
        m.cogthres = a.cogthres;
        m.cogoff = a.cogoff;
        m.pixthetax = a.pixthetax;
        m.pixthetay = a.pixthetay;
        m.sigtot = a.sigtot;
        m.mtche = a.mtche;
        m.i0sum = a.i0sum;
        m.saa = a.saa;
        m.saat = a.saat;
        m.ints = a.ints;
        m.grad = a.grad;
        m.i0acc = a.i0acc;
        m.dith = a.dith;
        m.noise = a.noise;
        m.dt = a.dt;
        m.llt_opd = h->d_llt_opd;
        m.llt_amp = h->d_llt_amp;
        m.llt_norm = (float)(1. / ((double)h->cfg.llt_nx * h->embfac * (double)h->cfg.llt_nx * h->embfac));
        m.lotf_full = lotf_cur;
        m.lotf_half = lotf_t_cur;
    }
    LltArgs &la = h->arg_llt;
    la = LltArgs{};
    la.nwvl = (int)h->nwvl;
    la.nwfs = h->nwfs;
    la.amp = h->d_llt_amp;
    la.opd = h->d_llt_opd;
    la.wv = h->d_wv;
    const double nlwvf = (double)h->cfg.llt_nx * h->embfac;
    la.norm = (float)(1. / (nlwvf * nlwvf));
    la.lotf = lotf_cur;
    la.lotf_t = lotf_t_cur;
    // The step in stream order: ray trace (subaperture and LLT grids), uplink
    // OTF (the MFMA kernel's uplink mode with the MFMA form, else the FP32 FFT
    // kernel), imaging.
    StepLaunch *L = sp.L;
    int nl = 0;
    const bool mf = variant == MX_VARIANT_MFMA_BF16 || f16;
    const void *kopd_t[2][2][2] = {
        {{reinterpret_cast<const void *>(&k_wfs_opd<16, false, false>),
          reinterpret_cast<const void *>(&k_wfs_opd<16, false, true>)},
         {reinterpret_cast<const void *>(&k_wfs_opd<16, true, false>),
          reinterpret_cast<const void *>(&k_wfs_opd<16, true, true>)}},
        {{reinterpret_cast<const void *>(&k_wfs_opd<32, false, false>),
          reinterpret_cast<const void *>(&k_wfs_opd<32, false, true>)},
         {reinterpret_cast<const void *>(&k_wfs_opd<32, true, false>),
          reinterpret_cast<const void *>(&k_wfs_opd<32, true, true>)}}};
    const void *kopd = kopd_t[opd_nx == 32][direct][o.tiles > 1];
    // MX_FUSED=1 in the environment: the MFMA imaging and uplink kernels trace
    // their own grids (no ray-trace launch, the OPD staged in LDS), for the
    // MFMA forms on the 32-point grid without DM surfaces or OPD inputs.
    // Slower than the separate ray trace (0.327 against 0.256 ms per lgs30m
    // step), so kept for comparison only.
    bool fused = false;
    {
        const char *e = getenv("MX_FUSED");
        fused = e && atoi(e) != 0 && mf && !h->lf && h->nxsa == 32 && N == 64 && !direct && !h->geom &&
                g_dm[0].ndm == 0 && !h->opdadd_set && !h->ptt_set && !pre_on &&
                (!llt || (h->cfg.llt_nx == 32 && !uplink_direct)) && g_atm[0].nps <= MX_MAXLAYER;
    }
    if (fused) {
        m.tr = o;
        m.opd_out = h->d_opd;
        m.llt_opd_out = h->d_llt_opd;
    } else {
        L[nl++] = {kopd,
                   dim3((unsigned)(h->nsa * h->nwfs * o.tiles * o.tiles + o.nllt)), dim3(MX_OPD_THREADS), &o};
    }
    // the DM surfaces (every realisation of the engine has the same count)
    const int ndm = g_dm[0].ndm;
    for (int r = 1; r < h->nreal; r++)
        if (g_dm[r].ndm != ndm)
            return set_err(MX_EINVAL, "realisation %d has %d DMs, realisation 0 %d", r, g_dm[r].ndm, ndm);
    if (ndm > 0) {
        DmArgs &dm = h->arg_dm;
        dm = DmArgs{};
        dm.nsa = (int)h->nsa;
        dm.nxsa = (int)h->nxsa;
        dm.nwfs = h->nwfs;
        dm.ndm = ndm;
        dm.ptsdx = h->cfg.dx;
        dm.ptsx = h->d_ptsx;
        dm.ptsy = h->d_ptsy;
        dm.dms = g_ddms;
        dm.wfs = h->d_wfs;
        dm.opd = h->d_opd;
        L[nl++] = {reinterpret_cast<const void *>(&k_dm_opd), dim3((unsigned)(h->nsa * h->nwfs)), dim3(64), &dm};
    }
    // static surface and tip/tilt inputs of the WFS OPD
    const bool opdin = h->opdadd_set || h->ptt_set;
    if (opdin) {
        if (h->ptt_dirty) {
            HIPCHK(hipMemcpyAsync(h->d_ptt, h->h_ptt.data(), sizeof(double) * h->h_ptt.size(), hipMemcpyHostToDevice,
                                  h->stream));
            h->ptt_dirty = false;
        }
        OpdAddArgs &q = h->arg_opdadd;
        q = OpdAddArgs{};
        q.nsa = (int)h->nsa;
        q.nxsa = (int)h->nxsa;
        q.nwfs = h->nwfs;
        q.dx = h->cfg.dx;
        q.ptsx = h->d_ptsx;
        q.ptsy = h->d_ptsy;
        q.opdadd = h->opdadd_set ? h->d_opdadd : nullptr;
        q.ptt = h->ptt_set ? h->d_ptt : nullptr;
        q.opd = h->d_opd;
        const long tot = h->nsa * h->nwfs * h->nxsa * h->nxsa;
        L[nl++] = {reinterpret_cast<const void *>(&k_opd_add), dim3((unsigned)std::min<long>((tot + 255) / 256, 65536)),
                   dim3(256), &q};
    }
    sp.slot_opd_last = nl - 1;
    if (h->geom == 2) {
        if (h->gs_nnz < 0) return set_err(MX_ESTATE, "G-tilt gradients need GS0 (mx_wfsgrad_set_gs0)");
        GtiltArgs &q = h->arg_gtilt;
        q = GtiltArgs{};
        q.nsa = (int)h->nsa;
        q.nwfs = h->nwfs;
        q.npts = h->nsa * h->nxsa * h->nxsa;
        q.rowp = h->d_gs_rowp;
        q.coli = h->d_gs_coli;
        q.val = h->d_gs_val;
        q.opd = h->d_opd;
        q.grad = h->d_grad;
        sp.slot_img = nl;
        L[nl++] = {reinterpret_cast<const void *>(&k_gtilt), dim3((unsigned)(2 * h->nsa)), dim3(64), &q};
        sp.nl = sp.nmain = nl;
        sp.key = (1 << 30) | (1 << 29) | (1 << 26) | (opdin ? 1 << 28 : 0) | (ndm ? 1 << 27 : 0) | (direct ? 8 : 0) |
                 (o.tiles > 1 ? 1 << 7 : 0) | (opd_nx << 8);
        sp.slot_llt = -1;
        sp.llt = false;
        return MX_OK;
    }
    if (h->geom) {
        ZtiltArgs &z = h->arg_ztilt;
        z = ZtiltArgs{};
        z.nsa = (int)h->nsa;
        z.nxsa = (int)h->nxsa;
        z.nwfs = h->nwfs;
        z.dx = h->cfg.dx;
        z.ptsx = h->d_ptsx;
        z.ptsy = h->d_ptsy;
        z.amp = h->d_amp;
        z.imcc = h->d_imcc;
        z.opd = h->d_opd;
        z.saa = h->d_saa;
        z.saat = (float)h->cfg.saat;
        z.grad = h->d_grad;
        sp.slot_img = nl;
        L[nl++] = {reinterpret_cast<const void *>(&k_ztilt), dim3((unsigned)(h->nsa * h->nwfs)), dim3(64), &z};
        sp.nl = sp.nmain = nl;
        sp.key = (1 << 30) | (1 << 29) | (opdin ? 1 << 28 : 0) | (ndm ? 1 << 27 : 0) | (direct ? 8 : 0) |
                 (o.tiles > 1 ? 1 << 7 : 0) | (opd_nx << 8);
        sp.slot_llt = -1;
        sp.llt = false;
        return MX_OK;
    }
    if (h->lf) {
        LfWvl hw[MX_MAXWVL] = {};
        for (int w = 0; w < h->nwvl; w++) {
            hw[w].inv_wvl = (float)(1. / h->wvl[w]);
            hw[w].nominal = h->d_nominal[w];
            hw[w].ny = h->lf_ny[w];
            hw[w].nx = h->lf_nx[w];
            hw[w].rows = h->d_lf_rows[w];
            hw[w].cols = h->d_lf_cols[w];
            hw[w].si_c = h->d_lf_sic[w];
            hw[w].si_w = h->d_si_w[w];
            hw[w].si_sa = h->dtf_ntab[w] > 1 ? (int)h->npix : 0;
            hw[w].nom_sa = h->dtf_ntab[w] > 1 ? (int)(N * N) : 0;
            hw[w].etf = h->d_etf[w];
            hw[w].etf2 = h->d_etf2[w];
        }
        if (std::memcmp(hw, h->h_lfw, sizeof(hw)) != 0) {
            HIPCHK(hipMemcpyAsync(h->d_lfw, hw, sizeof(hw), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            std::memcpy(h->h_lfw, hw, sizeof(hw));
        }
        const int nchain = (int)(h->nwfs * h->nsa * h->nwvl);
        LfArgs &f = h->arg_lf_base;
        f = LfArgs{};
        f.pn = h->lf_pn;
        f.pm = h->lf_pm;
        f.nwvf = (int)h->nwvf;
        f.notf = (int)N;
        f.nxsa = (int)h->nxsa;
        f.nsa = (int)h->nsa;
        f.nwfs = h->nwfs;
        f.nwvl = (int)h->nwvl;
        f.nchain = nchain;
        f.maxny = h->lf_maxny;
        f.maxnx = h->lf_maxnx;
        f.opd = h->d_opd;
        f.amp = h->d_amp;
        f.wv = h->d_lfw;
        f.b1 = h->d_b1;
        f.b2 = h->d_b2;
        f.b3 = h->d_b3;
        f.psf = h->d_psf;
        f.img = h->d_img;
        f.psf_scale = 1.f;
        f.lotf = llt ? h->d_lotf : nullptr;
        f.etf1wt = 1.f - h->etf2wt;
        f.etf2wt = h->etf2wt;
    }
    const int slot_llt = llt && !pre_hit ? nl : -1;
    if (llt && h->lf) {
        // uplink OTF of each (WFS, wavelength) by passes A-C and the uplink
        // column pass on the LLT grids (nlwvf == nwvf == notf); pass C takes
        // the PSF rows two at a time, the column passes the columns <= N/2
        const int Ln = lf_lines((int)h->nwvf), Lm = lf_lines((int)N), nch = (int)(h->nwfs * h->nwvl);
        const int nph = (int)(h->nwvf / 2), mh = (int)(N / 2 + 1);
        const int nlx = (int)h->cfg.llt_nx;
        const struct {
            int q, len, L, ngroup;
        } up[4] = {{0, (int)h->nwvf, Ln, (nlx + Ln - 1) / Ln},
                   {1, (int)h->nwvf, Ln, (int)((h->nwvf + Ln - 1) / Ln)},
                   {2, (int)N, Lm, (nph + Lm - 1) / Lm},
                   {5, (int)N, Lm, (mh + Lm - 1) / Lm}};
        for (int q = 0; q < 4; q++) {
            LfArgs &u = h->arg_lf_up[q];
            u = h->arg_lf_base;
            u.nsa = 1;
            u.nxsa = nlx;
            u.nchain = nch;
            u.opd = h->d_llt_opd;
            u.amp = h->d_llt_amp;
            u.psf_scale = (float)(1. / ((double)h->nwvf * (double)h->nwvf));
            u.L = up[q].L;
            u.lgL = ilog2c(up[q].L);
            L[nl++] = {lf_kernel(up[q].q, h->lf_threads, up[q].len), dim3((unsigned)(up[q].ngroup * nch)),
                       dim3((unsigned)h->lf_threads), &u};
        }
    } else if (llt && !pre_hit && uplink_direct) {
        LltAcorrArgs &u = h->arg_lltac;
        u = LltAcorrArgs{};
        u.nlx = (int)h->cfg.llt_nx;
        u.side = o.llt_tiles * opd_nx;
        u.nwvl = (int)h->nwvl;
        u.N = (int)N;
        for (int w = 0; w < h->nwvl; w++) u.inv_wvl[w] = (float)(1. / h->wvl[w]);
        u.opd = h->d_llt_opd;
        u.amp = h->d_llt_amp;
        u.lotf_full = lotf_cur;
        u.lotf_half = lotf_t_cur;
        L[nl++] = {reinterpret_cast<const void *>(&k_llt_acorr), dim3((unsigned)(h->nwfs * h->nwvl * N)), dim3(256),
                   &u};
    } else if (llt && !pre_hit) {
        if (mf && fused)
            L[nl++] = {f16 ? reinterpret_cast<const void *>(&k_shwfs_mfma<true, _Float16, true, true>)
                           : reinterpret_cast<const void *>(&k_shwfs_mfma<true, __bf16, true, true>),
                       dim3((unsigned)(h->nwfs * h->nwvl)), dim3(256), &m};
        else if (mf)
            L[nl++] = {f16 ? reinterpret_cast<const void *>(&k_shwfs_mfma<true, _Float16>)
                           : reinterpret_cast<const void *>(&k_shwfs_mfma<true, __bf16>),
                       dim3((unsigned)(h->nwfs * h->nwvl)),
                       dim3(256), &m};
        else if (fftform)
            L[nl++] = {reinterpret_cast<const void *>(&k_shwfs_fft<true>), dim3((unsigned)(h->nwfs * h->nwvl)),
                       dim3(MX_FFT_T), &m};
        else
            L[nl++] = {N == 64 ? reinterpret_cast<const void *>(&k_llt_otf<64, 32>)
                               : reinterpret_cast<const void *>(&k_llt_otf<32, 16>),
                       dim3((unsigned)(h->nwfs * h->nwvl)), dim3(64), &la};
    }
    const int slot_img = nl;
    if (h->lf) {
        const LfArgs &f = h->arg_lf_base;
        const int Ln = lf_lines((int)h->nwvf), Lm = lf_lines((int)N);
        const int nph = (int)(h->nwvf / 2), mh = (int)(N / 2 + 1);
        // pass: kernel, lines per block, line groups per chain (pass C: pairs
        // of PSF rows; pass D: the columns <= N/2)
        const struct {
            const void *k;
            int L, ngroup;
        } pass[5] = {{lf_kernel(0, h->lf_threads, (int)h->nwvf), Ln, (int)((h->nxsa + Ln - 1) / Ln)},
                     {lf_kernel(1, h->lf_threads, (int)h->nwvf), Ln, (int)((h->nwvf + Ln - 1) / Ln)},
                     {lf_kernel(2, h->lf_threads, (int)N), Lm, (nph + Lm - 1) / Lm},
                     {lf_kernel(3, h->lf_threads, (int)N), Lm, (mh + Lm - 1) / Lm},
                     {lf_kernel(4, h->lf_threads, (int)N), Lm, (h->lf_maxny + Lm - 1) / Lm}};
        const dim3 blk((unsigned)h->lf_threads);
        for (int q = 0; q < 5; q++) {
            h->arg_lf[q] = f;
            h->arg_lf[q].L = pass[q].L;
            h->arg_lf[q].lgL = ilog2c(pass[q].L);
            L[nl++] = {pass[q].k, dim3((unsigned)(pass[q].ngroup * f.nchain)), blk, &h->arg_lf[q]};
        }
        LfSampleArgs &g = h->arg_lfs;
        g = LfSampleArgs{};
        g.nsa = a.nsa;
        g.nwfs = a.nwfs;
        g.nwvl = a.nwvl;
        g.npix = a.npix;
        g.pixpsax = a.pixpsax;
        g.pixpsay = a.pixpsay;
        g.imgstride = h->lf_maxny * h->lf_maxnx;
        g.img = h->d_img;
        g.wv = h->d_lfw;
        g.wfs = h->d_wfs;
        g.phytype = a.phytype;
        g.sigmatch = a.sigmatch;
        g.cogthres = a.cogthres;
        g.cogoff = a.cogoff;
        g.pixthetax = a.pixthetax;
        g.pixthetay = a.pixthetay;
        g.sigtot = a.sigtot;
        g.mtche = a.mtche;
        g.i0sum = a.i0sum;
        g.saa = a.saa;
        g.saat = a.saat;
        g.ints = a.ints;
        g.grad = a.grad;
        g.i0acc = a.i0acc;
        g.dith = a.dith;
        g.noise = a.noise;
        g.dt = a.dt;
        L[nl++] = {reinterpret_cast<const void *>(&k_lf_sample), dim3((unsigned)(h->nwfs * h->nsa)),
                   dim3((unsigned)std::min(h->lf_threads, 256)), &g};
    } else if (mf || fftform) {
        bool two = false;
        for (int w = 0; w < h->nwvl; w++) two = two || h->d_etf2_t[w];
        if (two) {
            EtfBlendArgs &b = h->arg_etfb;
            b = EtfBlendArgs{};
            b.n = h->nsa * MX_HALF;
            b.nwvl = h->nwvl;
            b.w1 = 1.f - h->etf2wt;
            b.w2 = h->etf2wt;
            for (int w = 0; w < h->nwvl; w++) {
                b.e1[w] = h->d_etf_t[w];
                b.e2[w] = h->d_etf2_t[w];
                b.out[w] = h->d_etfb_t[w];
            }
            L[nl++] = {reinterpret_cast<const void *>(&k_etf_blend),
                       dim3((unsigned)std::min<long>((b.n * h->nwvl + 255) / 256, 8192)), dim3(256), &b};
        }
        const void *kimg = fftform ? reinterpret_cast<const void *>(&k_shwfs_fft<false>)
                           : fused ? (f16 ? reinterpret_cast<const void *>(&k_shwfs_mfma<false, _Float16, true, true>)
                                          : reinterpret_cast<const void *>(&k_shwfs_mfma<false, __bf16, true, true>))
                           : f16   ? (pp ? reinterpret_cast<const void *>(&k_shwfs_mfma<false, _Float16, true>)
                                         : reinterpret_cast<const void *>(&k_shwfs_mfma<false, _Float16, false>))
                                   : (pp ? reinterpret_cast<const void *>(&k_shwfs_mfma<false, __bf16, true>)
                                         : reinterpret_cast<const void *>(&k_shwfs_mfma<false, __bf16, false>));
        L[nl++] = {kimg, dim3((unsigned)(h->nsa * h->nwfs)), dim3(256), &m};
    } else {
        const long spw = 64 / N, ngroup = (h->nsa + spw - 1) / spw;
        L[nl++] = {N == 64 ? reinterpret_cast<const void *>(&k_shwfs_step<64, 32, 1>)
                           : reinterpret_cast<const void *>(&k_shwfs_step<32, 16, 2>),
                   dim3((unsigned)(ngroup * h->nwfs)), dim3(64), &a};
    }
    if (h->cfg.sigmatch == 2) {
        Sig2Args &q = h->arg_sig2;
        q = Sig2Args{};
        q.nsa = (int)h->nsa;
        q.npix = (int)h->npix;
        q.phytype = h->cfg.phytype;
        q.pixpsax = (int)h->cfg.pixpsax;
        q.pixpsay = (int)h->cfg.pixpsay;
        q.cogthres = (float)h->cfg.cogthres;
        q.pixthetax = (float)h->cfg.pixthetax;
        q.pixthetay = (float)h->cfg.pixthetay;
        q.saat = (float)h->cfg.saat;
        q.saasum = h->cfg.saasum;
        q.ints = h->d_ints;
        q.i0sum = h->d_i0sum;
        q.saa = h->d_saa;
        q.grad = h->d_grad;
        q.phase = a.dt.phase;
        L[nl++] = {reinterpret_cast<const void *>(&k_sigmatch2), dim3((unsigned)h->nwfs), dim3(256), &q};
    }
    if (h->pistat) {
        // the step's geometric gradients of every subaperture, then the
        // tip/tilt-removed PSFs shifted by them
        if (h->pistat_gtype == 2) {
            if (h->gs_nnz < 0) return set_err(MX_ESTATE, "pistat with G-tilt needs GS0 (mx_wfsgrad_set_gs0)");
            GtiltArgs &q = h->arg_pis_g;
            q = GtiltArgs{};
            q.nsa = (int)h->nsa;
            q.nwfs = h->nwfs;
            q.npts = h->nsa * h->nxsa * h->nxsa;
            q.rowp = h->d_gs_rowp;
            q.coli = h->d_gs_coli;
            q.val = h->d_gs_val;
            q.opd = h->d_opd;
            q.grad = h->d_gpis;
            L[nl++] = {reinterpret_cast<const void *>(&k_gtilt), dim3((unsigned)(2 * h->nsa)), dim3(64), &q};
        } else {
            ZtiltArgs &z = h->arg_pis_z;
            z = ZtiltArgs{};
            z.nsa = (int)h->nsa;
            z.nxsa = (int)h->nxsa;
            z.nwfs = h->nwfs;
            z.dx = h->cfg.dx;
            z.ptsx = h->d_ptsx;
            z.ptsy = h->d_ptsy;
            z.amp = h->d_amp;
            z.imcc = h->d_imcc;
            z.opd = h->d_opd;
            z.saa = h->d_saa;
            z.saat = (float)h->cfg.saat;
            z.grad = h->d_gpis;
            z.all = 1;
            L[nl++] = {reinterpret_cast<const void *>(&k_ztilt), dim3((unsigned)(h->nsa * h->nwfs)), dim3(64), &z};
        }
        PistatArgs &s = h->arg_pis;
        s = PistatArgs{};
        s.nsa = (int)h->nsa;
        s.nwfs = h->nwfs;
        s.nwvl = (int)h->nwvl;
        s.dx = (float)h->cfg.dx;
        const double norm_psf = std::sqrt(h->cfg.areascale) / ((double)h->nxsa * (double)h->nwvf);
        s.norm = (float)(norm_psf * norm_psf);
        for (int w = 0; w < h->nwvl; w++) s.inv_wvl[w] = (float)(1. / h->wvl[w]);
        s.opd = h->d_opd;
        s.amp = h->d_amp;
        s.gref = h->d_gpis;
        s.acc = h->d_pistat;
        L[nl++] = {reinterpret_cast<const void *>(&k_pistat), dim3((unsigned)(h->nsa * h->nwfs)), dim3(MX_FFT_T), &s};
    }
    sp.nmain = nl;
    if (pre_on) {
        // branch: the LLT grids of step isim + 1 (the ray trace's LLT blocks
        // alone), then their uplink OTF into the other buffer pair
        OpdArgs &q = h->arg_opd_pre;
        q = o;
        q.isim = isim + 1;
        q.llt_opd = h->d_llt_opd2;
        q.nllt = h->nwfs * o.llt_tiles * o.llt_tiles;
        L[nl] = {kopd, dim3((unsigned)q.nllt), dim3(MX_OPD_THREADS), &q};
        L[nl++].dep = -1;
        if (uplink_direct) {
            LltAcorrArgs &u = h->arg_lltac_pre;
            u = LltAcorrArgs{};
            u.nlx = (int)h->cfg.llt_nx;
            u.side = o.llt_tiles * opd_nx;
            u.nwvl = (int)h->nwvl;
            u.N = (int)N;
            for (int w = 0; w < h->nwvl; w++) u.inv_wvl[w] = (float)(1. / h->wvl[w]);
            u.opd = h->d_llt_opd2;
            u.amp = h->d_llt_amp;
            u.lotf_full = lotf_nxt;
            u.lotf_half = lotf_t_nxt;
            L[nl++] = {reinterpret_cast<const void *>(&k_llt_acorr), dim3((unsigned)(h->nwfs * h->nwvl * N)),
                       dim3(256), &u};
        } else if (mf || fftform) {
            MfmaArgs &u = h->arg_m_pre;
            u = m;
            u.llt_opd = h->d_llt_opd2;
            u.lotf_full = lotf_nxt;
            u.lotf_half = lotf_t_nxt;
            const void *k = fftform ? reinterpret_cast<const void *>(&k_shwfs_fft<true>)
                            : f16   ? reinterpret_cast<const void *>(&k_shwfs_mfma<true, _Float16>)
                                    : reinterpret_cast<const void *>(&k_shwfs_mfma<true, __bf16>);
            L[nl++] = {k, dim3((unsigned)(h->nwfs * h->nwvl)), dim3(256), &u};
        } else {
            LltArgs &u = h->arg_la_pre;
            u = la;
            u.opd = h->d_llt_opd2;
            u.lotf = lotf_nxt;
            u.lotf_t = lotf_t_nxt;
            L[nl++] = {N == 64 ? reinterpret_cast<const void *>(&k_llt_otf<64, 32>)
                               : reinterpret_cast<const void *>(&k_llt_otf<32, 16>),
                       dim3((unsigned)(h->nwfs * h->nwvl)), dim3(64), &u};
        }
        sp.pre_set = true;
        sp.pre_isim = isim + 1;
        sp.pre_dt = dt;
        sp.pre_form = upform;
        sp.pre_buf = 1 - cur;
        sp.pre_gen = g_atm_gen;
    }
    h->lotf_last = cur;
    if (nl > MX_MAXLAUNCH) return set_err(MX_EINVAL, "internal: %d kernels in a step", nl);
    sp.nl = nl;
    sp.key = (1 << 30) | (opdin ? 1 << 28 : 0) | (ndm ? 1 << 27 : 0) | (N == 64 ? 1 : 0) | (llt ? 2 : 0) | (h->lf ? 4 : 0) |
             (direct ? 8 : 0) | (variant << 4) | (o.tiles > 1 ? 1 << 7 : 0) | (opd_nx << 8) |
             (h->lf ? h->lf_threads << 14 : 0) | (h->cfg.sigmatch == 2 ? 1 << 25 : 0) |
             ((mf || fftform) && h->d_etf2_t[0] ? 1 << 24 : 0) | (fftform ? 1 << 6 : 0) |
             (mf && !pp ? 1 << 14 : 0) | // (lf_threads << 14 sets bits 21-23 only)
             (h->pistat ? 1 << 26 : 0) | (h->pistat && h->pistat_gtype == 2 ? 1 << 15 : 0) |
             (uplink_direct ? 1 << 16 : 0) | (pre_hit ? 1 << 17 : 0) | (pre_on ? 1 << 18 : 0) |
             (fused ? 1 << 19 : 0);
    sp.slot_llt = slot_llt;
    sp.slot_img = slot_img;
    sp.llt = llt;
    return MX_OK;
}

// kernel-node parameters of a plan; argv holds the argument-pointer arrays
void node_params(const StepPlan &sp, hipKernelNodeParams *kp, void *(*argv)[1]) {
    for (int i = 0; i < sp.nl; i++) {
        argv[i][0] = sp.L[i].arg;
        kp[i] = hipKernelNodeParams{};
        kp[i].func = const_cast<void *>(sp.L[i].f);
        kp[i].gridDim = sp.L[i].grid;
        kp[i].blockDim = sp.L[i].block;
        kp[i].sharedMemBytes = 0;
        kp[i].kernelParams = argv[i];
        kp[i].extra = nullptr;
    }
}

// the kernel nodes of a planned step, each after the launch it waits for
int add_step_nodes(hipGraph_t graph, hipGraphNode_t *nd, const StepPlan &sp, const hipKernelNodeParams *kp) {
    for (int i = 0; i < sp.nl; i++) {
        const int d = sp.L[i].dep == -2 ? i - 1 : sp.L[i].dep;
        HIPCHK(hipGraphAddKernelNode(&nd[i], graph, d >= 0 ? &nd[d] : nullptr, d >= 0 ? 1 : 0, &kp[i]));
    }
    return MX_OK;
}

int launch_chain(mx_wfsgrad_t *h, const StepPlan &sp);

// launch a planned step; with a second stream the next step's uplink branch
// runs there, beside the step's chain
int launch_step(mx_wfsgrad_t *h, const StepPlan &in) {
    if (!h->stream2 || in.nl == in.nmain) return launch_chain(h, in);
    StepPlan sp = in;
    sp.nl = sp.nmain;
    // the chain waits for the last branch (it reads that branch's buffers);
    // the branch waits for all work before the chain (the buffer pair it
    // overwrites was read there) and then runs beside the chain
    if (h->branch_pending) HIPCHK(hipStreamWaitEvent(h->stream, h->ev_branch, 0));
    HIPCHK(hipEventRecord(h->ev_main, h->stream));
    int rc = launch_chain(h, sp);
    if (rc) return rc;
    HIPCHK(hipStreamWaitEvent(h->stream2, h->ev_main, 0));
    for (int i = in.nmain; i < in.nl; i++) {
        void *argv[1] = {in.L[i].arg};
        HIPCHK(hipLaunchKernel(in.L[i].f, in.L[i].grid, in.L[i].block, argv, 0, h->stream2));
    }
    HIPCHK(hipEventRecord(h->ev_branch, h->stream2));
    h->branch_pending = true;
    return MX_OK;
}

// the step's launches in stream order: one graph launch, or the kernels one
// by one (with the timing events when timing is on)
int launch_chain(mx_wfsgrad_t *h, const StepPlan &sp) {
    const StepLaunch *L = sp.L;
    const int nl = sp.nl, slot_llt = sp.slot_llt, slot_img = sp.slot_img;
    const bool llt = sp.llt;
    if (h->use_graph && !h->timing) {
        // one graph launch per step: the graph is built for the kernel set and
        // its nodes get this step's arguments
        const int key = sp.key;
        hipKernelNodeParams kp[MX_MAXLAUNCH];
        void *argv[MX_MAXLAUNCH][1];
        node_params(sp, kp, argv);
        if (!h->gexec || h->gkey != key) {
            drop_graph(h);
            HIPCHK(hipGraphCreate(&h->graph, 0));
            int rc = add_step_nodes(h->graph, h->gnode, sp, kp);
            if (rc) return rc;
            HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
            h->gkey = key;
        } else {
            for (int i = 0; i < nl; i++) HIPCHK(hipGraphExecKernelNodeSetParams(h->gexec, h->gnode[i], &kp[i]));
        }
        HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    } else {
        // Timing events: e2/e3 bracket the ray trace, e0/e1 the uplink OTF,
        // e4/e5 the imaging kernels (one, or the passes of the large-field
        // form).
        hipEvent_t ev[6] = {};
        if (h->timing) {
            if (h->ev.size() < 6) {
                for (auto &e : h->ev) HIPCHK(hipEventDestroy(e));
                h->ev.assign(6, nullptr);
                for (auto &e : h->ev) HIPCHK(hipEventCreate(&e));
            }
            for (int i = 0; i < 6; i++) ev[i] = h->ev[i];
        }
        // (no ray-trace launch when the imaging kernels trace their grids:
        // the ray-trace interval is empty)
        if (h->timing && sp.slot_opd_last < 0) {
            HIPCHK(hipEventRecord(ev[2], h->stream));
            HIPCHK(hipEventRecord(ev[3], h->stream));
        }
        for (int i = 0; i < nl; i++) {
            hipEvent_t eb = (i == 0 && sp.slot_opd_last >= 0) ? ev[2]
                            : i == slot_llt                    ? ev[0]
                            : i == slot_img                    ? ev[4]
                                                               : nullptr;
            hipEvent_t ee = i == sp.slot_opd_last                     ? ev[3]
                            : (slot_llt >= 0 && i == slot_img - 1) ? ev[1]
                            : i == sp.nmain - 1                    ? ev[5]
                                                                   : nullptr;
            if (h->timing && eb) HIPCHK(hipEventRecord(eb, h->stream));
            void *argv[1] = {L[i].arg};
            HIPCHK(hipLaunchKernel(L[i].f, L[i].grid, L[i].block, argv, 0, h->stream));
            if (h->timing && ee) HIPCHK(hipEventRecord(ee, h->stream));
        }
        if (h->timing) {
            HIPCHK(hipEventSynchronize(ev[5]));
            float m1 = 0, m2 = 0, m3 = 0;
            // (no uplink in the chain when it was computed beside the last step)
            if (llt && slot_llt >= 0) HIPCHK(hipEventElapsedTime(&m1, ev[0], ev[1]));
            HIPCHK(hipEventElapsedTime(&m2, ev[2], ev[3]));
            HIPCHK(hipEventElapsedTime(&m3, ev[4], ev[5]));
            h->t_llt += m1;
            h->t_opd += m2;
            h->t_ints += m3;
            h->t_n++;
        }
    }
    return MX_OK;
}

// bookkeeping after a step was queued
void step_queued(mx_wfsgrad_t *h, const StepPlan &sp) {
    // what the step's branch computes ahead (valid once the step has run,
    // which every later step on the stream waits for)
    h->pre_valid = sp.pre_set;
    if (sp.pre_set) {
        h->pre_isim = sp.pre_isim;
        h->pre_dt = sp.pre_dt;
        h->pre_form = sp.pre_form;
        h->pre_buf = sp.pre_buf;
        h->pre_gen = sp.pre_gen;
    }
    // the i0 and dither sums take whole exposures
    if (h->i0_enable && h->last_out) h->i0_nframe++;
    if (h->dith_on && h->last_out) h->dith_nframe++;
    if (h->pistat) h->pistat_nstep++; // every step, as MAOS accumulates it
    h->queued = true;
}
} // namespace

extern "C" {

int mx_wfsgrad_queue(mx_wfsgrad_t *h, long isim, double dt) {
    StepPlan sp;
    int rc = plan_step(h, isim, dt, sp);
    if (rc) return rc;
    if ((rc = launch_step(h, sp))) return rc;
    step_queued(h, sp);
    return MX_OK;
}

} // extern "C"

// Several engines stepped by one graph launch: engine e's kernels form
// branch e of the graph, and the branches run concurrently.
struct mx_wfsgroup_s {
    int device = 0;
    std::vector<mx_wfsgrad_t *> h;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;            // the group's step finished
    std::vector<hipEvent_t> ready;        // engine e's earlier work finished
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<hipGraphNode_t> node;     // [e * MX_MAXLAUNCH + i]
    std::vector<int> key;                 // kernel-set key of each branch
};

extern "C" {

int mx_wfsgroup_create(mx_wfsgroup_t **out, mx_wfsgrad_t *const *h, int n) {
    if (!out || !h || n < 1) return set_err(MX_EINVAL, "mx_wfsgroup_create: bad argument");
    *out = nullptr;
    for (int e = 0; e < n; e++) {
        if (!h[e]) return set_err(MX_EINVAL, "mx_wfsgroup_create: engine %d is null", e);
        if (h[e]->device != h[0]->device) return set_err(MX_EINVAL, "engines of a group must share one GPU");
        for (int f = 0; f < e; f++)
            if (h[f] == h[e]) return set_err(MX_EINVAL, "engine %d appears twice in the group", e);
    }
    HIPCHK(hipSetDevice(h[0]->device));
    mx_wfsgroup_t *g = new mx_wfsgroup_t();
    g->device = h[0]->device;
    g->h.assign(h, h + n);
    g->ready.assign(n, nullptr);
    g->node.assign((size_t)n * MX_MAXLAUNCH, nullptr);
    g->key.assign(n, 0);
    bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&g->done, hipEventDisableTiming) == hipSuccess;
    for (int e = 0; ok && e < n; e++) ok = hipEventCreateWithFlags(&g->ready[e], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        mx_wfsgroup_free(g);
        return set_err(MX_EHIP, "mx_wfsgroup_create: stream or event creation failed");
    }
    *out = g;
    return MX_OK;
}

int mx_wfsgroup_queue(mx_wfsgroup_t *g, long isim, double dt) {
    if (!g) return set_err(MX_EINVAL, "mx_wfsgroup_queue: null group");
    HIPCHK(hipSetDevice(g->device));
    const int n = (int)g->h.size();
    std::vector<StepPlan> sp(n);
    for (int e = 0; e < n; e++) {
        int rc = plan_step(g->h[e], isim, dt, sp[e]);
        if (rc) return rc;
    }
    bool same = g->gexec != nullptr;
    for (int e = 0; e < n && same; e++) same = g->key[e] == sp[e].key;
    std::vector<hipKernelNodeParams> kp((size_t)n * MX_MAXLAUNCH);
    std::vector<void *> argv((size_t)n * MX_MAXLAUNCH);
    for (int e = 0; e < n; e++)
        node_params(sp[e], kp.data() + (size_t)e * MX_MAXLAUNCH,
                    reinterpret_cast<void *(*)[1]>(argv.data() + (size_t)e * MX_MAXLAUNCH));
    if (!same) {
        if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->gexec = nullptr;
        g->graph = nullptr;
        HIPCHK(hipGraphCreate(&g->graph, 0));
        for (int e = 0; e < n; e++) {
            hipGraphNode_t *nd = g->node.data() + (size_t)e * MX_MAXLAUNCH;
            int rc = add_step_nodes(g->graph, nd, sp[e], kp.data() + (size_t)e * MX_MAXLAUNCH);
            if (rc) return rc;
            g->key[e] = sp[e].key;
        }
        HIPCHK(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
    } else {
        for (int e = 0; e < n; e++)
            for (int i = 0; i < sp[e].nl; i++)
                HIPCHK(hipGraphExecKernelNodeSetParams(g->gexec, g->node[(size_t)e * MX_MAXLAUNCH + i],
                                                       &kp[(size_t)e * MX_MAXLAUNCH + i]));
    }
    // the step starts after each engine's earlier work; each engine's stream
    // then waits for the step, so mx_wfsgrad_sync / wait see its results
    for (int e = 0; e < n; e++) {
        HIPCHK(hipEventRecord(g->ready[e], g->h[e]->stream));
        HIPCHK(hipStreamWaitEvent(g->stream, g->ready[e], 0));
    }
    HIPCHK(hipGraphLaunch(g->gexec, g->stream));
    HIPCHK(hipEventRecord(g->done, g->stream));
    for (int e = 0; e < n; e++) {
        HIPCHK(hipStreamWaitEvent(g->h[e]->stream, g->done, 0));
        step_queued(g->h[e], sp[e]);
    }
    return MX_OK;
}

int mx_wfsgroup_free(mx_wfsgroup_t *g) {
    if (!g) return MX_OK;
    hipSetDevice(g->device);
    if (g->stream) hipStreamSynchronize(g->stream);
    if (g->gexec) hipGraphExecDestroy(g->gexec);
    if (g->graph) hipGraphDestroy(g->graph);
    for (hipEvent_t e : g->ready)
        if (e) hipEventDestroy(e);
    if (g->done) hipEventDestroy(g->done);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return MX_OK;
}

int mx_wfsgrad_wait(mx_wfsgrad_t *h) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->stream2) HIPCHK(hipStreamSynchronize(h->stream2));
    return MX_OK;
}

int mx_wfsgrad_sync(mx_wfsgrad_t *h, int iwfs, double *grad, double *ints) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_sync: bad handle or iwfs");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long nsa = h->nsa;
    if (grad) {
        h->hbuf.resize(2 * nsa);
        HIPCHK(hipMemcpy(h->hbuf.data(), h->d_grad + iwfs * 2 * nsa, sizeof(float) * 2 * nsa, hipMemcpyDeviceToHost));
        for (long i = 0; i < 2 * nsa; i++) grad[i] = h->hbuf[i];
    }
    if (ints) {
        const long n = nsa * h->npix;
        h->hbuf.resize(n);
        HIPCHK(hipMemcpy(h->hbuf.data(), h->d_ints + iwfs * n, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (long i = 0; i < n; i++) ints[i] = h->hbuf[i];
    }
    return MX_OK;
}

int mx_wfsgrad_i0_control(mx_wfsgrad_t *h, int enable, int reset) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    HIPCHK(hipSetDevice(h->device));
    if (reset) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemset(h->d_i0acc, 0, sizeof(double) * h->nwfs * h->nsa * h->npix));
        h->i0_nframe = 0;
    }
    h->i0_enable = enable != 0;
    return MX_OK;
}

int mx_wfsgrad_i0_get(mx_wfsgrad_t *h, int iwfs, double *out, long *nframe) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_i0_get: bad handle or iwfs");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long n = h->nsa * h->npix;
    if (out) HIPCHK(hipMemcpy(out, h->d_i0acc + iwfs * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (nframe) *nframe = h->i0_nframe;
    return MX_OK;
}

int mx_wfsgrad_pistat_control(mx_wfsgrad_t *h, int enable, int gtype, int reset) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    if (enable && gtype != 1 && gtype != 2)
        return set_err(MX_EINVAL, "mx_wfsgrad_pistat_control: gtype %d (1: Zernike tilt, 2: G-tilt)", gtype);
    if (enable && !(h->notf == 64 && h->nxsa == 32 && h->nwvf == 64))
        return set_err(MX_EINVAL, "pistat needs the 64-point subaperture geometry (nxsa 32, embfac 2, notf 64); "
                       "got nxsa %ld, notf %ld", h->nxsa, h->notf);
    if (enable && h->geom) return set_err(MX_ESTATE, "pistat needs the physical-optics step (geometric mode is on)");
    HIPCHK(hipSetDevice(h->device));
    const size_t n = (size_t)h->nwfs * h->nwvl * h->nsa * 4096;
    if (enable && !h->d_pistat) {
        int rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        if ((rc = dev_alloc(&h->d_pistat, n)) || (rc = dev_alloc(&h->d_gpis, (size_t)2 * h->nwfs * h->nsa)))
            return rc;
        HIPCHK(hipMemset(h->d_pistat, 0, sizeof(float) * n));
        h->pistat_nstep = 0;
    } else if (reset && h->d_pistat) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemset(h->d_pistat, 0, sizeof(float) * n));
        h->pistat_nstep = 0;
    }
    if (enable) h->pistat_gtype = gtype;
    h->pistat = enable != 0;
    return MX_OK;
}

int mx_wfsgrad_pistat_get(mx_wfsgrad_t *h, int iwfs, double *out, long *nstep) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_pistat_get: bad handle or iwfs");
    if (!h->d_pistat) return set_err(MX_ESTATE, "mx_wfsgrad_pistat_get: pistat was never enabled");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nwvl * h->nsa * 4096;
    if (out) {
        std::vector<float> t(n);
        HIPCHK(hipMemcpy(t.data(), h->d_pistat + iwfs * n, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) out[i] = t[i];
    }
    if (nstep) *nstep = h->pistat_nstep;
    return MX_OK;
}

int mx_wfsgrad_psfout(mx_wfsgrad_t *h, int iwfs, double *out) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs || !out) return set_err(MX_EINVAL, "mx_wfsgrad_psfout: bad argument");
    if (h->nwvf != h->notf)
        return set_err(MX_EINVAL, "mx_wfsgrad_psfout needs nxsa*embfac == notf (got %ld and %ld)", h->nwvf, h->notf);
    if (h->geom) return set_err(MX_ESTATE, "mx_wfsgrad_psfout: geometric mode keeps no physical-optics field");
    HIPCHK(hipSetDevice(h->device));
    const int nout = (int)(h->notf / 2 + 2);
    const size_t n = (size_t)h->nwvl * h->nsa * nout * nout;
    int rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    if ((rc = dev_alloc(&h->d_psfout, n))) return rc;
    PsfoutArgs a{};
    a.nsa = (int)h->nsa;
    a.nxsa = (int)h->nxsa;
    a.nout = nout;
    a.b = (int)((h->nwvf - h->nxsa) / 2 - (h->notf / 2 - nout / 2));
    a.norm = (float)(std::sqrt(h->cfg.areascale) / ((double)h->nxsa * (double)h->nwvf));
    for (int w = 0; w < h->nwvl; w++) a.inv_wvl[w] = (float)(1. / h->wvl[w]);
    a.opd = h->d_opd + (long)iwfs * h->nsa * h->nxsa * h->nxsa;
    a.amp = h->d_amp;
    a.out = h->d_psfout;
    void *argv[1] = {&a};
    HIPCHK(hipLaunchKernel(reinterpret_cast<const void *>(&k_psfout), dim3((unsigned)h->nsa, (unsigned)h->nwvl),
                           dim3(256), argv, 0, h->stream));
    std::vector<float2> t(n);
    HIPCHK(hipMemcpyAsync(t.data(), h->d_psfout, sizeof(float2) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; i++) {
        out[2 * i] = t[i].x;
        out[2 * i + 1] = t[i].y;
    }
    return MX_OK;
}

int mx_wfsgrad_dither_control(mx_wfsgrad_t *h, int enable, int reset) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    HIPCHK(hipSetDevice(h->device));
    const size_t n = (size_t)h->nwfs * h->nsa * h->npix;
    double **bufs[] = {&h->d_imb, &h->d_imx, &h->d_imy, &h->d_fi0, &h->d_fgx, &h->d_fgy,
                       &h->d_mi0, &h->d_mgx, &h->d_mgy};
    if (!h->d_imb) {
        int rc;
        for (double **b : bufs)
            if ((rc = dev_alloc(b, n))) return rc;
        if ((rc = dev_alloc(&h->d_nea, (size_t)4 * h->nwfs * h->nsa))) return rc;
        if ((rc = dev_alloc(&h->d_i0s64, (size_t)h->nwfs * h->nsa))) return rc;
        if ((rc = dev_alloc(&h->d_dcs, (size_t)2 * h->nwfs))) return rc;
        h->h_dcs.assign(2 * (size_t)h->nwfs, 0.);
        HIPCHK(hipStreamSynchronize(h->stream));
        for (double **b : bufs) HIPCHK(hipMemset(*b, 0, sizeof(double) * n));
        HIPCHK(hipMemset(h->d_dcs, 0, sizeof(double) * 2 * h->nwfs));
        h->dith_nframe = 0;
    } else if (reset) {
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int k = 0; k < 6; k++) HIPCHK(hipMemset(*bufs[k], 0, sizeof(double) * n));
        h->dith_nframe = 0;
    }
    h->dith_on = enable != 0;
    if (h->dith_on) h->dith_dirty = true;
    return MX_OK;
}

int mx_wfsgrad_set_dither(mx_wfsgrad_t *h, int iwfs, double cs, double ss) {
    if (!h || iwfs < -1 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_set_dither: bad handle or iwfs");
    if (!h->d_imb) return set_err(MX_ESTATE, "dither statistics not set up (mx_wfsgrad_dither_control)");
    if (!std::isfinite(cs) || !std::isfinite(ss)) return set_err(MX_EINVAL, "dither weights %g %g", cs, ss);
    const int w0 = iwfs < 0 ? 0 : iwfs, w1 = iwfs < 0 ? h->nwfs : iwfs + 1;
    for (int w = w0; w < w1; w++) {
        h->h_dcs[2 * w] = cs;
        h->h_dcs[2 * w + 1] = ss;
    }
    h->dith_dirty = true;
    return MX_OK;
}

int mx_wfsgrad_dither_fold(mx_wfsgrad_t *h, int iwfs, double s0, double sg) {
    if (!h || iwfs < -1 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_dither_fold: bad handle or iwfs");
    if (!h->d_imb) return set_err(MX_ESTATE, "dither statistics not set up (mx_wfsgrad_dither_control)");
    HIPCHK(hipSetDevice(h->device));
    const long per = h->nsa * h->npix, off = iwfs < 0 ? 0 : (long)iwfs * per;
    FoldArgs &f = h->arg_fold;
    f = FoldArgs{};
    f.n = iwfs < 0 ? per * h->nwfs : per;
    f.s0 = s0;
    f.sg = sg;
    f.imb = h->d_imb + off;
    f.imx = h->d_imx + off;
    f.imy = h->d_imy + off;
    f.fi0 = h->d_fi0 + off;
    f.fgx = h->d_fgx + off;
    f.fgy = h->d_fgy + off;
    const unsigned nb = (unsigned)std::min<long>((f.n + 255) / 256, 4096);
    k_dither_fold<<<nb, 256, 0, h->stream>>>(f);
    HIPCHK(hipGetLastError());
    return MX_OK;
}

int mx_wfsgrad_set_i0model(mx_wfsgrad_t *h, int iwfs, const double *i0, const double *gx, const double *gy) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_set_i0model: bad handle or iwfs");
    if (!h->d_imb) return set_err(MX_ESTATE, "dither statistics not set up (mx_wfsgrad_dither_control)");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long n = h->nsa * h->npix;
    const double *src[3] = {i0, gx, gy};
    double *dst[3] = {h->d_mi0, h->d_mgx, h->d_mgy};
    for (int k = 0; k < 3; k++) {
        if (src[k])
            HIPCHK(hipMemcpy(dst[k] + (long)iwfs * n, src[k], sizeof(double) * n, hipMemcpyHostToDevice));
        else
            HIPCHK(hipMemset(dst[k] + (long)iwfs * n, 0, sizeof(double) * n));
    }
    return MX_OK;
}

int mx_wfsgrad_mtch_update(mx_wfsgrad_t *h, int iwfs, double keep, double acc, double bkgrnd, double bkgrndc,
                           double rne, int mtchcr, double sigratio) {
    if (!h || iwfs < -1 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_mtch_update: bad handle or iwfs");
    if (!h->d_imb) return set_err(MX_ESTATE, "dither statistics not set up (mx_wfsgrad_dither_control)");
    if (mtchcr < 0 || mtchcr > 3 || !(rne >= 0.) || !(bkgrnd >= 0.) || !std::isfinite(keep) || !std::isfinite(acc))
        return set_err(MX_EINVAL, "mtch_update: mtchcr=%d rne=%g bkgrnd=%g", mtchcr, rne, bkgrnd);
    if (h->npix > 64 * MX_MTCH_THREADS) return set_err(MX_EINVAL, "mtch_update: npix=%ld", h->npix);
    HIPCHK(hipSetDevice(h->device));
    MtchArgs &m = h->arg_mtch;
    m = MtchArgs{};
    m.nsa = (int)h->nsa;
    m.npix = (int)h->npix;
    m.pixpsax = (int)h->cfg.pixpsax;
    m.pixpsay = (int)h->cfg.pixpsay;
    m.iwfs0 = iwfs < 0 ? 0 : iwfs;
    m.nmod = 3;
    if (mtchcr == 1 || mtchcr == 2) {
        m.crx = m.nmod;
        m.nmod += 2;
    }
    if (mtchcr == 1 || mtchcr == 3) {
        m.cry = m.nmod;
        m.nmod += 2;
    }
    m.keep = keep;
    m.acc = acc;
    m.bkres = bkgrnd - bkgrndc;
    m.bkgrnd = bkgrnd;
    m.rne = rne;
    m.rne2 = rne * rne;
    m.kpx = 1. / h->cfg.pixthetax;
    m.kpy = 1. / h->cfg.pixthetay;
    m.sigratio = sigratio ? sigratio : 1.;
    m.nea2thres = h->cfg.pixthetax * h->cfg.pixthetay * 100.;
    m.i0 = h->d_mi0;
    m.gx = h->d_mgx;
    m.gy = h->d_mgy;
    m.fi0 = h->d_fi0;
    m.fgx = h->d_fgx;
    m.fgy = h->d_fgy;
    m.nea = h->d_nea;
    m.i0sum = h->d_i0s64;
    m.mtche = h->d_mtche;
    m.i0sumf = h->d_i0sum;
    const unsigned nw = iwfs < 0 ? (unsigned)h->nwfs : 1u;
    k_mtch<<<nw * (unsigned)h->nsa, MX_MTCH_THREADS, 0, h->stream>>>(m);
    HIPCHK(hipGetLastError());
    k_mtch_fin<<<nw, MX_MTCH_THREADS, 0, h->stream>>>(m);
    HIPCHK(hipGetLastError());
    h->have_mtche = h->have_i0sum = true;
    return MX_OK;
}

int mx_wfsgrad_dither_get(mx_wfsgrad_t *h, int iwfs, double *imb, double *imx, double *imy, long *nframe) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_dither_get: bad handle or iwfs");
    if (!h->d_imb) return set_err(MX_ESTATE, "dither statistics not set up (mx_wfsgrad_dither_control)");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long n = h->nsa * h->npix;
    double *out[3] = {imb, imx, imy};
    const double *src[3] = {h->d_imb, h->d_imx, h->d_imy};
    for (int k = 0; k < 3; k++)
        if (out[k]) HIPCHK(hipMemcpy(out[k], src[k] + (long)iwfs * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (nframe) *nframe = h->dith_nframe;
    return MX_OK;
}

int mx_wfsgrad_mtch_get(mx_wfsgrad_t *h, int iwfs, double *mtche, double *i0sum, double *sanea, double *i0,
                        double *gx, double *gy) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_mtch_get: bad handle or iwfs");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long n = h->nsa * h->npix;
    if (mtche) {
        std::vector<float> t(2 * n);
        HIPCHK(hipMemcpy(t.data(), h->d_mtche + (long)iwfs * 2 * n, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
        for (long i = 0; i < 2 * n; i++) mtche[i] = t[i];
    }
    if ((i0sum || sanea || i0 || gx || gy) && !h->d_imb)
        return set_err(MX_ESTATE, "no GPU matched-filter update yet (mx_wfsgrad_dither_control)");
    if (i0sum) HIPCHK(hipMemcpy(i0sum, h->d_i0s64 + (long)iwfs * h->nsa, sizeof(double) * h->nsa, hipMemcpyDeviceToHost));
    if (sanea) {
        std::vector<double> t(4 * h->nsa);
        HIPCHK(hipMemcpy(t.data(), h->d_nea + (long)iwfs * 4 * h->nsa, sizeof(double) * 4 * h->nsa,
                         hipMemcpyDeviceToHost));
        for (long isa = 0; isa < h->nsa; isa++) {
            sanea[isa] = t[4 * isa];
            sanea[isa + h->nsa] = t[4 * isa + 3];
            sanea[isa + 2 * h->nsa] = t[4 * isa + 1];
        }
    }
    double *out[3] = {i0, gx, gy};
    const double *src[3] = {h->d_mi0, h->d_mgx, h->d_mgy};
    for (int k = 0; k < 3; k++)
        if (out[k]) HIPCHK(hipMemcpy(out[k], src[k] + (long)iwfs * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    return MX_OK;
}

int mx_wfsgrad_debug_opd(mx_wfsgrad_t *h, int iwfs, double *opd) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs || !opd) return set_err(MX_EINVAL, "mx_wfsgrad_debug_opd: bad argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long n = h->nsa * h->nxsa * h->nxsa;
    h->hbuf.resize(n);
    HIPCHK(hipMemcpy(h->hbuf.data(), h->d_opd + iwfs * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (long i = 0; i < n; i++) opd[i] = h->hbuf[i];
    return MX_OK;
}

int mx_wfsgrad_debug_lotf(mx_wfsgrad_t *h, int iwfs, int iwvl, double *lotf) {
    if (!h || iwfs < 0 || iwfs >= h->nwfs || iwvl < 0 || iwvl >= h->nwvl || !lotf)
        return set_err(MX_EINVAL, "mx_wfsgrad_debug_lotf: bad argument");
    if (!h->d_lotf) return set_err(MX_ESTATE, "powfs has no LLT");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long N = h->notf;
    std::vector<float2> t(N * N);
    const float2 *src = h->lotf_last ? h->d_lotf2 : h->d_lotf; // the pair the last step's imaging read
    HIPCHK(hipMemcpy(t.data(), src + ((long)iwfs * h->nwvl + iwvl) * N * N, sizeof(float2) * N * N,
                     hipMemcpyDeviceToHost));
    for (long ux = 0; ux < N; ux++)
        for (long uy = 0; uy < N; uy++) {
            // the large-field form stores ux <= N/2; lotf(-u) = conj lotf(u)
            const bool mirror = h->lf && ux > N / 2;
            const float2 v = mirror ? t[(N - ux) * N + (N - uy) % N] : t[ux * N + uy];
            lotf[2 * (ux + uy * N)] = v.x;
            lotf[2 * (ux + uy * N) + 1] = mirror ? v.y : -v.y; // stored conjugated
        }
    return MX_OK;
}

int mx_wfsgrad_set_variant(mx_wfsgrad_t *h, int variant) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    if (variant != MX_VARIANT_AUTO && variant != MX_VARIANT_FFT32 && variant != MX_VARIANT_MFMA_BF16 &&
        variant != MX_VARIANT_MFMA_FP16)
        return set_err(MX_EINVAL, "unknown variant %d", variant);
    if ((variant == MX_VARIANT_MFMA_BF16 || variant == MX_VARIANT_MFMA_FP16) && !h->mfma_size)
        return set_err(MX_EINVAL, "the MFMA form needs notf=64 and nxsa=32");
    if (variant == MX_VARIANT_MFMA_FP16 && h->amp_max > 1. + 1e-6)
        return set_err(MX_EINVAL, "the FP16 form's range plan needs amplitudes <= 1 (largest %g)", h->amp_max);
    if (variant == MX_VARIANT_MFMA_FP16 && h->llt_sum2 > 1024.)
        return set_err(MX_EINVAL, "the FP16 form's range plan needs lotf(0) = sum(llt_amp^2) <= 1024 (got %g)",
                       h->llt_sum2);
    h->variant = variant;
    return MX_OK;
}

int mx_wfsgrad_get_variant(mx_wfsgrad_t *h, int *variant) {
    if (!h || !variant) return set_err(MX_EINVAL, "null argument");
    *variant = effective_variant(h);
    return MX_OK;
}

int mx_wfsgrad_set_gs0(mx_wfsgrad_t *h, const long *rowp, const long *coli, const double *val, long nnz) {
    if (!h || !rowp || nnz < 0 || (nnz && (!coli || !val)) || nnz > (1L << 30))
        return set_err(MX_EINVAL, "mx_wfsgrad_set_gs0: bad argument");
    const long nrow = 2 * h->nsa, npts = h->nsa * h->nxsa * h->nxsa;
    // the kernel trusts the structure: check every row and column here
    if (rowp[0] != 0 || rowp[nrow] != nnz) return set_err(MX_EINVAL, "GS0: rowp[0]=%ld rowp[%ld]=%ld nnz=%ld",
                                                           rowp[0], nrow, rowp[nrow], nnz);
    for (long r = 0; r < nrow; r++)
        if (rowp[r + 1] < rowp[r]) return set_err(MX_EINVAL, "GS0: rowp decreases at row %ld", r);
    for (long k = 0; k < nnz; k++)
        if (coli[k] < 0 || coli[k] >= npts) return set_err(MX_EINVAL, "GS0: column %ld out of range", coli[k]);
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->d_gs_coli) {
        hipFree(h->d_gs_coli);
        hipFree(h->d_gs_val);
        h->d_gs_coli = nullptr;
        h->d_gs_val = nullptr;
    }
    int rc;
    if ((rc = dev_alloc(&h->d_gs_rowp, (size_t)nrow + 1))) return rc;
    if ((rc = dev_alloc(&h->d_gs_coli, (size_t)std::max<long>(nnz, 1)))) return rc;
    if ((rc = dev_alloc(&h->d_gs_val, (size_t)std::max<long>(nnz, 1)))) return rc;
    std::vector<int> ti(std::max<long>(nnz, nrow + 1));
    for (long r = 0; r <= nrow; r++) ti[r] = (int)rowp[r];
    HIPCHK(hipMemcpy(h->d_gs_rowp, ti.data(), sizeof(int) * (nrow + 1), hipMemcpyHostToDevice));
    for (long k = 0; k < nnz; k++) ti[k] = (int)coli[k];
    if (nnz) {
        HIPCHK(hipMemcpy(h->d_gs_coli, ti.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_gs_val, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
    h->gs_nnz = nnz;
    return MX_OK;
}

int mx_wfsgrad_set_geom(mx_wfsgrad_t *h, int mode) {
    if (!h || mode < 0 || mode > 2) return set_err(MX_EINVAL, "mx_wfsgrad_set_geom: bad handle or mode %d", mode);
    if (mode && h->dtrat > 1)
        return set_err(MX_EINVAL, "mx_wfsgrad_set_geom: geometric gradients are per step (set dtrat 1 first)");
    if (mode && h->pistat)
        return set_err(MX_ESTATE, "mx_wfsgrad_set_geom: pistat needs the physical-optics step (turn it off first)");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->geom = mode;
    return MX_OK;
}

int mx_wfsgrad_set_opdadd(mx_wfsgrad_t *h, int iwfs, const double *opd) {
    if (!h || iwfs < -1 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_set_opdadd: bad handle or iwfs");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const long n = h->nsa * h->nxsa * h->nxsa;
    int rc;
    if ((rc = dev_alloc(&h->d_opdadd, (size_t)h->nwfs * n))) return rc;
    if (!h->opdadd_set) HIPCHK(hipMemset(h->d_opdadd, 0, sizeof(float) * h->nwfs * n));
    std::vector<float> t(n, 0.f);
    if (opd)
        for (long i = 0; i < n; i++) t[i] = (float)opd[i];
    const int w0 = iwfs < 0 ? 0 : iwfs, w1 = iwfs < 0 ? h->nwfs : iwfs + 1;
    for (int w = w0; w < w1; w++)
        HIPCHK(hipMemcpy(h->d_opdadd + (long)w * n, t.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    h->opdadd_set = true;
    return MX_OK;
}

int mx_wfsgrad_set_ptt(mx_wfsgrad_t *h, int iwfs, const double *ptt) {
    if (!h || iwfs < -1 || iwfs >= h->nwfs) return set_err(MX_EINVAL, "mx_wfsgrad_set_ptt: bad handle or iwfs");
    int rc;
    if (!h->d_ptt) {
        HIPCHK(hipSetDevice(h->device));
        if ((rc = dev_alloc(&h->d_ptt, (size_t)3 * h->nwfs))) return rc;
        h->h_ptt.assign(3 * (size_t)h->nwfs, 0.);
    }
    const int w0 = iwfs < 0 ? 0 : iwfs, w1 = iwfs < 0 ? h->nwfs : iwfs + 1;
    for (int w = w0; w < w1; w++)
        for (int k = 0; k < 3; k++) h->h_ptt[3 * w + k] = ptt ? ptt[k] : 0.;
    h->ptt_set = true;
    h->ptt_dirty = true;
    return MX_OK;
}

int mx_wfsgrad_set_noise(mx_wfsgrad_t *h, int enable, double rne, double bkgrnd, double bkgrndc,
                         unsigned long long seed) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    if (enable && (!(rne >= 0.) || !(bkgrnd >= 0.) || !std::isfinite(bkgrndc)))
        return set_err(MX_EINVAL, "noise: rne=%g bkgrnd=%g bkgrndc=%g", rne, bkgrnd, bkgrndc);
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->noise = NoiseArgs{};
    h->noise.on = enable != 0;
    h->noise.rne = (float)rne;
    h->noise.bkgrnd = (float)bkgrnd;
    h->noise.bkgrndc = (float)bkgrndc;
    h->noise.key0 = (unsigned)(seed & 0xffffffffu);
    h->noise.key1 = (unsigned)(seed >> 32);
    return MX_OK;
}

int mx_wfsgrad_set_dtrat(mx_wfsgrad_t *h, int dtrat) {
    if (!h || dtrat < 1) return set_err(MX_EINVAL, "mx_wfsgrad_set_dtrat: bad handle or dtrat=%d", dtrat);
    if (h->geom && dtrat > 1)
        return set_err(MX_EINVAL, "mx_wfsgrad_set_dtrat: geometric gradients are per step (dtrat 1)");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream)); // a queued step may still use the frame sums
    if (dtrat > 1 && !h->d_dtacc) {
        int rc;
        if ((rc = dev_alloc(&h->d_dtacc, (size_t)h->nwfs * h->nsa * h->npix))) return rc;
    }
    h->dtrat = dtrat;
    return MX_OK;
}

int mx_wfsgrad_set_graph(mx_wfsgrad_t *h, int enable) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!enable) drop_graph(h);
    h->use_graph = enable != 0;
    return MX_OK;
}

int mx_wfsgrad_autotune(mx_wfsgrad_t *h, long isim, double dt, int nrep, char *report, long report_len) {
    if (!h || nrep < 1) return set_err(MX_EINVAL, "mx_wfsgrad_autotune: bad argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    struct Cand {
        int variant, tile, threads;
        double us;
    };
    std::vector<Cand> cand;
    const bool llt = h->cfg.llt_nx > 0;
    std::vector<int> tiles;
    if (llt) {
        tiles.push_back((int)h->nxsa); // the LLT grid is one tile
    } else {
        if (h->nxsa % 32 == 0) tiles.push_back(32);
        tiles.push_back(16);
    }
    std::vector<int> variants;
    if (h->lf) {
        variants.push_back(MX_VARIANT_FFT32);
    } else {
        const int saved = h->variant;
        h->variant = MX_VARIANT_MFMA_BF16;
        if (effective_variant(h) == MX_VARIANT_MFMA_BF16) {
            variants.push_back(MX_VARIANT_MFMA_BF16);
            if (h->amp_max <= 1. + 1e-6 && h->llt_sum2 <= 1024.) variants.push_back(MX_VARIANT_MFMA_FP16);
        }
        h->variant = saved;
        variants.push_back(MX_VARIANT_FFT32);
    }
    const std::vector<int> threads = h->lf ? std::vector<int>{128, 256, 512} : std::vector<int>{h->lf_threads};
    for (int v : variants)
        for (int t : tiles)
            for (int th : threads) cand.push_back({v, t, th, 0.});
    const int sv = h->variant, st = h->opd_tile, sth = h->lf_threads;
    const bool si0 = h->i0_enable, stim = h->timing, sdi = h->dith_on;
    const long sn = h->i0_nframe, sdn = h->dith_nframe;
    h->i0_enable = false; // tuning steps leave the i0 and dither sums alone
    h->dith_on = false;
    int rc = MX_OK;
    for (Cand &c : cand) {
        h->variant = c.variant;
        h->opd_tile = c.tile;
        h->lf_threads = c.threads;
        h->timing = false;
        for (int r = 0; r < 2 && !rc; r++) rc = mx_wfsgrad_queue(h, isim, dt); // warm the kernel set
        if (rc) break;
        mx_wfsgrad_timing_enable(h, 1);
        for (int r = 0; r < nrep && !rc; r++) rc = mx_wfsgrad_queue(h, isim + r, dt);
        if (rc) break;
        c.us = (h->t_llt + h->t_opd + h->t_ints) / (double)h->t_n * 1e3;
    }
    h->timing = stim;
    h->t_n = 0;
    h->t_llt = h->t_opd = h->t_ints = 0;
    h->i0_enable = si0;
    h->i0_nframe = sn;
    h->dith_on = sdi;
    h->dith_nframe = sdn;
    if (rc) {
        h->variant = sv;
        h->opd_tile = st;
        h->lf_threads = sth;
        return rc;
    }
    size_t best = 0;
    for (size_t i = 1; i < cand.size(); i++)
        if (cand[i].us < cand[best].us) best = i;
    h->variant = cand[best].variant;
    h->opd_tile = cand[best].tile;
    h->lf_threads = cand[best].threads;
    if (report && report_len > 0) {
        std::string txt;
        char line[160];
        for (size_t i = 0; i < cand.size(); i++) {
            std::snprintf(line, sizeof line, "%sform=%s opd_tile=%d threads=%d step_us=%.2f\n", i == best ? "*" : " ",
                          cand[i].variant == MX_VARIANT_MFMA_BF16   ? "mfma_bf16"
                          : cand[i].variant == MX_VARIANT_MFMA_FP16 ? "mfma_fp16"
                          : h->lf                                   ? "lf_fft32"
                                                                    : "fft32",
                          cand[i].tile, h->lf ? cand[i].threads : 0, cand[i].us);
            txt += line;
        }
        std::snprintf(report, (size_t)report_len, "%s", txt.c_str());
    }
    return MX_OK;
}

int mx_wfsgrad_timing_enable(mx_wfsgrad_t *h, int enable) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    h->timing = enable != 0;
    h->t_n = 0;
    h->t_llt = h->t_opd = h->t_ints = 0;
    return MX_OK;
}

int mx_wfsgrad_timing(mx_wfsgrad_t *h, long *n, double *ms_llt, double *ms_opd, double *ms_ints) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    if (n) *n = h->t_n;
    if (ms_llt) *ms_llt = h->t_llt;
    if (ms_opd) *ms_opd = h->t_opd;
    if (ms_ints) *ms_ints = h->t_ints;
    return MX_OK;
}

} // extern "C"

// science-field evaluation (shares this unit's screens, DM surfaces and helpers)
#include "evl.hip"
#include "pywfs.hip"
