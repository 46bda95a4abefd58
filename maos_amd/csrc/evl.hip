// Science-field evaluation: OPD of evaluation directions on the aperture
// points and the accumulated long-exposure PSF of each (direction,
// wavelength).  Part of the translation unit of shwfs.hip (included at its
// end): it shares the uploaded screens and DM surfaces, the error helpers and
// the frozen-flow coordinate map.
//
// PSF (locfft_psf, lib/locfft.c:110-188, in its GPU form psfcomp_r,
// cuda/sim/perfevl.cu:255-290): the field amp exp(i 2 pi opd / wvl) embedded
// at embed[iloc] of an nembed^2 grid, forward 2-D FFT (exp(-i...)), the
// centre psfsize^2 of it (corner2center_abs2_do, cuda/math/kernel.cu:292-307)
// taken as |.|^2 and added to an FP32 accumulator.  A psfsize of 1 is the
// Strehl-only form (strehlcomp_do, cuda/sim/perfevl.cu:193-219): the complex
// sum of the field, |.|^2 added.  The mean PSF is the sum divided by the
// number of steps and by (sum amp)^2 (psfnorm of locfft_psf with amp
// normalised to sum 1, maos/aper.c:109, and the 1/nacc scale of
// gpu_perfevl_save, cuda/sim/perfevl.cu:625-659).
//
// The 2-D FFT runs in two passes, both in LDS:
//   rows: one workgroup per R rows of the band of grid rows that hold points;
//         the field is scattered from a row-sorted point list straight into
//         LDS (no zero-filled grid in HBM), FFT along x, and only the psfsize
//         columns of the centre window are written, in blocks of C columns
//         ([dir][column block][band row][C]) so that the column pass reads
//         one contiguous block;
//   cols: one workgroup per C columns; the band rows are loaded, the rest of
//         the column is zero, FFT along y, and |.|^2 of the psfsize window
//         rows is added to the accumulator stored transposed
//         ([dir][x][y], contiguous along y for these writes).
// The FFT is a Stockham radix-16 autosort (a final radix-8, 4 or 2 stage
// when log2 N is not a multiple of 4), in place: every stage reads all its
// butterfly inputs into registers, waits, then writes.  One thread per
// radix-16 butterfly (64 to 1024 threads).  LDS element i lives at
// i + i/16 (one pad per 16 complex), which spreads the strided writes of the
// first two stages over the banks.

#define MX_EVL_MAXDIR 64
#define MX_EVL_LDS 8192 // complex elements per workgroup before padding (64 KB)

namespace {

// forward DFTs of 2, 4 and 8 points in registers, exp(-2 pi i jk / r)
__device__ __forceinline__ float2 mul_negi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ void dft_r(float2 (&v)[2]) {
    const float2 t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
}
__device__ __forceinline__ void dft4p(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_negi(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}
__device__ __forceinline__ void dft_r(float2 (&v)[4]) { dft4p(v[0], v[1], v[2], v[3]); }
// 16 points as 4 x 4: DFT4 over x[4 n1 + n2] for each n2, twiddle
// exp(-2 pi i n2 k1 / 16), DFT4 over n2 for each k1; output k1 + 4 k2
__device__ __forceinline__ void dft_r(float2 (&v)[16]) {
    constexpr float c1 = 0.92387953251128676f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
    float2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        a[n2][0] = v[n2];
        a[n2][1] = v[n2 + 4];
        a[n2][2] = v[n2 + 8];
        a[n2][3] = v[n2 + 12];
        dft4p(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
    }
    a[1][1] = cmul(a[1][1], make_float2(c1, -s1));
    a[1][2] = cmul(a[1][2], make_float2(h, -h));
    a[1][3] = cmul(a[1][3], make_float2(s1, -c1));
    a[2][1] = cmul(a[2][1], make_float2(h, -h));
    a[2][2] = mul_negi(a[2][2]);
    a[2][3] = cmul(a[2][3], make_float2(-h, -h));
    a[3][1] = cmul(a[3][1], make_float2(s1, -c1));
    a[3][2] = cmul(a[3][2], make_float2(-h, -h));
    a[3][3] = cmul(a[3][3], make_float2(-c1, s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        dft4p(a[0][k1], a[1][k1], a[2][k1], a[3][k1]);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = a[k2][k1];
    }
}
__device__ __forceinline__ void dft_r(float2 (&v)[8]) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4p(e0, e1, e2, e3);
    dft4p(o0, o1, o2, o3);
    constexpr float h = 0.70710678118654752f;
    // odd half times exp(-2 pi i k / 8), k = 1, 2, 3
    const float2 p1 = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
    const float2 p2 = mul_negi(o2);
    const float2 p3 = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, p1);
    v[5] = csub(e1, p1);
    v[2] = cadd(e2, p2);
    v[6] = csub(e2, p2);
    v[3] = cadd(e3, p3);
    v[7] = csub(e3, p3);
}

// One radix-R Stockham stage over NB rows of N points in LDS (rows at stride
// N).  Butterfly j of a row reads x[j + m N/R], m < R, twiddles input m by
// w^(m k0), w = exp(-2 pi i / (NS R)), k0 = j mod NS, and writes output m to
// (j / NS) NS R + k0 + m NS.  tw[t] = exp(-2 pi i t / N).
__device__ __forceinline__ int lp(int i) { return i + (i >> 4); }

template <int N, int NB, int T, int R, int NS>
__device__ __forceinline__ void fft_stage(float2 *s, const float2 *__restrict__ tw) {
    constexpr int NBF = NB * (N / R);
    constexpr int PER = (NBF + T - 1) / T;
    constexpr int TS = N / (NS * R); // twiddle index step per unit of m k0
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int q = (int)threadIdx.x + p * T;
        if (NBF % T == 0 || q < NBF) {
            const int row = q / (N / R), j = q % (N / R);
            const int src = row * N + j;
#pragma unroll
            for (int m = 0; m < R; m++) v[p][m] = s[lp(src + m * (N / R))];
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int q = (int)threadIdx.x + p * T;
        if (NBF % T == 0 || q < NBF) {
            const int row = q / (N / R), j = q % (N / R);
            const int k0 = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int m = 1; m < R; m++) v[p][m] = cmul(v[p][m], tw[m * k0 * TS]);
            }
            dft_r(v[p]);
            const int dst = row * N + (j / NS) * NS * R + k0;
#pragma unroll
            for (int m = 0; m < R; m++) s[lp(dst + m * NS)] = v[p][m];
        }
    }
    __syncthreads();
}

template <int N, int NB, int T, int NS>
__device__ __forceinline__ void fft_stages(float2 *s, const float2 *tw) {
    if constexpr (NS * 16 <= N) {
        fft_stage<N, NB, T, 16, NS>(s, tw);
        fft_stages<N, NB, T, NS * 16>(s, tw);
    } else if constexpr (NS * 8 == N) {
        fft_stage<N, NB, T, 8, NS>(s, tw);
    } else if constexpr (NS * 4 == N) {
        fft_stage<N, NB, T, 4, NS>(s, tw);
    } else if constexpr (NS * 2 == N) {
        fft_stage<N, NB, T, 2, NS>(s, tw);
    }
}

// forward FFT of NB rows of N points in LDS, in place, natural order; every
// thread of the workgroup takes part (it starts and ends with a barrier)
template <int N, int NB, int T>
__device__ __forceinline__ void lds_fft(float2 *s, const float2 *tw) {
    __syncthreads();
    fft_stages<N, NB, T, 1>(s, tw);
}

// rows (or columns) per workgroup, threads per workgroup (one per radix-16
// butterfly, 64 to 1024) and padded LDS elements
template <int N>
constexpr int evl_rows() {
    return MX_EVL_LDS / N < 16 ? MX_EVL_LDS / N : 16;
}
template <int N>
constexpr int evl_threads() {
    return evl_rows<N>() * N / 16 > 1024 ? 1024 : (evl_rows<N>() * N / 16 < 64 ? 64 : evl_rows<N>() * N / 16);
}
template <int N>
constexpr int evl_lds() {
    return evl_rows<N>() * N + evl_rows<N>() * N / 16;
}
// the column pass: half the LDS of the row pass (4096 points), so that four
// workgroups share a CU and hide each other's HBM round trips
template <int N>
constexpr int evl_cols() {
    return 4096 / N < 16 ? (4096 / N > 1 ? 4096 / N : 1) : 16;
}
template <int N>
constexpr int evl_cthreads() {
    return evl_cols<N>() * N / 16 > 1024 ? 1024 : (evl_cols<N>() * N / 16 < 64 ? 64 : evl_cols<N>() * N / 16);
}
template <int N>
constexpr int evl_clds() {
    return evl_cols<N>() * N + evl_cols<N>() * N / 16;
}

struct PsfRowArgs {
    const float *opd; // [ndir][nloc]
    const float *amp; // [nloc]
    const int2 *ent;  // (band position (iy - ylo) N + ix, point index), sorted by row
    const int *rowp;  // [nband + 1] entry range of each band row
    const float2 *tw; // [N]
    float2 *B;        // [ndir][P / C][nband][C]
    long nloc;
    int ndir, nband, ngroup, P, C;
    double k; // 2 pi / wvl
};

template <int N>
__global__ __launch_bounds__(evl_threads<N>()) void k_psf_rows(PsfRowArgs a) {
    constexpr int R = evl_rows<N>(), T = evl_threads<N>();
    __shared__ float2 s[evl_lds<N>()];
    const int unit = xcd_unit((int)blockIdx.x, a.ngroup * a.ndir);
    const int idir = unit / a.ngroup, b0 = (unit % a.ngroup) * R;
    const int nr = min(R, a.nband - b0);
    for (int i = threadIdx.x; i < evl_lds<N>(); i += T) s[i] = make_float2(0.f, 0.f);
    __syncthreads();
    const int e1 = a.rowp[b0 + nr];
    const float *op = a.opd + (long)idir * a.nloc;
    const double twopi = 6.283185307179586476925;
    // four entries per thread per round, their loads issued together
    for (int e0 = a.rowp[b0] + (int)threadIdx.x; e0 < e1; e0 += 4 * T) {
        int2 en[4];
        float o[4], am[4];
#pragma unroll
        for (int u = 0; u < 4; u++) en[u] = e0 + u * T < e1 ? a.ent[e0 + u * T] : make_int2(-1, 0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            o[u] = op[en[u].y];
            am[u] = a.amp[en[u].y];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (en[u].x < 0) continue;
            // phase in FP64, reduced to [-pi, pi] before the FP32 sincos
            double ph = (double)o[u] * a.k;
            ph -= rint(ph * (1. / twopi)) * twopi;
            float sn, cs;
            sincosf((float)ph, &sn, &cs);
            s[lp(en[u].x - b0 * N)] = make_float2(am[u] * cs, am[u] * sn);
        }
    }
    lds_fft<N, R, T>(s, a.tw);
    // centre window columns c < P: kx = c - P/2 (mod N)
    const int ncb = a.P / a.C, C = a.C;
    float2 *dst = a.B + (long)idir * ncb * a.nband * C;
    for (int i = threadIdx.x; i < ncb * nr * C; i += T) {
        const int cc = i % C, rr = (i / C) % nr, cb = i / (C * nr);
        const int kx = (cb * C + cc - a.P / 2 + N) & (N - 1);
        dst[((long)cb * a.nband + b0 + rr) * C + cc] = s[lp(rr * N + kx)];
    }
}

struct PsfColArgs {
    const float2 *B;  // [ndir][P / C][nband][C]
    const float2 *tw; // [N]
    float *acc;       // [ndir][P (x)][P (y)]
    int ndir, nband, ylo, P, C, ncb;
    int CK, nck; // columns per workgroup (a power of two dividing P, at most evl_cols), P / CK
};

template <int N>
__global__ __launch_bounds__(evl_cthreads<N>()) void k_psf_cols(PsfColArgs a) {
    constexpr int CM = evl_cols<N>(), T = evl_cthreads<N>();
    __shared__ float2 s[evl_clds<N>()];
    const int unit = xcd_unit((int)blockIdx.x, a.nck * a.ndir);
    const int idir = unit / a.nck, c0 = (unit % a.nck) * a.CK;
    const int C = a.C, CK = a.CK;
    // Both HBM round trips are started early: the column values before the
    // LDS is cleared, the accumulator values before the FFT.  Each lane
    // holds at most 16 of each (nband CK <= N CK = 16 T and CK P <= 16 T).
    const int nin = a.nband * CK, nout = CK * a.P;
    float2 pre[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int i = (int)threadIdx.x + u * T;
        if (i < nin) {
            const int c = c0 + i % CK, b = i / CK;
            pre[u] = a.B[(((long)idir * a.ncb + c / C) * a.nband + b) * C + c % C];
        }
    }
    for (int i = threadIdx.x; i < evl_clds<N>(); i += T) s[i] = make_float2(0.f, 0.f);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int i = (int)threadIdx.x + u * T;
        if (i < nin) s[lp((i % CK) * N + a.ylo + i / CK)] = pre[u];
    }
    float *acc = a.acc + ((long)idir * a.P + c0) * a.P; // column cc = i / P, row r = i % P
    float old[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int i = (int)threadIdx.x + u * T;
        if (i < nout) old[u] = acc[i];
    }
    lds_fft<N, CM, T>(s, a.tw);
    // centre window rows r < P: ky = r - P/2 (mod N)
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int i = (int)threadIdx.x + u * T;
        if (i >= nout) continue;
        const int cc = i / a.P, r = i % a.P;
        const float2 v = s[lp(cc * N + ((r - a.P / 2 + N) & (N - 1)))];
        acc[i] = old[u] + (v.x * v.x + v.y * v.y);
    }
}

#define MX_STREHL_PARTS 64
struct StrehlArgs {
    const float *opd;
    const float *amp;
    double *part; // [ndir][MX_STREHL_PARTS][2]
    double *acc1; // [ndir]
    long nloc;
    double k;
};

__device__ __forceinline__ double block_sum_d(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)blockDim.x / 64; i++) t += sh[i];
    return t;
}

// complex sum of amp exp(i k opd) over one slice of the points, FP64
__global__ __launch_bounds__(256) void k_strehl_part(StrehlArgs a) {
    __shared__ double sh[4];
    const int idir = blockIdx.y, part = blockIdx.x;
    const long per = (a.nloc + MX_STREHL_PARTS - 1) / MX_STREHL_PARTS;
    const long i0 = part * per, i1 = min(a.nloc, i0 + per);
    const float *op = a.opd + (long)idir * a.nloc;
    const double twopi = 6.283185307179586476925;
    double re = 0, im = 0;
    for (long i = i0 + threadIdx.x; i < i1; i += 256) {
        double ph = (double)op[i] * a.k;
        ph -= rint(ph * (1. / twopi)) * twopi;
        float sn, cs;
        sincosf((float)ph, &sn, &cs);
        re += (double)a.amp[i] * cs;
        im += (double)a.amp[i] * sn;
    }
    re = block_sum_d(re, sh);
    im = block_sum_d(im, sh);
    if (threadIdx.x == 0) {
        a.part[((long)idir * MX_STREHL_PARTS + part) * 2] = re;
        a.part[((long)idir * MX_STREHL_PARTS + part) * 2 + 1] = im;
    }
}
// |sum|^2 added to the direction's accumulator (parts summed in order)
__global__ __launch_bounds__(64) void k_strehl_fin(StrehlArgs a) {
    const int idir = blockIdx.x;
    if (threadIdx.x) return;
    double re = 0, im = 0;
    for (int p = 0; p < MX_STREHL_PARTS; p++) {
        re += a.part[((long)idir * MX_STREHL_PARTS + p) * 2];
        im += a.part[((long)idir * MX_STREHL_PARTS + p) * 2 + 1];
    }
    a.acc1[idir] += re * re + im * im;
}

// Piston/tip/tilt sums of one slice of the points, FP64 (calc_ptt,
// cuda/sim/perfevl.cu:60-103, in the FP64 form of loc_calc_ptt_stride,
// math/loc.c:653-704): sum amp opd, amp opd x, amp opd y, amp opd^2.
struct PttArgs {
    const float *opd;
    const double *amp, *locx, *locy;
    double *part; // [ndir][MX_STREHL_PARTS][4]
    long nloc;
};
__global__ __launch_bounds__(256) void k_evl_ptt(PttArgs a) {
    __shared__ double sh[4];
    const int idir = blockIdx.y, part = blockIdx.x;
    const long per = (a.nloc + MX_STREHL_PARTS - 1) / MX_STREHL_PARTS;
    const long i0 = part * per, i1 = min(a.nloc, i0 + per);
    const float *op = a.opd + (long)idir * a.nloc;
    double c0 = 0, cx = 0, cy = 0, tot = 0;
    for (long i = i0 + threadIdx.x; i < i1; i += 256) {
        const double o = op[i], w = a.amp[i] * o;
        c0 += w;
        cx += w * a.locx[i];
        cy += w * a.locy[i];
        tot += w * o;
    }
    double *dst = a.part + ((long)idir * MX_STREHL_PARTS + part) * 4;
    c0 = block_sum_d(c0, sh);
    cx = block_sum_d(cx, sh);
    cy = block_sum_d(cy, sh);
    tot = block_sum_d(tot, sh);
    if (threadIdx.x == 0) {
        dst[0] = c0;
        dst[1] = cx;
        dst[2] = cy;
        dst[3] = tot;
    }
}

// The four p/t/t sums of each direction from its MX_STREHL_PARTS partial
// sums, added in part order (the order mx_evl_wfe adds them on the host, so
// the totals are the same numbers): one thread per (direction, sum).
__global__ __launch_bounds__(256) void k_evl_ptt_fin(const double *part, double *tot, int ndir) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= ndir * 4) return;
    const int d = i / 4, k = i % 4;
    double s = 0;
    for (int p = 0; p < MX_STREHL_PARTS; p++) s += part[((long)d * MX_STREHL_PARTS + p) * 4 + k];
    tot[i] = s;
}

// OPD of the evaluation directions on the aperture points: the screens,
// bilinear with periodic wrap at the frozen-flow displacement of step isim
// (prop to aper->locs with alpha = atmscale, maos/perfevl.c:275-300 and
// gpu_atm2loc, cuda/sim/accphi.cu), then the DM surfaces subtracted
// (alpha = -1, no wrap; a point outside a surface takes nothing from it).
// The screen map is the one of the WFS ray trace with the point's own
// coordinate as origin.
struct EvlOpdArgs {
    const DevLayer *layers; // [MX_MAXLAYER]
    const DevLayer *dms;    // [MX_MAXDM]
    const double *locx, *locy;
    float *opd; // [ndir][nloc]
    long nloc, isim;
    double dt, atmscale;
    int nps, ndm;
    double thetax[MX_EVL_MAXDIR], thetay[MX_EVL_MAXDIR], hs[MX_EVL_MAXDIR];
};

__global__ __launch_bounds__(256) void k_evl_opd(EvlOpdArgs a) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int idir = blockIdx.y;
    if (i >= a.nloc) return;
    const double x = a.locx[i], y = a.locy[i];
    const double hs = a.hs[idir], ihs = hs != 0. ? 1. / hs : 0.;
    const double tx = a.thetax[idir], ty = a.thetay[idir];
    float sum = 0.f;
    for (int l = 0; l < a.nps; l++) {
        const DevLayer &L = a.layers[l];
        if (hs != 0. && L.ht > hs) continue;
        int ix, iy;
        float fx, fy;
        screen_coord(L, true, 0, ihs, tx, 0., 0., a.isim, a.dt, 0., x, ix, fx);
        screen_coord(L, false, 0, ihs, ty, 0., 0., a.isim, a.dt, 0., y, iy, fy);
        const int ix1 = ix + 1 == L.nx ? 0 : ix + 1, iy1 = iy + 1 == L.ny ? 0 : iy + 1;
        const float *r0 = L.p + (long)iy * L.nx, *r1 = L.p + (long)iy1 * L.nx;
        const float v0 = fmaf(fx, r0[ix1] - r0[ix], r0[ix]);
        const float v1 = fmaf(fx, r1[ix1] - r1[ix], r1[ix]);
        sum += fmaf(fy, v1 - v0, v0);
    }
    float v = (float)((double)sum * a.atmscale);
    for (int d = 0; d < a.ndm; d++) {
        const DevLayer &L = a.dms[d];
        if (hs != 0. && L.ht > hs) continue;
        const double scale = 1. - L.ht * ihs;
        const double X = (x * scale + L.ht * tx - L.ox) * L.dx1;
        const double Y = (y * scale + L.ht * ty - L.oy) * L.dy1;
        const int wrapx = L.nx - 1, wrapy = L.ny - 1;
        if (!(X >= 0. && X < (double)wrapx && Y >= 0. && Y <= (double)wrapy)) continue;
        const int nx0 = min((int)floor(X), wrapx - 1), ny0 = min((int)floor(Y), wrapy);
        const float fx = (float)(X - nx0), fy = (float)(Y - ny0);
        const float *r0 = L.p + (long)ny0 * L.nx, *r1 = L.p + (long)(ny0 == wrapy ? 0 : ny0 + 1) * L.nx;
        const float v0 = fmaf(fx, r0[nx0 + 1] - r0[nx0], r0[nx0]);
        const float v1 = fmaf(fx, r1[nx0 + 1] - r1[nx0], r1[nx0]);
        v -= fmaf(fy, v1 - v0, v0);
    }
    a.opd[(long)idir * a.nloc + i] = v;
}

struct EvlWvl {
    int N = 0, P = 0, C = 0, nband = 0, ylo = 0, ngroup = 0, ncb = 0, CK = 0, nck = 0;
    double wvl = 0;
    int2 *d_ent = nullptr;
    int *d_rowp = nullptr;
    float2 *d_tw = nullptr;
    float *d_acc = nullptr;   // [ndir][P][P], transposed
    double *d_acc1 = nullptr; // [ndir], psfsize 1
};

// the two passes of one wavelength; with ev (timing) each pass is bracketed
// by events and waited for, and its time added to ms[0] (rows) and ms[1]
// (columns)
template <int N>
int psf_launch(const EvlWvl &w, PsfRowArgs &ra, PsfColArgs &ca, hipStream_t st, hipEvent_t *ev, double *ms) {
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    k_psf_rows<N><<<dim3((unsigned)(w.ngroup * ra.ndir)), dim3(evl_threads<N>()), 0, st>>>(ra);
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    k_psf_cols<N><<<dim3((unsigned)(w.nck * ca.ndir)), dim3(evl_cthreads<N>()), 0, st>>>(ca);
    if (ev) {
        HIPCHK(hipEventRecord(ev[2], st));
        HIPCHK(hipEventSynchronize(ev[2]));
        float a = 0, b = 0;
        HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
        ms[0] += a;
        ms[1] += b;
    }
    return MX_OK;
}

} // namespace

struct mx_evl_s {
    int device = 0;
    hipStream_t stream = nullptr;
    long nloc = 0;
    int ndir = 0, nwvl = 0;
    double ampsum = 0;
    double thetax[MX_EVL_MAXDIR] = {}, thetay[MX_EVL_MAXDIR] = {}, hs[MX_EVL_MAXDIR] = {};
    float *d_amp = nullptr, *d_opd = nullptr;
    double *d_locx = nullptr, *d_locy = nullptr, *d_part = nullptr;
    double *d_amp64 = nullptr, *d_ptt = nullptr; // FP64 amplitude and p/t/t partial sums (with coordinates)
    double imcc[9] = {}, ipcc = 0;               // inverse of sum amp [1 x y]^T [1 x y], 1 / sum amp
    float2 *d_B = nullptr;
    EvlWvl w[MX_MAXWVL];
    long nacc = 0;
    int ireal = 0; // realisation whose screens and DM surfaces mx_evl_trace reads
    std::vector<float> hopd;
    bool timing = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    long t_n = 0;
    double t_trace = 0, t_pass[2] = {0, 0}; // ms: trace; row and column FFT passes
};

namespace {
// inverse of a symmetric 3x3 matrix by cofactors; 1 if singular
int inv3x3(const double *m, double *out) {
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const double scale = fabs(m[0] * m[4] * m[8]) + 1e-300;
    if (!(fabs(det) > 1e-12 * scale)) return 1;
    const double id = 1. / det;
    out[0] = c00 * id;
    out[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    out[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    out[3] = c01 * id;
    out[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    out[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    out[6] = c02 * id;
    out[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    out[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return 0;
}

// calc_ptt_post (cuda/sim/perfevl.cu:105-127) of the sums [ndir][4] =
// (sum amp opd, ... x, ... y, sum amp opd^2)
void wfe_post(const mx_evl_t *h, const double *sums, double *rms, double *coeff) {
    for (int d = 0; d < h->ndir; d++) {
        const double *c = sums + 4 * d, tot = c[3];
        double ic[3];
        for (int r = 0; r < 3; r++) ic[r] = h->imcc[r * 3] * c[0] + h->imcc[r * 3 + 1] * c[1] + h->imcc[r * 3 + 2] * c[2];
        if (coeff)
            for (int r = 0; r < 3; r++) coeff[d * 3 + r] = ic[r];
        if (rms) {
            const double pis = h->ipcc * c[0] * c[0], ptt = c[0] * ic[0] + c[1] * ic[1] + c[2] * ic[2];
            rms[d * 3] = tot - pis;     // piston removed
            rms[d * 3 + 1] = ptt - pis; // tip/tilt
            rms[d * 3 + 2] = tot - ptt; // piston, tip and tilt removed
        }
    }
}

void evl_release(mx_evl_t *h) {
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->d_amp, h->d_opd, h->d_locx, h->d_locy, h->d_part, h->d_B, h->d_amp64, h->d_ptt};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (int iw = 0; iw < MX_MAXWVL; iw++) {
        EvlWvl &w = h->w[iw];
        void *q[] = {w.d_ent, w.d_rowp, w.d_tw, w.d_acc, w.d_acc1};
        for (void *p : q)
            if (p) hipFree(p);
    }
    for (hipEvent_t e : h->ev)
        if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

int evl_psf_queue(mx_evl_t *h, const float *d_opd) {
    const hipStream_t st = h->stream;
    hipEvent_t *ev = h->timing ? h->ev : nullptr;
    for (int iw = 0; iw < h->nwvl; iw++) {
        EvlWvl &w = h->w[iw];
        const double k = 2. * M_PI / w.wvl;
        if (w.P == 1) {
            StrehlArgs sa{d_opd, h->d_amp, h->d_part, w.d_acc1, h->nloc, k};
            k_strehl_part<<<dim3(MX_STREHL_PARTS, (unsigned)h->ndir), dim3(256), 0, st>>>(sa);
            k_strehl_fin<<<dim3((unsigned)h->ndir), dim3(64), 0, st>>>(sa);
            continue;
        }
        PsfRowArgs ra{d_opd, h->d_amp, w.d_ent, w.d_rowp, w.d_tw, h->d_B, h->nloc, h->ndir, w.nband, w.ngroup,
                      w.P, w.C, k};
        PsfColArgs ca{h->d_B, w.d_tw, w.d_acc, h->ndir, w.nband, w.ylo, w.P, w.C, w.ncb, w.CK, w.nck};
        int rc;
        switch (w.N) {
        case 32: rc = psf_launch<32>(w, ra, ca, st, ev, h->t_pass); break;
        case 64: rc = psf_launch<64>(w, ra, ca, st, ev, h->t_pass); break;
        case 128: rc = psf_launch<128>(w, ra, ca, st, ev, h->t_pass); break;
        case 256: rc = psf_launch<256>(w, ra, ca, st, ev, h->t_pass); break;
        case 512: rc = psf_launch<512>(w, ra, ca, st, ev, h->t_pass); break;
        case 1024: rc = psf_launch<1024>(w, ra, ca, st, ev, h->t_pass); break;
        case 2048: rc = psf_launch<2048>(w, ra, ca, st, ev, h->t_pass); break;
        case 4096: rc = psf_launch<4096>(w, ra, ca, st, ev, h->t_pass); break;
        default: return set_err(MX_EINVAL, "internal: nembed %d", w.N);
        }
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    h->nacc++;
    return MX_OK;
}
} // namespace

// a failed HIP call inside mx_evl_init releases the partly built handle
#define EVCHK(call)                                                                   \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(set_err(MX_EHIP, "%s failed: %s", #call, hipGetErrorString(e_))); \
    } while (0)

extern "C" {

int mx_evl_init(mx_evl_t **out, long nloc, const double *locx, const double *locy, const double *amp, int ndir,
                const double *thetax, const double *thetay, const double *hs, int nwvl, const double *wvl,
                const long *nembed, const long *psfsize, const long *embed) {
    if (!out) return set_err(MX_EINVAL, "mx_evl_init: null output");
    *out = nullptr;
    if (nloc < 1 || nloc > (1L << 30) || !amp) return set_err(MX_EINVAL, "mx_evl_init: nloc=%ld", nloc);
    if (ndir < 1 || ndir > MX_EVL_MAXDIR)
        return set_err(MX_EINVAL, "mx_evl_init: ndir=%d (1..%d)", ndir, MX_EVL_MAXDIR);
    if (nwvl < 1 || nwvl > MX_MAXWVL || !wvl || !nembed || !psfsize)
        return set_err(MX_EINVAL, "mx_evl_init: nwvl=%d (1..%d) or a null table", nwvl, MX_MAXWVL);
    for (int iw = 0; iw < nwvl; iw++) {
        const long N = nembed[iw], P = psfsize[iw];
        if (!(wvl[iw] > 0)) return set_err(MX_EINVAL, "mx_evl_init: wvl[%d]=%g", iw, wvl[iw]);
        if (P != 1 && (N < 32 || N > 4096 || (N & (N - 1))))
            return set_err(MX_EINVAL,
                           "mx_evl_init: nembed[%d]=%ld; the PSF FFT takes powers of two from 32 to 4096 "
                           "(give the FFT size explicitly)", iw, N);
        if (P != 1 && (P < 2 || P > N || (P & 1)))
            return set_err(MX_EINVAL, "mx_evl_init: psfsize[%d]=%ld (1, or even and at most nembed)", iw, P);
        if (P != 1 && !embed) return set_err(MX_EINVAL, "mx_evl_init: null embed");
    }
    mx_evl_t *h = new mx_evl_t;
    auto fail = [&](int rc) {
        evl_release(h);
        return rc;
    };
    h->device = g_device;
    if (hipSetDevice(h->device) != hipSuccess) return fail(set_err(MX_EHIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(MX_EHIP, "hipStreamCreate failed"));
    for (hipEvent_t &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(set_err(MX_EHIP, "hipEventCreate failed"));
    h->nloc = nloc;
    h->ndir = ndir;
    h->nwvl = nwvl;
    for (int d = 0; d < ndir; d++) {
        h->thetax[d] = thetax ? thetax[d] : 0.;
        h->thetay[d] = thetay ? thetay[d] : 0.;
        h->hs[d] = hs ? hs[d] : 0.;
        if (h->hs[d] < 0 || !std::isfinite(h->hs[d]))
            return fail(set_err(MX_EINVAL, "mx_evl_init: hs[%d]=%g (0 or a finite height)", d, h->hs[d]));
    }
    std::vector<float> famp(nloc);
    double asum = 0;
    for (long i = 0; i < nloc; i++) {
        famp[i] = (float)amp[i];
        asum += amp[i];
    }
    h->ampsum = asum;
    int rc;
    if ((rc = dev_alloc(&h->d_amp, nloc)) || (rc = dev_alloc(&h->d_opd, (size_t)nloc * ndir)) ||
        (rc = dev_alloc(&h->d_part, (size_t)ndir * MX_STREHL_PARTS * 2)))
        return fail(rc);
    EVCHK(hipMemcpy(h->d_amp, famp.data(), sizeof(float) * nloc, hipMemcpyHostToDevice));
    EVCHK(hipMemset(h->d_opd, 0, sizeof(float) * nloc * ndir));
    if (locx && locy) {
        if ((rc = dev_alloc(&h->d_locx, nloc)) || (rc = dev_alloc(&h->d_locy, nloc))) return fail(rc);
        EVCHK(hipMemcpy(h->d_locx, locx, sizeof(double) * nloc, hipMemcpyHostToDevice));
        EVCHK(hipMemcpy(h->d_locy, locy, sizeof(double) * nloc, hipMemcpyHostToDevice));
        // the p/t/t weights: mcc = sum amp [1 x y]^T [1 x y] (loc_mcc_ptt,
        // math/loc.c:612-650), inverted in FP64
        double m[9] = {};
        for (long i = 0; i < nloc; i++) {
            const double v[3] = {1., locx[i], locy[i]};
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) m[r * 3 + c] += amp[i] * v[r] * v[c];
        }
        if (!(m[0] > 0)) return fail(set_err(MX_EINVAL, "mx_evl_init: the amplitudes sum to %g", m[0]));
        if (inv3x3(m, h->imcc)) return fail(set_err(MX_EINVAL, "mx_evl_init: the points span no plane"));
        h->ipcc = 1. / m[0];
        if ((rc = dev_alloc(&h->d_amp64, nloc)) || (rc = dev_alloc(&h->d_ptt, (size_t)ndir * MX_STREHL_PARTS * 4)))
            return fail(rc);
        EVCHK(hipMemcpy(h->d_amp64, amp, sizeof(double) * nloc, hipMemcpyHostToDevice));
    }
    size_t bmax = 0;
    for (int iw = 0; iw < nwvl; iw++) {
        EvlWvl &w = h->w[iw];
        w.wvl = wvl[iw];
        w.P = (int)psfsize[iw];
        if (w.P == 1) {
            if ((rc = dev_alloc(&w.d_acc1, ndir))) return fail(rc);
            EVCHK(hipMemset(w.d_acc1, 0, sizeof(double) * ndir));
            continue;
        }
        const long N = nembed[iw];
        w.N = (int)N;
        // the points sorted by (grid position, point index); where two points
        // share a position the later one wins, as in the reference's
        // sequential embedding
        const long *em = embed + (size_t)iw * nloc;
        std::vector<std::pair<long, int>> pe(nloc);
        for (long i = 0; i < nloc; i++) {
            if (em[i] < 0 || em[i] >= N * N)
                return fail(set_err(MX_EINVAL, "mx_evl_init: embed[%d][%ld]=%ld outside %ld^2", iw, i, em[i], N));
            pe[i] = {em[i], (int)i};
        }
        std::sort(pe.begin(), pe.end());
        std::vector<std::pair<long, int>> uq;
        uq.reserve(nloc);
        for (long i = 0; i < nloc; i++)
            if (i + 1 == nloc || pe[i + 1].first != pe[i].first) uq.push_back(pe[i]);
        const long ylo = uq.front().first / N, yhi = uq.back().first / N;
        w.ylo = (int)ylo;
        w.nband = (int)(yhi - ylo + 1);
        std::vector<int2> ent(uq.size());
        std::vector<int> rowp(w.nband + 1, 0);
        for (size_t i = 0; i < uq.size(); i++) {
            const long pos = uq[i].first - ylo * N;
            ent[i] = make_int2((int)pos, uq[i].second);
            rowp[pos / N + 1]++;
        }
        for (int b = 0; b < w.nband; b++) rowp[b + 1] += rowp[b];
        const int R = MX_EVL_LDS / w.N < 16 ? MX_EVL_LDS / w.N : 16;
        w.ngroup = (w.nband + R - 1) / R;
        // column blocks of the scratch array: the widest power of two up to
        // the row pass's rows that divides psfsize; columns per column-pass
        // workgroup: the same up to its LDS columns (4096 / N, at most 16).
        // Both are powers of two, so one divides the other.
        int C = R;
        while (w.P % C) C >>= 1;
        w.C = C;
        w.ncb = w.P / C;
        int CK = 4096 / w.N < 16 ? (4096 / w.N > 1 ? 4096 / w.N : 1) : 16;
        while (w.P % CK) CK >>= 1;
        w.CK = CK;
        w.nck = w.P / CK;
        {
            // the column pass keeps at most 16 values per lane of each kind
            const int cm = 4096 / w.N < 16 ? (4096 / w.N > 1 ? 4096 / w.N : 1) : 16;
            const int T = cm * w.N / 16 > 1024 ? 1024 : (cm * w.N / 16 < 64 ? 64 : cm * w.N / 16);
            if ((long)w.nband * CK > 16L * T || (long)CK * w.P > 16L * T)
                return fail(set_err(MX_EINVAL, "internal: column pass of nembed %d holds too many values", w.N));
        }
        std::vector<float2> tw(N);
        for (long t = 0; t < N; t++) {
            const double a = -2. * M_PI * (double)t / (double)N;
            tw[t] = make_float2((float)cos(a), (float)sin(a));
        }
        if ((rc = dev_alloc(&w.d_ent, ent.size())) || (rc = dev_alloc(&w.d_rowp, rowp.size())) ||
            (rc = dev_alloc(&w.d_tw, (size_t)N)) || (rc = dev_alloc(&w.d_acc, (size_t)ndir * w.P * w.P)))
            return fail(rc);
        EVCHK(hipMemcpy(w.d_ent, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice));
        EVCHK(hipMemcpy(w.d_rowp, rowp.data(), sizeof(int) * rowp.size(), hipMemcpyHostToDevice));
        EVCHK(hipMemcpy(w.d_tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
        EVCHK(hipMemset(w.d_acc, 0, sizeof(float) * ndir * w.P * w.P));
        bmax = std::max(bmax, (size_t)ndir * w.nband * w.P);
    }
    if (bmax && (rc = dev_alloc(&h->d_B, bmax))) return fail(rc);
    EVCHK(hipDeviceSynchronize());
    *out = h;
    return MX_OK;
}

int mx_evl_free(mx_evl_t *h) {
    evl_release(h);
    return MX_OK;
}

int mx_evl_psf_add(mx_evl_t *h, const double *opd) {
    if (!h || !opd) return set_err(MX_EINVAL, "mx_evl_psf_add: null argument");
    HIPCHK(hipSetDevice(h->device));
    const size_t n = (size_t)h->nloc * h->ndir;
    h->hopd.resize(n);
    for (size_t i = 0; i < n; i++) h->hopd[i] = (float)opd[i];
    HIPCHK(hipStreamSynchronize(h->stream)); // the previous step no longer reads the OPD buffer
    HIPCHK(hipMemcpy(h->d_opd, h->hopd.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    return evl_psf_queue(h, h->d_opd);
}

int mx_evl_psf_add_device(mx_evl_t *h, const float *d_opd) {
    if (!h || !d_opd) return set_err(MX_EINVAL, "mx_evl_psf_add_device: null argument");
    HIPCHK(hipSetDevice(h->device));
    return evl_psf_queue(h, d_opd);
}

int mx_evl_trace(mx_evl_t *h, long isim, double dt, int add_psf) {
    if (!h) return set_err(MX_EINVAL, "mx_evl_trace: null handle");
    if (!h->d_locx) return set_err(MX_ESTATE, "mx_evl_trace: the handle was made without point coordinates");
    const int r = h->ireal;
    if (g_atm[r].nps < 0) return set_err(MX_ESTATE, "mx_evl_trace: no screens uploaded for realisation %d", r);
    HIPCHK(hipSetDevice(h->device));
    EvlOpdArgs a{};
    a.layers = g_dlayers + (size_t)r * MX_MAXLAYER;
    a.dms = g_dm[r].ndm ? g_ddms + (size_t)r * MX_MAXDM : nullptr;
    a.locx = h->d_locx;
    a.locy = h->d_locy;
    a.opd = h->d_opd;
    a.nloc = h->nloc;
    a.isim = isim;
    a.dt = dt;
    a.atmscale = g_atm[r].atmscale;
    a.nps = g_atm[r].nps;
    a.ndm = g_dm[r].ndm;
    for (int d = 0; d < h->ndir; d++) {
        a.thetax[d] = h->thetax[d];
        a.thetay[d] = h->thetay[d];
        a.hs[d] = h->hs[d];
    }
    if (h->timing) HIPCHK(hipEventRecord(h->ev[0], h->stream));
    k_evl_opd<<<dim3((unsigned)((h->nloc + 255) / 256), (unsigned)h->ndir), dim3(256), 0, h->stream>>>(a);
    HIPCHK(hipGetLastError());
    if (h->timing) {
        HIPCHK(hipEventRecord(h->ev[1], h->stream));
        HIPCHK(hipEventSynchronize(h->ev[1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        h->t_trace += ms;
        h->t_n++;
    }
    return add_psf ? evl_psf_queue(h, h->d_opd) : MX_OK;
}

int mx_evl_set_real(mx_evl_t *h, int ireal) {
    if (!h) return set_err(MX_EINVAL, "mx_evl_set_real: null handle");
    if (ireal < 0 || ireal >= MX_MAXREAL) return set_err(MX_EINVAL, "mx_evl_set_real: ireal=%d", ireal);
    h->ireal = ireal;
    return MX_OK;
}

int mx_evl_opd_get(mx_evl_t *h, double *opd) {
    if (!h || !opd) return set_err(MX_EINVAL, "mx_evl_opd_get: null argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nloc * h->ndir;
    std::vector<float> t(n);
    HIPCHK(hipMemcpy(t.data(), h->d_opd, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) opd[i] = t[i];
    return MX_OK;
}

int mx_evl_wfe(mx_evl_t *h, double *rms, double *coeff) {
    if (!h || (!rms && !coeff)) return set_err(MX_EINVAL, "mx_evl_wfe: null argument");
    if (!h->d_ptt) return set_err(MX_ESTATE, "mx_evl_wfe: the handle was made without point coordinates");
    HIPCHK(hipSetDevice(h->device));
    PttArgs a{h->d_opd, h->d_amp64, h->d_locx, h->d_locy, h->d_ptt, h->nloc};
    k_evl_ptt<<<dim3(MX_STREHL_PARTS, (unsigned)h->ndir), dim3(256), 0, h->stream>>>(a);
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)h->ndir * MX_STREHL_PARTS * 4);
    HIPCHK(hipMemcpyAsync(part.data(), h->d_ptt, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> sums((size_t)h->ndir * 4, 0.);
    for (int d = 0; d < h->ndir; d++)
        for (int p = 0; p < MX_STREHL_PARTS; p++)
            for (int k = 0; k < 4; k++) sums[d * 4 + k] += part[((size_t)d * MX_STREHL_PARTS + p) * 4 + k];
    wfe_post(h, sums.data(), rms, coeff);
    return MX_OK;
}

int mx_evl_wfe_noalloc(mx_evl_t *h, const float *d_opd, double *rms, double *coeff) {
    if (!h || !d_opd || (!rms && !coeff)) return set_err(MX_EINVAL, "mx_evl_wfe_noalloc: null argument");
    if (!h->d_amp64) return set_err(MX_ESTATE, "mx_evl_wfe_noalloc: the handle was made without point coordinates");
    if ((long)h->ndir * (MX_STREHL_PARTS + 1) * 4 > MX_SCRATCH_DOUBLES)
        return set_err(MX_EINVAL, "mx_evl_wfe_noalloc: %d directions need more than the %d-double scratch array",
                       h->ndir, MX_SCRATCH_DOUBLES);
    HIPCHK(hipSetDevice(h->device));
    double *sc = nullptr;
    if (const int rc = scratch_on(h->device, &sc)) return rc;
    // scratch: partial sums [ndir][MX_STREHL_PARTS][4], then totals [ndir][4];
    // every entry read is written first by this call
    double *part = sc, *tot = sc + (size_t)h->ndir * MX_STREHL_PARTS * 4;
    PttArgs a{d_opd, h->d_amp64, h->d_locx, h->d_locy, part, h->nloc};
    k_evl_ptt<<<dim3(MX_STREHL_PARTS, (unsigned)h->ndir), dim3(256), 0, h->stream>>>(a);
    HIPCHK(hipGetLastError());
    k_evl_ptt_fin<<<dim3((unsigned)((h->ndir * 4 + 255) / 256)), dim3(256), 0, h->stream>>>(part, tot, h->ndir);
    HIPCHK(hipGetLastError());
    double sums[MX_EVL_MAXDIR * 4];
    HIPCHK(hipMemcpyAsync(sums, tot, sizeof(double) * 4 * h->ndir, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    wfe_post(h, sums, rms, coeff);
    return MX_OK;
}

int mx_evl_psf_get(mx_evl_t *h, int idir, int iwvl, double *psf, long *nacc) {
    if (!h || !psf) return set_err(MX_EINVAL, "mx_evl_psf_get: null argument");
    if (idir < 0 || idir >= h->ndir || iwvl < 0 || iwvl >= h->nwvl)
        return set_err(MX_EINVAL, "mx_evl_psf_get: idir=%d iwvl=%d", idir, iwvl);
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (nacc) *nacc = h->nacc;
    const EvlWvl &w = h->w[iwvl];
    const double sc = h->nacc ? 1. / ((double)h->nacc * h->ampsum * h->ampsum) : 0.;
    if (w.P == 1) {
        double v = 0;
        HIPCHK(hipMemcpy(&v, w.d_acc1 + idir, sizeof(double), hipMemcpyDeviceToHost));
        psf[0] = v * sc;
        return MX_OK;
    }
    const long P = w.P;
    std::vector<float> t((size_t)P * P);
    HIPCHK(hipMemcpy(t.data(), w.d_acc + (size_t)idir * P * P, sizeof(float) * P * P, hipMemcpyDeviceToHost));
    for (long ix = 0; ix < P; ix++)
        for (long iy = 0; iy < P; iy++) psf[ix + iy * P] = t[ix * P + iy] * sc;
    return MX_OK;
}

int mx_evl_psf_reset(mx_evl_t *h) {
    if (!h) return set_err(MX_EINVAL, "mx_evl_psf_reset: null handle");
    HIPCHK(hipSetDevice(h->device));
    for (int iw = 0; iw < h->nwvl; iw++) {
        const EvlWvl &w = h->w[iw];
        if (w.P == 1)
            HIPCHK(hipMemsetAsync(w.d_acc1, 0, sizeof(double) * h->ndir, h->stream));
        else
            HIPCHK(hipMemsetAsync(w.d_acc, 0, sizeof(float) * h->ndir * w.P * w.P, h->stream));
    }
    h->nacc = 0;
    return MX_OK;
}

int mx_evl_wait(mx_evl_t *h) {
    if (!h) return set_err(MX_EINVAL, "mx_evl_wait: null handle");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MX_OK;
}

int mx_evl_timing_enable(mx_evl_t *h, int enable) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    h->timing = enable != 0;
    h->t_n = 0;
    h->t_trace = h->t_pass[0] = h->t_pass[1] = 0;
    return MX_OK;
}

int mx_evl_timing(mx_evl_t *h, long *n, double *ms_trace, double *ms_rows, double *ms_cols) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    if (n) *n = h->t_n;
    if (ms_trace) *ms_trace = h->t_trace;
    if (ms_rows) *ms_rows = h->t_pass[0];
    if (ms_cols) *ms_cols = h->t_pass[1];
    return MX_OK;
}

} // extern "C"
