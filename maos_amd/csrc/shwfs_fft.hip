// FP32 radix-8 form of the subaperture image for the N = 64, nxsa = 32
// geometry (included by shwfs.hip after the MFMA form, whose argument block,
// half-plane tables and pixel epilogue it shares).
//
// The chain of wfsints (maos/wfsints.c:79-302) on the vector units:
//   field  amp exp(+i 2 pi (opd - mean) / wvl) on the 32 x 32 band of the
//          64 x 64 grid (cembed_wvf, math/matcomp.c:56-82);
//   A      64-point FFT of the 32 band rows along x;
//   B      64-point FFT of every column along y (band input), |.|^2 -> P;
//   C      FFT of the real P along y, only uy < 32 kept: the OTF is
//          Hermitian, and zero at lag 32 (the autocorrelation of a
//          32-sample support), so the half plane uy < 32 carries all of it;
//   D      FFT along x -> O[uy][ux], times conj(lotf) ETF (or the nominal),
//          the half-plane tables of the MFMA form (half_slot order);
//   E      inverse FFT along x, only the <= 16 image columns the detector
//          samples -> R[uy][xs];
//   F      I[ys][xs] = Re sum_{uy < 32} w(uy) R[uy][xs] exp(+2 pi i uy ys / 64),
//          w = 1 at uy = 0 and 2 elsewhere (the Hermitian half);
// then the si sampler and the pixel epilogue of the MFMA form.
//
// Every 64-point transform is two radix-8 stages (64 = 8 x 8) over eight
// lanes of one wave: lane k of a line takes x[k + 8j] (j < 8), an 8-point
// DFT, the twiddle w64^(k q), and writes Y_k[q] at line position 8k + q;
// lane q then reads Y_k[q] (k < 8) and a second 8-point DFT gives
// X[q + 8p].  The exchange stays inside a wave, so it needs no workgroup
// barrier, and the output of a forward transform (lane q holds q + 8p) is
// the input distribution of the next one along the same axis (C after B,
// E after D) with no exchange at all.  A workgroup of four waves runs one
// subaperture; each wave holds eight lines, so a pass over 32 lines is one
// round and over 64 lines two.
//
// Every pass works on rows of a 64 x 64 complex grid in LDS; a pass that
// changes axis writes its output transposed (after a workgroup barrier).
// Element (r, c) sits at r 72 + c + c/8: one pad after every eight
// elements and a row stride of 72.  Under it the lane patterns of the
// passes (eight rows x eight lanes reading elements k + 8j, writing 8k + q,
// reading 8j + q, and the transposed writes) each touch every bank pair
// exactly twice, the fewest a 64-lane 8-byte access can, and every address
// is a lane base plus a constant offset.

constexpr int MX_FFT_T = 256;
constexpr int MX_FFT_GS = 72; // grid row stride (complex)

// LDS index of grid element (r, c)
__device__ __forceinline__ int gidx(int r, int c) { return r * MX_FFT_GS + c + (c >> 3); }

// the lanes of a wave see each other's LDS writes issued before this point
// (DS instructions of one wave execute in order; this keeps the compiler
// from moving memory operations across it)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// v[m] <- sum_n v[n] exp(-/+ 2 pi i n m / 4) (forward / INV)
template <bool INV>
__device__ __forceinline__ void dft4(float2 &b0, float2 &b1, float2 &b2, float2 &b3) {
    const float2 s02 = cadd(b0, b2), d02 = csub(b0, b2), s13 = cadd(b1, b3), d13 = csub(b1, b3);
    // -i d13 (forward) or +i d13
    const float2 j13 = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
    b0 = cadd(s02, s13);
    b2 = csub(s02, s13);
    b1 = cadd(d02, j13);
    b3 = csub(d02, j13);
}

// v[m] <- sum_n v[n] exp(-/+ 2 pi i n m / 8): DFT4 of the even and odd
// samples, the odd ones times w8^m, then the butterflies
template <bool INV>
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
    constexpr float r = 0.70710678118654752f;
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    if (INV) {
        o1 = make_float2((o1.x - o1.y) * r, (o1.x + o1.y) * r);  // (1 + i) / sqrt 2
        o2 = make_float2(-o2.y, o2.x);                           // i
        o3 = make_float2(-(o3.x + o3.y) * r, (o3.x - o3.y) * r); // (-1 + i) / sqrt 2
    } else {
        o1 = make_float2((o1.x + o1.y) * r, (o1.y - o1.x) * r);  // (1 - i) / sqrt 2
        o2 = make_float2(o2.y, -o2.x);                           // -i
        o3 = make_float2((o3.y - o3.x) * r, -(o3.x + o3.y) * r); // (-1 - i) / sqrt 2
    }
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// forward 8-point DFT of samples that are zero but for v[2..5] (the band
// of a zero-padded line): dft8<false> with the zero terms left out
__device__ __forceinline__ void dft8_mid(float2 (&v)[8]) {
    constexpr float r = 0.70710678118654752f;
    const float2 a = v[2], b = v[3], c = v[4], d = v[5];
    // DFT4 of (0, a, c, 0) and of (0, b, d, 0)
    const float2 e0 = cadd(c, a), e2 = csub(c, a);
    const float2 e1 = make_float2(a.y - c.x, -a.x - c.y), e3 = make_float2(-c.x - a.y, a.x - c.y);
    float2 o0 = cadd(d, b), o2 = csub(d, b);
    float2 o1 = make_float2(b.y - d.x, -b.x - d.y), o3 = make_float2(-d.x - b.y, b.x - d.y);
    o1 = make_float2((o1.x + o1.y) * r, (o1.y - o1.x) * r);
    o2 = make_float2(o2.y, -o2.x);
    o3 = make_float2((o3.y - o3.x) * r, -(o3.x + o3.y) * r);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// forward 8-point DFT of real samples p[n]: the same steps with the
// imaginary parts known to be zero (y[8 - m] = conj y[m])
__device__ __forceinline__ void dft8_real(const float (&p)[8], float2 (&v)[8]) {
    constexpr float r = 0.70710678118654752f;
    const float s04 = p[0] + p[4], d04 = p[0] - p[4], s26 = p[2] + p[6], d26 = p[2] - p[6];
    const float s15 = p[1] + p[5], d15 = p[1] - p[5], s37 = p[3] + p[7], d37 = p[3] - p[7];
    // even DFT4: e0, e2 real; e1 = d04 - i d26, e3 = d04 + i d26
    const float e0 = s04 + s26, e2 = s04 - s26;
    // odd DFT4: o0, o2 real; o1 = d15 - i d37, o3 = d15 + i d37; times w8^m
    const float o0 = s15 + s37, o2 = s15 - s37;
    const float2 t1 = make_float2((d15 - d37) * r, -(d15 + d37) * r); // o1 (1 - i) / sqrt 2
    const float2 t3 = make_float2((d37 - d15) * r, -(d15 + d37) * r); // o3 (-1 - i) / sqrt 2
    v[0] = make_float2(e0 + o0, 0.f);
    v[4] = make_float2(e0 - o0, 0.f);
    v[2] = make_float2(e2, -o2);
    v[6] = make_float2(e2, o2);
    v[1] = make_float2(d04 + t1.x, -d26 + t1.y);
    v[5] = make_float2(d04 - t1.x, -d26 - t1.y);
    v[3] = make_float2(d04 + t3.x, d26 + t3.y);
    v[7] = make_float2(d04 - t3.x, d26 - t3.y);
}

// v[q] *= w64^(k q) (forward) or its conjugate (INV); tw[q] = (cos, sin) of
// 2 pi k q / 64
template <bool INV>
__device__ __forceinline__ void twiddle8(float2 (&v)[8], const float2 (&tw)[8]) {
#pragma unroll
    for (int q = 1; q < 8; q++) {
        const float c = tw[q].x, s = INV ? -tw[q].y : tw[q].y;
        const float2 a = v[q];
        v[q] = make_float2(fmaf(a.x, c, a.y * s), fmaf(a.y, c, -a.x * s));
    }
}

// The exchange of a 64-point transform along a grid row (row points at its
// element 0), by lane k of the row's eight: the first stage's outputs
// v[q] (DFT8 and twiddle done) go to elements 8k + q, and lane k's inputs
// of the second stage, elements 8j + k, come back in v[j].  The caller
// syncs the wave before (the row's earlier readers are done).
__device__ __forceinline__ void xchg8(float2 (&v)[8], float2 *row, int k) {
    float2 *w = row + 9 * k; // element 8k + q at 9k + q
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = v[q];
    wave_sync();
    const float2 *r = row + k; // element 8j + k at 9j + k
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = r[9 * j];
}

// 64-point transform of a grid row by lane k: inputs v[j] = element
// k + 8j, outputs v[p] = element k + 8p of the transform
template <bool INV>
__device__ __forceinline__ void row64(float2 (&v)[8], float2 *row, int k, const float2 (&tw)[8]) {
    dft8<INV>(v);
    twiddle8<INV>(v, tw);
    xchg8(v, row, k);
    dft8<INV>(v);
}

// the same for a band line: inputs v[2..5] (elements k + 16 .. k + 40)
__device__ __forceinline__ void row64_mid(float2 (&v)[8], float2 *row, int k, const float2 (&tw)[8]) {
    dft8_mid(v);
    twiddle8<false>(v, tw);
    xchg8(v, row, k);
    dft8<false>(v);
}

// UPLINK: passes A to D on the LLT grid give the uplink OTF
// lotfc = FFT(|FFT(embed)|^2) / nlwvf^2 (maos/wfsints.c:149-199); one
// workgroup per (WFS, wavelength) writes conj(lotfc) over the half plane
// (half_slot order) and the full plane, as the MFMA form's uplink mode.
template <bool UPLINK>
__global__ __launch_bounds__(MX_FFT_T) __attribute__((amdgpu_waves_per_eu(4))) void k_shwfs_fft(MfmaArgs a) {
    __shared__ float2 G[64 * MX_FFT_GS];
    __shared__ float img[MX_MFMA_NS * MX_MFMA_NS];
    __shared__ float red[4];
    float2 *const Fy = G + 32 * MX_FFT_GS; // pass F's table, in grid rows that pass D no longer needs
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int k = lane & 7, li = lane >> 3; // lane of a line, line of the wave's eight
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
    // stage twiddles w64^(k q) of this lane: cos and sin of 2 pi k q / 64
    // from the FP64-built table of 2 pi m / 256
    float2 tw[8];
#pragma unroll
    for (int q = 0; q < 8; q++) tw[q] = make_float2(kTwCos[4 * k * q], kTwSin[4 * k * q]);
    // OPD and amplitude: thread (ix = tid % 32, iy = 4 (tid / 32) + j)
    constexpr int NX = 32, B0 = 16;
    const int ix = tid % 32, iy0 = 4 * (tid / 32);
    const float *og = UPLINK ? a.llt_opd + (long)iwfs * NX * NX : a.opd + ((long)iwfs * a.nsa + isa) * NX * NX;
    const float *ag = UPLINK ? a.llt_amp : a.amp + (long)isa * NX * NX;
    float mean;
    {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) s += og[(iy0 + j) * NX + ix];
        seg_sum<64>(s);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        mean = (red[0] + red[1] + red[2] + red[3]) * (1.f / (NX * NX));
    }
    float pv = 0.f; // pixel tid (< npix <= 64, wave 0)
    for (int iwvl = iwvl0; iwvl < iwvl1; iwvl++) {
        const MfmaWvl &WV = a.wv[iwvl];
        // the field on the band: element (y, x)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float re, im;
            cis_rev((og[(iy0 + j) * NX + ix] - mean) * WV.inv_wvl, ag[(iy0 + j) * NX + ix], re, im);
            G[gidx(B0 + iy0 + j, B0 + ix)] = make_float2(re, im);
        }
        __syncthreads();
        float2 v[8];
        // A: rows y of the band, FFT along x (input columns 16..47: j = 2..5),
        // written transposed: element (kx, y)
        {
            const int y = B0 + 8 * wave + li;
            float2 *row = G + y * MX_FFT_GS;
#pragma unroll
            for (int j = 2; j < 6; j++) v[j] = row[k + 9 * j];
            wave_sync();
            row64_mid(v, row, k, tw);
            __syncthreads(); // every row read before the transposed writes
            float2 *col = G + k * MX_FFT_GS + y + (y >> 3);
#pragma unroll
            for (int p = 0; p < 8; p++) col[8 * MX_FFT_GS * p] = v[p];
        }
        __syncthreads();
        // B, C: rows kx (two rounds of eight per wave), FFT along y (input
        // columns 16..47), |.|^2, FFT of that real row along y, uy < 32 kept
        float2 qk[2][4];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int kx = 16 * wave + 8 * t + li;
            float2 *row = G + kx * MX_FFT_GS;
#pragma unroll
            for (int j = 2; j < 6; j++) v[j] = row[k + 9 * j];
            wave_sync();
            row64_mid(v, row, k, tw);
            float P[8];
#pragma unroll
            for (int p = 0; p < 8; p++) P[p] = v[p].x * v[p].x + v[p].y * v[p].y;
            // lane k holds P[k + 8p]: the first stage of C takes it as is
            dft8_real(P, v);
            twiddle8<false>(v, tw);
            wave_sync();
            xchg8(v, row, k);
            dft8<false>(v);
#pragma unroll
            for (int p = 0; p < 4; p++) qk[t][p] = v[p];
        }
        __syncthreads(); // every row read before the transposed writes
        // Q[kx][uy = k + 8p] -> element (uy, kx), uy < 32
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int kx = 16 * wave + 8 * t + li;
            float2 *col = G + k * MX_FFT_GS + kx + (kx >> 3);
#pragma unroll
            for (int p = 0; p < 4; p++) col[8 * MX_FFT_GS * p] = qk[t][p];
        }
        __syncthreads();
        // D, E: rows uy < 32, FFT along x, times the transfer functions,
        // inverse FFT along x at the sampled columns
        {
            const int uy = 8 * wave + li;
            float2 *row = G + uy * MX_FFT_GS;
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = row[k + 9 * j];
            // pass F's table for this wavelength's sampled rows (grid rows
            // 32 and on are free from here)
            if (!UPLINK) {
                Fy[tid] = WV.fy[tid];
                Fy[tid + MX_FFT_T] = WV.fy[tid + MX_FFT_T];
            }
            wave_sync();
            row64<false>(v, row, k, tw);
            if constexpr (UPLINK) {
                // conj(lotfc) = conj(O) / nlwvf^2: the half plane uy < 32 in
                // slot order, and the full plane [ux][uy] with uy > 32 from
                // conj(lotfc)(-u) = lotfc(u) and the lag-32 row zero (the LLT
                // field fills 32 of 64 samples)
                float2 *hp = a.lotf_half + ((long)iwfs * a.nwvl + iwvl) * MX_HALF;
                float2 *fp = a.lotf_full + ((long)iwfs * a.nwvl + iwvl) * 64 * 64;
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const int ux = k + 8 * p;
                    const float re = v[p].x * a.llt_norm, im = v[p].y * a.llt_norm;
                    hp[half_slot(ux, uy)] = make_float2(re, -im);
                    fp[ux * 64 + uy] = make_float2(re, -im);
                    if (uy) fp[((64 - ux) & 63) * 64 + (64 - uy)] = make_float2(re, im);
                    if (uy == 0) fp[ux * 64 + 32] = make_float2(0.f, 0.f);
                }
                continue;
            }
            // v[p] = O[uy][ux], ux = k + 8p
            if (a.lotf_t && WV.etf_t) {
                const float2 *lt = a.lotf_t + ((long)iwfs * a.nwvl + iwvl) * MX_HALF;
                const float2 *et = WV.etf_t + (long)isa * MX_HALF;
                float2 tl[8], te[8];
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const int s = half_slot(k + 8 * p, uy);
                    tl[p] = lt[s];
                    te[p] = et[s];
                }
#pragma unroll
                for (int p = 0; p < 8; p++) v[p] = cmul(v[p], cmul(tl[p], te[p]));
            } else if (WV.nominal_t) {
                float tn[8];
#pragma unroll
                for (int p = 0; p < 8; p++) tn[p] = WV.nominal_t[isa * WV.nom_sa + half_slot(k + 8 * p, uy)];
#pragma unroll
                for (int p = 0; p < 8; p++) v[p] = make_float2(v[p].x * tn[p], v[p].y * tn[p]);
            }
            // inverse along x: lane k holds X[k + 8j] in v[j], the first
            // stage's input; its outputs go to elements 8k + q
            dft8<true>(v);
            twiddle8<true>(v, tw);
            wave_sync();
#pragma unroll
            for (int q = 0; q < 8; q++) row[9 * k + q] = v[q];
            wave_sync();
            // sampled columns xs = q + 8p (entries k and k + 8 of the list):
            // sum_kk Y_kk[q] exp(+2 pi i kk p / 8) by Horner in z = exp(i pi p / 4)
            float2 rs[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int xs = WV.fxs[k + 8 * h], q = xs & 7, p = xs >> 3;
                const float zc = kTwCos[32 * p], zs = kTwSin[32 * p]; // exp(i pi p / 4)
                float2 acc = row[63 + q]; // element 56 + q
#pragma unroll
                for (int kk = 6; kk >= 0; kk--) {
                    const float2 y = row[9 * kk + q];
                    acc = make_float2(fmaf(acc.x, zc, fmaf(-acc.y, zs, y.x)), fmaf(acc.x, zs, fmaf(acc.y, zc, y.y)));
                }
                rs[h] = acc;
            }
            wave_sync();
            // R[uy][s] at the row's first sixteen slots
            row[k] = rs[0];
            row[k + 8] = rs[1];
        }
        __syncthreads();
        // F: I[sy][sx] = Re sum_uy R[uy][sx] Fy[uy][sy]
        {
            const int sy = tid >> 4, sx = tid & 15;
            float s = 0.f;
#pragma unroll 8
            for (int uy = 0; uy < 32; uy++) {
                const float2 r = G[uy * MX_FFT_GS + sx], f = Fy[uy * MX_MFMA_NS + sy];
                s = fmaf(r.x, f.x, fmaf(-r.y, f.y, s));
            }
            img[sy * MX_MFMA_NS + sx] = s;
        }
        __syncthreads();
        if (tid < a.npix) {
            const int4 id = WV.si_c[isa * WV.si_sa + tid];
            const float4 w = WV.si_w[isa * WV.si_sa + tid];
            pv += W.alpha[iwvl] * (w.x * img[id.x] + w.y * img[id.y] + w.z * img[id.z] + w.w * img[id.w]);
        }
        __syncthreads();
    }
    if constexpr (UPLINK) return;
    if (wave == 0) shwfs_pixels(a, iwfs, isa, lane, pv, pix_pre(a, iwfs, isa, lane));
}

// Tip/tilt-removed PSF statistics (pistatout of wfsints, maos/wfsints.c:
// 262-275; GPU form cuda/sim/wfsints.cu:458-472) for the N = 64, nxsa = 32
// geometry.  MAOS tilts the downlink OTF by the step's geometric gradient g
// and transforms back.  The OTF of a 32-sample field vanishes at lag 32 of
// 64, so that is the PSF sampled at k + g dtheta1 (dtheta1 = 64 dx / wvl),
// which equals |FFT(amp exp(i 2 pi (opd - g.x) / wvl))|^2, the PSF of the
// field with the tilt g removed.  One workgroup per (WFS, subaperture) forms
// it per wavelength with passes A and B of k_shwfs_fft and adds
// norm_psf^2 |.|^2, the peak moved to the centre (x fastest), to an FP32
// accumulator.
struct PistatArgs {
    int nsa, nwfs, nwvl;
    float dx;
    float norm;             // norm_psf^2
    float inv_wvl[MX_MAXWVL];
    const float *opd, *amp; // [nwfs][nsa][32*32], [nsa][32*32]
    const float *gref;      // [nwfs][2*nsa] geometric gradients (radians)
    float *acc;             // [nwfs][nwvl][nsa][64*64]
};
__global__ __launch_bounds__(MX_FFT_T) __attribute__((amdgpu_waves_per_eu(4))) void k_pistat(PistatArgs a) {
    __shared__ float2 G[64 * MX_FFT_GS];
    __shared__ float red[4];
    float *const S = reinterpret_cast<float *>(G); // |.|^2 at Y 65 + X after pass B
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int k = lane & 7, li = lane >> 3;
    const int unit = xcd_unit(blockIdx.x, a.nsa * a.nwfs);
    const int iwfs = unit % a.nwfs, isa = unit / a.nwfs;
    float2 tw[8];
#pragma unroll
    for (int q = 0; q < 8; q++) tw[q] = make_float2(kTwCos[4 * k * q], kTwSin[4 * k * q]);
    constexpr int NX = 32, B0 = 16;
    const int ix = tid % 32, iy0 = 4 * (tid / 32);
    const float *og = a.opd + ((long)iwfs * a.nsa + isa) * NX * NX;
    const float *ag = a.amp + (long)isa * NX * NX;
    const float gx = a.gref[(long)iwfs * 2 * a.nsa + isa], gy = a.gref[(long)iwfs * 2 * a.nsa + a.nsa + isa];
    // the OPD less the tilt (coordinates from the grid centre) and its mean
    float o[4];
    float mean;
    {
        const float cx = ((float)ix - 15.5f) * a.dx;
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float cy = ((float)(iy0 + j) - 15.5f) * a.dx;
            o[j] = og[(iy0 + j) * NX + ix] - fmaf(gx, cx, gy * cy);
            s += o[j];
        }
        seg_sum<64>(s);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        mean = (red[0] + red[1] + red[2] + red[3]) * (1.f / (NX * NX));
    }
    float am[4];
#pragma unroll
    for (int j = 0; j < 4; j++) am[j] = ag[(iy0 + j) * NX + ix];
    for (int iwvl = 0; iwvl < a.nwvl; iwvl++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float re, im;
            cis_rev((o[j] - mean) * a.inv_wvl[iwvl], am[j], re, im);
            G[gidx(B0 + iy0 + j, B0 + ix)] = make_float2(re, im);
        }
        __syncthreads();
        float2 v[8];
        // A: band rows y, FFT along x, written transposed: element (kx, y)
        {
            const int y = B0 + 8 * wave + li;
            float2 *row = G + y * MX_FFT_GS;
#pragma unroll
            for (int j = 2; j < 6; j++) v[j] = row[k + 9 * j];
            wave_sync();
            row64_mid(v, row, k, tw);
            __syncthreads();
            float2 *col = G + k * MX_FFT_GS + y + (y >> 3);
#pragma unroll
            for (int p = 0; p < 8; p++) col[8 * MX_FFT_GS * p] = v[p];
        }
        __syncthreads();
        // B: rows kx, FFT along y (band input), |.|^2 of Y[kx][ky = k + 8p]
        float P[2][8];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int kx = 16 * wave + 8 * t + li;
            float2 *row = G + kx * MX_FFT_GS;
#pragma unroll
            for (int j = 2; j < 6; j++) v[j] = row[k + 9 * j];
            wave_sync();
            row64_mid(v, row, k, tw);
#pragma unroll
            for (int p = 0; p < 8; p++) P[t][p] = v[p].x * v[p].x + v[p].y * v[p].y;
        }
        __syncthreads(); // every row read before S overwrites the grid
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int X = (16 * wave + 8 * t + li + 32) & 63;
#pragma unroll
            for (int p = 0; p < 8; p++) S[((k + 8 * p + 32) & 63) * 65 + X] = P[t][p];
        }
        __syncthreads();
        float *acc = a.acc + (((long)iwfs * a.nwvl + iwvl) * a.nsa + isa) * 4096;
#pragma unroll 4
        for (int e = tid; e < 4096; e += MX_FFT_T) acc[e] += a.norm * S[(e >> 6) * 65 + (e & 63)];
        __syncthreads(); // S read before the next wavelength's field
    }
}
