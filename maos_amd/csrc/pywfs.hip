// Pyramid WFS imaging on the GPU (included by shwfs.hip).
//
// One step of a pyramid WFS (pywfs_ints and pywfs_grad, lib/pywfs.c:370-522;
// the reference's GPU form is cuda/sim/pywfs.cu:160-280):
//   field   amp exp(i 2 pi opd / wvl) on the pupil points, embedded in an
//           N x N grid, forward 2-D FFT, centred and scaled so that the PSF
//           sums to one;
//   modul.  for each modulation position: the centred PSF shifted by the
//           position (rounded to a sample) times the pyramid phase mask,
//           inverse 2-D FFT, |.|^2 added to the pupil intensity;
//   DTF     pupil intensity -> FFT -> x detector transfer function ->
//           inverse FFT;
//   sample  bilinear at every subaperture centre of each pupil image;
//   grads   the quad-cell (or three-facet) combination, flux-normalised.
// Every 2-D transform is a pass of line FFTs along the contiguous axis
// (Stockham stages in LDS, radices 8, 4, 2, 3, 5, 7: lds_fft_rt), a 32 x 32
// tiled transpose, and a second line pass; the orientation alternates so
// that each pass reads and writes whole rows.  Layouts are [row][column]
// with the column contiguous; "transposed" arrays hold [x][y].
constexpr int MX_PY_MAXN = 4096;
constexpr int MX_PY_T = 1024;
constexpr int MX_PY_MAXSIDE = 4;

enum PyMode { PY_EMBED = 0, PY_PSF, PY_MOD, PY_ACC, PY_REAL, PY_DTF, PY_IMG };

struct PyRowArgs {
    LfPlan plan;              // line length n
    int n, L, S, nrows;       // lines per block, LDS line stride, rows of the pass
    // PY_EMBED: CSR of the band rows: (x, point) entries, duplicates resolved
    const int *rowp;
    const int2 *ent;
    const float *opd, *amp;
    float inv_wvl;
    const float2 *in;         // input rows (all modes but PY_EMBED, PY_REAL)
    const float *inr;         // PY_REAL: pupil intensity rows
    float2 *out;              // output rows
    float *outr;              // PY_ACC: pupil intensity (+=); PY_IMG: image
    const float2 *mask;       // PY_MOD: pyramid mask, transposed [ux][uy]
    int ox, oy;               // PY_MOD: modulation offset (samples) of the window in the PSF
    float scale;              // PY_PSF: PSF normalisation; PY_ACC: weight of |.|^2
    const float *nomt;        // PY_DTF: detector transfer function, transposed [kx][ky]
};

// L rows of the pass through the LDS line FFT, one block per L rows; the
// two line buffers (L S complex each) are the block's dynamic LDS.  1024
// threads take at most one butterfly each per stage at the 2916-point lines
// of a 30 m pyramid, so the stage's twiddle and line loads overlap across
// 16 waves rather than queue within a thread.
template <int MODE>
__global__ __launch_bounds__(MX_PY_T) __attribute__((amdgpu_waves_per_eu(8))) void k_py_rows(PyRowArgs a) {
    extern __shared__ float2 pyl[];
    const int n = a.n, L = a.L, S = a.S, tid = threadIdx.x;
    float2 *const X = pyl, *const Y = pyl + L * S;
    const int r0 = (int)blockIdx.x * L;
    for (int l = 0; l < L; l++) {
        const int r = r0 + l;
        float2 *x = X + l * S;
        if (MODE == PY_EMBED) {
            for (int e = tid; e < n; e += MX_PY_T) x[lf_pos(e)] = make_float2(0.f, 0.f);
        } else if (r >= a.nrows) {
            for (int e = tid; e < n; e += MX_PY_T) x[lf_pos(e)] = make_float2(0.f, 0.f);
        } else if (MODE == PY_REAL) {
            const float *src = a.inr + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) x[lf_pos(e)] = make_float2(src[e], 0.f);
        } else if (MODE == PY_MOD) {
            // row ux of the window: PSF row ux + ox (x), columns uy + oy (y)
            const int sx = r + a.ox;
            const bool rin = sx >= 0 && sx < n;
            const float2 *src = a.in + (long)(rin ? sx : 0) * n;
            const float2 *m = a.mask + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) {
                const int sy = e + a.oy;
                x[lf_pos(e)] = rin && sy >= 0 && sy < n ? cmul(src[sy], m[e]) : make_float2(0.f, 0.f);
            }
        } else {
            const float2 *src = a.in + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) x[lf_pos(e)] = src[e];
        }
    }
    if (MODE == PY_EMBED) {
        __syncthreads(); // the zeroed lines before the scatter
        for (int l = 0; l < L; l++) {
            const int r = r0 + l;
            if (r >= a.nrows) continue;
            float2 *x = X + l * S;
            for (int k = a.rowp[r] + tid; k < a.rowp[r + 1]; k += MX_PY_T) {
                const int2 en = a.ent[k];
                float re, im;
                cis_rev(a.opd[en.y] * a.inv_wvl, a.amp[en.y], re, im);
                x[lf_pos(en.x)] = make_float2(re, im);
            }
        }
    }
    __syncthreads();
    constexpr bool INV = MODE == PY_MOD || MODE == PY_ACC || MODE == PY_IMG;
    float2 *res = lds_fft_rt<INV, MX_PY_T>(X, Y, a.plan, L, S);
    if (MODE == PY_DTF) {
        // along ky: times the transfer function, then back to y
        for (int l = 0; l < L; l++) {
            const int r = r0 + l;
            if (r >= a.nrows) continue;
            float2 *x = res + l * S;
            const float *t = a.nomt + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) {
                const float2 v = x[lf_pos(e)];
                x[lf_pos(e)] = make_float2(v.x * t[e], v.y * t[e]);
            }
        }
        __syncthreads();
        res = lds_fft_rt<true, MX_PY_T>(res, res == X ? Y : X, a.plan, L, S);
    }
    const int n2 = n / 2;
    for (int l = 0; l < L; l++) {
        const int r = r0 + l;
        if (r >= a.nrows) continue;
        const float2 *x = res + l * S;
        if (MODE == PY_PSF) {
            // centred (cfftshift) and scaled: row x -> (x + n/2) % n, ky -> (ky + n/2) % n
            float2 *dst = a.out + (long)((r + n2) % n) * n;
            for (int e = tid; e < n; e += MX_PY_T) {
                const float2 v = x[lf_pos(e)];
                dst[(e + n2) % n] = make_float2(v.x * a.scale, v.y * a.scale);
            }
        } else if (MODE == PY_ACC) {
            float *dst = a.outr + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) {
                const float2 v = x[lf_pos(e)];
                dst[e] += a.scale * (v.x * v.x + v.y * v.y);
            }
        } else if (MODE == PY_IMG) {
            float *dst = a.outr + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) dst[e] = x[lf_pos(e)].x;
        } else {
            float2 *dst = a.out + (long)r * n;
            for (int e = tid; e < n; e += MX_PY_T) dst[e] = x[lf_pos(e)];
        }
    }
}

// out[c][r] = in[r][c] for an n x n array in 32 x 32 tiles; rows of in
// outside [y0, y1) are read as zero (the field rows without points).
struct PyTrArgs {
    const float2 *in;
    float2 *out;
    int n, y0, y1;
};
__global__ __launch_bounds__(256) void k_py_transpose(PyTrArgs a) {
    __shared__ float2 t[32][33];
    const int bx = (int)blockIdx.x * 32, by = (int)blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int j = ty; j < 32; j += 8) {
        const int r = by + j, c = bx + tx;
        t[j][tx] = r < a.n && c < a.n && r >= a.y0 && r < a.y1 ? a.in[(long)r * a.n + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = bx + j, c = by + tx;
        if (r < a.n && c < a.n) a.out[(long)r * a.n + c] = t[tx][j];
    }
}

__global__ __launch_bounds__(256) void k_py_zero(float *p, long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0.f;
}

// ints[side][isa] = alpha * (bilinear sample of the image at the
// subaperture centre of that pupil image), with mkht's rules (weights <=
// 1e-15 and points off the grid dropped, no renormalisation)
struct PySampleArgs {
    const float *img;
    int n, nsa, nside;
    const double *sax, *say;        // subaperture centres (m)
    double inv_dx;                  // 1 / pupil-image sampling
    double offx[MX_PY_MAXSIDE], offy[MX_PY_MAXSIDE]; // image centre of each side (samples), minus 1/2
    float alpha;
    float *ints;
};
__global__ __launch_bounds__(256) void k_py_sample(PySampleArgs a) {
    const int i = (int)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.nsa * a.nside) return;
    const int side = i / a.nsa, isa = i - side * a.nsa;
    const double fx = a.sax[isa] * a.inv_dx + a.offx[side], fy = a.say[isa] * a.inv_dx + a.offy[side];
    const double ix = floor(fx), iy = floor(fy);
    const float wx1 = (float)(fx - ix), wy1 = (float)(fy - iy);
    const float wx[2] = {1.f - wx1, wx1}, wy[2] = {1.f - wy1, wy1};
    float s = 0.f;
    for (int q = 0; q < 4; q++) {
        const long x = (long)ix + (q & 1), y = (long)iy + (q >> 1);
        const float w = wx[q & 1] * wy[q >> 1];
        if (w > 1e-15f && x >= 0 && x < a.n && y >= 0 && y < a.n) s += w * a.img[y * a.n + x];
    }
    a.ints[i] = a.alpha * s;
}

// gradients (pywfs_grad, lib/pywfs.c:457-522), one block
struct PyGradArgs {
    const float *ints;
    int nsa, nside, sigmatch, raw;
    float gain, siglev;
    const float *saa, *gradoff; // gradoff may be null
    float *grad;
};
__global__ __launch_bounds__(256) void k_py_grad(PyGradArgs a) {
    __shared__ double red[256];
    const int tid = threadIdx.x, nsa = a.nsa;
    double s = 0;
    if (a.sigmatch == 2)
        for (int i = tid; i < nsa * a.nside; i += 256) s += (double)a.ints[i];
    red[tid] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) red[tid] += red[tid + k];
        __syncthreads();
    }
    const float imean = (float)(red[0] / nsa);
    const bool raw = a.raw || a.nside < 3;
    const int ng = raw ? a.nside : 2;
    for (int isa = tid; isa < nsa; isa += 256) {
        const float *I = a.ints + isa;
        float g[MX_PY_MAXSIDE] = {};
        if (raw) {
            for (int sd = 0; sd < a.nside; sd++) g[sd] = I[sd * nsa] * (a.gain / a.siglev);
        } else {
            float isum = 0.f;
            if (a.sigmatch == 0) isum = a.siglev * a.saa[isa];
            else if (a.sigmatch == 1)
                for (int sd = 0; sd < a.nside; sd++) isum += I[sd * nsa];
            else isum = imean * a.saa[isa];
            const float al = a.gain / isum;
            if (a.nside == 3) {
                g[0] = (I[nsa] - I[2 * nsa]) * al * 0.86602540378443865f;
                g[1] = (I[0] - 0.5f * (I[nsa] + I[2 * nsa])) * al;
            } else {
                g[0] = (I[0] - I[nsa] + I[2 * nsa] - I[3 * nsa]) * al;
                g[1] = (I[0] + I[nsa] - I[2 * nsa] - I[3 * nsa]) * al;
            }
        }
        for (int k = 0; k < ng; k++) a.grad[isa + k * nsa] = g[k] - (a.gradoff ? a.gradoff[isa + k * nsa] : 0.f);
    }
}

struct mx_pywfs_s {
    int device = 0;
    hipStream_t stream = nullptr;
    long nloc = 0, n = 0, nsa = 0;
    int nwvl = 0, nside = 4, ng = 2, sigmatch = 2, raw = 0;
    double dx = 0, siglev = 1, gain = 1, thetax = 0, thetay = 0, hs = 0;
    double wvl[MX_MAXWVL] = {}, wvlwts[MX_MAXWVL] = {};
    double ampnorm = 0;
    double modulate = 0;
    int modulpos = 0, modulring = 1;
    double sioff[2 * MX_PY_MAXSIDE] = {};
    int ylo = 0, nband = 0;  // grid rows holding points
    int ireal = 0;
    LfPlan plan{};
    float2 *d_tw = nullptr;
    int *d_rowp = nullptr;
    int2 *d_ent = nullptr;
    double *d_locx = nullptr, *d_locy = nullptr, *d_sax = nullptr, *d_say = nullptr;
    float *d_amp = nullptr, *d_opd = nullptr, *d_saa = nullptr, *d_gradoff = nullptr;
    float2 *d_mask = nullptr;   // [nwvl][n][n] transposed
    float *d_nomt = nullptr;    // [n][n] transposed
    float2 *d_A = nullptr, *d_B = nullptr, *d_P = nullptr; // work, work, centred PSF [nwvl] (transposed)
    float *d_pup = nullptr, *d_img = nullptr;
    float *d_ints = nullptr, *d_grad = nullptr;
    bool have_gradoff = false;
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    long t_n = 0;
    double t_ms = 0;
    std::vector<float> hbuf;
};

namespace {
void pywfs_release(mx_pywfs_t *h) {
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->d_tw, h->d_rowp, h->d_ent, h->d_locx, h->d_locy, h->d_sax, h->d_say, h->d_amp, h->d_opd,
                    h->d_saa, h->d_gradoff, h->d_mask, h->d_nomt, h->d_A, h->d_B, h->d_P, h->d_pup, h->d_img,
                    h->d_ints, h->d_grad};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (hipEvent_t e : h->ev)
        if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

// Stockham plan of a pyramid grid side n <= MX_PY_MAXN with factors 2, 3, 5, 7
bool py_plan(long n, LfPlan &P) {
    P = LfPlan{};
    if (n < 2 || n > MX_PY_MAXN) return false;
    P.n = (int)n;
    long prod = 1;
    while (P.nrad < MX_LF_MAXRAD && lf_radix((int)n, P.nrad)) {
        P.rad[P.nrad] = lf_radix((int)n, P.nrad);
        prod *= P.rad[P.nrad++];
    }
    return prod == n;
}

int py_rows(mx_pywfs_t *h, int mode, PyRowArgs a, int nrows) {
    a.plan = h->plan;
    a.plan.tw = h->d_tw;
    a.n = (int)h->n;
    a.S = lf_stride(a.n);
    a.L = lf_lines(a.n);
    a.nrows = nrows;
    const dim3 grid((unsigned)((nrows + a.L - 1) / a.L)), block(MX_PY_T);
    const size_t lds = sizeof(float2) * 2 * (size_t)a.L * a.S; // two line buffers, <= 2 MX_LF_LDS complex
    switch (mode) {
    case PY_EMBED: k_py_rows<PY_EMBED><<<grid, block, lds, h->stream>>>(a); break;
    case PY_PSF: k_py_rows<PY_PSF><<<grid, block, lds, h->stream>>>(a); break;
    case PY_MOD: k_py_rows<PY_MOD><<<grid, block, lds, h->stream>>>(a); break;
    case PY_ACC: k_py_rows<PY_ACC><<<grid, block, lds, h->stream>>>(a); break;
    case PY_REAL: k_py_rows<PY_REAL><<<grid, block, lds, h->stream>>>(a); break;
    case PY_DTF: k_py_rows<PY_DTF><<<grid, block, lds, h->stream>>>(a); break;
    default: k_py_rows<PY_IMG><<<grid, block, lds, h->stream>>>(a); break;
    }
    return MX_OK;
}

void py_transpose(mx_pywfs_t *h, const float2 *in, float2 *out, int y0, int y1) {
    const unsigned nt = (unsigned)((h->n + 31) / 32);
    k_py_transpose<<<dim3(nt, nt), dim3(256), 0, h->stream>>>(PyTrArgs{in, out, (int)h->n, y0, y1});
}

// pywfs_ints and pywfs_grad on the OPD in d_opd (queued on the stream)
int py_image(mx_pywfs_t *h, double siglev) {
    const long n = h->n, nn = n * n;
    const int n2 = (int)(n / 2);
    if (h->timing) HIPCHK(hipEventRecord(h->ev[0], h->stream));
    k_py_zero<<<dim3(1024), dim3(256), 0, h->stream>>>(h->d_pup, nn);
    for (int iw = 0; iw < h->nwvl; iw++) {
        float2 *P = h->d_P + (size_t)iw * nn;
        // field rows (band only) -> A; transpose -> B [x][y]; FFT along y,
        // centred and scaled -> P [x][y]
        PyRowArgs e{};
        e.rowp = h->d_rowp;
        e.ent = h->d_ent;
        e.opd = h->d_opd;
        e.amp = h->d_amp;
        e.inv_wvl = (float)(1. / h->wvl[iw]);
        e.out = h->d_A + (size_t)h->ylo * n;
        py_rows(h, PY_EMBED, e, h->nband);
        py_transpose(h, h->d_A, h->d_B, h->ylo, h->ylo + h->nband);
        PyRowArgs f{};
        f.in = h->d_B;
        f.out = P;
        f.scale = (float)(1. / (std::sqrt(h->ampnorm) * (double)n));
        py_rows(h, PY_PSF, f, (int)n);
        // modulation positions (lib/pywfs.c:390-430)
        const double dtheta = h->wvl[iw] / (h->dx * (double)n);
        for (int ir = 0; ir < h->modulring; ir++) {
            const double ri = h->modulate * (ir + 1) / h->modulring;
            const int npos = h->modulpos * (ir + 1) / h->modulring;
            for (int ip = 0; ip < npos; ip++) {
                const double th = 2. * M_PI * ((double)ip / npos);
                PyRowArgs m{};
                m.in = P;
                m.mask = h->d_mask + (size_t)iw * nn;
                m.ox = (int)std::lround(std::cos(th) * ri / dtheta);
                m.oy = (int)std::lround(std::sin(th) * ri / dtheta);
                m.out = h->d_B;
                py_rows(h, PY_MOD, m, (int)n);
                py_transpose(h, h->d_B, h->d_A, 0, (int)n);
                PyRowArgs c{};
                c.in = h->d_A;
                c.outr = h->d_pup;
                c.scale = (float)(h->wvlwts[iw] / ((double)nn * npos * h->modulring));
                py_rows(h, PY_ACC, c, (int)n);
            }
        }
    }
    // detector transfer function: rows of the pupil intensity, transpose,
    // FFT x DTF x inverse FFT along y, transpose, inverse FFT along x
    PyRowArgs r1{};
    r1.inr = h->d_pup;
    r1.out = h->d_A;
    py_rows(h, PY_REAL, r1, (int)n);
    py_transpose(h, h->d_A, h->d_B, 0, (int)n);
    PyRowArgs r2{};
    r2.in = h->d_B;
    r2.nomt = h->d_nomt;
    r2.out = h->d_A;
    py_rows(h, PY_DTF, r2, (int)n);
    py_transpose(h, h->d_A, h->d_B, 0, (int)n);
    PyRowArgs r3{};
    r3.in = h->d_B;
    r3.outr = h->d_img;
    py_rows(h, PY_IMG, r3, (int)n);
    // samples and gradients
    PySampleArgs s{};
    s.img = h->d_img;
    s.n = (int)n;
    s.nsa = (int)h->nsa;
    s.nside = h->nside;
    s.sax = h->d_sax;
    s.say = h->d_say;
    s.inv_dx = 1. / h->dx;
    for (int k = 0; k < h->nside; k++) {
        s.offx[k] = h->sioff[2 * k] * n2 + n2 - 0.5;
        s.offy[k] = h->sioff[2 * k + 1] * n2 + n2 - 0.5;
    }
    s.alpha = (float)((double)h->nsa * siglev / (double)nn);
    s.ints = h->d_ints;
    k_py_sample<<<dim3((unsigned)((h->nsa * h->nside + 255) / 256)), dim3(256), 0, h->stream>>>(s);
    PyGradArgs g{h->d_ints, (int)h->nsa, h->nside, h->sigmatch, h->raw, (float)h->gain, (float)h->siglev,
                 h->d_saa, h->have_gradoff ? h->d_gradoff : nullptr, h->d_grad};
    k_py_grad<<<dim3(1), dim3(256), 0, h->stream>>>(g);
    HIPCHK(hipGetLastError());
    if (h->timing) {
        HIPCHK(hipEventRecord(h->ev[1], h->stream));
        HIPCHK(hipEventSynchronize(h->ev[1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        h->t_ms += ms;
        h->t_n++;
    }
    return MX_OK;
}
} // namespace

#define PYCHK(call)                                                                   \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(set_err(MX_EHIP, "%s failed: %s", #call, hipGetErrorString(e_))); \
    } while (0)

extern "C" {

int mx_pywfs_init(mx_pywfs_t **out, const mx_pywfs_cfg_t *c) {
    if (!out || !c) return set_err(MX_EINVAL, "mx_pywfs_init: null argument");
    *out = nullptr;
    const long n = c->nembed, nn = n * n;
    if (c->nloc < 1 || !c->amp || !c->embed) return set_err(MX_EINVAL, "mx_pywfs_init: no pupil points");
    if (c->nwvl < 1 || c->nwvl > MX_MAXWVL || !c->wvl) return set_err(MX_EINVAL, "mx_pywfs_init: nwvl=%d", c->nwvl);
    if (c->nside < 2 || c->nside > 4)
        return set_err(MX_EINVAL, "mx_pywfs_init: nside=%d (2, 3 or 4 pupil images)", c->nside);
    if (n % 2)
        return set_err(MX_EINVAL, "mx_pywfs_init: the FFT grid side %ld must be even", n);
    LfPlan plan;
    if (!py_plan(n, plan))
        return set_err(MX_EINVAL, "mx_pywfs_init: FFT grid side %ld is not a product of 2, 3, 5 and 7 up to %d", n,
                       MX_PY_MAXN);
    if (c->nsa < 1 || !c->salocx || !c->salocy || !c->sioff || !c->mask || !c->nominal)
        return set_err(MX_EINVAL, "mx_pywfs_init: missing tables");
    if (!(c->dx > 0) || c->modulpos < 1 || c->modulring < 1 || c->modulate < 0)
        return set_err(MX_EINVAL, "mx_pywfs_init: bad sampling or modulation");
    for (long i = 0; i < c->nloc; i++)
        if (c->embed[i] < 0 || c->embed[i] >= nn)
            return set_err(MX_EINVAL, "mx_pywfs_init: embed[%ld]=%ld outside the %ld^2 grid", i, c->embed[i], n);
    mx_pywfs_t *h = new mx_pywfs_t();
    auto fail = [h](int rc) {
        pywfs_release(h);
        return rc;
    };
    h->device = g_device;
    h->nloc = c->nloc;
    h->n = n;
    h->nsa = c->nsa;
    h->nwvl = c->nwvl;
    h->nside = c->nside;
    h->raw = c->raw;
    h->ng = (c->raw || c->nside < 3) ? c->nside : 2;
    h->sigmatch = c->sigmatch;
    h->dx = c->dx;
    h->siglev = c->siglev;
    h->gain = 1.;
    h->modulate = c->modulate;
    h->modulpos = c->modulpos;
    h->modulring = c->modulring;
    h->thetax = c->thetax;
    h->thetay = c->thetay;
    h->hs = std::isfinite(c->hs) ? c->hs : 0.;
    for (int iw = 0; iw < c->nwvl; iw++) {
        h->wvl[iw] = c->wvl[iw];
        h->wvlwts[iw] = c->wvlwts ? c->wvlwts[iw] : 1. / c->nwvl;
    }
    for (int k = 0; k < 2 * c->nside; k++) h->sioff[k] = c->sioff[k];
    for (long i = 0; i < c->nloc; i++) h->ampnorm += c->amp[i] * c->amp[i];
    h->plan = plan;
    PYCHK(hipSetDevice(h->device));
    {
        // the row passes' two line buffers may exceed the default dynamic LDS limit
        const int lds = (int)(sizeof(float2) * 2 * MX_LF_LDS);
        const void *k[] = {reinterpret_cast<const void *>(&k_py_rows<PY_EMBED>),
                           reinterpret_cast<const void *>(&k_py_rows<PY_PSF>),
                           reinterpret_cast<const void *>(&k_py_rows<PY_MOD>),
                           reinterpret_cast<const void *>(&k_py_rows<PY_ACC>),
                           reinterpret_cast<const void *>(&k_py_rows<PY_REAL>),
                           reinterpret_cast<const void *>(&k_py_rows<PY_DTF>),
                           reinterpret_cast<const void *>(&k_py_rows<PY_IMG>)};
        for (const void *f : k) PYCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    }
    PYCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    PYCHK(hipEventCreate(&h->ev[0]));
    PYCHK(hipEventCreate(&h->ev[1]));
    // band rows and their entries; a later point at one index replaces the
    // earlier one (locfft_psf's sequential embedding)
    long ylo = n, yhi = -1;
    for (long i = 0; i < c->nloc; i++) {
        ylo = std::min(ylo, c->embed[i] / n);
        yhi = std::max(yhi, c->embed[i] / n);
    }
    h->ylo = (int)ylo;
    h->nband = (int)(yhi - ylo + 1);
    std::vector<std::vector<std::pair<int, int>>> rows(h->nband);
    {
        // points sorted by grid index, the last of a run of equal indices kept
        std::vector<std::pair<long, long>> idx(c->nloc);
        for (long i = 0; i < c->nloc; i++) idx[i] = {c->embed[i], i};
        std::stable_sort(idx.begin(), idx.end(), [](const std::pair<long, long> &a, const std::pair<long, long> &b) {
            return a.first < b.first;
        });
        for (size_t k = 0; k < idx.size(); k++) {
            if (k + 1 < idx.size() && idx[k + 1].first == idx[k].first) continue; // a later point follows
            const long e = idx[k].first;
            rows[e / n - ylo].push_back({(int)(e % n), (int)idx[k].second});
        }
    }
    std::vector<int> rowp(h->nband + 1, 0);
    std::vector<int2> ent;
    for (int r = 0; r < h->nband; r++) {
        for (auto &p : rows[r]) ent.push_back(make_int2(p.first, p.second));
        rowp[r + 1] = (int)ent.size();
    }
    std::vector<float2> tw;
    lf_twiddles(h->plan, tw);
    std::vector<float> amp(c->nloc);
    for (long i = 0; i < c->nloc; i++) amp[i] = (float)c->amp[i];
    std::vector<float2> mt((size_t)c->nwvl * nn);
    for (int iw = 0; iw < c->nwvl; iw++)
        for (long y = 0; y < n; y++)
            for (long x = 0; x < n; x++) {
                const double *m = c->mask + 2 * ((size_t)iw * nn + x + y * n);
                mt[(size_t)iw * nn + x * n + y] = make_float2((float)m[0], (float)m[1]);
            }
    std::vector<float> nt(nn);
    for (long y = 0; y < n; y++)
        for (long x = 0; x < n; x++) nt[x * n + y] = (float)c->nominal[x + y * n];
    auto up = [&](auto **d, const auto *src, size_t cnt) -> hipError_t {
        hipError_t e = hipMalloc((void **)d, sizeof(**d) * cnt + 16);
        if (e == hipSuccess && src) e = hipMemcpy(*d, src, sizeof(**d) * cnt, hipMemcpyHostToDevice);
        return e;
    };
    PYCHK(up(&h->d_tw, tw.data(), tw.size()));
    PYCHK(up(&h->d_rowp, rowp.data(), rowp.size()));
    PYCHK(up(&h->d_ent, ent.data(), std::max<size_t>(ent.size(), 1)));
    PYCHK(up(&h->d_amp, amp.data(), amp.size()));
    PYCHK(up(&h->d_mask, mt.data(), mt.size()));
    PYCHK(up(&h->d_nomt, nt.data(), nt.size()));
    PYCHK(up(&h->d_sax, c->salocx, (size_t)c->nsa));
    PYCHK(up(&h->d_say, c->salocy, (size_t)c->nsa));
    if (c->locx && c->locy) {
        PYCHK(up(&h->d_locx, c->locx, (size_t)c->nloc));
        PYCHK(up(&h->d_locy, c->locy, (size_t)c->nloc));
    }
    std::vector<float> ones(c->nsa, 1.f);
    PYCHK(up(&h->d_saa, ones.data(), ones.size()));
    PYCHK(up(&h->d_gradoff, (const float *)nullptr, (size_t)h->ng * c->nsa));
    PYCHK(up(&h->d_opd, (const float *)nullptr, (size_t)c->nloc));
    PYCHK(up(&h->d_A, (const float2 *)nullptr, (size_t)nn));
    PYCHK(up(&h->d_B, (const float2 *)nullptr, (size_t)nn));
    PYCHK(up(&h->d_P, (const float2 *)nullptr, (size_t)nn * c->nwvl));
    PYCHK(up(&h->d_pup, (const float *)nullptr, (size_t)nn));
    PYCHK(up(&h->d_img, (const float *)nullptr, (size_t)nn));
    PYCHK(up(&h->d_ints, (const float *)nullptr, (size_t)c->nsa * c->nside));
    PYCHK(up(&h->d_grad, (const float *)nullptr, (size_t)h->ng * c->nsa));
    *out = h;
    return MX_OK;
}

int mx_pywfs_free(mx_pywfs_t *h) {
    pywfs_release(h);
    return MX_OK;
}

int mx_pywfs_set_calib(mx_pywfs_t *h, double gain, const double *saa, const double *gradoff) {
    if (!h || !(gain > 0)) return set_err(MX_EINVAL, "mx_pywfs_set_calib: bad argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream)); // a queued step may still read the tables replaced here
    h->gain = gain;
    std::vector<float> t((size_t)h->ng * h->nsa);
    if (saa) {
        for (long i = 0; i < h->nsa; i++) t[i] = (float)saa[i];
        HIPCHK(hipMemcpy(h->d_saa, t.data(), sizeof(float) * h->nsa, hipMemcpyHostToDevice));
    }
    h->have_gradoff = gradoff != nullptr;
    if (gradoff) {
        for (size_t i = 0; i < t.size(); i++) t[i] = (float)gradoff[i];
        HIPCHK(hipMemcpy(h->d_gradoff, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
    }
    return MX_OK;
}

int mx_pywfs_ints_opd(mx_pywfs_t *h, const double *opd, double siglev) {
    if (!h || !opd) return set_err(MX_EINVAL, "mx_pywfs_ints_opd: bad argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->hbuf.resize(h->nloc);
    for (long i = 0; i < h->nloc; i++) h->hbuf[i] = (float)opd[i];
    HIPCHK(hipMemcpy(h->d_opd, h->hbuf.data(), sizeof(float) * h->nloc, hipMemcpyHostToDevice));
    return py_image(h, siglev);
}

int mx_pywfs_queue(mx_pywfs_t *h, long isim, double dt, double siglev) {
    if (!h) return set_err(MX_EINVAL, "mx_pywfs_queue: null handle");
    if (!h->d_locx) return set_err(MX_ESTATE, "mx_pywfs_queue: the handle was made without point coordinates");
    const int r = h->ireal;
    if (g_atm[r].nps < 0) return set_err(MX_ESTATE, "mx_pywfs_queue: no screens uploaded for realisation %d", r);
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
    a.thetax[0] = h->thetax;
    a.thetay[0] = h->thetay;
    a.hs[0] = h->hs;
    k_evl_opd<<<dim3((unsigned)((h->nloc + 255) / 256), 1), dim3(256), 0, h->stream>>>(a);
    HIPCHK(hipGetLastError());
    return py_image(h, siglev);
}

int mx_pywfs_sync(mx_pywfs_t *h, double *grad, double *ints) {
    if (!h) return set_err(MX_EINVAL, "mx_pywfs_sync: null handle");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (grad) {
        const size_t m = (size_t)h->ng * h->nsa;
        h->hbuf.resize(m);
        HIPCHK(hipMemcpy(h->hbuf.data(), h->d_grad, sizeof(float) * m, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < m; i++) grad[i] = h->hbuf[i];
    }
    if (ints) {
        const size_t m = (size_t)h->nside * h->nsa;
        h->hbuf.resize(m);
        HIPCHK(hipMemcpy(h->hbuf.data(), h->d_ints, sizeof(float) * m, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < m; i++) ints[i] = h->hbuf[i];
    }
    return MX_OK;
}

int mx_pywfs_opd_get(mx_pywfs_t *h, double *opd) {
    if (!h || !opd) return set_err(MX_EINVAL, "mx_pywfs_opd_get: bad argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->hbuf.resize(h->nloc);
    HIPCHK(hipMemcpy(h->hbuf.data(), h->d_opd, sizeof(float) * h->nloc, hipMemcpyDeviceToHost));
    for (long i = 0; i < h->nloc; i++) opd[i] = h->hbuf[i];
    return MX_OK;
}

int mx_pywfs_set_real(mx_pywfs_t *h, int ireal) {
    if (!h || ireal < 0 || ireal >= MX_MAXREAL) return set_err(MX_EINVAL, "mx_pywfs_set_real: bad argument");
    h->ireal = ireal;
    return MX_OK;
}

int mx_pywfs_timing_enable(mx_pywfs_t *h, int enable) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    h->timing = enable != 0;
    h->t_n = 0;
    h->t_ms = 0;
    return MX_OK;
}

int mx_pywfs_timing(mx_pywfs_t *h, long *nstep, double *ms) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    if (nstep) *nstep = h->t_n;
    if (ms) *ms = h->t_ms;
    return MX_OK;
}

int mx_pywfs_wait(mx_pywfs_t *h) {
    if (!h) return set_err(MX_EINVAL, "null handle");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MX_OK;
}

} // extern "C"
