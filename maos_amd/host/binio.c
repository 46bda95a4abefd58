/*
 * MAOS .bin telemetry and table files, written and read from C.
 *
 * Format (sys/bin.c:693-800, sys/bin.h:39-84): an optional comment block
 * (M_COMMENT, its length rounded up to 8 bytes, the NUL-padded text, the
 * length again, M_COMMENT again), then M_SKIP, the magic number of the data
 * type, nx and ny as 64-bit counts, then nx*ny elements, x fastest.  A cell
 * (MCC_ANY) header is followed by its nx*ny elements, each with its own
 * header; an empty element is an empty MCC_ANY header.  A zfarr
 * (math/zfarr.c:33-110) is a cell file of nstep x 1 elements pushed one per
 * step.  Files whose names end in ".gz" are gzip-compressed, as zfopen
 * (sys/bin.c:301) does; reading accepts either.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "../../include/maos_setup.h"

struct mxs_bin_s {
    gzFile gz;
    int writing;
};

static int ends_with(const char *s, const char *t) {
    const size_t n = strlen(s), m = strlen(t);
    return n >= m && strcmp(s + n - m, t) == 0;
}

int mxs_bin_open(mxs_bin_t **out, const char *fn, int write) {
    if (!out || !fn) return -1;
    *out = NULL;
    const char *mode = write ? (ends_with(fn, ".gz") ? "wb6" : "wbT") : "rb";
    gzFile gz = gzopen(fn, mode);
    if (!gz) return -2;
    mxs_bin_t *f = (mxs_bin_t *)calloc(1, sizeof(mxs_bin_t));
    if (!f) {
        gzclose(gz);
        return -3;
    }
    f->gz = gz;
    f->writing = write;
    *out = f;
    return 0;
}

int mxs_bin_close(mxs_bin_t *f) {
    if (!f) return 0;
    const int rc = gzclose(f->gz) == Z_OK ? 0 : -2;
    free(f);
    return rc;
}

static int put(mxs_bin_t *f, const void *p, size_t n) {
    return n == 0 || gzwrite(f->gz, p, (unsigned)n) == (int)n ? 0 : -2;
}
static int get(mxs_bin_t *f, void *p, size_t n) {
    return n == 0 || gzread(f->gz, p, (unsigned)n) == (int)n ? 0 : -2;
}

int mxs_bin_write_header(mxs_bin_t *f, uint32_t magic, uint64_t nx, uint64_t ny, const char *keywords) {
    if (!f || !f->writing) return -1;
    int rc = 0;
    if (keywords && keywords[0]) {
        const uint32_t mc = MXS_M_COMMENT;
        const uint64_t nlen8 = ((strlen(keywords) + 8) >> 3) << 3; /* text and its NUL, to 8 bytes */
        char *buf = (char *)calloc(nlen8, 1);
        if (!buf) return -3;
        memcpy(buf, keywords, strlen(keywords));
        rc = put(f, &mc, 4) || put(f, &nlen8, 8) || put(f, buf, nlen8) || put(f, &nlen8, 8) || put(f, &mc, 4);
        free(buf);
        if (rc) return -2;
    }
    const uint32_t skip = MXS_M_SKIP;
    return put(f, &skip, 4) || put(f, &magic, 4) || put(f, &nx, 8) || put(f, &ny, 8) ? -2 : 0;
}

int mxs_bin_write_data(mxs_bin_t *f, const void *data, uint64_t bytes) {
    if (!f || !f->writing || (bytes && !data)) return -1;
    return put(f, data, (size_t)bytes);
}

int mxs_bin_read_header(mxs_bin_t *f, uint32_t *magic, uint64_t *nx, uint64_t *ny, char *keywords,
                        uint64_t klen) {
    if (!f || f->writing || !magic || !nx || !ny) return -1;
    if (keywords && klen) keywords[0] = 0;
    for (;;) {
        uint32_t m;
        if (get(f, &m, 4)) return 1; /* end of file */
        if (m == MXS_M_EOD) return 1;
        if ((m & MXS_M_SKIP) == MXS_M_SKIP) continue;
        if (m == MXS_M_COMMENT) {
            uint64_t n = 0, n2 = 0;
            uint32_t m2 = 0;
            if (get(f, &n, 8) || n > (1u << 30)) return -2;
            char *buf = (char *)malloc(n + 1);
            if (!buf) return -3;
            if (get(f, buf, n)) {
                free(buf);
                return -2;
            }
            buf[n] = 0;
            if (keywords && klen) {
                const size_t have = strlen(keywords);
                if (have + 1 < klen) strncat(keywords, buf, (size_t)(klen - have - 1));
            }
            free(buf);
            if (get(f, &n2, 8) || n2 != n || get(f, &m2, 4) || m2 != m) return -2;
            continue;
        }
        if ((m & 0x6400) != 0x6400) return -2;
        *magic = m;
        return get(f, nx, 8) || get(f, ny, 8) ? -2 : 0;
    }
}

int mxs_bin_read_data(mxs_bin_t *f, void *data, uint64_t bytes) {
    if (!f || f->writing || (bytes && !data)) return -1;
    return get(f, data, (size_t)bytes);
}

struct mxs_zfarr_s {
    mxs_bin_t *f;
    long cur, tot;
};

int mxs_zfarr_open(mxs_zfarr_t **out, const char *fn, long nstep, const char *keywords) {
    if (!out || !fn || nstep < 0) return -1;
    *out = NULL;
    mxs_zfarr_t *z = (mxs_zfarr_t *)calloc(1, sizeof(mxs_zfarr_t));
    if (!z) return -3;
    int rc = mxs_bin_open(&z->f, fn, 1);
    if (!rc) rc = mxs_bin_write_header(z->f, MXS_MCC_ANY, (uint64_t)nstep, 1, keywords);
    if (rc) {
        mxs_bin_close(z->f);
        free(z);
        return rc;
    }
    z->tot = nstep;
    *out = z;
    return 0;
}

/* steps skipped since the last push are written as empty elements */
int mxs_zfarr_push(mxs_zfarr_t *z, long istep, uint32_t magic, uint64_t nx, uint64_t ny, const void *data,
                   uint64_t elsize) {
    if (!z || istep < z->cur || (z->tot && istep >= z->tot)) return -1;
    for (; z->cur < istep; z->cur++)
        if (mxs_bin_write_header(z->f, MXS_MCC_ANY, 0, 0, NULL)) return -2;
    if (mxs_bin_write_header(z->f, magic, nx, ny, NULL) || mxs_bin_write_data(z->f, data, nx * ny * elsize))
        return -2;
    z->cur++;
    return 0;
}

int mxs_zfarr_close(mxs_zfarr_t *z) {
    if (!z) return 0;
    int rc = 0;
    for (; z->cur < z->tot; z->cur++)
        if (mxs_bin_write_header(z->f, MXS_MCC_ANY, 0, 0, NULL)) rc = -2;
    if (mxs_bin_close(z->f)) rc = -2;
    free(z);
    return rc;
}
