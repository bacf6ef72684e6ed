/* sdpa.c - reader of SDPA sparse files (.dat-s) into the numbers brisk_solve_data takes.
 * Format (as the reference's read_sdpa, problem.c:426): comment lines starting with '"' or '*'
 * before the header; m; nblk; the block sizes (separators ",(){}" allowed, LP blocks negative);
 * the m entries of c; then lines "mat blk i j value". */
#include <ctype.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "brisk_amd.h"

typedef struct { char *buf; size_t len, pos; } Rd;

static int next_num(Rd *r, double *out, int header) {
    for (;;) {
        while (r->pos < r->len) {
            const char ch = r->buf[r->pos];
            if (isspace((unsigned char)ch) || (header && strchr(",(){}=", ch))) { r->pos++; continue; }
            break;
        }
        if (r->pos >= r->len) return -1;
        char *end;
        const double v = strtod(r->buf + r->pos, &end);
        if (end == r->buf + r->pos) {
            /* a word (for example "=mDIM"): skip it */
            while (r->pos < r->len && !isspace((unsigned char)r->buf[r->pos])) r->pos++;
            if (!header) return -2;
            continue;
        }
        r->pos = end - r->buf;
        *out = v;
        return 0;
    }
}

/* owned: one allocation holding every array of p (free it after the solve) */
__attribute__((visibility("default"))) int brisk_read_sdpa(const char *fname, BriskProblem *p, void **owned) {
    FILE *f = fopen(fname, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) { fclose(f); return -1; }
    Rd r = { malloc(sz + 1), (size_t)sz, 0 };
    if (!r.buf || fread(r.buf, 1, sz, f) != (size_t)sz) { fclose(f); free(r.buf); return -1; }
    fclose(f);
    r.buf[sz] = 0;
    /* comments: lines starting with " or * at the top */
    while (r.pos < r.len) {
        size_t q = r.pos;
        while (q < r.len && (r.buf[q] == ' ' || r.buf[q] == '\t')) q++;
        if (q < r.len && (r.buf[q] == '"' || r.buf[q] == '*')) {
            while (q < r.len && r.buf[q] != '\n') q++;
            r.pos = q + 1;
            continue;
        }
        break;
    }
    double x;
    if (next_num(&r, &x, 1)) { free(r.buf); return -1; }
    const int m = (int)x;
    if (next_num(&r, &x, 1)) { free(r.buf); return -1; }
    const int nblk = (int)x;
    if (m < 0 || nblk < 1) { free(r.buf); return -1; }
    /* count entry lines (upper bound: number of newlines) */
    size_t lines = 1;
    for (size_t q = r.pos; q < r.len; q++) lines += r.buf[q] == '\n';
    const size_t bytes = sizeof(int) * nblk + sizeof(double) * (m + 1) + lines * (4 * sizeof(int) + sizeof(double)) + 64;
    char *mem = malloc(bytes);
    if (!mem) { free(r.buf); return -1; }
    int *bs = (int *)mem;
    double *c = (double *)(mem + ((sizeof(int) * nblk + 7) & ~(size_t)7));
    double *v = c + (m + 1);
    int *mat = (int *)(v + lines), *blk = mat + lines, *ii = blk + lines, *jj = ii + lines;
    for (int k = 0; k < nblk; k++) { if (next_num(&r, &x, 1)) goto bad; bs[k] = (int)x; }
    for (int k = 0; k < m; k++) { if (next_num(&r, &x, 1)) goto bad; c[k] = x; }
    /* the separators that close the vector c ("}" after its last entry) */
    while (r.pos < r.len && (isspace((unsigned char)r.buf[r.pos]) || strchr(",(){}=", r.buf[r.pos]))) r.pos++;
    int64_t nnz = 0;
    for (;;) {
        double e[5];
        int got = 0;
        for (; got < 5; got++) { const int s = next_num(&r, &e[got], 0); if (s) break; }
        if (got == 0) break;
        if (got < 5 || (size_t)nnz >= lines) goto bad;
        mat[nnz] = (int)e[0]; blk[nnz] = (int)e[1]; ii[nnz] = (int)e[2]; jj[nnz] = (int)e[3]; v[nnz] = e[4];
        nnz++;
    }
    free(r.buf);
    p->m = m; p->nblk = nblk; p->bs = bs; p->c = c; p->nnz = nnz; p->mat = mat; p->blk = blk; p->i = ii; p->j = jj; p->v = v;
    *owned = mem;
    return 0;
bad:
    free(r.buf); free(mem);
    return -1;
}

__attribute__((visibility("default"))) void brisk_sdpa_free(void *owned) { free(owned); }
