/* sdpa.c - reader of SDPA sparse files (.dat-s) into the numbers brisk_solve_data takes.
 *
 * The rules of the reference's read_sdpa (problem.c:426-566):
 *   - the characters , { } ( ) = are separators everywhere; a Fortran exponent letter (1.5D+03,
 *     2d-1: D or d after a digit or '.', before a digit or a signed digit) reads as 'e'
 *     (problem.c:438-446);
 *   - m is the first number of the first line that is not a comment (a line whose first
 *     non-blank character is '"' or '*'), nblk the first number of the next such line; the rest
 *     of those two lines is ignored (line_num, problem.c:75-90);
 *   - from there on comment lines are blank (problem.c:457-469); the nblk block sizes (nonzero
 *     integers, LP blocks negative) and the m entries of c follow, words among them skipped;
 *   - then entries "mat blk i j value": the indices integers, the value finite; a zero value is
 *     skipped; a last entry of fewer than five numbers is ignored (problem.c:531-535);
 *   - text after the last entry ends the data: at an entry boundary, a token that does not start
 *     a number, with no later line of five numbers, stops the reading (problem.c:497-529); any
 *     other word in the data section, or a number glued to text, makes the file malformed
 *     (problem.c:563-567).
 * Range checks of the entries (indices within their blocks, LP entries diagonal) are
 * hprob_build's, on the numbers as read. */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "brisk_amd.h"

typedef struct { char *buf; size_t len, pos; long words; } Rd;

static int is_sep(char ch) { return ch == ',' || ch == '{' || ch == '}' || ch == '(' || ch == ')' || ch == '='; }

/* separators to blanks and Fortran exponents to 'e', over the whole buffer */
static void normalise(char *b, size_t len) {
    for (size_t i = 0; i < len; i++) {
        const char ch = b[i];
        if (is_sep(ch)) b[i] = ' ';
        else if ((ch == 'D' || ch == 'd') && i > 0 && i + 1 < len && (isdigit((unsigned char)b[i - 1]) || b[i - 1] == '.') &&
                 (isdigit((unsigned char)b[i + 1]) ||
                  ((b[i + 1] == '+' || b[i + 1] == '-') && i + 2 < len && isdigit((unsigned char)b[i + 2]))))
            b[i] = 'e';
    }
}
static int comment_at(const Rd *r, size_t q) {
    while (q < r->len && (r->buf[q] == ' ' || r->buf[q] == '\t' || r->buf[q] == '\r')) q++;
    return q < r->len && (r->buf[q] == '"' || r->buf[q] == '*');
}
/* the next number; words are skipped and counted (a number glued to text is taken and counted).
 * Returns 0, or -1 at the end of the buffer */
static int next_num(Rd *r, double *out) {
    for (;;) {
        while (r->pos < r->len && isspace((unsigned char)r->buf[r->pos])) r->pos++;
        if (r->pos >= r->len) return -1;
        char *s = r->buf + r->pos, *end;
        const double v = strtod(s, &end);
        if (end != s) {
            if (end < r->buf + r->len && !isspace((unsigned char)*end)) r->words++;
            r->pos = end - r->buf;
            *out = v;
            return 0;
        }
        while (r->pos < r->len && !isspace((unsigned char)r->buf[r->pos])) r->pos++;
        r->words++;
    }
}
/* the first number of the next line that is not a comment; the reading goes on at the line after */
static int line_num(Rd *r, double *out) {
    while (r->pos < r->len) {
        size_t eol = r->pos;
        while (eol < r->len && r->buf[eol] != '\n') eol++;
        int ok = 0;
        if (!comment_at(r, r->pos)) {
            const char keep = eol < r->len ? r->buf[eol] : 0;
            if (eol < r->len) r->buf[eol] = 0;
            Rd line = { r->buf + r->pos, eol - r->pos, 0, 0 };
            ok = next_num(&line, out) == 0;
            if (eol < r->len) r->buf[eol] = keep;
        }
        r->pos = eol < r->len ? eol + 1 : r->len;
        if (ok) return 0;
    }
    return -1;
}
/* at an entry boundary: 1 when the next token is text and no later line holds five numbers (the
 * data has ended), else 0 */
static int data_ended(const Rd *r) {
    size_t q = r->pos;
    while (q < r->len && isspace((unsigned char)r->buf[q])) q++;
    if (q >= r->len) return 0;
    const char ch = r->buf[q];
    if (isdigit((unsigned char)ch) || ch == '-' || ch == '+' || ch == '.') return 0;
    while (q < r->len) {
        size_t eol = q;
        while (eol < r->len && r->buf[eol] != '\n') eol++;
        int nnum = 0;
        char *t = r->buf + q, *e = r->buf + eol;
        while (t < e) {
            while (t < e && isspace((unsigned char)*t)) t++;
            if (t >= e) break;
            char *x;
            (void)strtod(t, &x);
            if (x == t || x > e) break;
            nnum++;
            t = x;
        }
        if (nnum >= 5) return 0;
        q = eol + 1;
    }
    return 1;
}

/* owned: one allocation holding every array of p (free it after the solve) */
__attribute__((visibility("default"))) int brisk_read_sdpa(const char *fname, BriskProblem *p, void **owned) {
    FILE *f = fopen(fname, "rb");
    if (!f) return -1;
    if (fseek(f, 0, SEEK_END)) { fclose(f); return -1; }
    const long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) { fclose(f); return -1; }
    Rd r = { malloc(sz + 1), (size_t)sz, 0, 0 };
    if (!r.buf || fread(r.buf, 1, sz, f) != (size_t)sz) { fclose(f); free(r.buf); return -1; }
    fclose(f);
    r.buf[sz] = 0;
    normalise(r.buf, r.len);
    double x;
    if (line_num(&r, &x)) { free(r.buf); return -1; }
    const double mv = x;
    if (line_num(&r, &x)) { free(r.buf); return -1; }
    const double nbv = x;
    if (!(mv >= 0 && mv <= 1e8) || !(nbv >= 1 && nbv <= 1e9 && nbv == floor(nbv))) { free(r.buf); return -1; }
    const int m = (int)mv, nblk = (int)nbv;
    /* comment lines are blank from here on */
    for (size_t q = r.pos; q < r.len;) {
        size_t eol = q;
        while (eol < r.len && r.buf[eol] != '\n') eol++;
        if (comment_at(&r, q)) memset(r.buf + q, ' ', eol - q);
        q = eol + 1;
    }
    /* entries: at most one per five numbers of what is left */
    size_t toks = 1;
    for (size_t q = r.pos; q < r.len; q++) toks += !isspace((unsigned char)r.buf[q]) && (q == 0 || isspace((unsigned char)r.buf[q - 1]));
    const size_t lines = toks / 5 + 1;
    const size_t bytes = sizeof(int) * nblk + sizeof(double) * (m + 1) + lines * (4 * sizeof(int) + sizeof(double)) + 64;
    char *mem = malloc(bytes);
    if (!mem) { free(r.buf); return -1; }
    int *bs = (int *)mem;
    double *c = (double *)(mem + ((sizeof(int) * nblk + 7) & ~(size_t)7));
    double *v = c + (m + 1);
    int *mat = (int *)(v + lines), *blk = mat + lines, *ii = blk + lines, *jj = ii + lines;
    for (int k = 0; k < nblk; k++) {
        if (next_num(&r, &x) || !(fabs(x) >= 1 && fabs(x) <= 2e9) || x != floor(x)) goto bad;
        bs[k] = (int)x;
    }
    for (int k = 0; k < m; k++) {
        if (next_num(&r, &x) || !isfinite(x)) goto bad;
        c[k] = x;
    }
    r.words = 0;                          /* words in the header are allowed */
    int64_t nnz = 0;
    for (;;) {
        if (data_ended(&r)) break;
        double e[5];
        int got = 0;
        while (got < 5 && next_num(&r, &e[got]) == 0) got++;
        if (got < 5) break;               /* an incomplete last entry is ignored */
        for (int q = 0; q < 4; q++) if (e[q] != floor(e[q]) || fabs(e[q]) > 2e9) goto bad;
        if (!isfinite(e[4])) goto bad;
        if (e[4] == 0.0) continue;
        if ((size_t)nnz >= lines) goto bad;
        mat[nnz] = (int)e[0]; blk[nnz] = (int)e[1]; ii[nnz] = (int)e[2]; jj[nnz] = (int)e[3]; v[nnz] = e[4];
        nnz++;
    }
    if (r.words) goto bad;                /* a word in the data section */
    free(r.buf);
    p->m = m; p->nblk = nblk; p->bs = bs; p->c = c; p->nnz = nnz; p->mat = mat; p->blk = blk; p->i = ii; p->j = jj; p->v = v;
    *owned = mem;
    return 0;
bad:
    free(r.buf); free(mem);
    return -1;
}

__attribute__((visibility("default"))) void brisk_sdpa_free(void *owned) { free(owned); }
