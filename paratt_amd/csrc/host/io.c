/*
 * io.c -- the TT folder format of the reference (paratt.c:1621-1892) for
 * device-resident cores.
 *
 * A TT with d modes is a folder of text files core_1.cr .. core_d.cr.  File
 * k+1 holds core k: a header line "r_k,n_k,r_{k+1}" and then one value per
 * line, printed "%+1.16e" (17 significant digits: every double survives a
 * write/read round trip exactly), in the core's column-major order (left
 * rank index fastest, then the mode index, then the right rank index).
 *
 * Cores live in HBM.  Writing downloads a core once into a host staging
 * buffer and formats it in large chunks; loading parses a file into a host
 * staging buffer and uploads it with one copy.  At one rank the reference's
 * MPI token passing (paratt.c:1651-1687) and per-rank mode slices collapse
 * to one writer and the full mode range [0, n_k).
 *
 * Error behaviour follows the reference: a file that cannot be opened for
 * writing is fatal (exit 1); read errors print a message and return, leaving
 * what was loaded so far.
 */
#define _POSIX_C_SOURCE 200809L /* getc_unlocked, mkdir */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "paratt.h"
#include "ptt_hip.h"
#include "ptt_host.h"

#define CR_NAME_MAX 4096
#define CR_LINE_MAX 128 /* the reference's line buffer (paratt.c:1780) */
#define CR_CHUNK (1L << 16)

static void core_file_name(char *out, const char *folder, int k1) {
    int len = snprintf(out, CR_NAME_MAX, "%s/core_%d.cr", folder, k1);
    if (len < 0 || len >= CR_NAME_MAX) ptt_host_fatal("paratt_amd", "TT folder name too long");
}

/* Reads one line (without its '\n') into buf.  Returns its length, or -1 at
 * end of file before a '\n' or when the line does not fit the buffer. */
static int read_line(FILE *f, char *buf, int cap) {
    int len = 0;
    for (;;) {
        int ch = getc_unlocked(f);
        if (ch == EOF) return -1;
        if (ch == '\n') break;
        if (len + 1 >= cap) return -1;
        buf[len++] = (char)ch;
    }
    buf[len] = '\0';
    return len;
}

/* Parses the header "rl,N,rr" the way the reference does (atoi of the
 * comma-separated fields, paratt.c:1756-1767).  Returns 0 on success. */
static int read_header(FILE *f, const char *who, int sz[3]) {
    char line[CR_LINE_MAX];
    if (read_line(f, line, CR_LINE_MAX) < 0) {
        fprintf(stderr, "\n%s error: can't read line 1.\n", who);
        return 1;
    }
    char *p = line;
    for (int t = 0; t < 3; t++) {
        sz[t] = atoi(p);
        char *c = strchr(p, ',');
        p = c ? c + 1 : p + strlen(p);
    }
    return 0;
}

void read_core_size(char *fname, int *rl, int *N, int *rr) {
    FILE *f = fopen(fname, "r");
    if (!f) {
        fprintf(stderr, "\nread_core_size error: can't open %s\n", fname);
        return;
    }
    int sz[3];
    if (read_header(f, "read_core_size", sz) == 0) {
        *rl = sz[0];
        *N = sz[1];
        *rr = sz[2];
    }
    fclose(f);
}

void load_core_slice(char *fname, double **cr, int *cr_sz, int *rl, int *N, int *rr, int N_beg,
                     int N_end) {
    FILE *f = fopen(fname, "r");
    if (!f) {
        fprintf(stderr, "\nload_core_slice error: can't open %s\n", fname);
        return;
    }
    int sz[3];
    if (read_header(f, "load_core_slice", sz)) {
        fclose(f);
        return;
    }
    if (N_end > sz[1] || N_beg < 0 || N_end < N_beg || sz[0] < 0 || sz[2] < 0) {
        fclose(f);
        fprintf(stderr,
                "\nload_core_slice error: slice indexes out of bounds.\n"
                "N_beg = %d, N_end = %d, core_sz[1] = %d\n",
                N_beg, N_end, sz[1]);
        return;
    }
    const long want = (long)sz[0] * (N_end - N_beg) * sz[2];
    if (want > 2147483647L) {
        fclose(f);
        ptt_host_fatal("load_core_slice", "core larger than 2^31-1 entries");
    }
    /* The device buffer grows like the reference's realloc; its old
     * contents are overwritten below, so nothing is copied over. */
    if (*cr_sz < want) {
        if (*cr) ptt_hip_free(*cr);
        *cr = (double *)ptt_hip_malloc(sizeof(double) * (size_t)(want > 0 ? want : 1));
        *cr_sz = (int)want;
    }
    *rl = sz[0];
    *N = N_end - N_beg;
    *rr = sz[2];
    double *host = (double *)ptt_host_malloc(sizeof(double) * (size_t)(want > 0 ? want : 1));
    long got = 0, line_no = 1;
    char line[CR_LINE_MAX];
    int ok = 1;
    for (int k = 0; k < sz[2] && ok; k++)
        for (int j = 0; j < sz[1] && ok; j++)
            for (int i = 0; i < sz[0]; i++) {
                if (read_line(f, line, CR_LINE_MAX) < 0) {
                    fprintf(stderr, "\nload_core_slice error: can't read line %ld.\n", line_no);
                    ok = 0;
                    break;
                }
                if (N_beg <= j && j < N_end) host[got++] = strtod(line, NULL);
                line_no++;
            }
    fclose(f);
    if (got > 0) ptt_hip_h2d(*cr, host, sizeof(double) * (size_t)got);
    free(host);
}

paratt *pattac_load_tt_folder(char *folder_name, int ndims) {
    if (ndims < 1) ptt_host_fatal("pattac_load_tt_folder", "ndims must be >= 1");
    char fname[CR_NAME_MAX];
    int *dim = (int *)ptt_host_calloc((size_t)ndims, sizeof(int));
    int *ttr = (int *)ptt_host_calloc((size_t)ndims + 1, sizeof(int));
    int rl = 0, rr = 0;
    /* Sizes from the headers; rank mismatches are reported and the file's
     * own sizes win (paratt.c:1705-1735). */
    for (int k = 0; k < ndims; k++) {
        core_file_name(fname, folder_name, k + 1);
        int prev_rr = rr;
        rl = rr = -1;
        read_core_size(fname, &rl, &dim[k], &rr);
        if (rl < 0) {
            fprintf(stderr, "pattac_load_tt_folder: cannot read the size of %s\n", fname);
            free(dim);
            free(ttr);
            return NULL;
        }
        if (k == 0 && rl != 1)
            fprintf(stderr, "%s:Line %d %s::pattac_load_tt_folder:: Bad left TT rank in file %d.\n",
                    __FILE__, __LINE__, __func__, k + 1);
        if (k > 0 && rl != prev_rr)
            fprintf(stderr, "%s:Line %d %s::pattac_load_tt_folder:: Bad left TT rank in file %d.\n",
                    __FILE__, __LINE__, __func__, k + 1);
        ttr[k] = (k == 0) ? 1 : rl;
    }
    if (rr != 1)
        fprintf(stderr, "%s:Line %d %s::pattac_load_tt_folder:: Bad right TT rank in file %d.\n",
                __FILE__, __LINE__, __func__, ndims);
    ttr[ndims] = rr;
    paratt *p = alloc_ptt(ndims, dim, ttr);
    for (int k = 0; k < ndims; k++) {
        core_file_name(fname, folder_name, k + 1);
        load_core_slice(fname, &p->x->coreList[k], &p->x->coreMem[k], &p->x->ttRank[k],
                        &p->x->dim[k], &p->x->ttRank[k + 1], 0, p->dim[k]);
    }
    free(dim);
    free(ttr);
    return p;
}

void pattac_write_tt_folder(char *folder_name, paratt *f_ptt) {
    struct stat st;
    if (stat(folder_name, &st) == -1) mkdir(folder_name, 0777);
    attac_tt *x = f_ptt->x;
    long maxc = 1;
    for (int k = 0; k < x->numDim; k++) {
        long c = (long)x->ttRank[k] * x->dim[k] * x->ttRank[k + 1];
        if (c > maxc) maxc = c;
    }
    double *host = (double *)ptt_host_malloc(sizeof(double) * (size_t)maxc);
    /* 25 bytes bound one "%+1.16e\n" line: sign, 17 digits, point, 'e',
     * exponent sign and up to 3 exponent digits, newline. */
    char *text = (char *)ptt_host_malloc((size_t)CR_CHUNK * 26);
    char fname[CR_NAME_MAX];
    for (int k = 0; k < x->numDim; k++) {
        core_file_name(fname, folder_name, k + 1);
        FILE *f = fopen(fname, "w+");
        if (!f) {
            fprintf(stderr, "\nTT write error: can't open %s\n", fname);
            exit(1);
        }
        fprintf(f, "%d,%d,%d\n", x->ttRank[k], x->dim[k], x->ttRank[k + 1]);
        long c = (long)x->ttRank[k] * x->dim[k] * x->ttRank[k + 1];
        if (c > 0) ptt_hip_d2h(host, x->coreList[k], sizeof(double) * (size_t)c);
        for (long e0 = 0; e0 < c; e0 += CR_CHUNK) {
            long e1 = e0 + CR_CHUNK < c ? e0 + CR_CHUNK : c;
            size_t len = 0;
            for (long e = e0; e < e1; e++) len += (size_t)sprintf(text + len, "%+1.16e\n", host[e]);
            if (fwrite(text, 1, len, f) != len) {
                fprintf(stderr, "\nTT write error: can't write %s (%s)\n", fname, strerror(errno));
                exit(1);
            }
        }
        if (fclose(f) != 0) {
            fprintf(stderr, "\nTT write error: can't close %s (%s)\n", fname, strerror(errno));
            exit(1);
        }
    }
    free(text);
    free(host);
}
