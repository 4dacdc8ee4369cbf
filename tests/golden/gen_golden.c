/*
 * gen_golden.c -- writes golden vectors from the COMPILED REFERENCE
 * (oracle/_ref/libparatt_ref.so, built from the reference's own sources by
 * oracle/build_ref.sh) at one MPI rank.  Run by gen_golden.sh; the JSON it
 * prints is committed as tests/golden/golden_ref.json and pins the CPU
 * oracle (tests/test_oracle_golden.py).
 *
 * Inputs come from the project's seeded generator (same algorithm as
 * oracle/ptt_oracle.c o_gauss_fill); small cases also store their inputs so
 * the generator itself is pinned.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <sys/stat.h>
#include <mpi.h>
#include "paratt.h"
#include "paratt_ode.h"

static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static void gauss_fill(uint64_t seed, int core, double scale, long count,
                       double *out) {
    uint64_t h = mix64(mix64(seed) + (uint64_t)core);
    for (long e = 0; e < count; e += 2) {
        uint64_t p = (uint64_t)(e >> 1);
        uint64_t a = mix64(h + 2 * p), b = mix64(h + 2 * p + 1);
        double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
        double u2 = (double)(b >> 11) * 0x1.0p-53;
        double rad = sqrt(-2.0 * log(u1));
        double th = 6.283185307179586476925286766559 * u2;
        out[e] = rad * cos(th) * scale;
        if (e + 1 < count) out[e + 1] = rad * sin(th) * scale;
    }
}

static paratt *rand_ptt(int d, int *n, int *r, uint64_t seed, double scale) {
    paratt *p = alloc_ptt(d, n, r);
    for (int k = 0; k < d; k++)
        gauss_fill(seed, k, scale, (long)r[k] * n[k] * r[k + 1], p->x->coreList[k]);
    return p;
}
static paratt *clone_ptt(paratt *s) {
    paratt *p = alloc_ptt(s->numDim, s->dim, s->x->ttRank);
    for (int k = 0; k < s->numDim; k++)
        memcpy(p->x->coreList[k], s->x->coreList[k],
               sizeof(double) * (size_t)s->x->ttRank[k] * s->dim[k] * s->x->ttRank[k + 1]);
    return p;
}
/* Full-tensor entry by a left-to-right chain of slices. */
static double entry(paratt *p, const int *idx) {
    int d = p->numDim, rmax = 1;
    for (int k = 0; k <= d; k++) rmax = rmax > p->x->ttRank[k] ? rmax : p->x->ttRank[k];
    double *v = calloc(rmax, sizeof(double)), *w = calloc(rmax, sizeof(double));
    v[0] = 1.0;
    for (int k = 0; k < d; k++) {
        int r0 = p->x->ttRank[k], r1 = p->x->ttRank[k + 1], nk = p->dim[k];
        for (int b = 0; b < r1; b++) {
            double s = 0;
            for (int a = 0; a < r0; a++)
                s += v[a] * p->x->coreList[k][a + (long)r0 * (idx[k] + (long)nk * b)];
            w[b] = s;
        }
        memcpy(v, w, sizeof(double) * r1);
    }
    double res = v[0];
    free(v); free(w);
    return res;
}

static int first_item = 1;
static void key(const char *k) {
    printf("%s\n  \"%s\": ", first_item ? "" : ",", k);
    first_item = 0;
}
static void put_ints(const int *a, int n) {
    printf("[");
    for (int i = 0; i < n; i++) printf("%s%d", i ? "," : "", a[i]);
    printf("]");
}
static void put_dbls(const double *a, long n) {
    printf("[");
    for (long i = 0; i < n; i++) printf("%s%.17g", i ? "," : "", a[i]);
    printf("]");
}
static void put_tt(paratt *p) {
    int d = p->numDim;
    printf("{\"d\": %d, \"n\": ", d);
    put_ints(p->dim, d);
    printf(", \"r\": ");
    put_ints(p->x->ttRank, d + 1);
    printf(", \"cores\": [");
    for (int k = 0; k < d; k++) {
        if (k) printf(",");
        put_dbls(p->x->coreList[k], (long)p->x->ttRank[k] * p->dim[k] * p->x->ttRank[k + 1]);
    }
    printf("]}");
}
static void put_rank_only(paratt *p) { put_ints(p->x->ttRank, p->numDim + 1); }

/* Sampled entries at seeded multi-indices. */
static void put_entries(paratt *p, uint64_t seed, int count) {
    int d = p->numDim;
    int *idx = malloc(sizeof(int) * d);
    printf("{\"idx\": [");
    for (int s = 0; s < count; s++) {
        if (s) printf(",");
        for (int k = 0; k < d; k++) idx[k] = (int)(mix64(seed * 1000003ULL + s * 97ULL + k) % (uint64_t)p->dim[k]);
        put_ints(idx, d);
    }
    printf("], \"val\": [");
    for (int s = 0; s < count; s++) {
        for (int k = 0; k < d; k++) idx[k] = (int)(mix64(seed * 1000003ULL + s * 97ULL + k) % (uint64_t)p->dim[k]);
        printf("%s%.17g", s ? "," : "", entry(p, idx));
    }
    printf("]}");
    free(idx);
}

/* <p, probe_s> with rank-1 seeded probes (paratt_dotprod). */
static void put_probes(paratt *p, uint64_t seed, int count) {
    int d = p->numDim;
    int *one = malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) one[k] = 1;
    printf("[");
    for (int s = 0; s < count; s++) {
        paratt *q = rand_ptt(d, p->dim, one, seed + (uint64_t)s, 1.0);
        paratt *c = clone_ptt(p);
        printf("%s%.17g", s ? "," : "", paratt_dotprod(c, q));
        free_ptt(q); free_ptt(c);
    }
    printf("]");
    free(one);
}

static double norm_of_copy(paratt *p) {
    paratt *c = clone_ptt(p);
    double v = paratt_norm(c);
    free_ptt(c);
    return v;
}

/* ---------------- rounding cases ---------------- */
static void round_case_summary(const char *name, int d, int nn, int rk,
                               uint64_t sx, uint64_t sy, double beta,
                               double eps) {
    int *n = malloc(sizeof(int) * d), *r = malloc(sizeof(int) * (d + 1));
    for (int k = 0; k < d; k++) n[k] = nn;
    for (int k = 0; k <= d; k++) r[k] = (k == 0 || k == d) ? 1 : rk;
    double sc = 1.0 / sqrt((double)nn * rk);
    paratt *x = rand_ptt(d, n, r, sx, sc), *y = rand_ptt(d, n, r, sy, sc);
    int *rz = malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) rz[k] = (k == 0 || k == d) ? 1 : 2 * rk;
    paratt *z = alloc_ptt(d, n, rz);
    attac_tt_axpby(1.0, x->x, beta, y->x, z->x);
    double nz = norm_of_copy(z);
    double tol = eps * nz;
    paratt_truncate(z, tol, z->x->ttRank);
    key(name);
    printf("{\"d\": %d, \"n\": %d, \"rank\": %d, \"seed_x\": %llu, \"seed_y\": %llu, "
           "\"beta\": %.17g, \"eps\": %.17g, \"norm_in\": %.17g, \"tol\": %.17g, \"ranks\": ",
           d, nn, rk, (unsigned long long)sx, (unsigned long long)sy, beta, eps, nz, tol);
    put_rank_only(z);
    printf(", \"norm_out\": %.17g, \"entries\": ", norm_of_copy(z));
    put_entries(z, 7, 16);
    printf(", \"probes\": ");
    put_probes(z, 1000, 4);
    printf("}");
    free_ptt(x); free_ptt(y); free_ptt(z);
    free(n); free(r); free(rz);
}

/* Small case, full dumps: z = a*x + b*y, truncate(tol_abs, maxr) */
static void round_case_full(const char *name, int d, int *n, int *rx, int *ry,
                            uint64_t sx, uint64_t sy, double a, double b,
                            double tol_rel, double tol_abs_extra, int *maxr) {
    paratt *x = rand_ptt(d, n, rx, sx, 1.0), *y = rand_ptt(d, n, ry, sy, 1.0);
    int *rz = malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) rz[k] = (k == 0 || k == d) ? 1 : rx[k] + ry[k];
    paratt *z = alloc_ptt(d, n, rz);
    attac_tt_axpby(a, x->x, b, y->x, z->x);
    double nx = norm_of_copy(x);
    double tol = tol_rel * norm_of_copy(z) + tol_abs_extra * nx;
    key(name);
    printf("{\"x\": ");
    put_tt(x);
    printf(", \"y\": ");
    put_tt(y);
    printf(", \"a\": %.17g, \"b\": %.17g, \"z\": ", a, b);
    put_tt(z);
    printf(", \"tol\": %.17g, \"maxr\": ", tol);
    int *mr = malloc(sizeof(int) * (d + 1));
    for (int k = 0; k <= d; k++) mr[k] = maxr ? maxr[k] : z->x->ttRank[k];
    put_ints(mr, d + 1);
    paratt_truncate(z, tol, mr);
    printf(", \"out\": ");
    put_tt(z);
    printf(", \"norm_out\": %.17g}", norm_of_copy(z));
    free(mr); free(rz);
    free_ptt(x); free_ptt(y); free_ptt(z);
}

/* ---------------- operator cases ---------------- */
static void ops_cases(void) {
    int d = 3, n[3] = {8, 7, 9}, r[4] = {1, 2, 3, 1}, ry[4] = {1, 3, 2, 1};
    paratt *x = rand_ptt(d, n, r, 11, 1.0), *y = rand_ptt(d, n, ry, 12, 1.0);
    key("ops_x"); put_tt(x);
    key("ops_y"); put_tt(y);
    {   /* axpby */
        int rz[4] = {1, 5, 5, 1};
        paratt *z = alloc_ptt(d, n, rz);
        attac_tt_axpby(0.5, x->x, -2.0, y->x, z->x);
        key("ops_axpby"); put_tt(z);
        free_ptt(z);
    }
    {   /* Hadamard: caller sizes z with rz = rx*ry */
        int rz[4] = {1, 6, 6, 1};
        paratt *z = alloc_ptt(d, n, rz);
        pattacHadamard(d, n, x->x->ttRank, x->x->coreList, y->x->ttRank,
                       y->x->coreList, z->x->ttRank, z->x->coreList);
        key("ops_hadamard"); put_tt(z);
        free_ptt(z);
    }
    double sten5[5] = {-1.0 / 12, 4.0 / 3, -2.5, 4.0 / 3, -1.0 / 12};
    double sten3[3] = {0.3, -1.1, 0.7};
    char nm[64];
    for (int mode = 0; mode < d; mode++) {
        for (int s = 0; s < 2; s++) {
            double *st = s ? sten5 : sten3;
            int sl = s ? 5 : 3;
            paratt *c = clone_ptt(x);
            pattac_apply_periodic_stencil(c, mode, st, sl);
            snprintf(nm, sizeof nm, "ops_pstencil_m%d_s%d", mode, sl);
            key(nm); put_tt(c); free_ptt(c);
            c = clone_ptt(x);
            attac_apply_periodic_stencil(c->x, mode, st, sl);
            snprintf(nm, sizeof nm, "ops_sstencil_m%d_s%d", mode, sl);
            key(nm); put_tt(c); free_ptt(c);
            c = clone_ptt(x);
            pattac_apply_cdf_stencil(c, mode, st, sl);
            snprintf(nm, sizeof nm, "ops_cdfstencil_m%d_s%d", mode, sl);
            key(nm); put_tt(c); free_ptt(c);
        }
    }
    {   /* even slen is decremented: only the first slen-1 entries are used */
        double st4[4] = {0.25, -0.5, 0.75, 9.0};
        paratt *c = clone_ptt(x);
        pattac_apply_periodic_stencil(c, 1, st4, 4);
        key("ops_pstencil_m1_s4"); put_tt(c); free_ptt(c);
    }
    for (int mode = 0; mode < d; mode++) {
        int nn = n[mode];
        double *ws = malloc(sizeof(double) * nn), *src = malloc(sizeof(double) * nn);
        double *bw = malloc(sizeof(double) * (nn + 1)), gcw[2] = {0.3, -0.45};
        for (int i = 0; i < nn; i++) { ws[i] = 0.2 + 0.1 * i; src[i] = 0.05 * i - 0.1; }
        for (int i = 0; i <= nn; i++) bw[i] = 1.0 - 0.03 * i;
        paratt *c = clone_ptt(x);
        pattac_apply_cdf_lf_flux(c, mode, ws, gcw, src, 0.01, 0.125);
        snprintf(nm, sizeof nm, "ops_lfflux_m%d", mode);
        key(nm); put_tt(c); free_ptt(c);
        c = clone_ptt(x);
        pattac_apply_cdf_lw_rhs(c, mode, ws, gcw, bw, src, 0.01, 0.125);
        snprintf(nm, sizeof nm, "ops_lwrhs_m%d", mode);
        key(nm); put_tt(c); free_ptt(c);
        c = clone_ptt(x);
        attac_mult_diag_mat(c->x, mode, ws);
        snprintf(nm, sizeof nm, "ops_diag_m%d", mode);
        key(nm); put_tt(c); free_ptt(c);
        free(ws); free(src); free(bw);
    }
    {
        paratt *c = clone_ptt(x);
        attac_tt_scale(-1.75, c->x);
        key("ops_scale"); put_tt(c); free_ptt(c);
    }
    {
        paratt *c = clone_ptt(x), *e = clone_ptt(y);
        key("ops_dotprod_xy"); printf("%.17g", paratt_dotprod(c, e));
        free_ptt(c); free_ptt(e);
        c = clone_ptt(x);
        key("ops_dotprod_xx"); printf("%.17g", paratt_dotprod(c, c));
        free_ptt(c);
        c = clone_ptt(x);
        key("ops_norm_x"); printf("%.17g", paratt_norm(c));
        key("ops_norm_x_after"); put_tt(c);
        free_ptt(c);
    }
    free_ptt(x); free_ptt(y);
}

/* ---------------- Euler step with a heat-equation vector field ---------- */
static double g_h2inv;
static void heat_rhs(paratt *f, paratt *out, int nargs, ...) {
    (void)nargs;
    int d = f->numDim;
    double st[3] = {g_h2inv, -2.0 * g_h2inv, g_h2inv};
    paratt *t = clone_ptt(f);
    pattac_apply_periodic_stencil(t, 0, st, 3);
    copy_rank_change_tt(out->x, t->x);
    free_ptt(t);
    for (int i = 1; i < d; i++) {
        t = clone_ptt(f);
        pattac_apply_periodic_stencil(t, i, st, 3);
        int *rz = malloc(sizeof(int) * (d + 1));
        for (int k = 0; k <= d; k++) rz[k] = (k == 0 || k == d) ? 1 : out->x->ttRank[k] + t->x->ttRank[k];
        paratt *acc = alloc_ptt(d, f->dim, rz);
        attac_tt_axpby(1.0, out->x, 1.0, t->x, acc->x);
        copy_rank_change_tt(out->x, acc->x);
        paratt_truncate(out, 1e-12, out->x->ttRank);
        free_ptt(acc); free_ptt(t); free(rz);
    }
}

static void euler_case(void) {
    int d = 4, n[4] = {16, 16, 16, 16}, r[5] = {1, 2, 2, 2, 1};
    g_h2inv = 1.0 / ((2 * M_PI / 16) * (2 * M_PI / 16));
    paratt *f = rand_ptt(d, n, r, 21, 0.5);
    key("euler_in"); put_tt(f);
    paratt *fn = alloc_ptt(d, n, r), *wk = alloc_ptt(d, n, r);
    for (int s = 0; s < 3; s++) {
        pattac_exp_euler_st(f, fn, wk, (void (*)(paratt *, paratt *, int, ...))heat_rhs,
                            1e-3, 1e-10, 1e-12);
        copy_rank_change_tt(f->x, fn->x);
    }
    key("euler_out"); put_tt(f);
    free_ptt(f); free_ptt(fn); free_ptt(wk);
}

/* ---------------- quadrature, marginals, entries, TT folders ---------------- */
static void put_text_file(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) { fprintf(stderr, "cannot read %s\n", path); exit(1); }
    printf("\"");
    int ch;
    while ((ch = fgetc(f)) != EOF) {
        if (ch == '\n') printf("\\n");
        else if (ch == '"' || ch == '\\') printf("\\%c", ch);
        else putchar(ch);
    }
    printf("\"");
    fclose(f);
}

static void quad_cases(void) {
    int d = 4, n[4] = {5, 6, 4, 3}, r[5] = {1, 3, 2, 4, 1};
    paratt *x = rand_ptt(d, n, r, 61, 1.0);
    key("quad_x"); put_tt(x);
    char name[64];
    for (int mode = 0; mode < d; mode++) {
        double w[8], s[8];
        gauss_fill(62, mode, 1.0, n[mode], w);
        gauss_fill(63, mode, 1.0, n[mode], s);
        snprintf(name, sizeof name, "quad_w_%d", mode); key(name); put_dbls(w, n[mode]);
        snprintf(name, sizeof name, "quad_s_%d", mode); key(name); put_dbls(s, n[mode]);
        long cnt = (long)r[mode] * n[mode] * r[mode + 1];
        paratt *c = clone_ptt(x);
        pattac_marginal_trap(c, mode, w);
        snprintf(name, sizeof name, "quad_trap_%d", mode); key(name);
        put_dbls(c->x->coreList[mode], cnt);
        free_ptt(c);
        c = clone_ptt(x);
        pattac_marginal_cusum(c, mode, w);
        snprintf(name, sizeof name, "quad_cusum_%d", mode); key(name);
        put_dbls(c->x->coreList[mode], cnt);
        free_ptt(c);
        c = clone_ptt(x);
        pattac_scale_weight_sum(c, mode, s, w);
        snprintf(name, sizeof name, "quad_sws_%d", mode); key(name);
        put_dbls(c->x->coreList[mode], cnt);
        free_ptt(c);
        c = clone_ptt(x);
        paratt *m = NULL;
        pattac_apply_marginal(c, &m, mode, w);
        snprintf(name, sizeof name, "quad_marginal_%d", mode); key(name);
        put_tt(m);
        free_ptt(m);
        free_ptt(c);
    }
    /* entries through the library's own paratt_entry */
    {
        int idx[4];
        key("quad_entry");
        printf("{\"idx\": [");
        for (int t = 0; t < 12; t++) {
            for (int k = 0; k < d; k++) idx[k] = (int)(mix64(7000ULL + 31ULL * t + k) % (uint64_t)n[k]);
            if (t) printf(",");
            put_ints(idx, d);
        }
        printf("], \"val\": [");
        for (int t = 0; t < 12; t++) {
            for (int k = 0; k < d; k++) idx[k] = (int)(mix64(7000ULL + 31ULL * t + k) % (uint64_t)n[k]);
            printf("%s%.17g", t ? "," : "", paratt_entry(x, idx));
        }
        printf("]}");
    }
    /* TT folders.  The reference's writer deadlocks at one MPI rank as soon
     * as a core has right rank > 1 (blocking MPI_Send to itself,
     * paratt.c:1684-1686), so it writes a rank-1 TT here; the rank > 1
     * case is written in the documented format by write_plain() below and
     * read back with the reference's loader, which must reproduce x. */
    {
        char dir[] = "/tmp/ptt_golden_ioXXXXXX";
        if (!mkdtemp(dir)) { perror("mkdtemp"); exit(1); }
        char folder[128], path[192];
        int n1[3] = {5, 4, 3}, r1[4] = {1, 1, 1, 1};
        paratt *y = rand_ptt(3, n1, r1, 64, 1.0);
        key("io_y_rank1"); put_tt(y);
        snprintf(folder, sizeof folder, "%s/rank1", dir);
        pattac_write_tt_folder(folder, y);
        key("io_files_rank1");
        printf("[");
        for (int k = 0; k < 3; k++) {
            snprintf(path, sizeof path, "%s/core_%d.cr", folder, k + 1);
            if (k) printf(",");
            put_text_file(path);
            remove(path);
        }
        printf("]");
        rmdir(folder);
        free_ptt(y);
        snprintf(folder, sizeof folder, "%s/plain", dir);
        mkdir(folder, 0777);
        for (int k = 0; k < d; k++) {
            snprintf(path, sizeof path, "%s/core_%d.cr", folder, k + 1);
            FILE *f = fopen(path, "w");
            fprintf(f, "%d,%d,%d\n", r[k], n[k], r[k + 1]);
            for (long e = 0; e < (long)r[k] * n[k] * r[k + 1]; e++)
                fprintf(f, "%+1.16e\n", x->x->coreList[k][e]);
            fclose(f);
        }
        paratt *z = pattac_load_tt_folder(folder, d);
        int same = 1;
        for (int k = 0; k <= d; k++) same &= (z->x->ttRank[k] == r[k]);
        for (int k = 0; k < d && same; k++)
            same &= memcmp(z->x->coreList[k], x->x->coreList[k],
                           sizeof(double) * (size_t)r[k] * n[k] * r[k + 1]) == 0;
        key("io_reload_identical"); printf("%d", same);
        /* a slice [1, 4) of mode 1 through the reference's load_core_slice */
        {
            double *buf = NULL;
            int cap = 0, rl = 0, nn = 0, rr = 0;
            snprintf(path, sizeof path, "%s/core_2.cr", folder);
            load_core_slice(path, &buf, &cap, &rl, &nn, &rr, 1, 4);
            key("io_slice_core2_1_4");
            printf("{\"r0\": %d, \"n\": %d, \"r1\": %d, \"v\": ", rl, nn, rr);
            put_dbls(buf, (long)rl * nn * rr);
            printf("}");
            free(buf);
        }
        free_ptt(z);
        for (int k = 0; k < d; k++) {
            snprintf(path, sizeof path, "%s/core_%d.cr", folder, k + 1);
            remove(path);
        }
        rmdir(folder);
        rmdir(dir);
    }
    free_ptt(x);
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    printf("{");
    key("generator_check");
    {
        double v[6];
        gauss_fill(5, 3, 1.0, 6, v);
        put_dbls(v, 6);
    }
    /* config 1 and its eps-binding variant */
    round_case_summary("cfg1", 20, 32, 16, 1, 2, 1.0, 1e-10);
    round_case_summary("cfg1_epsbind", 20, 32, 16, 1, 2, 1e-12, 1e-10);
    {
        int d = 5, n[5] = {4, 5, 3, 6, 4};
        int rx[6] = {1, 3, 4, 5, 2, 1}, ry[6] = {1, 2, 3, 2, 3, 1};
        int cap2[6] = {1, 2, 2, 2, 2, 1};
        round_case_full("small_tight", d, n, rx, ry, 31, 32, 0.7, -1.3, 1e-13, 0, NULL);
        round_case_full("small_loose", d, n, rx, ry, 31, 32, 0.7, -1.3, 0.2, 0, NULL);
        round_case_full("small_cap2", d, n, rx, ry, 31, 32, 0.7, -1.3, 0.0, 0, cap2);
        round_case_full("small_x_plus_x", d, n, rx, rx, 33, 33, 1.0, 1.0, 1e-10, 0, NULL);
        round_case_full("small_x_minus_x", d, n, rx, rx, 33, 33, 1.0, -1.0, 0.0, 1e-10, NULL);
    }
    {
        int d = 4, n[4] = {2, 3, 2, 2};
        int rx[5] = {1, 4, 5, 3, 1}, ry[5] = {1, 1, 2, 2, 1};
        round_case_full("small_wide", d, n, rx, ry, 41, 42, 1.0, 0.5, 1e-12, 0, NULL);
    }
    {
        int d = 6, n[6] = {16, 16, 16, 16, 16, 16};
        int rx[7] = {1, 5, 5, 5, 5, 5, 1}, ry[7] = {1, 4, 4, 4, 4, 4, 1};
        round_case_full("mid_eps_bind", d, n, rx, ry, 51, 52, 1.0, 1e-7, 1e-6, 0, NULL);
        round_case_full("mid_eps_keep", d, n, rx, ry, 51, 52, 1.0, 1e-7, 1e-9, 0, NULL);
    }
    ops_cases();
    euler_case();
    quad_cases();
    printf("\n}\n");
    MPI_Finalize();
    return 0;
}
