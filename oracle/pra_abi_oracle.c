/*
 * pra_abi_oracle.c -- the subset of the pra_* C ABI (include/pra.h) that the
 * C++ facade (include/parallel_RA_inc.h) calls, implemented on the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Linking a program written against the
 * reference's C++ API with this library instead of the HIP backend runs it on
 * the oracle, so that whole programs (many relations, SCCs and rules) have a
 * CPU answer to compare the device run with.  Only tests/ link it.
 *
 * Semantics follow the backend's entry points one for one: versions and
 * predicate codes are the PRA_* constants, rows are arity+1 u64 words, files
 * use the tsv_to_bin layout (<path> = rows of u64, <path>.size =
 * "rows\ncols", src/IO/parallel_io.cpp:115-235).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pra.h"

typedef uint64_t u64;
typedef struct ora_ctx ora_ctx;

ora_ctx *ora_ctx_create(int nbuckets);
void ora_ctx_destroy(ora_ctx *c);
int ora_rel_create(ora_ctx *c, int arity, int jcc, int tag);
int ora_rel_load_rows(ora_ctx *c, int rel, const u64 *rows, u64 n, int version);
int ora_scc_create(ora_ctx *c, int fixpoint, int id);
int ora_scc_add_relation(ora_ctx *c, int scc, int rel, int move_full_to_delta);
int ora_scc_add_join(ora_ctx *c, int scc, int out, int in0, int v0, int in1, int v1, const int *proj, int nproj);
int ora_scc_add_copy(ora_ctx *c, int scc, int out, int in, int v, const int *map, int n);
int ora_scc_add_acopy(ora_ctx *c, int scc, int out, int in, int v, const int *map, int n);
int ora_scc_add_copy_filter(ora_ctx *c, int scc, int out, int in, int v, const int *map, int n, int op);
int ora_scc_run(ora_ctx *c, int scc, int batch_size);
int ora_scc_begin(ora_ctx *c, int scc, int move_status);
u64 ora_scc_step(ora_ctx *c, int scc, int batch_size);
int ora_scc_end(ora_ctx *c, int scc);
u64 ora_rel_count(ora_ctx *c, int rel, int version);
u64 ora_rel_rows(ora_ctx *c, int rel, int version, u64 *out, u64 cap);
int ora_scc_history(ora_ctx *c, int scc, u64 *out, int cap);

struct pra_ctx {
    ora_ctx *o;
    int nb;
    uint32_t batch;
    char err[512];
};
struct pra_rel {
    pra_ctx *c;
    int id, arity;
};
struct pra_scc {
    pra_ctx *c;
    int id, ran, running;
};

static int fail(pra_ctx *c, int code, const char *msg)
{
    if (c) snprintf(c->err, sizeof(c->err), "%s", msg);
    return code;
}

int pra_ctx_create(int device, int nbuckets, uint64_t hbm_budget, pra_ctx **out)
{
    (void)device;
    (void)hbm_budget;
    if (!out || nbuckets < 1) return PRA_E_ARG;
    pra_ctx *c = (pra_ctx *)calloc(1, sizeof(pra_ctx));
    if (!c) return PRA_E_ARG;
    c->o = ora_ctx_create(nbuckets);
    c->nb = nbuckets;
    c->batch = 10;
    *out = c;
    return PRA_OK;
}

int pra_ctx_destroy(pra_ctx *c)
{
    if (!c) return PRA_E_ARG;
    ora_ctx_destroy(c->o);
    free(c);
    return PRA_OK;
}

const char *pra_last_error(pra_ctx *c) { return c ? c->err : "no context"; }

/* the CPU backend times no kernels: no kernel classes to report */
int pra_ctx_set_profiling(pra_ctx *c, int mask)
{
    (void)mask;
    return c ? PRA_OK : PRA_E_ARG;
}

int pra_ctx_kernel_stat(pra_ctx *c, int idx, const char **name, uint64_t *launches, double *ms, double *bytes)
{
    (void)c; (void)idx; (void)name; (void)launches; (void)ms; (void)bytes;
    return PRA_E_ARG;
}

int pra_ctx_set_batch_size(pra_ctx *c, uint32_t batch)
{
    if (!c || batch < 1) return PRA_E_ARG;
    c->batch = batch;
    return PRA_OK;
}

int pra_rel_create(pra_ctx *c, uint32_t arity, uint32_t jcc, uint32_t tag, int canonical, pra_rel **out)
{
    (void)canonical;
    if (!c || !out) return PRA_E_ARG;
    int id = ora_rel_create(c->o, (int)arity, (int)jcc, (int)tag);
    if (id < 0) return fail(c, PRA_E_ARG, "bad relation shape");
    pra_rel *r = (pra_rel *)calloc(1, sizeof(pra_rel));
    r->c = c;
    r->id = id;
    r->arity = (int)arity;
    *out = r;
    return PRA_OK;
}

int pra_rel_destroy(pra_rel *r)
{
    if (!r) return PRA_E_ARG;
    free(r);
    return PRA_OK;
}

int pra_rel_load_rows(pra_rel *r, const uint64_t *rows, uint64_t n, int version)
{
    if (!r || (n && !rows)) return PRA_E_ARG;
    if (ora_rel_load_rows(r->c->o, r->id, rows, n, version) != 0) return fail(r->c, PRA_E_ARG, "bad version");
    return PRA_OK;
}

int pra_rel_load_bin(pra_rel *r, const char *path, int version)
{
    if (!r || !path) return PRA_E_ARG;
    char sp[4096];
    snprintf(sp, sizeof(sp), "%s.size", path);
    FILE *f = fopen(sp, "r");
    if (!f) return fail(r->c, PRA_E_IO, "cannot open size file");
    long long rows = -1, cols = -1;
    int got = fscanf(f, "%lld\n%lld", &rows, &cols);
    fclose(f);
    if (got != 2 || rows < 0) return fail(r->c, PRA_E_IO, "malformed size file");
    if (cols != r->arity + 1) return fail(r->c, PRA_E_IO, "column count does not match arity+1");
    u64 *buf = (u64 *)malloc((size_t)(rows ? rows : 1) * (size_t)cols * 8);
    if (rows) {
        FILE *d = fopen(path, "rb");
        if (!d) {
            free(buf);
            return fail(r->c, PRA_E_IO, "cannot open data file");
        }
        size_t n = fread(buf, 8, (size_t)rows * (size_t)cols, d);
        fclose(d);
        if (n != (size_t)rows * (size_t)cols) {
            free(buf);
            return fail(r->c, PRA_E_IO, "short read");
        }
    }
    int rc = pra_rel_load_rows(r, buf, (u64)rows, version);
    free(buf);
    return rc;
}

int pra_rel_count(pra_rel *r, int version, uint64_t *out)
{
    if (!r || !out || version < 0 || version > 2) return PRA_E_ARG;
    *out = ora_rel_count(r->c->o, r->id, version);
    return PRA_OK;
}

int pra_rel_download(pra_rel *r, int version, uint64_t *rows, uint64_t cap, uint64_t *nrows)
{
    if (!r || !nrows || version < 0 || version > 2) return PRA_E_ARG;
    *nrows = ora_rel_rows(r->c->o, r->id, version, rows, rows ? cap : 0);
    return PRA_OK;
}

int pra_rel_write_bin(pra_rel *r, const char *path, int version)
{
    if (!r || !path || version < 0 || version > 2) return PRA_E_ARG;
    int w = r->arity + 1;
    u64 n = ora_rel_count(r->c->o, r->id, version);
    u64 *buf = (u64 *)malloc((size_t)(n ? n : 1) * w * 8);
    ora_rel_rows(r->c->o, r->id, version, buf, n);
    FILE *f = fopen(path, "wb");
    if (!f) {
        free(buf);
        return fail(r->c, PRA_E_IO, "cannot create data file");
    }
    size_t wrote = n ? fwrite(buf, 8, (size_t)n * w, f) : 0;
    free(buf);
    if (fclose(f) != 0 || wrote != (size_t)n * w) return fail(r->c, PRA_E_IO, "short write");
    char sp[4096];
    snprintf(sp, sizeof(sp), "%s.size", path);
    FILE *m = fopen(sp, "w");
    if (!m) return fail(r->c, PRA_E_IO, "cannot create size file");
    fprintf(m, "%llu\n%d", (unsigned long long)n, w);
    return fclose(m) ? fail(r->c, PRA_E_IO, "short write") : PRA_OK;
}

int pra_scc_create(pra_ctx *c, int fixpoint, int id, pra_scc **out)
{
    if (!c || !out) return PRA_E_ARG;
    pra_scc *s = (pra_scc *)calloc(1, sizeof(pra_scc));
    s->c = c;
    s->id = ora_scc_create(c->o, fixpoint, id);
    *out = s;
    return PRA_OK;
}

int pra_scc_destroy(pra_scc *s)
{
    if (!s) return PRA_E_ARG;
    free(s);
    return PRA_OK;
}

int pra_scc_add_relation(pra_scc *s, pra_rel *r, int move_full_to_delta)
{
    if (!s || !r) return PRA_E_ARG;
    return ora_scc_add_relation(s->c->o, s->id, r->id, move_full_to_delta) ? fail(s->c, PRA_E_ARG, "too many relations")
                                                                             : PRA_OK;
}

int pra_scc_add_join(pra_scc *s, pra_rel *out, pra_rel *in0, int v0, pra_rel *in1, int v1, const int *proj, int n)
{
    if (!s || !out || !in0 || !in1 || !proj) return PRA_E_ARG;
    return ora_scc_add_join(s->c->o, s->id, out->id, in0->id, v0, in1->id, v1, proj, n) ? fail(s->c, PRA_E_ARG, "bad join")
                                                                                         : PRA_OK;
}

int pra_scc_add_copy(pra_scc *s, pra_rel *out, pra_rel *in, int v, const int *map, int n)
{
    if (!s || !out || !in || !map) return PRA_E_ARG;
    return ora_scc_add_copy(s->c->o, s->id, out->id, in->id, v, map, n) ? fail(s->c, PRA_E_ARG, "bad copy") : PRA_OK;
}

int pra_scc_add_acopy(pra_scc *s, pra_rel *out, pra_rel *in, int v, const int *map, int n)
{
    if (!s || !out || !in || !map) return PRA_E_ARG;
    return ora_scc_add_acopy(s->c->o, s->id, out->id, in->id, v, map, n) ? fail(s->c, PRA_E_ARG, "bad acopy") : PRA_OK;
}

int pra_scc_add_copy_filter(pra_scc *s, pra_rel *out, pra_rel *in, int v, const int *map, int n, int op)
{
    if (!s || !out || !in || !map) return PRA_E_ARG;
    return ora_scc_add_copy_filter(s->c->o, s->id, out->id, in->id, v, map, n, op) ? fail(s->c, PRA_E_ARG, "bad filter")
                                                                                   : PRA_OK;
}

int pra_scc_run(pra_scc *s, uint64_t *delta_hist, uint64_t cap, uint64_t *nbatches)
{
    if (!s) return PRA_E_ARG;
    if (s->ran) return fail(s->c, PRA_E_STATE, "SCC already ran");
    ora_scc_run(s->c->o, s->id, (int)s->c->batch);
    s->ran = 1;
    int n = ora_scc_history(s->c->o, s->id, delta_hist, delta_hist ? (int)cap : 0);
    if (nbatches) *nbatches = (uint64_t)n;
    return PRA_OK;
}

int pra_scc_begin(pra_scc *s, int move_full_to_delta)
{
    if (!s) return PRA_E_ARG;
    if (s->ran || s->running) return fail(s->c, PRA_E_STATE, "SCC already started");
    ora_scc_begin(s->c->o, s->id, move_full_to_delta);
    s->running = 1;
    return PRA_OK;
}

int pra_scc_step(pra_scc *s, uint64_t *delta)
{
    if (!s) return PRA_E_ARG;
    if (!s->running) return fail(s->c, PRA_E_STATE, "SCC not started");
    u64 d = ora_scc_step(s->c->o, s->id, (int)s->c->batch);
    if (delta) *delta = d;
    return PRA_OK;
}

int pra_scc_end(pra_scc *s)
{
    if (!s) return PRA_E_ARG;
    if (!s->running) return fail(s->c, PRA_E_STATE, "SCC not started");
    ora_scc_end(s->c->o, s->id);
    s->running = 0;
    s->ran = 1;
    return PRA_OK;
}
