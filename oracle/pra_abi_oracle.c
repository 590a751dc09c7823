/*
 * pra_abi_oracle.c -- the subset of the pra_* C ABI (include/pra.h) that the
 * C++ facade (include/parallel_RA_inc.h) and the serving loop
 * (parallel-ra_amd/csrc/pra_server.cpp) call, implemented on the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Linking a program written against the
 * reference's C++ API with this library instead of the HIP backend runs it on
 * the oracle, so that whole programs (many relations, SCCs and rules) have a
 * CPU answer to compare the device run with; linking the serving loop with it
 * exercises admission, preemption and out-of-memory recovery on the CPU.
 * Only tests/ link it.
 *
 * Semantics follow the backend's entry points one for one: versions and
 * predicate codes are the PRA_* constants, rows are arity+1 u64 words, files
 * use the tsv_to_bin layout (<path> = rows of u64, <path>.size =
 * "rows\ncols", src/IO/parallel_io.cpp:115-235).
 *
 * Program instances (pra_ctx_set_group): each instance is an oracle context
 * of its own, i.e. an independent LIE; pra_ctx_step runs one batch of every
 * instance's SCC in progress (one iteration at batch size 1, the size the
 * serving loop uses), starting its next runnable SCC first, as the device's
 * lock step does.  Device bytes are modelled as the bytes of the rows the
 * instance's relations hold (all three versions); when the environment sets
 * PRA_ORACLE_DEVICE_BYTES, a step after which the context holds more than
 * that fails with PRA_E_NOMEM, standing in for a device that ran out of
 * memory.
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
int ora_add_dependency(ora_ctx *c, int before, int after);
int ora_scc_run(ora_ctx *c, int scc, int batch_size);
int ora_scc_begin(ora_ctx *c, int scc, int move_status);
u64 ora_scc_step(ora_ctx *c, int scc, int batch_size);
int ora_scc_end(ora_ctx *c, int scc);
int ora_next_runnable(ora_ctx *c);
u64 ora_rel_count(ora_ctx *c, int rel, int version);
u64 ora_rel_rows(ora_ctx *c, int rel, int version, u64 *out, u64 cap);
int ora_scc_history(ora_ctx *c, int scc, u64 *out, int cap);
u64 ora_joined_tuples(ora_ctx *c);

/* one program instance: its oracle context and the handles made for it */
typedef struct {
    ora_ctx *o;
    pra_rel **rels;
    int nrel, caprel;
    pra_scc **sccs;   /* creation order; index = oracle SCC id */
    int nscc, capscc;
    int running;      /* oracle id of the SCC in progress, or -1 */
    u64 iterations;
} group_t;

struct pra_ctx {
    int nb;
    uint32_t batch;
    char err[512];
    group_t *g;
    int ng, cur;
    u64 capacity, peak;   /* modelled device bytes: limit (0 = none), peak held */
};
struct pra_rel {
    pra_ctx *c;
    ora_ctx *o;
    int group, id, arity;
};
struct pra_scc {
    pra_ctx *c;
    ora_ctx *o;
    int group, id, fixpoint, ran, running;
};

static int fail(pra_ctx *c, int code, const char *msg)
{
    if (c) snprintf(c->err, sizeof(c->err), "%s", msg);
    return code;
}

static void *grow(void *p, int *cap, int need, size_t elem)
{
    if (need <= *cap) return p;
    int nc = *cap ? 2 * *cap : 16;
    while (nc < need) nc *= 2;
    void *q = realloc(p, (size_t)nc * elem);
    if (!q) abort();
    *cap = nc;
    return q;
}

/* instance g, its oracle context created on first use */
static group_t *group(pra_ctx *c, int g)
{
    if (g >= c->ng) {
        group_t *ng = (group_t *)realloc(c->g, (size_t)(g + 1) * sizeof(group_t));
        if (!ng) abort();
        memset(ng + c->ng, 0, (size_t)(g + 1 - c->ng) * sizeof(group_t));
        for (int i = c->ng; i <= g; i++) ng[i].running = -1;
        c->g = ng;
        c->ng = g + 1;
    }
    group_t *G = &c->g[g];
    if (!G->o) {
        G->o = ora_ctx_create(c->nb);
        G->running = -1;
        G->iterations = 0;
    }
    return G;
}

static u64 group_bytes(group_t *G)
{
    u64 b = 0;
    for (int i = 0; i < G->nrel; i++)
        for (int v = 0; v < 3; v++) b += ora_rel_count(G->o, G->rels[i]->id, v) * (u64)(G->rels[i]->arity + 1) * 8;
    return b;
}

static u64 ctx_bytes(pra_ctx *c)
{
    u64 b = 0;
    for (int g = 0; g < c->ng; g++)
        if (c->g[g].o) b += group_bytes(&c->g[g]);
    return b;
}

int pra_ctx_create(int device, int nbuckets, uint64_t hbm_budget, pra_ctx **out)
{
    (void)device;
    (void)hbm_budget;
    if (!out || nbuckets < 1) return PRA_E_ARG;
    pra_ctx *c = (pra_ctx *)calloc(1, sizeof(pra_ctx));
    if (!c) return PRA_E_ARG;
    c->nb = nbuckets;
    c->batch = 10;
    const char *cap = getenv("PRA_ORACLE_DEVICE_BYTES");
    c->capacity = cap ? strtoull(cap, NULL, 10) : 0;
    group(c, 0);
    *out = c;
    return PRA_OK;
}

/* the handles of the program stay valid (and owned by the caller) */
int pra_ctx_destroy(pra_ctx *c)
{
    if (!c) return PRA_E_ARG;
    for (int g = 0; g < c->ng; g++) {
        ora_ctx_destroy(c->g[g].o);
        free(c->g[g].rels);
        free(c->g[g].sccs);
    }
    free(c->g);
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

int pra_ctx_set_group(pra_ctx *c, int g)
{
    if (!c || g < 0 || g >= PRA_MAX_GROUPS) return PRA_E_ARG;
    group(c, g);
    c->cur = g;
    return PRA_OK;
}

int pra_ctx_group_joined(pra_ctx *c, int g, uint64_t *out)
{
    if (!c || !out || g < 0 || g >= c->ng) return PRA_E_ARG;
    *out = c->g[g].o ? ora_joined_tuples(c->g[g].o) : 0;
    return PRA_OK;
}

/* LIE::execute's choice for instance g (src/lie/lie.cpp:27-60) */
static int runnable(pra_ctx *c, int g)
{
    group_t *G = &c->g[g];
    return G->o && G->nscc ? ora_next_runnable(G->o) : -1;
}

int pra_ctx_step(pra_ctx *c, uint64_t *active)
{
    if (!c || !active) return PRA_E_ARG;
    for (int g = 0; g < c->ng; g++) {
        group_t *G = &c->g[g];
        if (G->running < 0) {
            const int s = runnable(c, g);
            if (s < 0) continue;
            ora_scc_begin(G->o, s, 1);
            G->sccs[s]->running = 1;
            G->running = s;
        }
        pra_scc *S = G->sccs[G->running];
        const u64 d = ora_scc_step(G->o, S->id, (int)c->batch);
        G->iterations += S->fixpoint ? c->batch : 1;
        if (!S->fixpoint || d == 0) {
            ora_scc_end(G->o, S->id);
            S->running = 0;
            S->ran = 1;
            G->running = -1;
        }
    }
    u64 left = 0;
    for (int g = 0; g < c->ng; g++)
        if (c->g[g].running >= 0 || runnable(c, g) >= 0) left++;
    *active = left;
    const u64 held = ctx_bytes(c);
    if (held > c->peak) c->peak = held;
    if (c->capacity && held > c->capacity) {
        char m[160];
        snprintf(m, sizeof(m), "modelled device out of memory: %llu bytes held > %llu", (unsigned long long)held,
                 (unsigned long long)c->capacity);
        return fail(c, PRA_E_NOMEM, m);
    }
    return PRA_OK;
}

int pra_ctx_group_status(pra_ctx *c, int g, int *done, uint64_t *iterations, uint64_t *bytes)
{
    if (!c || g < 0 || g >= c->ng) return PRA_E_ARG;
    group_t *G = &c->g[g];
    if (done) *done = G->running < 0 && runnable(c, g) < 0;
    if (iterations) *iterations = G->iterations;
    if (bytes) *bytes = G->o ? group_bytes(G) : 0;
    return PRA_OK;
}

int pra_ctx_release_group(pra_ctx *c, int g)
{
    if (!c || g < 0 || g >= c->ng) return PRA_E_ARG;
    group_t *G = &c->g[g];
    for (int i = 0; i < G->nrel; i++) free(G->rels[i]);
    for (int i = 0; i < G->nscc; i++) free(G->sccs[i]);
    free(G->rels);
    free(G->sccs);
    ora_ctx_destroy(G->o);
    memset(G, 0, sizeof(*G));
    G->running = -1;
    return PRA_OK;
}

/* the oracle keeps no scratch between iterations */
int pra_ctx_trim(pra_ctx *c) { return c ? PRA_OK : PRA_E_ARG; }

int pra_ctx_device_bytes(pra_ctx *c, uint64_t *live, uint64_t *peak)
{
    if (!c) return PRA_E_ARG;
    const u64 held = ctx_bytes(c);
    if (held > c->peak) c->peak = held;
    if (live) *live = held;
    if (peak) *peak = c->peak;
    return PRA_OK;
}

int pra_rel_create(pra_ctx *c, uint32_t arity, uint32_t jcc, uint32_t tag, int canonical, pra_rel **out)
{
    (void)canonical;
    if (!c || !out) return PRA_E_ARG;
    group_t *G = group(c, c->cur);
    int id = ora_rel_create(G->o, (int)arity, (int)jcc, (int)tag);
    if (id < 0) return fail(c, PRA_E_ARG, "bad relation shape");
    pra_rel *r = (pra_rel *)calloc(1, sizeof(pra_rel));
    r->c = c;
    r->o = G->o;
    r->group = c->cur;
    r->id = id;
    r->arity = (int)arity;
    G->rels = (pra_rel **)grow(G->rels, &G->caprel, G->nrel + 1, sizeof(pra_rel *));
    G->rels[G->nrel++] = r;
    *out = r;
    return PRA_OK;
}

int pra_rel_destroy(pra_rel *r)
{
    if (!r) return PRA_E_ARG;
    pra_ctx *c = r->c;
    if (r->group < c->ng) {
        group_t *G = &c->g[r->group];
        for (int i = 0; i < G->nrel; i++)
            if (G->rels[i] == r) {
                G->rels[i] = G->rels[--G->nrel];
                break;
            }
    }
    free(r);
    return PRA_OK;
}

int pra_rel_load_rows(pra_rel *r, const uint64_t *rows, uint64_t n, int version)
{
    if (!r || (n && !rows)) return PRA_E_ARG;
    if (ora_rel_load_rows(r->o, r->id, rows, n, version) != 0) return fail(r->c, PRA_E_ARG, "bad version");
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
    *out = ora_rel_count(r->o, r->id, version);
    return PRA_OK;
}

int pra_rel_download(pra_rel *r, int version, uint64_t *rows, uint64_t cap, uint64_t *nrows)
{
    if (!r || !nrows || version < 0 || version > 2) return PRA_E_ARG;
    *nrows = ora_rel_rows(r->o, r->id, version, rows, rows ? cap : 0);
    return PRA_OK;
}

int pra_rel_write_bin(pra_rel *r, const char *path, int version)
{
    if (!r || !path || version < 0 || version > 2) return PRA_E_ARG;
    int w = r->arity + 1;
    u64 n = ora_rel_count(r->o, r->id, version);
    u64 *buf = (u64 *)malloc((size_t)(n ? n : 1) * w * 8);
    ora_rel_rows(r->o, r->id, version, buf, n);
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
    group_t *G = group(c, c->cur);
    pra_scc *s = (pra_scc *)calloc(1, sizeof(pra_scc));
    s->c = c;
    s->o = G->o;
    s->group = c->cur;
    s->fixpoint = fixpoint;
    s->id = ora_scc_create(G->o, fixpoint, id);
    G->sccs = (pra_scc **)grow(G->sccs, &G->capscc, G->nscc + 1, sizeof(pra_scc *));
    G->sccs[G->nscc++] = s;
    *out = s;
    return PRA_OK;
}

int pra_scc_destroy(pra_scc *s)
{
    if (!s) return PRA_E_ARG;
    pra_ctx *c = s->c;
    if (s->group < c->ng) {
        group_t *G = &c->g[s->group];
        for (int i = 0; i < G->nscc; i++)
            if (G->sccs[i] == s) G->sccs[i] = NULL;
    }
    free(s);
    return PRA_OK;
}

int pra_scc_add_relation(pra_scc *s, pra_rel *r, int move_full_to_delta)
{
    if (!s || !r || r->o != s->o) return PRA_E_ARG;
    return ora_scc_add_relation(s->o, s->id, r->id, move_full_to_delta) ? fail(s->c, PRA_E_ARG, "too many relations")
                                                                         : PRA_OK;
}

int pra_scc_add_join(pra_scc *s, pra_rel *out, pra_rel *in0, int v0, pra_rel *in1, int v1, const int *proj, int n)
{
    if (!s || !out || !in0 || !in1 || !proj) return PRA_E_ARG;
    return ora_scc_add_join(s->o, s->id, out->id, in0->id, v0, in1->id, v1, proj, n) ? fail(s->c, PRA_E_ARG, "bad join")
                                                                                     : PRA_OK;
}

int pra_scc_add_copy(pra_scc *s, pra_rel *out, pra_rel *in, int v, const int *map, int n)
{
    if (!s || !out || !in || !map) return PRA_E_ARG;
    return ora_scc_add_copy(s->o, s->id, out->id, in->id, v, map, n) ? fail(s->c, PRA_E_ARG, "bad copy") : PRA_OK;
}

int pra_scc_add_acopy(pra_scc *s, pra_rel *out, pra_rel *in, int v, const int *map, int n)
{
    if (!s || !out || !in || !map) return PRA_E_ARG;
    return ora_scc_add_acopy(s->o, s->id, out->id, in->id, v, map, n) ? fail(s->c, PRA_E_ARG, "bad acopy") : PRA_OK;
}

int pra_scc_add_copy_filter(pra_scc *s, pra_rel *out, pra_rel *in, int v, const int *map, int n, int op)
{
    if (!s || !out || !in || !map) return PRA_E_ARG;
    return ora_scc_add_copy_filter(s->o, s->id, out->id, in->id, v, map, n, op) ? fail(s->c, PRA_E_ARG, "bad filter")
                                                                                : PRA_OK;
}

/* `before` runs to its end before `after` starts (one instance only) */
int pra_scc_add_dependency(pra_scc *before, pra_scc *after)
{
    if (!before || !after || before->o != after->o) return PRA_E_ARG;
    ora_add_dependency(before->o, before->id, after->id);
    return PRA_OK;
}

int pra_scc_run(pra_scc *s, uint64_t *delta_hist, uint64_t cap, uint64_t *nbatches)
{
    if (!s) return PRA_E_ARG;
    if (s->ran) return fail(s->c, PRA_E_STATE, "SCC already ran");
    ora_scc_run(s->o, s->id, (int)s->c->batch);
    s->ran = 1;
    int n = ora_scc_history(s->o, s->id, delta_hist, delta_hist ? (int)cap : 0);
    if (nbatches) *nbatches = (uint64_t)n;
    return PRA_OK;
}

int pra_scc_begin(pra_scc *s, int move_full_to_delta)
{
    if (!s) return PRA_E_ARG;
    if (s->ran || s->running) return fail(s->c, PRA_E_STATE, "SCC already started");
    ora_scc_begin(s->o, s->id, move_full_to_delta);
    s->running = 1;
    return PRA_OK;
}

int pra_scc_step(pra_scc *s, uint64_t *delta)
{
    if (!s) return PRA_E_ARG;
    if (!s->running) return fail(s->c, PRA_E_STATE, "SCC not started");
    u64 d = ora_scc_step(s->o, s->id, (int)s->c->batch);
    if (delta) *delta = d;
    return PRA_OK;
}

int pra_scc_end(pra_scc *s)
{
    if (!s) return PRA_E_ARG;
    if (!s->running) return fail(s->c, PRA_E_STATE, "SCC not started");
    ora_scc_end(s->o, s->id);
    s->running = 0;
    s->ran = 1;
    return PRA_OK;
}
