/*
 * pra_oracle.c -- CPU restatement of parallel-RA's semi-naive evaluation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP backend is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * It restates, with plain sorted arrays and hash sets instead of the
 * reference's nested B-tree tries and MPI ranks, the exact observable behaviour
 * of the reference on k ranks ("buckets"), including the intern-id column:
 *
 *   tuple_hash (FNV-1a-64 over LE bytes)      src/hash/hash.h:11-30
 *   relation versions, find/insert           src/relation/balanced_hash_relation.cpp:1015-1115
 *   insert_delta_in_full / full_in_delta     src/relation/balanced_hash_relation.cpp:1120-1179
 *   newt -> delta move                        src/relation/balanced_hash_relation.cpp:1183-1231
 *   lexicographic trie walk order            src/relation/shmap_relation_exp.cpp:77-112
 *   copy / copy_filter / acopy emitters      src/relation/shmap_relation_exp.cpp:116-265
 *   join emitters (LEFT / RIGHT)             src/relation/shmap_relation_exp.cpp:326-455
 *   local_join, 20M threshold + offset       src/RA/parallel_join.cpp:157-230
 *   local_compute version dispatch           src/RAM/RA_tasks.cpp:305-555
 *   exchange arrival order (src-major, RA)   src/comm/all_to_allv_comm.cpp:64-142
 *   local_insert_in_newt + intern ids        src/RAM/RA_tasks.cpp:824-905
 *   local_insert_in_full, fixed-point test   src/RAM/RA_tasks.cpp:1044-1085
 *   execute_in_batches                       src/RAM/RA_tasks.cpp:1121-1269
 *   LIE scheduling and SCC prologue/epilogue src/lie/lie.cpp:27-102, 197-448
 *   input routing of loaded rows             src/IO/parallel_io.cpp:239-297,
 *                                            src/relation/balanced_hash_relation.cpp:773-815
 *
 * Rank r of the reference owns bucket r (one sub-bucket per bucket,
 * src/relation/balanced_hash_relation.cpp:665,718-749), so "source rank s" in
 * the reference is "source bucket s" here.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;

#define ORA_DELTA 0
#define ORA_FULL 1
#define ORA_NEWT 2

#define RA_COPY 0
#define RA_COPY_FILTER 1
#define RA_ACOPY 3
#define RA_JOIN 4

#define F_EQ 0
#define F_NEQ 1
#define F_LT 2
#define F_LE 3
#define F_GT 4
#define F_GE 5
#define F_TRUE 6
#define F_FALSE 7

#define MAXW 64

/* ------------------------------------------------------------------ hash -- */

/* FNV-1a-64 over the 8 little-endian bytes of each word (src/hash/hash.h:11-30). */
u64 ora_tuple_hash(const u64 *t, u64 n)
{
    u64 h = 14695981039346656037ULL;
    for (u64 i = 0; i < n; i++) {
        u64 w = t[i];
        for (int b = 0; b < 8; b++) {
            h ^= (w >> (8 * b)) & 255ULL;
            h *= 1099511628211ULL;
        }
    }
    return h;
}

static inline u64 mix_words(const u64 *t, int w)
{
    u64 h = 0x243F6A8885A308D3ULL ^ (u64)w;
    for (int i = 0; i < w; i++) {
        h ^= t[i] + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
        h ^= h >> 33;
        h *= 0xff51afd7ed558ccdULL;
        h ^= h >> 33;
    }
    return h;
}

/* ---------------------------------------------------- fixed-width tuple set */

typedef struct {
    int w;
    u64 cap, n;
    u64 *keys;
    unsigned char *used;
} tset;

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "pra_oracle: out of memory\n"); abort(); }
    return p;
}

static void *xrealloc(void *p, size_t s)
{
    void *q = realloc(p, s ? s : 1);
    if (!q) { fprintf(stderr, "pra_oracle: out of memory\n"); abort(); }
    return q;
}

static void tset_init(tset *s, int w)
{
    s->w = w; s->cap = 0; s->n = 0; s->keys = NULL; s->used = NULL;
}

static void tset_free(tset *s)
{
    free(s->keys); free(s->used);
    s->keys = NULL; s->used = NULL; s->cap = s->n = 0;
}

static int tset_find(const tset *s, const u64 *t)
{
    if (s->cap == 0) return 0;
    u64 m = s->cap - 1, i = mix_words(t, s->w) & m;
    for (;;) {
        if (!s->used[i]) return 0;
        if (memcmp(s->keys + i * s->w, t, sizeof(u64) * s->w) == 0) return 1;
        i = (i + 1) & m;
    }
}

static void tset_grow(tset *s);

/* returns 1 if inserted, 0 if already present */
static int tset_insert(tset *s, const u64 *t)
{
    if (2 * (s->n + 1) > s->cap) tset_grow(s);
    u64 m = s->cap - 1, i = mix_words(t, s->w) & m;
    for (;;) {
        if (!s->used[i]) {
            s->used[i] = 1;
            memcpy(s->keys + i * s->w, t, sizeof(u64) * s->w);
            s->n++;
            return 1;
        }
        if (memcmp(s->keys + i * s->w, t, sizeof(u64) * s->w) == 0) return 0;
        i = (i + 1) & m;
    }
}

static void tset_grow(tset *s)
{
    u64 oc = s->cap;
    u64 *ok = s->keys;
    unsigned char *ou = s->used;
    s->cap = oc ? oc * 2 : 64;
    s->keys = (u64 *)xcalloc(s->cap * (s->w ? s->w : 1), sizeof(u64));
    s->used = (unsigned char *)xcalloc(s->cap, 1);
    s->n = 0;
    for (u64 i = 0; i < oc; i++)
        if (ou[i]) tset_insert(s, ok + i * s->w);
    free(ok); free(ou);
}

/* ------------------------------------------------------ u64 -> u64 map ---- */

typedef struct {
    u64 cap, n;
    u64 *k, *v;
    unsigned char *used;
} umap;

static void umap_grow(umap *m);

static u64 *umap_slot(umap *m, u64 key, int create)
{
    if (create && 2 * (m->n + 1) > m->cap) umap_grow(m);
    if (m->cap == 0) return NULL;
    u64 mask = m->cap - 1, i = mix_words(&key, 1) & mask;
    for (;;) {
        if (!m->used[i]) {
            if (!create) return NULL;
            m->used[i] = 1; m->k[i] = key; m->v[i] = 0; m->n++;
            return &m->v[i];
        }
        if (m->k[i] == key) return &m->v[i];
        i = (i + 1) & mask;
    }
}

static void umap_grow(umap *m)
{
    u64 oc = m->cap, *ok = m->k, *ov = m->v;
    unsigned char *ou = m->used;
    m->cap = oc ? oc * 2 : 64;
    m->k = (u64 *)xcalloc(m->cap, sizeof(u64));
    m->v = (u64 *)xcalloc(m->cap, sizeof(u64));
    m->used = (unsigned char *)xcalloc(m->cap, 1);
    m->n = 0;
    for (u64 i = 0; i < oc; i++)
        if (ou[i]) *umap_slot(m, ok[i], 1) = ov[i];
    free(ok); free(ov); free(ou);
}

/* ----------------------------------------------------------- row vector --- */

typedef struct {
    u64 *rows;
    u64 n, cap;
    u64 nsorted;    /* rows[0, nsorted) are in lexicographic order */
} rowvec;

static void rv_push(rowvec *v, const u64 *row, int w)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 16;
        v->rows = (u64 *)xrealloc(v->rows, v->cap * w * sizeof(u64));
    }
    memcpy(v->rows + v->n * w, row, w * sizeof(u64));
    v->n++;
}

static void rv_free(rowvec *v)
{
    free(v->rows);
    v->rows = NULL; v->n = v->cap = v->nsorted = 0;
}

/* Lexicographic row order over all columns: the order in which the nested
 * B-tree trie is walked (src/relation/shmap_relation_exp.cpp:77-112). */
static int g_cmp_w;
static int cmp_rows(const void *a, const void *b)
{
    const u64 *x = (const u64 *)a, *y = (const u64 *)b;
    for (int i = 0; i < g_cmp_w; i++) {
        if (x[i] < y[i]) return -1;
        if (x[i] > y[i]) return 1;
    }
    return 0;
}

static void rv_sort(rowvec *v, int w)
{
    if (v->nsorted == v->n) return;
    g_cmp_w = w;
    u64 a = v->nsorted, n = v->n;
    /* sort the rows appended since the last sort, then merge the two runs */
    qsort(v->rows + a * w, n - a, w * sizeof(u64), cmp_rows);
    if (a) {
        /* merge from the back: the tail run moves to a temporary, the head
         * run stays in place (rows of one version are distinct) */
        u64 d = n - a;
        u64 *t = (u64 *)xrealloc(NULL, d * w * sizeof(u64));
        memcpy(t, v->rows + a * w, d * w * sizeof(u64));
        u64 i = a, j = d, k = n;
        while (j > 0) {
            if (i > 0 && cmp_rows(v->rows + (i - 1) * w, t + (j - 1) * w) > 0)
                memcpy(v->rows + (--k) * w, v->rows + (--i) * w, w * sizeof(u64));
            else
                memcpy(v->rows + (--k) * w, t + (--j) * w, w * sizeof(u64));
        }
        free(t);
    }
    v->nsorted = n;
}

/* first index whose first p columns are >= key */
static u64 rv_lower(const rowvec *v, int w, const u64 *key, int p)
{
    u64 lo = 0, hi = v->n;
    while (lo < hi) {
        u64 mid = (lo + hi) / 2;
        const u64 *r = v->rows + mid * w;
        int c = 0;
        for (int i = 0; i < p; i++) {
            if (r[i] < key[i]) { c = -1; break; }
            if (r[i] > key[i]) { c = 1; break; }
        }
        if (c < 0) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static u64 rv_upper(const rowvec *v, int w, const u64 *key, int p)
{
    u64 lo = 0, hi = v->n;
    while (lo < hi) {
        u64 mid = (lo + hi) / 2;
        const u64 *r = v->rows + mid * w;
        int c = 0;
        for (int i = 0; i < p; i++) {
            if (r[i] < key[i]) { c = -1; break; }
            if (r[i] > key[i]) { c = 1; break; }
        }
        if (c <= 0) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* --------------------------------------------------------------- model --- */

typedef struct {
    rowvec *b;      /* one row vector per bucket */
    tset rows;      /* all arity+1 columns (trie insert dedup) */
    tset prefix;    /* first arity columns (find_in_* with length = arity) */
    u64 count;
} version;

typedef struct {
    int arity, jcc, tag;
    version v[3];
} relation_t;

typedef struct {
    int type;
    int out, in0, v0, in1, v1;
    int map[MAXW];
    int nmap;
    int fop;
} ra_t;

typedef struct {
    int fixpoint, id;
    int nrel;
    int rel[64];
    int status[64];
    int nra;
    ra_t *ra;
    u64 *hist;      /* per executed batch: global delta sum */
    int nhist, caphist;
    u64 *bhist;     /* per iteration: [relation in SCC][bucket] delta counts */
    int nbhist, capbhist;
    int done;
} scc_t;

typedef struct {
    int nb;
    long long threshold;
    int nrel, caprel;
    relation_t *rel;
    int nscc, capscc;
    scc_t *scc;
    int ndep, capdep;
    int *dep;       /* pairs (before, after) */
    umap intern;
    u64 joined;     /* join pairs, counted before any dedup */
    u64 max_iters;  /* 0 = unlimited (for bounded timing samples) */
    u64 iters_run;
    char err[256];
} ora_ctx;

static u64 fnv_bucket(const ora_ctx *c, const u64 *t, int n)
{
    return ora_tuple_hash(t, (u64)n) % (u64)c->nb;
}

/* ------------------------------------------------------------ versions --- */

static void ver_init(ora_ctx *c, relation_t *r, int v)
{
    version *x = &r->v[v];
    x->b = (rowvec *)xcalloc(c->nb, sizeof(rowvec));
    tset_init(&x->rows, r->arity + 1);
    tset_init(&x->prefix, r->arity);
    x->count = 0;
}

static void ver_free(ora_ctx *c, relation_t *r, int v)
{
    version *x = &r->v[v];
    for (int i = 0; i < c->nb; i++) rv_free(&x->b[i]);
    free(x->b);
    x->b = NULL;
    tset_free(&x->rows);
    tset_free(&x->prefix);
    x->count = 0;
}

/* insert_in_{full,delta,newt}: trie insert of all arity+1 columns into the
 * bucket of the first jcc columns; returns 1 when the row was new
 * (src/relation/balanced_hash_relation.cpp:1039-1115). */
static int ver_insert(ora_ctx *c, relation_t *r, int v, const u64 *row)
{
    version *x = &r->v[v];
    if (!tset_insert(&x->rows, row)) return 0;
    tset_insert(&x->prefix, row);
    rv_push(&x->b[fnv_bucket(c, row, r->jcc)], row, r->arity + 1);
    x->count++;
    return 1;
}

/* find_in_{full,delta,newt}(t, length): prefix lookup of `length` columns
 * (src/relation/balanced_hash_relation.cpp:1015-1035). */
static int ver_find(relation_t *r, int v, const u64 *t, int len)
{
    if (len == r->arity + 1) return tset_find(&r->v[v].rows, t);
    if (len == r->arity) return tset_find(&r->v[v].prefix, t);
    fprintf(stderr, "pra_oracle: unsupported find length %d\n", len);
    abort();
}

static void ver_clear(ora_ctx *c, relation_t *r, int v)
{
    ver_free(c, r, v);
    ver_init(c, r, v);
}

/* full -> delta at SCC start (src/relation/balanced_hash_relation.cpp:1152-1179) */
static void rel_full_to_delta(ora_ctx *c, relation_t *r)
{
    int w = r->arity + 1;
    for (int b = 0; b < c->nb; b++) {
        rowvec *v = &r->v[ORA_FULL].b[b];
        rv_sort(v, w);
        for (u64 i = 0; i < v->n; i++) ver_insert(c, r, ORA_DELTA, v->rows + i * w);
    }
    ver_clear(c, r, ORA_FULL);
}

/* delta -> full (src/relation/balanced_hash_relation.cpp:1120-1148) */
static void rel_delta_to_full(ora_ctx *c, relation_t *r)
{
    int w = r->arity + 1;
    for (int b = 0; b < c->nb; b++) {
        rowvec *v = &r->v[ORA_DELTA].b[b];
        for (u64 i = 0; i < v->n; i++) ver_insert(c, r, ORA_FULL, v->rows + i * w);
    }
    ver_clear(c, r, ORA_DELTA);
}

/* delta = newt; newt = {} (src/relation/balanced_hash_relation.cpp:1183-1231) */
static void rel_newt_to_delta(ora_ctx *c, relation_t *r)
{
    ver_free(c, r, ORA_DELTA);
    r->v[ORA_DELTA] = r->v[ORA_NEWT];
    ver_init(c, r, ORA_NEWT);
}

/* ----------------------------------------------------------- emitters ---- */

typedef struct {
    rowvec *buf;    /* [src][ra][dst] */
    int nra, nb;
} sendbufs;

static rowvec *sb_at(sendbufs *s, int src, int ra, int dst)
{
    return &s->buf[((size_t)src * s->nra + ra) * s->nb + dst];
}

static int filter_ok(int op, u64 a, u64 b)
{
    switch (op) {
    case F_EQ: return a == b;
    case F_NEQ: return a != b;
    case F_LT: return a < b;
    case F_LE: return a <= b;
    case F_GT: return a > b;
    case F_GE: return a >= b;
    case F_TRUE: return 1;
    default: return 0;
    }
}

static int ra_out_width(ora_ctx *c, const ra_t *ra)
{
    (void)c;
    return ra->nmap;
}

/* copy / copy_filter / acopy over the source bucket in trie order
 * (src/relation/shmap_relation_exp.cpp:116-265) */
static void emit_copylike(ora_ctx *c, const ra_t *ra, int ri, int s, sendbufs *sb)
{
    relation_t *in = &c->rel[ra->in0], *out = &c->rel[ra->out];
    int w = in->arity + 1, ow = ra_out_width(c, ra);
    rowvec *v = &in->v[ra->v0].b[s];
    rv_sort(v, w);
    u64 o[MAXW];
    for (u64 i = 0; i < v->n; i++) {
        const u64 *row = v->rows + i * w;
        if (ra->type == RA_COPY_FILTER && !filter_ok(ra->fop, row[0], row[1])) continue;
        for (int j = 0; j < ow; j++) o[j] = row[ra->map[j]];
        int dst = (int)fnv_bucket(c, o, out->jcc);
        rv_push(sb_at(sb, s, ri, dst), o, ow);
    }
}

typedef struct {
    relation_t *outer, *inner;
    int ov, iv;
    int left;   /* LEFT: concat = outer ++ inner[jcc:];  RIGHT: inner ++ outer[jcc:] */
} join_sides;

static join_sides join_plan(ora_ctx *c, const ra_t *ra)
{
    join_sides js;
    /* src/RAM/RA_tasks.cpp:438-517: DELTA on input0 -> LEFT (outer = input0);
     * FULL on input0 -> RIGHT (outer = input1, probed trie = input0 full). */
    if (ra->v0 == ORA_DELTA) {
        js.outer = &c->rel[ra->in0]; js.ov = ORA_DELTA;
        js.inner = &c->rel[ra->in1]; js.iv = ra->v1;
        js.left = 1;
    } else {
        js.outer = &c->rel[ra->in1]; js.ov = ra->v1;
        js.inner = &c->rel[ra->in0]; js.iv = ORA_FULL;
        js.left = 0;
    }
    return js;
}

/* pre-dedup join pairs of one source bucket: the joined-tuples unit of work */
static u64 join_pairs(ora_ctx *c, const ra_t *ra, int s)
{
    join_sides js = join_plan(c, ra);
    int jcc = c->rel[ra->in0].jcc;
    int wo = js.outer->arity + 1, wi = js.inner->arity + 1;
    rowvec *ov = &js.outer->v[js.ov].b[s], *iv = &js.inner->v[js.iv].b[s];
    rv_sort(ov, wo); rv_sort(iv, wi);
    u64 n = 0;
    for (u64 i = 0; i < ov->n; i++) {
        const u64 *d = ov->rows + i * wo;
        if ((int)fnv_bucket(c, d, jcc) != s) continue;
        n += rv_upper(iv, wi, d, jcc) - rv_lower(iv, wi, d, jcc);
    }
    return n;
}

/* parallel_join::local_join for one source bucket: walks the outer rows from
 * *offset (in words), emits per-call deduplicated projections, and stops once
 * more than `threshold` outputs were emitted (src/RA/parallel_join.cpp:157-230,
 * src/relation/shmap_relation_exp.cpp:326-455).  Returns 1 when complete. */
static int local_join(ora_ctx *c, const ra_t *ra, int ri, int s, long long *offset, sendbufs *sb)
{
    join_sides js = join_plan(c, ra);
    relation_t *out = &c->rel[ra->out];
    int jcc = c->rel[ra->in0].jcc;
    int wo = js.outer->arity + 1, wi = js.inner->arity + 1, ow = ra->nmap;
    rowvec *ov = &js.outer->v[js.ov].b[s], *iv = &js.inner->v[js.iv].b[s];
    rv_sort(ov, wo); rv_sort(iv, wi);
    long long words = (long long)ov->n * wo;
    if (*offset > words || ov->n == 0 || iv->n == 0) return 1;

    tset dedup;
    tset_init(&dedup, ow);
    long long cnt = 0;
    u64 cat[2 * MAXW + 2], o[MAXW];
    for (u64 i = (u64)(*offset / wo); i < ov->n; i++) {
        const u64 *d = ov->rows + i * wo;
        if ((int)fnv_bucket(c, d, jcc) == s) {
            u64 lo = rv_lower(iv, wi, d, jcc), hi = rv_upper(iv, wi, d, jcc);
            for (u64 j = lo; j < hi; j++) {
                const u64 *f = iv->rows + j * wi;
                int n = 0;
                if (js.left) {
                    for (int x = 0; x < wo; x++) cat[n++] = d[x];
                    for (int x = jcc; x < wi; x++) cat[n++] = f[x];
                } else {
                    for (int x = 0; x < wi; x++) cat[n++] = f[x];
                    for (int x = jcc; x < wo; x++) cat[n++] = d[x];
                }
                for (int x = 0; x < ow; x++) o[x] = cat[ra->map[x]];
                if (tset_insert(&dedup, o)) {
                    int dst = (int)fnv_bucket(c, o, out->jcc);
                    rv_push(sb_at(sb, s, ri, dst), o, ow);
                    cnt++;
                }
            }
        }
        if (cnt > c->threshold) {
            *offset = (long long)(i + 1) * wo;
            tset_free(&dedup);
            return 0;
        }
    }
    tset_free(&dedup);
    return 1;
}

/* local_insert_in_newt for one received segment
 * (src/RAM/RA_tasks.cpp:824-905) */
static void insert_segment(ora_ctx *c, const ra_t *ra, rowvec *seg)
{
    relation_t *out = &c->rel[ra->out];
    int ow = ra->nmap;
    u64 t[MAXW + 1];
    for (u64 i = 0; i < seg->n; i++) {
        const u64 *x = seg->rows + i * ow;
        if (ra->type == RA_ACOPY) {
            if (!ver_find(out, ORA_FULL, x, ow) && !ver_find(out, ORA_DELTA, x, ow))
                ver_insert(c, out, ORA_NEWT, x);
        } else {
            if (ver_find(out, ORA_FULL, x, ow) || ver_find(out, ORA_DELTA, x, ow) ||
                ver_find(out, ORA_NEWT, x, ow))
                continue;
            memcpy(t, x, ow * sizeof(u64));
            u64 key = ((u64)out->tag << 46) | (fnv_bucket(c, t, out->jcc) << 28);
            u64 *slot = umap_slot(&c->intern, key, 1);
            u64 seq = (*slot)++;
            t[ow] = key | seq;
            ver_insert(c, out, ORA_NEWT, t);
        }
    }
}

/* one semi-naive iteration: local compute rounds + exchange + newt insert,
 * then local_insert_in_full (src/RAM/RA_tasks.cpp:1121-1269) */
static void run_iteration(ora_ctx *c, scc_t *sc)
{
    int nb = c->nb, nra = sc->nra;
    long long *offset = (long long *)xcalloc((size_t)nb * (nra ? nra : 1), sizeof(long long));

    for (int s = 0; s < nb; s++)
        for (int r = 0; r < nra; r++)
            if (sc->ra[r].type == RA_JOIN) c->joined += join_pairs(c, &sc->ra[r], s);

    int all_done;
    do {
        all_done = 1;
        sendbufs sb;
        sb.nra = nra; sb.nb = nb;
        sb.buf = (rowvec *)xcalloc((size_t)nb * nra * nb + 1, sizeof(rowvec));
        for (int s = 0; s < nb; s++) {
            for (int r = 0; r < nra; r++) {
                const ra_t *ra = &sc->ra[r];
                if (ra->type == RA_JOIN)
                    all_done &= local_join(c, ra, r, s, &offset[(size_t)s * nra + r], &sb);
                else
                    emit_copylike(c, ra, r, s, &sb);
            }
        }
        /* receiver bucket b: source-major, RA-minor, emission order */
        for (int b = 0; b < nb; b++)
            for (int s = 0; s < nb; s++)
                for (int r = 0; r < nra; r++)
                    insert_segment(c, &sc->ra[r], sb_at(&sb, s, r, b));
        for (size_t i = 0; i < (size_t)nb * nra * nb; i++) rv_free(&sb.buf[i]);
        free(sb.buf);
    } while (!all_done);
    free(offset);

    for (int i = 0; i < sc->nrel; i++) {
        relation_t *r = &c->rel[sc->rel[i]];
        rel_delta_to_full(c, r);
        rel_newt_to_delta(c, r);
    }
    int row = sc->nrel * nb;
    if (sc->nbhist == sc->capbhist) {
        sc->capbhist = sc->capbhist ? 2 * sc->capbhist : 64;
        sc->bhist = (u64 *)xrealloc(sc->bhist, (size_t)sc->capbhist * (row ? row : 1) * sizeof(u64));
    }
    for (int i = 0; i < sc->nrel; i++)
        for (int b = 0; b < nb; b++)
            sc->bhist[(size_t)sc->nbhist * row + i * nb + b] = c->rel[sc->rel[i]].v[ORA_DELTA].b[b].n;
    sc->nbhist++;
    c->iters_run++;
}

static void hist_push(scc_t *sc, u64 v)
{
    if (sc->nhist == sc->caphist) {
        sc->caphist = sc->caphist ? 2 * sc->caphist : 64;
        sc->hist = (u64 *)xrealloc(sc->hist, sc->caphist * sizeof(u64));
    }
    sc->hist[sc->nhist++] = v;
}

static u64 scc_delta_sum(ora_ctx *c, scc_t *sc)
{
    u64 d = 0;
    for (int i = 0; i < sc->nrel; i++) d += c->rel[sc->rel[i]].v[ORA_DELTA].count;
    return d;
}

/* first SCC in add order that no unfinished SCC points to (src/lie/lie.cpp:27-60) */
static int one_runnable(ora_ctx *c)
{
    for (int i = 0; i < c->nscc; i++) {
        if (c->scc[i].done) continue;
        int blocked = 0;
        for (int e = 0; e < c->ndep && !blocked; e++) {
            int a = c->dep[2 * e], b = c->dep[2 * e + 1];
            if (b == i && !c->scc[a].done) blocked = 1;
        }
        if (!blocked) return i;
    }
    return -1;
}

static int limit_hit(ora_ctx *c)
{
    return c->max_iters && c->iters_run >= c->max_iters;
}

/* ---------------------------------------------------------------- API ---- */

ora_ctx *ora_ctx_create(int nbuckets)
{
    if (nbuckets < 1) return NULL;
    ora_ctx *c = (ora_ctx *)xcalloc(1, sizeof(ora_ctx));
    c->nb = nbuckets;
    c->threshold = 20000000;   /* src/RAM/RA_tasks.cpp:312 */
    return c;
}

void ora_ctx_destroy(ora_ctx *c)
{
    if (!c) return;
    for (int i = 0; i < c->nrel; i++)
        for (int v = 0; v < 3; v++) ver_free(c, &c->rel[i], v);
    free(c->rel);
    for (int i = 0; i < c->nscc; i++) { free(c->scc[i].ra); free(c->scc[i].hist); free(c->scc[i].bhist); }
    free(c->scc);
    free(c->dep);
    free(c->intern.k); free(c->intern.v); free(c->intern.used);
    free(c);
}

void ora_set_join_threshold(ora_ctx *c, long long t) { c->threshold = t; }
void ora_set_max_iterations(ora_ctx *c, u64 n) { c->max_iters = n; }
u64 ora_joined_tuples(ora_ctx *c) { return c->joined; }
u64 ora_iterations_run(ora_ctx *c) { return c->iters_run; }

int ora_rel_create(ora_ctx *c, int arity, int jcc, int tag)
{
    if (arity < 1 || arity + 1 > MAXW || jcc < 0 || jcc > arity) return -1;
    if (c->nrel == c->caprel) {
        c->caprel = c->caprel ? 2 * c->caprel : 16;
        c->rel = (relation_t *)xrealloc(c->rel, c->caprel * sizeof(relation_t));
    }
    relation_t *r = &c->rel[c->nrel];
    memset(r, 0, sizeof(*r));
    r->arity = arity; r->jcc = jcc; r->tag = tag;
    for (int v = 0; v < 3; v++) ver_init(c, r, v);
    return c->nrel++;
}

/* populate_full / populate_delta of loaded rows (arity+1 columns each),
 * src/relation/balanced_hash_relation.cpp:773-815 */
int ora_rel_load_rows(ora_ctx *c, int rel, const u64 *rows, u64 n, int version)
{
    if (rel < 0 || rel >= c->nrel || (version != ORA_FULL && version != ORA_DELTA)) return -1;
    relation_t *r = &c->rel[rel];
    for (u64 i = 0; i < n; i++) ver_insert(c, r, version, rows + i * (r->arity + 1));
    return 0;
}

int ora_scc_create(ora_ctx *c, int fixpoint, int id)
{
    if (c->nscc == c->capscc) {
        c->capscc = c->capscc ? 2 * c->capscc : 16;
        c->scc = (scc_t *)xrealloc(c->scc, c->capscc * sizeof(scc_t));
    }
    scc_t *s = &c->scc[c->nscc];
    memset(s, 0, sizeof(*s));
    s->fixpoint = fixpoint; s->id = id;
    return c->nscc++;
}

int ora_scc_add_relation(ora_ctx *c, int scc, int rel, int move_full_to_delta)
{
    scc_t *s = &c->scc[scc];
    if (s->nrel >= 64) return -1;
    s->rel[s->nrel] = rel;
    s->status[s->nrel] = move_full_to_delta;
    s->nrel++;
    return 0;
}

static ra_t *scc_new_ra(scc_t *s)
{
    s->ra = (ra_t *)xrealloc(s->ra, (s->nra + 1) * sizeof(ra_t));
    ra_t *r = &s->ra[s->nra++];
    memset(r, 0, sizeof(*r));
    return r;
}

int ora_scc_add_join(ora_ctx *c, int scc, int out, int in0, int v0, int in1, int v1, const int *proj, int nproj)
{
    relation_t *o = &c->rel[out];
    if (nproj != o->arity || nproj > MAXW) return -1;
    /* input 0's join column count is the one matched (src/RAM/RA_tasks.cpp:448);
     * with a different count on input 1, local_join below still walks each
     * bucket's own rows, which is what the reference's ranks see */
    if (c->rel[in0].jcc > c->rel[in1].arity) return -1;
    int span = c->rel[in0].arity + 1 + c->rel[in1].arity + 1 - c->rel[in0].jcc;
    for (int i = 0; i < nproj; i++) if (proj[i] < 0 || proj[i] >= span) return -1;
    ra_t *r = scc_new_ra(&c->scc[scc]);
    r->type = RA_JOIN; r->out = out; r->in0 = in0; r->v0 = v0; r->in1 = in1; r->v1 = v1;
    r->nmap = nproj;
    memcpy(r->map, proj, nproj * sizeof(int));
    return 0;
}

static int add_copylike(ora_ctx *c, int scc, int type, int out, int in, int v, const int *map, int n, int op)
{
    relation_t *o = &c->rel[out];
    int want = type == RA_ACOPY ? o->arity + 1 : o->arity;
    if (n != want || n > MAXW) return -1;
    for (int i = 0; i < n; i++) if (map[i] < 0 || map[i] > c->rel[in].arity) return -1;
    ra_t *r = scc_new_ra(&c->scc[scc]);
    r->type = type; r->out = out; r->in0 = in; r->v0 = v; r->nmap = n; r->fop = op;
    memcpy(r->map, map, n * sizeof(int));
    return 0;
}

int ora_scc_add_copy(ora_ctx *c, int scc, int out, int in, int v, const int *map, int n)
{
    return add_copylike(c, scc, RA_COPY, out, in, v, map, n, 0);
}

int ora_scc_add_acopy(ora_ctx *c, int scc, int out, int in, int v, const int *map, int n)
{
    return add_copylike(c, scc, RA_ACOPY, out, in, v, map, n, 0);
}

int ora_scc_add_copy_filter(ora_ctx *c, int scc, int out, int in, int v, const int *map, int n, int op)
{
    return add_copylike(c, scc, RA_COPY_FILTER, out, in, v, map, n, op);
}

int ora_add_dependency(ora_ctx *c, int before, int after)
{
    if (c->ndep == c->capdep) {
        c->capdep = c->capdep ? 2 * c->capdep : 16;
        c->dep = (int *)xrealloc(c->dep, 2 * c->capdep * sizeof(int));
    }
    c->dep[2 * c->ndep] = before;
    c->dep[2 * c->ndep + 1] = after;
    c->ndep++;
    return 0;
}

/* Run one SCC the way LIE::execute does: status-true relations move full to
 * delta, then either one iteration or batches until the global delta is empty,
 * then delta is folded into full (src/lie/lie.cpp:286-441).  The three steps
 * are also callable one at a time. */
int ora_scc_begin(ora_ctx *c, int scc, int move_status)
{
    scc_t *sc = &c->scc[scc];
    if (move_status)
        for (int i = 0; i < sc->nrel; i++)
            if (sc->status[i]) rel_full_to_delta(c, &c->rel[sc->rel[i]]);
    return 0;
}

/* one batch (one iteration for a non-fixpoint SCC); returns the global delta */
u64 ora_scc_step(ora_ctx *c, int scc, int batch_size)
{
    scc_t *sc = &c->scc[scc];
    int iters = sc->fixpoint ? batch_size : 1;
    for (int b = 0; b < iters && !limit_hit(c); b++) run_iteration(c, sc);
    u64 d = scc_delta_sum(c, sc);
    hist_push(sc, d);
    return d;
}

int ora_scc_end(ora_ctx *c, int scc)
{
    scc_t *sc = &c->scc[scc];
    for (int i = 0; i < sc->nrel; i++) rel_delta_to_full(c, &c->rel[sc->rel[i]]);
    sc->done = 1;
    return 0;
}

int ora_scc_run(ora_ctx *c, int scc, int batch_size)
{
    scc_t *sc = &c->scc[scc];
    ora_scc_begin(c, scc, 1);
    if (!sc->fixpoint) {
        ora_scc_step(c, scc, batch_size);
    } else {
        u64 d;
        do {
            d = ora_scc_step(c, scc, batch_size);
        } while (d != 0 && !limit_hit(c));
    }
    return ora_scc_end(c, scc);
}

/* the SCC LIE::execute would run next, or -1 when none is left */
int ora_next_runnable(ora_ctx *c) { return one_runnable(c); }

int ora_execute(ora_ctx *c, int batch_size)
{
    int t;
    while ((t = one_runnable(c)) >= 0) {
        ora_scc_run(c, t, batch_size);
        if (limit_hit(c)) break;
    }
    return 0;
}

u64 ora_rel_count(ora_ctx *c, int rel, int version)
{
    return c->rel[rel].v[version].count;
}

u64 ora_rel_bucket_count(ora_ctx *c, int rel, int version, int bucket)
{
    return c->rel[rel].v[version].b[bucket].n;
}

/* rows of one version, bucket-major and lexicographic inside each bucket */
u64 ora_rel_rows(ora_ctx *c, int rel, int version, u64 *out, u64 cap)
{
    relation_t *r = &c->rel[rel];
    int w = r->arity + 1;
    u64 n = 0;
    for (int b = 0; b < c->nb; b++) {
        rowvec *v = &r->v[version].b[b];
        rv_sort(v, w);
        for (u64 i = 0; i < v->n; i++, n++)
            if (n < cap) memcpy(out + n * w, v->rows + i * w, w * sizeof(u64));
    }
    return n;
}

int ora_scc_history(ora_ctx *c, int scc, u64 *out, int cap)
{
    scc_t *sc = &c->scc[scc];
    for (int i = 0; i < sc->nhist && i < cap; i++) out[i] = sc->hist[i];
    return sc->nhist;
}


/* per-iteration delta counts of every SCC relation in every bucket:
 * out[iter][relation index in SCC][bucket]; returns the iteration count */
int ora_scc_bucket_history(ora_ctx *c, int scc, u64 *out, u64 cap)
{
    scc_t *sc = &c->scc[scc];
    u64 row = (u64)sc->nrel * c->nb;
    for (u64 i = 0; i < (u64)sc->nbhist * row && i < cap; i++) out[i] = sc->bhist[i];
    return sc->nbhist;
}
