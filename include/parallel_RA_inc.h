/*
 * parallel_RA_inc.h -- drop-in C++ facade of parallel-RA's program surface over
 * the MI355X backend (pra.h).
 *
 * Generated Datalog programs (e.g. the reference's tests/transitive_closure.cpp)
 * include "../src/parallel_RA_inc.h" and use relation, RAM, parallel_join,
 * parallel_copy, parallel_acopy, parallel_copy_filter, LIE and mpi_comm.  This
 * header declares the same names with the same constructor signatures and
 * ownership rules, so such a program compiles unchanged and runs on one GPU.
 *
 * One process drives one GPU.  The rank layout of an `mpirun -n k` run of the
 * reference (bucket = tuple_hash(key) % k) is reproduced with k virtual buckets,
 * taken from the PRA_BUCKETS environment variable (default 1), which makes the
 * intern-id column identical to the k-rank reference run.  PRA_DEVICE selects
 * the GPU (default 0).
 *
 * Interface correspondence (reference file:line):
 *   relation            src/relation/balanced_hash_relation.h:17-233
 *   parallel_RA         src/RA/parallel_RA.h:11-72
 *   parallel_join       src/RA/parallel_join.h:10-58
 *   parallel_copy       src/RA/parallel_copy.h:28-32
 *   parallel_acopy      src/RA/parallel_acopy.h:45-49
 *   parallel_copy_filter src/RA/parallel_copy_filter.h:29-33
 *   parallel_copy_generate src/RA/parallel_copy_generate.h:29-33 (rejected at execute)
 *   RAM                 src/RAM/RA_tasks.h:13-165
 *   LIE                 src/lie/lie.h:13-153
 *   mpi_comm            src/comm/comm.h:8-76
 */
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "pra.h"
#include "pra_mpi_single.h"

typedef int8_t s8;
typedef int16_t s16;
typedef int32_t s32;
typedef int64_t s64;
typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef char c8;
typedef wchar_t c16;

/* src/relation/balanced_hash_relation.h:12-15 */
enum { LEFT = 0, RIGHT };
enum { DELTA = 0, FULL, FULL_AND_DELTA };
enum { COPY = 0, COPY_FILTER, COPY_GENERATE, ACOPY, JOIN };
enum { STATIC = 0, DYNAMIC };

/* FNV-1a-64 over LE bytes (src/hash/hash.h:11-30) */
inline u64 tuple_hash(const u64* start_ptr, u64 prefix_len)
{
    u64 h = 14695981039346656037ULL;
    for (u64 i = 0; i < prefix_len; ++i)
        for (int b = 0; b < 8; b++) {
            h ^= (start_ptr[i] >> (8 * b)) & 255ULL;
            h *= 1099511628211ULL;
        }
    return h;
}

namespace pra_facade {

inline void die(pra_ctx* ctx, int rc, const char* what)
{
    fprintf(stderr, "parallel-RA (MI355X backend): %s failed (%d): %s\n", what, rc,
            ctx ? pra_last_error(ctx) : "");
    exit(1);
}

inline void chk(pra_ctx* ctx, int rc, const char* what)
{
    if (rc != PRA_OK) die(ctx, rc, what);
}

inline bool file_exists(const std::string& p) { return access(p.c_str(), F_OK) == 0; }

/* a relation file as the reference writes it: rows (when any) and the size
 * record "rows\ncols" */
inline void write_rows_file(const std::string& path, const u64* rows, size_t words, long long size_rows, int size_cols,
                            bool data)
{
    if (data) {
        FILE* f = fopen(path.c_str(), "wb");
        if (!f || fwrite(rows, 8, words, f) != words) {
            std::cout << path << " Wrong IO" << std::endl;
            MPI_Abort(MPI_COMM_WORLD, -1);
        }
        fclose(f);
    }
    FILE* m = fopen((path + ".size").c_str(), "w");
    if (!m) {
        std::cout << path << ".size Wrong IO" << std::endl;
        MPI_Abort(MPI_COMM_WORLD, -1);
    }
    fprintf(m, "%lld\n%d", size_rows, size_cols);
    fclose(m);
}

inline int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

/* Map a copy_filter predicate over columns 0 and 1 to a backend predicate code
 * by evaluating it on probe rows: every pair of a set of edge values (0, 1,
 * 2, 5, 2^63 - 1, 2^63, 2^64 - 1), then 400 pseudo-random rows (equal, adjacent
 * and unrelated column-0/1 values), every row with pseudo-random values in
 * columns 2..7 (every other row below 16).  A predicate that disagrees with all of the comparisons below
 * on any probe -- one that is not such a comparison, or reads other columns --
 * is rejected (a host function cannot run on the GPU). */
inline u64 splitmix64(u64& x)
{
    u64 z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline int classify_filter(bool (*f)(const u64* const))
{
    static const u64 edge[] = {0, 1, 2, 5, 0x7FFFFFFFFFFFFFFFULL, 0x8000000000000000ULL, 0xFFFFFFFFFFFFFFFFULL};
    std::vector<std::array<u64, 8>> probes;
    u64 seed = 0x5EED5EED5EEDULL;
    auto row = [&](u64 a, u64 b) {
        std::array<u64, 8> r{};
        r[0] = a;
        r[1] = b;
        const bool small = probes.size() & 1;
        for (int k = 2; k < 8; k++) r[k] = small ? splitmix64(seed) % 16 : splitmix64(seed);
        probes.push_back(r);
    };
    for (u64 a : edge)
        for (u64 b : edge) row(a, b);
    for (int i = 0; i < 400; i++) {
        const u64 a = (i & 3) == 0 ? splitmix64(seed) % 16 : splitmix64(seed);
        const u64 kind = splitmix64(seed) % 4;
        row(a, kind == 0 ? a : kind == 1 ? a + 1 : kind == 2 ? a - 1 : splitmix64(seed));
    }
    const int codes[] = {PRA_FILTER_EQ, PRA_FILTER_NEQ, PRA_FILTER_LT, PRA_FILTER_LE,
                         PRA_FILTER_GT, PRA_FILTER_GE, PRA_FILTER_TRUE, PRA_FILTER_FALSE};
    for (int code : codes) {
        bool ok = true;
        for (const auto& p : probes) {
            bool want;
            switch (code) {
            case PRA_FILTER_EQ: want = p[0] == p[1]; break;
            case PRA_FILTER_NEQ: want = p[0] != p[1]; break;
            case PRA_FILTER_LT: want = p[0] < p[1]; break;
            case PRA_FILTER_LE: want = p[0] <= p[1]; break;
            case PRA_FILTER_GT: want = p[0] > p[1]; break;
            case PRA_FILTER_GE: want = p[0] >= p[1]; break;
            case PRA_FILTER_TRUE: want = true; break;
            default: want = false; break;
            }
            if (f(p.data()) != want) {
                ok = false;
                break;
            }
        }
        if (ok) return code;
    }
    return -1;
}

/* tuple_hash (src/hash/hash.h:11-30): FNV-1a-64 over the little-endian bytes
 * of the first n words */
inline u64 tuple_hash_host(const u64* t, int n)
{
    u64 h = 14695981039346656037ULL;
    for (int i = 0; i < n; i++)
        for (int b = 0; b < 8; b++) {
            h ^= (t[i] >> (8 * b)) & 255ULL;
            h *= 1099511628211ULL;
        }
    return h;
}

} // namespace pra_facade

/* ------------------------------------------------------------- mpi_comm --- */

class mpi_comm
{
    int rank = 0, nprocs = 1, local_rank = 0, local_nprocs = 1;
    MPI_Comm world_comm = MPI_COMM_WORLD, local_comm = MPI_COMM_WORLD;

public:
    mpi_comm() {}
    MPI_Comm get_comm() { return world_comm; }
    int get_nprocs() { return nprocs; }
    int get_rank() { return rank; }
    int get_local_nprocs() { return local_nprocs; }
    int get_local_rank() { return local_rank; }
    MPI_Comm get_local_comm() { return local_comm; }
    void set_local_comm(MPI_Comm c) { local_comm = c; }
    void create(int, char**) {}
    void destroy() {}
};

/* ------------------------------------------------------------- relation --- */

class relation
{
    u32 join_column_count;
    bool is_canonical;
    u32 arity;
    u32 intern_tag;
    std::string debug_id;
    int initailization_type = -1;
    std::string filename;
    pra_rel* handle = nullptr;

public:
    relation(u32 jcc, bool is_c, u32 ar, u32 tg, std::string fname, int version)
        : join_column_count(jcc), is_canonical(is_c), arity(ar), intern_tag(tg), initailization_type(version),
          filename(fname)
    {
    }
    relation(u32 jcc, bool is_c, u32 ar, u32 tg, std::string did, std::string fname, int version)
        : join_column_count(jcc), is_canonical(is_c), arity(ar), intern_tag(tg), debug_id(did),
          initailization_type(version), filename(fname)
    {
    }
    ~relation()
    {
        if (handle) pra_rel_destroy(handle);
    }

    u32 get_arity() { return arity; }
    int get_join_column_count() { return (int)join_column_count; }
    bool get_is_canonical() { return is_canonical; }
    u32 get_intern_tag() { return intern_tag; }
    std::string get_debug_id() { return debug_id; }
    std::string get_filename() { return filename; }
    void set_filename(std::string f) { filename = f; }
    void set_initailization_type(int x) { initailization_type = x; }
    /* IO modes of src/relation/balanced_hash_relation.h, set by LIE::execute */
    bool restart_flag = false, offset_io = false, share_io = false, separate_io = false;
    void set_restart_flag(bool b) { restart_flag = b; }
    void set_offset_io(bool b) { offset_io = b; }
    void set_share_io(bool b) { share_io = b; }
    void set_separate_io(bool b) { separate_io = b; }

    /* counts of this relation's versions */
    int get_full_element_count()
    {
        uint64_t n = 0;
        if (handle) pra_rel_count(handle, PRA_FULL, &n);
        return (int)n;
    }
    int get_delta_element_count()
    {
        uint64_t n = 0;
        if (handle) pra_rel_count(handle, PRA_DELTA, &n);
        return (int)n;
    }
    u32 get_global_full_element_count() { return (u32)get_full_element_count(); }
    u32 get_global_delta_element_count() { return (u32)get_delta_element_count(); }

    /* rows of a version in trie order (bucket-major, lexicographic) */
    std::vector<u64> rows(int version = FULL)
    {
        std::vector<u64> out;
        if (!handle) return out;
        uint64_t n = 0;
        pra_rel_count(handle, version, &n);
        out.resize((size_t)n * (arity + 1));
        uint64_t got = 0;
        pra_rel_download(handle, version, out.data(), n, &got);
        return out;
    }

    /* relation::print (src/relation/balanced_hash_relation.cpp:433-522): per
     * bucket with rows, "<name> <rank> FULL Rows <n> columns <w>" and one
     * "F [<j>] <words>" line per row in trie order, then the same for delta
     * ("D <words>"); bucket b lives on rank b of the reference's layout */
    void print(int nbuckets)
    {
        const int w = (int)arity + 1;
        for (int version : {FULL, DELTA}) {
            std::vector<u64> r = rows(version);
            const size_t n = r.size() / (size_t)w;
            size_t i = 0;
            for (int b = 0; b < nbuckets; b++) {
                size_t j = i;
                while (j < n && (int)(pra_facade::tuple_hash_host(&r[j * w], (int)join_column_count) %
                                      (u64)nbuckets) == b)
                    j++;
                if (j > i)
                    std::cout << debug_id << " " << b << (version == FULL ? " FULL" : " DELTA") << " Rows " << j - i
                              << " columns " << w << std::endl;
                for (size_t k = i; k < j; k++) {
                    if (version == FULL)
                        std::cout << "F [" << k - i << "] ";
                    else
                        std::cout << "D ";
                    for (int c = 0; c < w; c++) std::cout << r[k * w + c] << " ";
                    std::cout << std::endl;
                }
                i = j;
            }
        }
    }

    /* backend side: created by LIE::execute */
    pra_rel* backend() { return handle; }

    /* one relation file into a version (tsv_to_bin layout) */
    void load_file(pra_ctx* ctx, const std::string& path, int version)
    {
        int rc = pra_rel_load_bin(handle, path.c_str(), version == DELTA ? PRA_DELTA : PRA_FULL);
        if (rc != PRA_OK) {
            std::cout << path << " Wrong IO: " << pra_last_error(ctx) << std::endl;
            MPI_Abort(MPI_COMM_WORLD, -1);
        }
    }

    /* relation::initialize_relation (src/relation/balanced_hash_relation.cpp:658-793):
     * on a restart with per-rank files, rank r reads <file>_<r>; here the
     * buckets of all ranks are read */
    void initialize_relation_on(pra_ctx* ctx, int nbuckets)
    {
        if (handle) return;
        pra_facade::chk(ctx, pra_rel_create(ctx, arity, join_column_count, intern_tag, is_canonical, &handle),
                        "relation create");
        if (initailization_type != DELTA && initailization_type != FULL) return;
        if (restart_flag && separate_io) {
            for (int b = 0; b < nbuckets; b++) load_file(ctx, filename + "_" + std::to_string(b), initailization_type);
        } else {
            load_file(ctx, filename, initailization_type);
        }
    }

    /* the delta of a restart (src/lie/lie.cpp:304-333): <dir>/<name>_delta, or
     * <dir>/<name>_delta_<rank> per rank; missing files are skipped */
    void load_restart_delta(pra_ctx* ctx, const std::string& dir, bool separate, int nbuckets)
    {
        std::string base = dir + "/" + debug_id + "_delta";
        if (separate) {
            for (int b = 0; b < nbuckets; b++) {
                std::string p = base + "_" + std::to_string(b);
                if (pra_facade::file_exists(p + ".size")) load_file(ctx, p, DELTA);
            }
        } else if (pra_facade::file_exists(base + ".size")) {
            load_file(ctx, base, DELTA);
        }
    }

    /* relation::parallel_IO (src/relation/balanced_hash_relation.cpp:109-430):
     * <dir>/<name>_full and _delta, either one file of all ranks' rows in rank
     * order with "rows\narity+1" (written only when non-empty), or per rank
     * <name>_full_<r> (data when non-empty) with the size record the reference
     * writes for separate files, "bytes/24\n3" */
    void parallel_IO(const std::string& dir, bool separate, int nbuckets)
    {
        const int w = (int)arity + 1;
        for (int version : {FULL, DELTA}) {
            std::string base = dir + "/" + debug_id + (version == FULL ? "_full" : "_delta");
            std::vector<u64> all = rows(version);
            size_t n = all.size() / w;
            if (!separate) {
                if (n) pra_facade::write_rows_file(base, all.data(), all.size(), (long long)n, w, true);
                continue;
            }
            std::vector<std::vector<u64>> per(nbuckets);
            for (size_t i = 0; i < n; i++) {
                const u64* r = all.data() + i * w;
                u64 b = tuple_hash(r, join_column_count) % (u64)nbuckets;
                per[b].insert(per[b].end(), r, r + w);
            }
            for (int b = 0; b < nbuckets; b++) {
                size_t bytes = per[b].size() * 8;
                pra_facade::write_rows_file(base + "_" + std::to_string(b), per[b].data(), per[b].size(),
                                            (long long)(int)(bytes / 24), 3, bytes > 0);
            }
        }
    }
};

/* ----------------------------------------------------------- RA operators - */

class parallel_RA
{
protected:
    u32 RA_type = 0;
    mpi_comm mcomm;

public:
    parallel_RA() {}
    virtual ~parallel_RA() {}
    u32 get_RA_type() { return RA_type; }
    void set_comm(mpi_comm& m) { mcomm = m; }
    virtual void add_to(pra_scc* scc, pra_ctx* ctx) = 0;
};

class parallel_join : public parallel_RA
{
    relation* join_input0_table;
    int join_input0_graph_type;
    relation* join_input1_table;
    int join_input1_graph_type;
    relation* join_output_table;
    std::vector<int> projection_reorder_index_array;

public:
    parallel_join(relation* output, relation* G, int G_type, relation* T, int T_type, std::vector<int> proj)
        : join_input0_table(G), join_input0_graph_type(G_type), join_input1_table(T), join_input1_graph_type(T_type),
          join_output_table(output), projection_reorder_index_array(proj)
    {
        RA_type = JOIN;
    }
    relation* get_join_input0() { return join_input0_table; }
    int get_join_input0_graph_type() { return join_input0_graph_type; }
    relation* get_join_input1() { return join_input1_table; }
    int get_join_input1_graph_type() { return join_input1_graph_type; }
    relation* get_join_output() { return join_output_table; }
    void add_to(pra_scc* scc, pra_ctx* ctx) override
    {
        pra_facade::chk(ctx,
                        pra_scc_add_join(scc, join_output_table->backend(), join_input0_table->backend(),
                                         join_input0_graph_type, join_input1_table->backend(), join_input1_graph_type,
                                         projection_reorder_index_array.data(),
                                         (int)projection_reorder_index_array.size()),
                        "parallel_join");
    }
};

class parallel_copy : public parallel_RA
{
    relation* copy_input0_table;
    int copy_input0_graph_type;
    relation* copy_output_table;
    std::vector<int> copy_reorder_index_array;

public:
    parallel_copy(relation* dest, relation* src, int src_version, std::vector<int> reorder)
        : copy_input0_table(src), copy_input0_graph_type(src_version), copy_output_table(dest),
          copy_reorder_index_array(reorder)
    {
        RA_type = COPY;
    }
    relation* get_copy_input() { return copy_input0_table; }
    relation* get_copy_output() { return copy_output_table; }
    void add_to(pra_scc* scc, pra_ctx* ctx) override
    {
        pra_facade::chk(ctx,
                        pra_scc_add_copy(scc, copy_output_table->backend(), copy_input0_table->backend(),
                                         copy_input0_graph_type, copy_reorder_index_array.data(),
                                         (int)copy_reorder_index_array.size()),
                        "parallel_copy");
    }
};

class parallel_acopy : public parallel_RA
{
    relation* acopy_input0_table;
    int acopy_input0_graph_type;
    relation* acopy_output_table;
    std::vector<int> acopy_reorder_index_array;

public:
    parallel_acopy(relation* dest, relation* src, int src_version, std::vector<int> reorder)
        : acopy_input0_table(src), acopy_input0_graph_type(src_version), acopy_output_table(dest),
          acopy_reorder_index_array(reorder)
    {
        RA_type = ACOPY;
    }
    relation* get_acopy_input() { return acopy_input0_table; }
    relation* get_acopy_output() { return acopy_output_table; }
    void add_to(pra_scc* scc, pra_ctx* ctx) override
    {
        pra_facade::chk(ctx,
                        pra_scc_add_acopy(scc, acopy_output_table->backend(), acopy_input0_table->backend(),
                                          acopy_input0_graph_type, acopy_reorder_index_array.data(),
                                          (int)acopy_reorder_index_array.size()),
                        "parallel_acopy");
    }
};

class parallel_copy_filter : public parallel_RA
{
    relation* copy_filter_input0_table;
    int copy_filter_input0_graph_type;
    relation* copy_filter_output_table;
    std::vector<int> copy_filter_reorder_index_array;
    bool (*lambda)(const u64* const);

public:
    parallel_copy_filter(relation* dest, relation* src, int src_version, std::vector<int> reorder,
                         bool (*func)(const u64* const))
        : copy_filter_input0_table(src), copy_filter_input0_graph_type(src_version), copy_filter_output_table(dest),
          copy_filter_reorder_index_array(reorder), lambda(func)
    {
        RA_type = COPY_FILTER;
    }
    relation* get_copy_filter_input() { return copy_filter_input0_table; }
    relation* get_copy_filter_output() { return copy_filter_output_table; }
    void add_to(pra_scc* scc, pra_ctx* ctx) override
    {
        int op = pra_facade::classify_filter(lambda);
        if (op < 0) {
            fprintf(stderr, "parallel_copy_filter: predicate is not a comparison of columns 0 and 1\n");
            exit(1);
        }
        pra_facade::chk(ctx,
                        pra_scc_add_copy_filter(scc, copy_filter_output_table->backend(),
                                                copy_filter_input0_table->backend(), copy_filter_input0_graph_type,
                                                copy_filter_reorder_index_array.data(),
                                                (int)copy_filter_reorder_index_array.size(), op),
                        "parallel_copy_filter");
    }
};

class parallel_copy_generate : public parallel_RA
{
public:
    parallel_copy_generate(relation*, relation*, int, int (*)(const u64* const, u64* const))
    {
        RA_type = COPY_GENERATE;
    }
    void add_to(pra_scc*, pra_ctx*) override
    {
        /* the reference's RAM cannot route it either (src/RAM/RA_tasks.cpp:844-845) */
        fprintf(stderr, "parallel_copy_generate is not supported by this backend\n");
        exit(1);
    }
};

/* ------------------------------------------------------------------ RAM --- */

class RAM
{
    int ram_id;
    bool fixpoint;
    u32 ram_relation_count = 0;
    relation* ram_relations[64];
    bool ram_relation_status[64];
    std::vector<parallel_RA*> RA_list;
    bool logging = false;
    mpi_comm mcomm;

public:
    RAM(bool ic, int r_id) : ram_id(r_id), fixpoint(ic) {}
    ~RAM()
    {
        for (auto* r : RA_list) delete r;
    }
    void add_relation(relation*& G, bool i_status)
    {
        ram_relations[ram_relation_count] = G;
        ram_relation_status[ram_relation_count] = i_status;
        ram_relation_count++;
    }
    void add_rule(parallel_RA* pj) { RA_list.push_back(pj); }
    void set_comm(mpi_comm& m) { mcomm = m; }
    void enable_logging() { logging = true; }
    bool logging_enabled() const { return logging; }
    /* RAM::print_all_relation (src/RAM/RA_tasks.cpp:56-60) */
    void print_all_relation(int nbuckets)
    {
        for (u32 i = 0; i < ram_relation_count; i++) ram_relations[i]->print(nbuckets);
    }
    int get_id() { return ram_id; }
    int get_iteration_count() { return fixpoint ? -1 : 1; }
    u32 get_ram_relation_count() { return ram_relation_count; }
    relation** get_RAM_relations() { return ram_relations; }
    bool* get_RAM_relations_status() { return ram_relation_status; }
    u32 get_bucket_count() { return 1; }

    pra_scc* build_on(pra_ctx* ctx)
    {
        pra_scc* s = nullptr;
        pra_facade::chk(ctx, pra_scc_create(ctx, fixpoint ? 1 : 0, ram_id, &s), "RAM");
        for (u32 i = 0; i < ram_relation_count; i++)
            pra_facade::chk(ctx, pra_scc_add_relation(s, ram_relations[i]->backend(), ram_relation_status[i]),
                            "RAM::add_relation");
        for (auto* r : RA_list) r->add_to(s, ctx);
        return s;
    }
};

/* ------------------------------------------------------------------ LIE --- */

class LIE
{
    u32 batch_size = 10;
    double batch_time = 0;
    mpi_comm mcomm;
    std::string app_name;
    u32 lie_relation_count = 0;
    std::vector<relation*> lie_relations;
    u32 lie_sccs_count = 0;
    std::vector<RAM*> lie_sccs;
    std::map<RAM*, std::set<RAM*>> taskgraph;
    std::vector<int> executed_scc_id;
    int loop_counter = 0;
    std::string output_dir, restart_dir_name;
    int cp_iteration = 0;
    bool enable_io = false, separate_io = false, share_io = false, offset_io = false, restart_flag = false;
    bool all_to_all_meta_data_dump = false;
    int nbuckets = 1;
    pra_ctx* ctx = nullptr;

    /* first SCC in add order that no remaining SCC points to (src/lie/lie.cpp:27-60) */
    RAM* one_runnable_tasks()
    {
        for (u32 i = 0; i < lie_sccs_count; i++) {
            if (!lie_sccs[i]) continue;
            bool blocked = false;
            for (auto& e : taskgraph)
                if (e.second.count(lie_sccs[i])) blocked = true;
            if (!blocked) return lie_sccs[i];
        }
        return nullptr;
    }
    void update_task_graph(RAM* t)
    {
        for (u32 i = 0; i < lie_sccs_count; i++)
            if (lie_sccs[i] == t) {
                taskgraph.erase(t);
                delete t;
                lie_sccs[i] = nullptr;
            }
    }

public:
    LIE() {}
    ~LIE()
    {
        for (u32 i = 0; i < lie_relation_count; i++) delete lie_relations[i];
        for (u32 i = 0; i < lie_sccs_count; i++) delete lie_sccs[i];
        if (ctx) pra_ctx_destroy(ctx);
    }

    void add_relation(relation* rel)
    {
        lie_relations.push_back(rel);
        lie_relation_count++;
    }
    void add_scc(RAM* ra)
    {
        lie_sccs.push_back(ra);
        lie_sccs_count++;
    }
    void add_scc_dependance(RAM* src_task, RAM* destination_task) { taskgraph[src_task].insert(destination_task); }
    void set_comm(mpi_comm comm) { mcomm = comm; }
    void set_name(std::string name) { app_name = name; }
    void set_batch_size(u32 bs) { batch_size = bs; }
    u32 get_batch_size() { return batch_size; }
    void set_batch_time(double bt) { batch_time = bt; }
    double get_batch_time() { return batch_time; }

    /* checkpoint / restart controls (src/lie/lie.h:65-110) */
    void set_cp_iteration(int it) { cp_iteration = it; }
    void set_output_dir(std::string o) { output_dir = o; }
    void set_executed_scc_id(std::vector<int> id) { executed_scc_id = id; }
    std::vector<int> get_executed_scc_id() { return executed_scc_id; }
    void set_loop_counter(int c) { loop_counter = c; }
    int get_loop_counter() { return loop_counter; }
    void enable_separate_io() { separate_io = true; }
    /* the reference's per-iteration send-count record (a2a_<rank>) is not
     * written: this backend has no sender-side buffers to count */
    void enable_all_to_all_dump() { all_to_all_meta_data_dump = true; }
    void enable_offset_io() { offset_io = true; }
    void enable_share_io() { share_io = true; }
    void set_restart_dir_name(std::string path) { restart_dir_name = path; }
    void set_restart_flag(bool flag) { restart_flag = flag; }
    void enable_IO() { enable_io = true; }

    /* LIE::create_checkpoint_dump (src/lie/lie.cpp:182-193) */
    void create_checkpoint_dump(int loop, int scc_id)
    {
        std::string dir = output_dir + "/checkpoint-" + std::to_string(scc_id) + "-" + std::to_string(loop);
        mkdir(dir.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    }

    /* LIE::write_checkpoint_dump (src/lie/lie.cpp:154-178): scc_metadata (the
     * executed SCC ids, one per line) and every relation's full and delta */
    void write_checkpoint_dump(int loop, std::vector<int> executed, int scc_id = 0)
    {
        std::string dir = output_dir + "/checkpoint-" + std::to_string(scc_id) + "-" + std::to_string(loop);
        FILE* fp = fopen((dir + "/scc_metadata").c_str(), "w");
        if (!fp) {
            std::cout << dir << "/scc_metadata Wrong IO" << std::endl;
            MPI_Abort(MPI_COMM_WORLD, -1);
        }
        for (int id : executed) fprintf(fp, "%d\n", id);
        fclose(fp);
        for (u32 i = 0; i < lie_relation_count; i++) lie_relations[i]->parallel_IO(dir, separate_io, nbuckets);
    }

    pra_ctx* backend() { return ctx; }

    /* cumulative device milliseconds per kernel class (zero unless the
     * backend times its kernels, PRA_PHASE_TIMES=1) */
    std::vector<double> kernel_ms()
    {
        std::vector<double> v;
        for (int i = 0;; i++) {
            double ms = 0;
            if (pra_ctx_kernel_stat(ctx, i, nullptr, nullptr, &ms, nullptr) != PRA_OK) break;
            v.push_back(ms);
        }
        return v;
    }

    /* one row of the reference's per-iteration timing table (rows
     * src/RAM/RA_tasks.cpp:1198-1235, header src/lie/lie.cpp:387-388), per
     * batch, in seconds: Total_time is the batch's wall time; with
     * PRA_PHASE_TIMES=1 the device time of the kernels doing each part --
     * computation (join counts and expansions, copies), all-to-all (the
     * stable bucket partition that replaces the exchange), insert in newt
     * (membership claims and decisions, tables, scans, sorts) and insert in
     * full (merges) -- fills the other columns; buffer creation and freeing
     * and the intra-bucket exchange have no per-iteration counterpart here */
    void timing_row(int loop, const std::vector<double>& k0, double wall)
    {
        const std::vector<double> k1 = kernel_ms();
        double comp = 0, a2a = 0, newt = 0, full = 0;
        for (size_t i = 0; i < k1.size() && i < k0.size(); i++) {
            const char* name = nullptr;
            pra_ctx_kernel_stat(ctx, (int)i, &name, nullptr, nullptr, nullptr);
            const std::string n = name ? name : "";
            const double d = (k1[i] - k0[i]) * 1e-3;
            if (n == "join_count" || n == "join_expand" || n == "copy_expand" || n == "index_build")
                comp += d;
            else if (n == "dedup_scatter")
                a2a += d;
            else if (n == "merge_full")
                full += d;
            else
                newt += d;
        }
        std::cout << loop << "\t" << 0.0 << "\t" << comp << "\t" << a2a << "\t" << 0.0 << "\t" << newt << "\t" << 0.0
                  << "\t" << full << "\t" << wall << std::endl;
    }

    /* LIE::execute (src/lie/lie.cpp:197-448), one SCC at a time in steps of
     * one batch so that checkpoints are written where the reference writes
     * them (every cp_iteration loops of an SCC). */
    bool execute()
    {
        nbuckets = pra_facade::env_int("PRA_BUCKETS", 1);
        int dev = pra_facade::env_int("PRA_DEVICE", 0);
        int rc = pra_ctx_create(dev, nbuckets, 0, &ctx);
        if (rc != PRA_OK) pra_facade::die(nullptr, rc, "GPU context");
        pra_facade::chk(ctx, pra_ctx_set_batch_size(ctx, batch_size), "set_batch_size");
        if (pra_facade::env_int("PRA_PHASE_TIMES", 0)) pra_facade::chk(ctx, pra_ctx_set_profiling(ctx, -1), "profiling");
        for (u32 i = 0; i < lie_relation_count; i++) {
            relation* r = lie_relations[i];
            r->set_restart_flag(restart_flag);
            r->set_share_io(share_io);
            r->set_separate_io(separate_io);
            r->set_offset_io(offset_io);
            r->initialize_relation_on(ctx, nbuckets);
        }
        if (enable_io) mkdir(output_dir.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
        std::cout << "----------------- Initialization Complete ---------------------" << std::endl << std::endl;

        if (restart_flag) {
            std::ifstream meta(restart_dir_name + "/scc_metadata");
            int a;
            while (meta >> a) executed_scc_id.push_back(a);
        }
        const bool io = enable_io && cp_iteration > 0;
        RAM* task;
        while ((task = one_runnable_tasks()) != nullptr) {
            const int id = task->get_id();
            if (std::find(executed_scc_id.begin(), executed_scc_id.end(), id) != executed_scc_id.end()) {
                update_task_graph(task);
                continue;
            }
            std::cout << "-------------------Executing SCC " << id << "------------------" << std::endl;
            pra_scc* s = task->build_on(ctx);
            if (restart_flag) {
                relation** rels = task->get_RAM_relations();
                for (u32 i = 0; i < task->get_ram_relation_count(); i++)
                    rels[i]->load_restart_delta(ctx, restart_dir_name, separate_io, nbuckets);
            }
            pra_facade::chk(ctx, pra_scc_begin(s, restart_flag ? 0 : 1), "SCC start");
            if (task->get_iteration_count() != 1)
                std::cout << "Iteration#\tBuffer_creation_time\tComputation_time\tAll_to_all_time\tBuffer_free_time"
                             "\tInsert_in_newt_time\tIntra_comm_time\tInsert_in_full_time\tTotal_time"
                          << std::endl;
            int loop = 0;
            uint64_t delta = 0;
            do {
                if (io && loop % cp_iteration == 0) create_checkpoint_dump(loop, id);
                const std::vector<double> k0 = kernel_ms();
                const double t0 = MPI_Wtime();
                pra_facade::chk(ctx, pra_scc_step(s, &delta), "SCC execute");
                const double t1 = MPI_Wtime();
                timing_row(loop, k0, t1 - t0);
                if (task->logging_enabled()) task->print_all_relation(nbuckets);
                if (task->get_iteration_count() == 1 || delta == 0) executed_scc_id.push_back(id);
                if (io && loop % cp_iteration == 0) write_checkpoint_dump(loop, executed_scc_id, id);
                loop++;
            } while (task->get_iteration_count() != 1 && delta != 0);
            pra_facade::chk(ctx, pra_scc_end(s), "SCC finish");
            if (task->get_iteration_count() == 1) print_all_relation_size();
            loop_counter = loop;
            pra_scc_destroy(s);
            update_task_graph(task);
        }
        /* PRA_DUMP_DIR: write every relation's full version as
         * <dir>/<name>_full (+ .size), the shared-file layout of
         * relation::parallel_IO (src/relation/balanced_hash_relation.cpp:109-235) */
        const char* dump = getenv("PRA_DUMP_DIR");
        if (dump && *dump)
            for (u32 i = 0; i < lie_relation_count; i++) {
                relation* r = lie_relations[i];
                std::string name = r->get_debug_id().empty() ? "relation_" + std::to_string(i) : r->get_debug_id();
                std::string path = std::string(dump) + "/" + name + "_full";
                pra_facade::chk(ctx, pra_rel_write_bin(r->backend(), path.c_str(), PRA_FULL), "relation dump");
            }
        return true;
    }

    /* LIE::print_all_relation_size (src/lie/lie.cpp:128-150) */
    void print_all_relation_size()
    {
        u64 total = 0;
        for (u32 i = 0; i < lie_relation_count; i++) {
            relation* r = lie_relations[i];
            u64 n = (u64)(u32)r->get_full_element_count();
            std::cout << r->get_debug_id() << ": {" << r->get_arity() << "}. (" << n << " total facts)" << std::endl;
            total += n;
        }
        std::cout << "Total facts across all relations " << total << std::endl << std::endl;
    }
};
