// Host engine of the MI355X semi-naive RA backend: device-resident relation
// versions, the per-iteration pipeline (join / copy emit -> arrival-ordered
// candidates -> dedup + intern -> stable bucket partition -> LSD sort) and the
// LIE-style SCC scheduler, exported through the C ABI of include/pra.h.
//
// Layout in HBM: every version (full, delta, newt) of a relation is one array
// of rows of w = arity+1 u64 words.  delta and newt are kept sorted by
// (bucket, lexicographic row); full is an append log that is re-sorted only
// when an ordered walk of it is needed (copy from FULL, join probe side).
// Membership of full u delta is an open-addressing table keyed by the dedup
// columns; join probe sides get a (key -> first row, row count) index.
#include "../../include/pra.h"
#include "pra_blockcache.hpp"
#include "pra_tune.hpp"
#include "pra_kernels.hpp"
#include "pra_partition.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

namespace {

enum { RA_COPY = 0, RA_COPY_FILTER = 1, RA_ACOPY = 3, RA_JOIN = 4 };

struct PraError : std::runtime_error {
    int code;
    PraError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            throw PraError(PRA_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));            \
    } while (0)

static void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw PraError(PRA_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Large blocks (>= 64 MiB) go through the process-wide block cache of
// pra_blockcache.hpp (its ownership rule: a block released while its stream
// runs is reused only on that stream until a host synchronization of the
// stream declares it idle).  Out of memory: idle blocks are freed, oldest
// first, until the request fits; then the requesting stream is synchronized,
// its own blocks declared idle and the stream-ordered pool trimmed before one
// more attempt.
namespace {
constexpr size_t kBigBlock = 64ULL << 20;
pra::BlockCache g_big(kBigBlock);

void* hip_block_alloc(size_t bytes, void*)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}
void hip_block_free(void* p, void*) { (void)hipFree(p); }
const pra::BlockAllocator kHipBlocks{hip_block_alloc, hip_block_free, nullptr};

// Device bytes held by each context, keyed by its stream (every buffer of a
// context is ordered on the context's stream): live bytes, their peak, and the
// context's cap (0 = none).
struct MemAcct {
    long long live = 0, peak = 0;
    u64 budget = 0;
};
std::mutex g_acct_mu;
std::map<hipStream_t, MemAcct> g_acct;

// charge `bytes` more to the context of stream s (fails past its budget while
// `check`), or credit them back when negative
void acct_charge(hipStream_t s, long long bytes, bool check)
{
    std::lock_guard<std::mutex> lk(g_acct_mu);
    auto it = g_acct.find(s);
    if (it == g_acct.end()) return;
    MemAcct& a = it->second;
    if (check && bytes > 0 && a.budget && (u64)(a.live + bytes) > a.budget)
        throw PraError(PRA_E_BUDGET, "device memory budget exceeded: " + std::to_string(a.live) + " bytes held + " +
                                         std::to_string(bytes) + " requested > budget " + std::to_string(a.budget));
    a.live += bytes;
    if (a.live > a.peak) a.peak = a.live;
}

// the stream-ordered pool keeps freed memory for reuse; give it back
void trim_pool(int dev)
{
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) (void)hipMemPoolTrimTo(pool, 0);
}

void* big_get(size_t bytes, hipStream_t st, size_t* got)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    void* p = g_big.get(dev, st, bytes, got, kHipBlocks);
    if (!p) {
        HIPCHK(hipStreamSynchronize(st));
        g_big.idle(st);
        trim_pool(dev);
        p = g_big.get(dev, st, bytes, got, kHipBlocks);
    }
    if (!p)
        throw PraError(PRA_E_NOMEM, "hipMalloc: out of memory (" + std::to_string(g_big.size_class(bytes)) + " bytes)");
    return p;
}

void big_put(void* p, size_t size, hipStream_t st)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    g_big.put(dev, st, p, size);
}

// every piece of work enqueued on st so far has completed
void stream_idle(hipStream_t st) { g_big.idle(st); }

// Small blocks (below 64 MiB) released by a buffer are kept for the stream
// they were released on and handed to later requests of the same
// quarter-octave class on that stream: work enqueued there after the release
// runs after every earlier use of the block, so stream order alone makes the
// reuse safe, and no hipFreeAsync / hipMallocAsync pair is paid per buffer
// (about 0.1 ms of host time each).  At most kSmallKeep bytes are kept per
// stream; the rest, and a stream's blocks when its context goes away or the
// device runs out of memory, are freed on their stream.
constexpr size_t kSmallKeep = 2ULL << 30;
struct SmallFree {
    std::map<size_t, std::vector<void*>> by_class;
    size_t bytes = 0;
};
std::mutex g_small_mu;
std::map<hipStream_t, SmallFree> g_small;

size_t small_class(size_t bytes)
{
    size_t oct = 256;
    while (oct * 2 <= bytes) oct <<= 1;
    for (size_t q = 4; q <= 8; q++)
        if (oct / 4 * q >= bytes) return oct / 4 * q;
    return oct * 2;
}

void* small_take(hipStream_t st, size_t cls)
{
    std::lock_guard<std::mutex> lk(g_small_mu);
    auto it = g_small.find(st);
    if (it == g_small.end()) return nullptr;
    auto c = it->second.by_class.find(cls);
    if (c == it->second.by_class.end() || c->second.empty()) return nullptr;
    void* p = c->second.back();
    c->second.pop_back();
    it->second.bytes -= cls;
    return p;
}

void small_put(hipStream_t st, void* p, size_t cls)
{
    {
        std::lock_guard<std::mutex> lk(g_small_mu);
        SmallFree& f = g_small[st];
        if (f.bytes + cls <= kSmallKeep) {
            f.by_class[cls].push_back(p);
            f.bytes += cls;
            return;
        }
    }
    (void)hipFreeAsync(p, st);
}

// free the small blocks kept for a stream, on that stream
void small_flush(hipStream_t st)
{
    SmallFree f;
    {
        std::lock_guard<std::mutex> lk(g_small_mu);
        auto it = g_small.find(st);
        if (it == g_small.end()) return;
        f = std::move(it->second);
        g_small.erase(it);
    }
    for (auto& c : f.by_class)
        for (void* p : c.second) (void)hipFreeAsync(p, st);
}

// Autotuned knobs, per device and band of log2(work): the dense claim
// table's size (2, 4 or 8 slots per candidate of capacity, so a load of at
// most 1/2, 1/4 or 1/8) and the newt scatter's block width (256 or 1,024
// threads).  Neither changes a result: which arrival wins a key, and where
// every row and id goes, are the same under every variant.  Variants are
// timed with events on single-output iterations, per candidate.
enum { TUNE_DENSE_LOAD, TUNE_SCATTER_WIDE, TUNE_KNOBS };
const char* const kTuneNames[TUNE_KNOBS] = {"dense_table_slots_log2_minus_1", "newt_scatter_wide"};
const int kTuneVariants[TUNE_KNOBS] = {3, 2};
constexpr u64 kTuneMinCandidates = 1ULL << 16;
pra::Tuner g_tune(3);

// Tuning cache: with PRA_AUTOTUNE_FILE set, the settled choices are read from
// that file when the first context is created (those cells run without
// trials) and the file is rewritten whenever a cell settles.
std::once_flag g_tune_once;
std::mutex g_tune_file_mu;
std::string g_tune_file;

void tune_init()
{
    std::call_once(g_tune_once, [] {
        const char* f = getenv("PRA_AUTOTUNE_FILE");
        if (!f || !*f) return;
        g_tune_file = f;
        FILE* in = fopen(f, "r");
        if (!in) return;
        std::string text;
        char buf[4096];
        size_t got;
        while ((got = fread(buf, 1, sizeof(buf), in)) > 0) text.append(buf, got);
        fclose(in);
        g_tune.load(text, kTuneNames, TUNE_KNOBS);
    });
}

void tune_save()
{
    if (g_tune_file.empty()) return;
    const std::string r = g_tune.report(kTuneNames, TUNE_KNOBS);
    std::lock_guard<std::mutex> lk(g_tune_file_mu);
    FILE* out = fopen(g_tune_file.c_str(), "w");
    if (!out) return;
    fwrite(r.data(), 1, r.size(), out);
    fclose(out);
}

int band_of(u64 n)
{
    int b = 0;
    while (n > 1) {
        n >>= 1;
        b++;
    }
    return b;
}
} // namespace

// Device buffer owned by one context and ordered on its stream: small blocks
// come from the stream-ordered pool, large ones from the block cache above;
// growth and release never synchronize the device, so contexts driven from
// different host threads overlap.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipStream_t st = nullptr;
    bool big = false;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), cap(o.cap), st(o.st), big(o.big) { o.p = nullptr; o.cap = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            cap = o.cap;
            st = o.st;
            big = o.big;
            o.p = nullptr;
            o.cap = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void release()
    {
        if (p) {
            if (big)
                big_put(p, cap, st);
            else
                small_put(st, p, cap);
            acct_charge(st, -(long long)cap, false);
        }
        p = nullptr;
        cap = 0;
        big = false;
    }
    // grow to at least `bytes`, preserving the first `used` bytes
    void reserve(size_t bytes, hipStream_t s, size_t used = 0)
    {
        if (bytes <= cap) return;
        // grow by doubling (fewer growth copies than growing by half, which
        // measured slower); the quarter-octave block classes keep the
        // rounding on top of that within 25 %
        size_t nc = std::max(bytes, 2 * cap);
        void* np = nullptr;
        bool nbig = nc >= kBigBlock;
        nc = nbig ? g_big.size_class(nc) : small_class(nc);
        acct_charge(s, (long long)nc, true);
        try {
            if (nbig) {
                const size_t want = nc;
                np = big_get(want, s, &nc);
                acct_charge(s, (long long)nc - (long long)want, false);   // a cached block of the next class
            } else if (!(np = small_take(s, nc)) && hipMallocAsync(&np, nc, s) != hipSuccess) {
                int dev = 0;
                (void)hipGetLastError();
                HIPCHK(hipGetDevice(&dev));
                small_flush(s);
                g_big.flush(dev, kHipBlocks, nc);
                if (hipMallocAsync(&np, nc, s) != hipSuccess) {
                    (void)hipGetLastError();
                    HIPCHK(hipStreamSynchronize(s));
                    g_big.idle(s);
                    g_big.flush(dev, kHipBlocks);
                    trim_pool(dev);
                    const hipError_t e = hipMallocAsync(&np, nc, s);
                    if (e == hipErrorOutOfMemory) {
                        (void)hipGetLastError();
                        throw PraError(PRA_E_NOMEM, "hipMallocAsync: out of memory (" + std::to_string(nc) + " bytes)");
                    }
                    HIPCHK(e);
                }
            }
        } catch (...) {
            acct_charge(s, -(long long)nc, false);
            throw;
        }
        if (p && used) HIPCHK(hipMemcpyAsync(np, p, used, hipMemcpyDeviceToDevice, s));
        if (p) {
            if (big)
                big_put(p, cap, s);
            else
                small_put(s, p, cap);
            acct_charge(st ? st : s, -(long long)cap, false);
        }
        p = np;
        cap = nc;
        st = s;
        big = nbig;
    }
    template <typename T> T* as() { return static_cast<T*>(p); }
    void swap(DBuf& o)
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(st, o.st);
        std::swap(big, o.big);
    }
};

struct Rows {
    DBuf buf;
    u64 n = 0;
    bool sorted = true;
    u64* d() { return buf.as<u64>(); }
};

static u64 pow2_at_least(u64 x)
{
    u64 c = 1;
    while (c < x) c <<= 1;
    return c;
}

} // namespace

struct pra_scc;

struct pra_rel {
    pra_ctx* ctx = nullptr;
    int group = 0;
    int arity = 0, jcc = 0, canonical = 1, w = 0;
    u32 tag = 0;
    Rows ver[3];  // PRA_DELTA, PRA_FULL, PRA_NEWT
    // membership table of full u delta keyed on mt_kw leading columns
    DBuf mt;
    u64 mt_cap = 0, mt_entries = 0;
    int mt_kw = -1;
    bool mt_valid = false;
    // dense membership (interned outputs with dictionary-ranked keys): bitmap
    // over the key's rank index, built for dictionary generation bm_gen and
    // kept current by the iterations that take the dense path; the
    // per-iteration claim table (empty between iterations)
    DBuf bm, dtab;
    u64 bm_gen = ~0ULL, dtab_cap = 0, dtab_use = 0;   // slots allocated, slots in use this iteration
    int dtab_eb = -1;      // epoch bits of the claim table's slots (0: winners empty their slots)
    u32 dtab_epoch = 0;    // this iteration's epoch (slots of any other count as empty)
    int bm_kw = 0;
    bool bm_valid = false;
    // join index (key -> first row, count) per version
    DBuf idx[2];
    u64 idx_cap[2] = {0, 0};
    u64 idx_maxlen[2] = {0, 0};   // longest run of rows sharing a join key
    u64 jhint = 0;                // candidates of the previous iteration producing this relation
    bool idx_valid[2] = {false, false};
    // packed sort keys of newt rows, the offsets of their top-byte bins, and
    // the bins of (key, newt position) pairs the scatter fills
    DBuf nkeys, kgofs, bin_keys, bin_idx, bin_cur;
    u32 bin_cap = 0;
    // bucket bounds per version
    DBuf bnd[2];
    bool bnd_valid[2] = {false, false};
    // set while an SCC runs: the relation is produced by one of its rules
    // (changes every iteration), and its full version is read by one of them
    // (so it is kept sorted by merging each delta into it)
    bool scc_changing = false, merge_full = false;
    // set while an SCC runs: a rule of it reads the relation's full version
    bool full_read = false;
    // changes of the delta and full versions so far
    u64 gen[2] = {0, 0};
    // a view: the rows of relation view_of that a join keyed on this
    // relation's jcc columns can match (built from view_of's versions of
    // generation view_gen)
    pra_rel* view_of = nullptr;
    u64 view_gen[2] = {~0ULL, ~0ULL};
    void touched(int v)
    {
        if (v == PRA_DELTA || v == PRA_FULL) {
            idx_valid[v] = false;
            bnd_valid[v] = false;
            gen[v]++;
        }
    }
};

namespace {

struct RADesc {
    int type = 0;
    pra_rel *out = nullptr, *in0 = nullptr, *in1 = nullptr;
    int v0 = 0, v1 = 0, op = PRA_F_TRUE;
    std::vector<int> map;
    // per-iteration state
    DBuf off, istart, posadj;
    u64 n_outer = 0;
    bool active = false;   // outer (and inner) non-empty this iteration
    bool sorted_probe = false;   // inner changes inside the SCC: binary search, no hash index
    pra_rel* outer = nullptr;
    int outer_v = 0;
    pra_rel* inner = nullptr;
    int inner_v = 0;
    int left = 1;
};

struct KStat {
    const char* name;
    uint64_t launches = 0;
    double ms = 0, bytes = 0;
};

enum {
    K_JOIN_COUNT, K_JOIN_EXPAND, K_COPY_EXPAND, K_DEDUP_CLAIM, K_DEDUP_DECIDE, K_DEDUP_VERIFY, K_DEDUP_SCATTER,
    K_RADIX_HIST, K_RADIX_SCATTER, K_TABLE_BUILD, K_INDEX_BUILD, K_SCAN, K_GATHER, K_MERGE, K_BIN_SORT, K_SMALL_SORT,
    K_BLOCK_SORT, K_COUNT
};
static const char* kNames[K_COUNT] = {"join_count", "join_expand", "copy_expand", "dedup_claim", "dedup_decide",
                                      "dedup_verify_hist", "dedup_scatter", "radix_hist", "radix_scatter",
                                      "table_build", "index_build", "scan", "gather_rows", "merge_full",
                                      "bin_sort", "small_sort", "block_sort"};

} // namespace

// One program instance (group) of a context.  A context holds one or more
// independent programs whose SCCs run in lock step (pra_ctx_set_group); each
// has its own intern counters (the reference keeps them per LIE,
// src/lie/lie.h:36), value dictionary, join-pair count and iteration count.
struct Group {
    std::map<u32, std::unique_ptr<DBuf>> seq;  // per intern tag: nb counters
    // order-preserving value dictionary for packed sort keys
    DBuf dict;
    u64 dict_cap = 0, dict_n = 0;
    int dict_state = 0, dict_rebuilds = 0;   // state: 0 stale, 1 valid, 2 disabled
    u64 dict_gen = 0;                        // dictionary builds so far
    DBuf joined;                             // join pairs so far (one device word)
    u64 iterations = 0;
    // lock-step execution: the instance's SCC in progress and the iterations
    // of its current batch
    pra_scc* running = nullptr;
    u32 batch_it = 0;
};

struct pra_scc {
    pra_ctx* ctx = nullptr;
    int group = 0;
    int fixpoint = 1, id = 0;
    std::vector<pra_rel*> rels;
    std::vector<int> status;
    std::vector<RADesc> ras;
    bool done = false, running = false;
    std::vector<u64> hist;
};

// Job-table staging slots.  A slot's pinned buffer is rewritten only after the
// host has waited for a kernel enqueued behind the slot's previous copy: every
// table but the early appends' is staged before an iteration's summary
// publication, or after its wait and before the next one's; the early appends
// are staged between a publication and its wait, so they alternate between
// two slots, and the wait of the iteration in between covers each copy.
enum { JOB_EXPAND, JOB_SCATTER, JOB_DENSE, JOB_SCAN, JOB_FOLD, JOB_SORT, JOB_EARLY_APPEND, JOB_EARLY_APPEND2,
       JOB_APPEND, JOB_PLAN, PRA_JOB_SLOTS };

struct pra_ctx {
    int device = 0, nb = 1;
    hipStream_t st = nullptr;
    u64 budget = 0;
    u32 batch_size = 10;  // LIE::batch_size default (src/lie/lie.h:17)
    long long threshold = 20000000;
    std::vector<pra_rel*> rels;
    std::vector<pra_scc*> sccs;
    std::vector<std::pair<pra_scc*, pra_scc*>> deps;
    std::vector<std::unique_ptr<Group>> groups;   // program instances; relations and SCCs name theirs
    int cur_group = 0;                            // instance of the next pra_rel_create / pra_scc_create
    // scratch
    DBuf hist, offs, sort_tmp, flags_buf, stage, stage2, orand, tmp64a, tmp64b, tmp64c;
    DBuf cand, dest, slot_of, keep;
    DBuf ghist, gofs, lb_status, lb_ctr;
    DBuf part_scratch;   // pra_partition_rows
    // order-preserving value dictionary for packed sort keys
    DBuf ktmp, idxa, idxb, posadj_dev, rep_words, rctl, ttab, kp, jobtab[PRA_JOB_SLOTS];
    int rep_sets = 0;              // summary replica sets in rep_words (one per output of a batched partition)
    // pinned staging of batched kernels' job tables, one slot per table of an
    // iteration (expansions, partitions, dense claims, histogram scans, folds,
    // bin sorts, early appends, appends)
    char* job_stage[PRA_JOB_SLOTS] = {};
    size_t job_stage_cap[PRA_JOB_SLOTS] = {};
    int early_parity = 0;          // early appends' slot of the last iteration
    bool dense_on = true;        // dense membership for outputs that allow it
    bool tune_on = true;         // autotuned knobs (PRA_AUTOTUNE=0: the fixed defaults)
    bool seg_sort_on = true;     // segmented sort of newts too large for the bins (PRA_SEG_SORT=0: byte passes)
    u64 bin_cap_test = 0;        // PRA_SORT_TEST_BIN_CAP: a lower bin capacity, so tests reach the large-newt paths
    int dense_epoch_bits = 16;   // PRA_DENSE_EPOCH_BITS: at most this many epoch bits in claim slots (0: none)
    DBuf segb;                   // segment bounds of the segmented sort, then its largest segment
    // timed launches of an iteration, resolved per candidate after its readback
    struct TuneSample {
        int knob, band, var;
        hipEvent_t a, b;
        size_t oi;
        u64 cap;
    };
    std::vector<TuneSample> tune_pending;
    size_t lb_zeroed = 0;        // words of lb_status known to be zero-initialised
    u32 sort_epoch = 0;
    u64 dedup_epoch = 0;
    u64* host_stage = nullptr;  // pinned
    size_t host_stage_cap = 0;
    u64* mailbox = nullptr;     // pinned, written by the publish kernel
    size_t mailbox_cap = 0;
    u64 pub_seq = 0;            // sequence number of the last publication
    hipEvent_t pub_ev = nullptr;
    std::string err;
    // profiling
    u32 prof_mask = 0;           // bit k: time kernel class k
    KStat kst[K_COUNT];
    struct Pending { int k; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
    // (index into pending, bytes per candidate, output index) to fix up once counts are read back
    std::vector<std::tuple<size_t, double, size_t>> per_candidate;
    std::vector<hipEvent_t> evpool;

    pra_ctx()
    {
        for (int i = 0; i < K_COUNT; i++) kst[i].name = kNames[i];
    }
    ~pra_ctx()
    {
        resolve_prof();
        // stream-ordered buffers are returned before their stream goes away
        if (st) (void)hipStreamSynchronize(st);
        groups.clear();
        for (DBuf* b : {&hist, &offs, &sort_tmp, &flags_buf, &stage, &stage2, &orand, &tmp64a, &tmp64b,
                        &tmp64c, &cand, &dest, &slot_of, &keep, &ghist, &gofs, &lb_status, &lb_ctr, &ktmp,
                        &idxa, &idxb, &posadj_dev, &rep_words, &rctl, &ttab, &kp, &part_scratch, &segb})
            b->release();
        for (DBuf& b : jobtab) b.release();
        if (st) small_flush(st);   // the small blocks kept for this stream, freed on it
        if (st) (void)hipStreamSynchronize(st);
        for (auto e : evpool) (void)hipEventDestroy(e);
        if (st) {
            stream_idle(st);
            std::lock_guard<std::mutex> lk(g_acct_mu);
            g_acct.erase(st);
        }
        if (host_stage) (void)hipHostFree(host_stage);
        if (mailbox) (void)hipHostFree(mailbox);
        for (char* js : job_stage)
            if (js) (void)hipHostFree(js);
        if (pub_ev) (void)hipEventDestroy(pub_ev);
        if (st) (void)hipStreamDestroy(st);
    }

    hipEvent_t ev()
    {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    int prof_begin(int k, hipEvent_t* a)
    {
        if (!(prof_mask & (1u << k))) return -1;
        *a = ev();
        HIPCHK(hipEventRecord(*a, st));
        return k;
    }
    // event opening a timed launch of a tuned knob, or null when not timing
    hipEvent_t tune_begin(bool timing)
    {
        if (!timing) return nullptr;
        hipEvent_t a = ev();
        HIPCHK(hipEventRecord(a, st));
        return a;
    }
    void tune_end(hipEvent_t a, int knob, int band, int var, size_t oi, u64 cap)
    {
        if (!a) return;
        hipEvent_t b = ev();
        HIPCHK(hipEventRecord(b, st));
        tune_pending.push_back({knob, band, var, a, b, oi, cap});
    }
    void prof_end(int k, hipEvent_t a, double bytes) noexcept
    {
        if (k < 0) return;
        hipEvent_t b = nullptr;
        if (!evpool.empty()) {
            b = evpool.back();
            evpool.pop_back();
        } else if (hipEventCreate(&b) != hipSuccess) {
            return;
        }
        if (hipEventRecord(b, st) != hipSuccess) return;
        pending.push_back({k, a, b, bytes});
    }
    void resolve_prof()
    {
        if (pending.empty()) return;
        if (hipStreamSynchronize(st) != hipSuccess) return;
        for (auto& p : pending) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                kst[p.k].ms += ms;
                kst[p.k].launches++;
                kst[p.k].bytes += p.bytes;
            }
            evpool.push_back(p.a);
            evpool.push_back(p.b);
        }
        pending.clear();
    }

    u64* host_buf(size_t words)
    {
        if (words > host_stage_cap) {
            if (host_stage) HIPCHK(hipHostFree(host_stage));
            host_stage = nullptr;
            size_t c = std::max<size_t>(words, 1024);
            HIPCHK(hipHostMalloc((void**)&host_stage, c * sizeof(u64)));
            host_stage_cap = c;
        }
        return host_stage;
    }
    // pinned staging for job tables copied to the device: a slot is rewritten
    // only after the host has waited for a kernel enqueued behind its previous
    // copy (see the slot list), and a larger buffer waits for the stream too
    char* job_stage_buf(int slot, size_t bytes)
    {
        if (bytes > job_stage_cap[slot]) {
            if (job_stage[slot]) {
                HIPCHK(hipStreamSynchronize(st));
                HIPCHK(hipHostFree(job_stage[slot]));
            }
            job_stage[slot] = nullptr;
            const size_t c = std::max<size_t>(bytes, 64 << 10);
            HIPCHK(hipHostMalloc((void**)&job_stage[slot], c, hipHostMallocDefault));
            job_stage_cap[slot] = c;
        }
        return job_stage[slot];
    }
    u64* mailbox_buf(size_t words)
    {
        if (words > mailbox_cap) {
            if (mailbox) HIPCHK(hipHostFree(mailbox));
            mailbox = nullptr;
            size_t c = std::max<size_t>(words, 1024);
            HIPCHK(hipHostMalloc((void**)&mailbox, c * sizeof(u64), hipHostMallocCoherent));
            memset(mailbox, 0, c * sizeof(u64));
            mailbox_cap = c;
        }
        return mailbox;
    }
    u32* dflags() { return flags_buf.as<u32>(); }
};

namespace {

// instance g of a context (created on first use)
static Group& grp(pra_ctx* c, int g)
{
    while ((int)c->groups.size() <= g) c->groups.emplace_back(new Group());
    return *c->groups[(size_t)g];
}

struct Prof {
    pra_ctx* c;
    int k;
    hipEvent_t a;
    double bytes;
    Prof(pra_ctx* c_, int k_, double b) : c(c_), k(-1), a(nullptr), bytes(b) { k = c->prof_begin(k_, &a); }
    ~Prof() { c->prof_end(k, a, bytes); }
};

// a timed launch whose byte count is (bytes per candidate) x (candidates of
// output `oi`), known only after the iteration's readback
struct ProfPerCand {
    pra_ctx* c;
    int k;
    hipEvent_t a;
    double per;
    size_t oi;
    ProfPerCand(pra_ctx* c_, int k_, double per_, size_t oi_) : c(c_), k(-1), a(nullptr), per(per_), oi(oi_)
    {
        k = c->prof_begin(k_, &a);
    }
    ~ProfPerCand()
    {
        if (k < 0) return;
        c->prof_end(k, a, 0.0);
        c->per_candidate.emplace_back(c->pending.size() - 1, per, oi);
    }
};

// A batched launch's job table: through its pinned staging slot into device
// memory, ordered on the context stream; returns the device copy.
static char* stage_table(pra_ctx* c, int slot, const void* data, size_t bytes)
{
    char* hs = c->job_stage_buf(slot, bytes);
    memcpy(hs, data, bytes);
    c->jobtab[slot].reserve(bytes, c->st);
    HIPCHK(hipMemcpyAsync(c->jobtab[slot].p, hs, bytes, hipMemcpyHostToDevice, c->st));
    return c->jobtab[slot].as<char>();
}

// D2D copies of several ranges: one copy when there is one, else one launch
static void copy_ranges(pra_ctx* c, int slot, const std::vector<CopyJob>& jobs)
{
    if (jobs.empty()) return;
    if (jobs.size() == 1) {
        HIPCHK(hipMemcpyAsync(jobs[0].dst, jobs[0].src, jobs[0].words * 8, hipMemcpyDeviceToDevice, c->st));
        return;
    }
    std::vector<char> t(jobs.size() * sizeof(CopyJob) + (jobs.size() + 1) * 4);
    memcpy(t.data(), jobs.data(), jobs.size() * sizeof(CopyJob));
    u32* bb = (u32*)(t.data() + jobs.size() * sizeof(CopyJob));
    bb[0] = 0;
    for (size_t k = 0; k < jobs.size(); k++) bb[k + 1] = bb[k] + (u32)((jobs[k].words + 2047) / 2048);
    const char* d = stage_table(c, slot, t.data(), t.size());
    pra_launch_copy_many(c->st, (const CopyJob*)d, (int)jobs.size(), bb[jobs.size()]);
    check_launch("copy_many");
}

// ----------------------------------------------------------------- scans ---

static void prepare_lookback(pra_ctx* c, u64 n);
static u32 next_sort_epoch(pra_ctx* c);

// exclusive scans on the context stream (one-pass look-back above
// PRA_SMALL_SCAN elements, over the sort's status words)
static void scan_u64(pra_ctx* c, const u64* in, u64* out, u64 n, u64* total)
{
    Prof p(c, K_SCAN, 16.0 * n);
    u32 ep = 0;
    if (n > PRA_SMALL_SCAN) {
        prepare_lookback(c, n);
        ep = next_sort_epoch(c);
    }
    pra_scan_u64(c->st, in, out, n, c->lb_status.as<u64>(), c->lb_ctr.as<u32>(), ep, c->dflags(), total);
    check_launch("scan_u64");
}

static void scan_u32(pra_ctx* c, const u32* in, u32* out, u64 n, u32* total)
{
    Prof p(c, K_SCAN, 8.0 * n);
    u32 ep = 0;
    if (n > PRA_SMALL_SCAN) {
        prepare_lookback(c, n);
        ep = next_sort_epoch(c);
    }
    pra_scan_u32(c->st, in, out, n, c->lb_status.as<u64>(), c->lb_ctr.as<u32>(), ep, c->dflags(), total);
    check_launch("scan_u32");
}

static void check_flags(u32 f)
{
    if (f & (PRA_FLAG_COLLISION | PRA_FLAG_INDEX_COLLISION))
        throw PraError(PRA_E_COLLISION, "64-bit key hash collision between distinct keys");
    if (f & PRA_FLAG_TABLE_FULL) throw PraError(PRA_E_HIP, "hash table overflow");
    if (f & PRA_FLAG_SORT_STALL) throw PraError(PRA_E_HIP, "sort or scan look-back did not complete");
    if (f & PRA_FLAG_BIN_OVERFLOW) throw PraError(PRA_E_HIP, "sort bin exceeds its capacity");
    if (f & PRA_FLAG_SCAN_OVERFLOW) throw PraError(PRA_E_HIP, "prefix sum exceeds 2^46");
}

// Read back `words` u64 from device memory in one synchronization, together
// with the device error flags.  Flags are sticky: kernels only ever set them
// on a detected error, and every synchronization point checks them.
static u64* readback(pra_ctx* c, const void* src, size_t words)
{
    u64* h = c->host_buf(words + 1);
    if (words) HIPCHK(hipMemcpyAsync(h, src, words * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(h + words, c->flags_buf.p, 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    stream_idle(c->st);
    check_flags((u32)h[words]);
    return h;
}

static void sync_check(pra_ctx* c) { readback(c, nullptr, 0); }

// The per-iteration readback: one kernel copies the summary words and the
// error flags into pinned host memory (no copy-engine transfers), then, once
// they are visible to the host, a sequence number in the word before them.  The host spins
// on that word instead of sleeping in a blocking synchronization or waiting
// for the kernel's completion to be signalled; an event recorded behind the
// kernel reports errors, and after a while the host falls back to a blocking
// wait on it.
// fold: a newt fold still to run, done by the publishing kernel itself
static u64* publish_begin(pra_ctx* c, const u64* src, size_t words, const NewtFoldArgs* fold = nullptr)
{
    // mailbox word 0: the sequence number; the summary and flags after it
    u64* h = c->mailbox_buf(words + 2) + 1;
    const u64 seq = ++c->pub_seq;
    if (fold)
        pra_launch_newt_fold_publish(c->st, *fold, src, words, c->flags_buf.as<u32>(), h, seq);
    else
        pra_launch_publish(c->st, src, words, c->flags_buf.as<u32>(), h, seq);
    check_launch("publish");
    if (!c->pub_ev) HIPCHK(hipEventCreateWithFlags(&c->pub_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(c->pub_ev, c->st));
    return h;
}

// waits for the summary of publish_begin (work enqueued after it keeps the
// device busy meanwhile) and checks the error flags
static u64* publish_wait(pra_ctx* c, u64* h, size_t words)
{
    auto t0 = std::chrono::steady_clock::now();
    for (u32 k = 0;; k++) {
        if (__atomic_load_n(h - 1, __ATOMIC_ACQUIRE) == c->pub_seq) break;
        if ((k & 63) == 63) {
            const hipError_t q = hipEventQuery(c->pub_ev);
            if (q != hipSuccess && q != hipErrorNotReady) HIPCHK(q);
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
                HIPCHK(hipEventSynchronize(c->pub_ev));
                break;
            }
        }
    }
    check_flags((u32)h[words]);
    return h;
}

static u64* publish_readback(pra_ctx* c, const u64* src, size_t words)
{
    return publish_wait(c, publish_begin(c, src, words), words);
}

// ------------------------------------------------------------------ sort ---

// look-back status words for `n` rows, zero-initialised once per allocation
static void prepare_lookback(pra_ctx* c, u64 n)
{
    size_t need = (size_t)pra_tiles(n) * PRA_RADIX;
    if (need > c->lb_zeroed) {
        c->lb_status.reserve(need * 8, c->st);
        size_t words = c->lb_status.cap / 8;
        pra_launch_fill_u64(c->st, c->lb_status.as<u64>(), words, 0);
        c->lb_zeroed = words;
        c->sort_epoch = 0;
    }
    if (!c->lb_ctr.p) {   // per-pass tile counters; each sort pass returns its own to zero
        c->lb_ctr.reserve(PRA_MAXPASS * 4 + 8, c->st);
        pra_launch_fill_u64(c->st, c->lb_ctr.as<u64>(), (PRA_MAXPASS * 4 + 8) / 8, 0);
    }
    c->ghist.reserve(PRA_MAXPASS * PRA_RADIX * 4, c->st);
    c->gofs.reserve(PRA_MAXPASS * PRA_RADIX * 4, c->st);
}

static u32 next_sort_epoch(pra_ctx* c)
{
    if (++c->sort_epoch >= 0xFFFF) {
        pra_launch_fill_u64(c->st, c->lb_status.as<u64>(), c->lb_zeroed, 0);
        c->sort_epoch = 1;
    }
    return c->sort_epoch;
}

// One-sweep LSD radix sort of rows by (bucket, col0 .. col[keycols-1]).
// `orm` holds the OR (words 0..15) and AND (16..31) of the first 16 key
// columns: digits equal in every row are skipped.  When the bucket digit is
// sorted and nb <= 256, the bucket bounds are written to `bnd` (nb+1 words)
// and true is returned.
// With `defer`, a sort that takes the one-block path is not launched: its job
// (into `defer_out`, which the caller swaps with r's buffer once the batch
// has been launched) is returned in *defer, with defer->n = 0 otherwise.
static bool sort_rows_masked(pra_ctx* c, Rows& r, int w, int jcc, int keycols, const u64* orm, u64* bnd,
                             bool by_bucket = true, SmallSortJob* defer = nullptr, DBuf* defer_out = nullptr)
{
    if (defer) defer->n = 0;
    u64 n = r.n;
    if (n <= 1) {
        r.sorted = true;
        return false;
    }
    std::vector<DigitSpec> passes;
    for (int col = keycols - 1; col >= 0; col--) {
        u64 vary = (orm && col < 16) ? (orm[col] ^ orm[16 + col]) : ~0ULL;
        for (int sh = 0; sh < 64; sh += PRA_RADIX_BITS)
            if ((vary >> sh) & (PRA_RADIX - 1)) passes.push_back(DigitSpec{col, sh, 0, 0, 0});
    }
    bool bucket_last = false;
    if (c->nb > 1 && by_bucket) {
        for (int sh = 0; ((u64)(c->nb - 1) >> sh) != 0; sh += PRA_RADIX_BITS)
            passes.push_back(DigitSpec{-1, sh, jcc, 0, (u64)c->nb});
        bucket_last = c->nb <= PRA_RADIX;
    }
    if (passes.empty()) {
        r.sorted = true;
        return false;
    }
    // small inputs with several varying bytes: one block sorts them in LDS
    const bool bucketed = c->nb > 1 && by_bucket;
    const int kw = keycols + (bucketed ? 1 : 0);
    if (n <= PRA_SMALL_SORT && n * kw <= (u64)PRA_SMALL_SORT * PRA_SMALL_SORT_KW && passes.size() >= 2) {
        u64* bout = (bnd && bucketed && bucket_last) ? bnd : nullptr;
        if (defer) {
            defer_out->reserve(n * w * 8, c->st);
            *defer = SmallSortJob{r.d(), defer_out->as<u64>(), bout, n, w, jcc, keycols, bucketed ? 1 : 0};
            r.sorted = true;
            return bout != nullptr;
        }
        c->sort_tmp.reserve(n * w * 8, c->st);
        {
            Prof p(c, K_SMALL_SORT, 8.0 * n * kw + 16.0 * n * w);
            pra_launch_small_sort(c->st, r.d(), n, w, jcc, keycols, (u64)c->nb, bucketed ? 1 : 0,
                                  c->sort_tmp.as<u64>(), bout);
            check_launch("small_sort");
        }
        r.buf.swap(c->sort_tmp);
        r.sorted = true;
        return bout != nullptr;
    }
    // up to PRA_BLOCK_SORT rows: one block runs every pass in one launch
    if (n <= PRA_BLOCK_SORT && passes.size() >= 2 && passes.size() <= PRA_MAXPASS) {
        PassList pl;
        memset(&pl, 0, sizeof(pl));
        pl.np = (int)passes.size();
        for (size_t i = 0; i < passes.size(); i++) pl.p[i] = passes[i];
        c->sort_tmp.reserve(n * w * 8, c->st);
        u64* bout = (bnd && bucket_last) ? bnd : nullptr;
        {
            Prof p(c, K_BLOCK_SORT, 8.0 * n * pl.np + 16.0 * n * w);
            pra_launch_block_sort(c->st, r.d(), n, w, pl, c->sort_tmp.as<u64>(), bout, jcc, (u64)c->nb);
            check_launch("block_sort");
        }
        r.buf.swap(c->sort_tmp);
        r.sorted = true;
        return bout != nullptr;
    }
    prepare_lookback(c, n);
    c->sort_tmp.reserve(n * w * 8, c->st);
    bool wrote_bnd = false;
    for (size_t first = 0; first < passes.size(); first += PRA_MAXPASS) {
        size_t np = std::min<size_t>(PRA_MAXPASS, passes.size() - first);
        PassList pl;
        memset(&pl, 0, sizeof(pl));
        pl.np = (int)np;
        for (size_t i = 0; i < np; i++) pl.p[i] = passes[first + i];
        bool last_chunk = first + np == passes.size();
        pra_launch_fill_u64(c->st, c->ghist.as<u64>(), np * PRA_RADIX / 2, 0);
        {
            Prof p(c, K_RADIX_HIST, 8.0 * n * keycols);
            pra_launch_sort_hist(c->st, r.d(), n, w, pl, c->ghist.as<u32>());
            check_launch("sort_hist");
        }
        u64* bout = (bnd && bucket_last && last_chunk) ? bnd : nullptr;
        pra_launch_sort_scan(c->st, c->ghist.as<u32>(), (int)np, c->gofs.as<u32>(), bout, (u32)c->nb, n);
        check_launch("sort_scan");
        wrote_bnd = wrote_bnd || bout != nullptr;
        for (size_t i = 0; i < np; i++) {
            u32 ep = next_sort_epoch(c);
            Prof p(c, K_RADIX_SCATTER, 16.0 * n * w);
            pra_launch_sort_scatter(c->st, r.d(), c->sort_tmp.as<u64>(), n, w, pl.p[i],
                                    c->gofs.as<u32>() + i * PRA_RADIX, c->lb_status.as<u64>(),
                                    c->lb_ctr.as<u32>() + i, ep, c->dflags(), nullptr, nullptr);
            check_launch("sort_scatter");
            r.buf.swap(c->sort_tmp);
        }
    }
    r.sorted = true;
    return wrote_bnd;
}

static void sort_rows(pra_ctx* c, Rows& r, int w, int jcc, int keycols, bool by_bucket = true)
{
    if (r.n <= 1) {
        r.sorted = true;
        return;
    }
    int kc = std::min(keycols, 16);
    c->orand.reserve(32 * 8, c->st);
    u64* om = c->orand.as<u64>();
    pra_launch_fill_u64(c->st, om, 16, 0);
    pra_launch_fill_u64(c->st, om + 16, 16, ~0ULL);
    pra_launch_col_orand2(c->st, r.d(), r.n, w, kc, om);
    check_launch("col_orand");
    u64* h = readback(c, om, 32);
    u64 m[32];
    std::copy(h, h + 32, m);
    sort_rows_masked(c, r, w, jcc, keycols, m, nullptr, by_bucket);
}

static int ceil_log2(u64 x);

// Packed sort key of an interned relation's newt rows: bucket * D^arity + the
// column ranks as digits of radix D (the dictionary size), which orders rows
// by (bucket, col0, col1, ..), shifted left by ls so that its top bit is bit
// 63.  bmul = D^arity.  False when the key range does not fit 64 bits.
static bool packed_layout(u64 nb, u64 D, int arity, u64* bmul, int* ls)
{
    unsigned __int128 m = 1;
    for (int i = 0; i < arity; i++) {
        m *= D;
        if (m >> 64) return false;
    }
    unsigned __int128 range = m * nb;
    if (range >> 64) return false;
    int bits = ceil_log2((u64)range);
    *bmul = (u64)m;
    *ls = bits ? 64 - bits : 0;
    return true;
}

// Sort a relation's newt rows by their packed keys (all distinct; R->nkeys in
// newt order, R->kgofs the exclusive offsets of the keys' top-byte bins, and
// the bins of (key, position) pairs filled by the newt scatter).  When the
// largest bin fits its capacity and one block, k_bin_sort orders every bin in
// LDS, gathers the rows and writes the bucket bounds (nb + 1 words);
// otherwise every varying byte gets a one-sweep pass over R->nkeys, followed
// by a row gather and a bounds search.
// With `defer`, the bin sort is not launched: its job (writing into
// `defer_out`, which the caller swaps with newt's buffer once the batch is
// launched) goes to *defer and true is returned.
static bool sort_by_packed_key(pra_ctx* c, pra_rel* R, u64 kor, u64 kand, u64 maxbin, u64 bmul, int ls, u64* bnd,
                               BinSortJob* defer = nullptr, DBuf* defer_out = nullptr)
{
    Rows& r = R->ver[PRA_NEWT];
    const int w = R->w;
    DBuf& keys = R->nkeys;
    const u32* top_ofs = R->kgofs.as<u32>();
    u64 n = r.n;
    u64 vary = kor ^ kand;
    const u64 bin_limit = std::min<u64>(std::min<u64>(R->bin_cap, PRA_BIN_CAP), c->bin_cap_test ? c->bin_cap_test : ~0ULL);
    if (n > 1 && maxbin <= bin_limit) {
        u32 bytes = 0;
        for (int b = 0; b < 7; b++)
            if ((vary >> (8 * b)) & (PRA_RADIX - 1)) bytes |= 1u << b;
        if (defer) {
            defer_out->reserve(n * w * 8, c->st);
            *defer = BinSortJob{R->bin_keys.as<u64>(), R->bin_idx.as<u32>(), top_ofs, r.d(), defer_out->as<u64>(), bnd,
                                n, bmul, (u64)c->nb, R->bin_cap, bytes, w, ls};
            r.sorted = true;
            return true;
        }
        c->sort_tmp.reserve(n * w * 8, c->st);
        {
            Prof p(c, K_BIN_SORT, 12.0 * n + 16.0 * n * w);
            pra_launch_bin_sort(c->st, R->bin_keys.as<u64>(), R->bin_idx.as<u32>(), R->bin_cap, n, top_ofs, bytes,
                                r.d(), w, c->sort_tmp.as<u64>(), bmul, ls, (u64)c->nb, bnd, c->dflags(), maxbin);
            check_launch("bin_sort");
        }
        r.buf.swap(c->sort_tmp);
        r.sorted = true;
        return false;
    }
    // LSD passes over the keys (and the permutation in idxa, from `permuted`
    // on): the byte-pass sort, and the first stage of the segmented one
    bool permuted = false;
    auto run_passes = [&](const std::vector<DigitSpec>& ps) {
        if (ps.empty()) return;
        prepare_lookback(c, n);
        c->ktmp.reserve(n * 8, c->st);
        c->idxa.reserve(n * 4, c->st);
        c->idxb.reserve(n * 4, c->st);
        const int np = (int)ps.size();
        PassList pl;
        memset(&pl, 0, sizeof(pl));
        pl.np = np;
        for (int i = 0; i < np; i++) pl.p[i] = ps[i];
        pra_launch_fill_u64(c->st, c->ghist.as<u64>(), np * PRA_RADIX / 2, 0);
        {
            Prof p(c, K_RADIX_HIST, 8.0 * n);
            pra_launch_sort_hist(c->st, keys.as<u64>(), n, 1, pl, c->ghist.as<u32>());
            check_launch("sort_hist");
        }
        pra_launch_sort_scan(c->st, c->ghist.as<u32>(), np, c->gofs.as<u32>(), nullptr, (u32)c->nb, n);
        check_launch("sort_scan");
        for (int i = 0; i < np; i++) {
            u32 ep = next_sort_epoch(c);
            Prof p(c, K_RADIX_SCATTER, 24.0 * n);
            pra_launch_sort_scatter(c->st, keys.as<u64>(), c->ktmp.as<u64>(), n, 1, ps[i],
                                    c->gofs.as<u32>() + (size_t)i * PRA_RADIX, c->lb_status.as<u64>(),
                                    c->lb_ctr.as<u32>() + i, ep, c->dflags(), permuted ? c->idxa.as<u32>() : nullptr,
                                    c->idxb.as<u32>());
            check_launch("key_sort_scatter");
            keys.swap(c->ktmp);
            c->idxa.swap(c->idxb);
            permuted = true;
        }
    };
    // Segmented sort: two counting passes over the keys' top two bytes
    // order the pairs by their top 16 bits; then the segments of equal top
    // `sb` bits (sized to average 2,048 pairs) are each sorted in LDS by
    // their lower varying bytes, rows gathered in the same kernel.  A segment
    // over PRA_BIN_CAP pairs sends the sort on to the byte passes, which take
    // the permutation as it stands.
    const u64 seg_min = c->bin_cap_test ? 2 : (1ULL << 16);
    std::vector<DigitSpec> top2;
    for (int sh : {48, 56})
        if ((vary >> sh) & (PRA_RADIX - 1)) top2.push_back(DigitSpec{0, sh, 0, 0, 0});
    if (c->seg_sort_on && n >= seg_min && n < (1ULL << 32) && !top2.empty()) {
        run_passes(top2);
        // segments of about 2,048 pairs on average (the keys of one iteration
        // are skewed: with smaller segments the largest still overflows the
        // narrow blocks, and the wide ones then sort many tiny segments)
        const int sb = std::min(16, std::max(9, ceil_log2(std::max<u64>(1, n / 2048))));
        const int top = 64 - sb;
        const u32 nseg = 1u << sb;
        // segment bounds, then the largest segment and the number of those
        // over PRA_SEG_SMALL pairs, then their list
        c->segb.reserve(((size_t)nseg + 3) * 8 + (size_t)nseg * 4, c->st);
        u64* seg = c->segb.as<u64>();
        u64* sst = seg + nseg + 1;
        u32* large = (u32*)(seg + nseg + 3);
        pra_launch_fill_u64(c->st, sst, 2, 0);
        pra_launch_seg_bounds(c->st, keys.as<u64>(), n, top, nseg, seg, sst, large);
        check_launch("seg_bounds");
        const u64* sstat = readback(c, sst, 2);
        const u64 maxseg = sstat[0], nlarge = sstat[1];
        if (maxseg <= PRA_BIN_CAP) {
            const u64 low = vary & ((1ULL << top) - 1);
            u32 bytes = 0;
            for (int b = 0; b < 8; b++)
                if ((low >> (8 * b)) & (PRA_RADIX - 1)) bytes |= 1u << b;
            c->sort_tmp.reserve(n * w * 8, c->st);
            {
                Prof p(c, K_BIN_SORT, 12.0 * n + 16.0 * n * w);
                pra_launch_seg_sort(c->st, keys.as<u64>(), c->idxa.as<u32>(), seg, top, nseg, n, bytes, r.d(), w,
                                    c->sort_tmp.as<u64>(), bmul, ls, (u64)c->nb, bnd, c->dflags(), maxseg, large,
                                    nlarge);
                check_launch("seg_sort");
            }
            r.buf.swap(c->sort_tmp);
            r.sorted = true;
            return false;
        }
    }
    std::vector<DigitSpec> passes;
    for (int sh = 0; sh < 64; sh += PRA_RADIX_BITS)
        if ((vary >> sh) & (PRA_RADIX - 1)) passes.push_back(DigitSpec{0, sh, 0, 0, 0});
    if (n > 1 && !passes.empty()) {
        run_passes(passes);
        c->sort_tmp.reserve(n * w * 8, c->st);
        {
            Prof p(c, K_GATHER, 16.0 * n * w + 4.0 * n);
            pra_launch_gather_rows(c->st, r.d(), c->idxa.as<u32>(), n, w, c->sort_tmp.as<u64>());
            check_launch("gather_rows");
        }
        r.buf.swap(c->sort_tmp);
    }
    r.sorted = true;
    if (c->nb > 1) {
        pra_launch_key_bounds(c->st, keys.as<u64>(), n, bmul, ls, (u64)c->nb, bnd);
        check_launch("key_bounds");
    } else {
        pra_launch_fill_u64(c->st, bnd, 1, 0);
        pra_launch_fill_u64(c->st, bnd + 1, 1, n);
    }
    return false;
}

static int ceil_log2(u64 x)
{
    int b = 0;
    while (b < 64 && (1ULL << b) < x) b++;
    return b;
}

// remove rows equal (in all w columns) to their predecessor; rows sorted
static void unique_rows(pra_ctx* c, Rows& r, int w)
{
    u64 n = r.n;
    if (n <= 1) return;
    c->tmp64a.reserve(n * 8, c->st);
    c->tmp64b.reserve(n * 8 + 8, c->st);
    c->sort_tmp.reserve(n * w * 8, c->st);
    pra_launch_group_flags(c->st, r.d(), n, w, w, c->tmp64a.as<u64>());
    scan_u64(c, c->tmp64a.as<u64>(), c->tmp64b.as<u64>(), n, c->tmp64b.as<u64>() + n);
    pra_launch_compact_rows(c->st, r.d(), n, w, c->tmp64a.as<u64>(), c->tmp64b.as<u64>(), c->sort_tmp.as<u64>());
    check_launch("compact_rows");
    u64* h = readback(c, c->tmp64b.as<u64>() + n, 1);
    r.buf.swap(c->sort_tmp);
    r.n = h[0];
}

// (Re)build instance gi's value dictionary from the key columns of every
// relation of that instance.
static void build_dictionary(pra_ctx* c, int gi)
{
    Group& G = grp(c, gi);
    u64 total = 0;
    for (pra_rel* R : c->rels)
        if (R->group == gi && !R->view_of)
            for (int v : {PRA_FULL, PRA_DELTA}) total += R->ver[v].n * (u64)R->arity;
    if (total == 0) return;
    Rows V;
    V.buf.reserve(total * 8, c->st);
    u64 off = 0;
    for (pra_rel* R : c->rels)
        for (int v : {PRA_FULL, PRA_DELTA}) {
            if (R->group != gi || R->view_of) continue;
            Rows& r = R->ver[v];
            pra_launch_collect_cols(c->st, r.d(), r.n, R->w, R->arity, V.d() + off);
            off += r.n * (u64)R->arity;
        }
    check_launch("collect_cols");
    V.n = total;
    sort_rows(c, V, 1, 0, 1, false);
    unique_rows(c, V, 1);
    u64 cap = pow2_at_least(std::max<u64>(64, 2 * V.n));
    G.dict.reserve(cap * 16, c->st);
    pra_launch_fill_u64(c->st, G.dict.as<u64>(), cap * 2, 0);
    pra_launch_dict_insert(c->st, V.d(), V.n, G.dict.as<u64>(), cap, c->dflags());
    check_launch("dict_insert");
    sync_check(c);
    G.dict_cap = cap;
    G.dict_n = V.n;
    G.dict_state = 1;
    G.dict_gen++;
}

static void ensure_sorted(pra_ctx* c, pra_rel* R, int v)
{
    Rows& r = R->ver[v];
    if (r.sorted) return;
    sort_rows(c, r, R->w, R->jcc, R->w);
    R->touched(v);
}

static const u64* bucket_bounds(pra_ctx* c, pra_rel* R, int v)
{
    Rows& r = R->ver[v];
    R->bnd[v].reserve((size_t)(c->nb + 1) * 8, c->st);
    if (!R->bnd_valid[v]) {
        if (c->nb == 1) {
            pra_launch_fill_u64(c->st, R->bnd[v].as<u64>(), 1, 0);
            pra_launch_fill_u64(c->st, R->bnd[v].as<u64>() + 1, 1, r.n);
            check_launch("bucket_bounds");
        } else {
            pra_launch_bucket_bounds(c->st, r.d(), r.n, R->w, R->jcc, (u64)c->nb, R->bnd[v].as<u64>());
            check_launch("bucket_bounds");
        }
        R->bnd_valid[v] = true;
    }
    return R->bnd[v].as<u64>();
}

// (key -> first row, count) index over the sorted rows of a version
static void ensure_index(pra_ctx* c, pra_rel* R, int v)
{
    if (R->idx_valid[v]) return;
    ensure_sorted(c, R, v);
    Rows& r = R->ver[v];
    u64 n = r.n;
    if (n == 0) {
        R->idx_cap[v] = 0;
        R->idx_maxlen[v] = 0;
        R->idx_valid[v] = true;
        return;
    }
    Prof p(c, K_INDEX_BUILD, 8.0 * n * R->jcc);
    c->tmp64a.reserve(n * 8, c->st);
    c->tmp64b.reserve((n + 1) * 8, c->st);
    c->tmp64c.reserve((n + 1) * 8, c->st);
    pra_launch_group_flags(c->st, r.d(), n, R->w, R->jcc, c->tmp64a.as<u64>());
    scan_u64(c, c->tmp64a.as<u64>(), c->tmp64b.as<u64>(), n, c->tmp64b.as<u64>() + n);
    pra_launch_group_compact(c->st, c->tmp64a.as<u64>(), c->tmp64b.as<u64>(), n, c->tmp64c.as<u64>());
    check_launch("group_compact");
    c->orand.reserve(32 * 8, c->st);
    pra_launch_fill_u64(c->st, c->orand.as<u64>(), 1, 0);
    pra_launch_group_maxlen(c->st, c->tmp64a.as<u64>(), c->tmp64b.as<u64>(), c->tmp64c.as<u64>(), n,
                            c->orand.as<u64>());
    check_launch("group_maxlen");
    HIPCHK(hipMemcpyAsync(c->orand.as<u64>() + 1, c->tmp64b.as<u64>() + n, 8, hipMemcpyDeviceToDevice, c->st));
    u64* hm = readback(c, c->orand.p, 2);
    R->idx_maxlen[v] = hm[0];
    u64 ng = hm[1];
    u64 cap = pow2_at_least(std::max<u64>(64, 2 * ng));
    R->idx[v].reserve(cap * 3 * 8, c->st);
    pra_launch_table_clear(c->st, R->idx[v].as<u64>(), cap, 3);
    pra_launch_index_insert(c->st, r.d(), R->w, R->jcc, c->tmp64c.as<u64>(), ng, R->idx[v].as<u64>(), cap,
                            c->dflags());
    check_launch("index_insert");
    sync_check(c);
    R->idx_cap[v] = cap;
    R->idx_valid[v] = true;
}

// membership table of full u delta keyed on kw leading columns, with room for
// `incoming` more keys at load factor <= 1/2
static void ensure_member_table(pra_ctx* c, pra_rel* R, int kw, u64 incoming)
{
    int stride = 2 + kw;
    if (!R->mt_valid || R->mt_kw != kw) {
        Prof p(c, K_TABLE_BUILD, 8.0 * (R->ver[PRA_FULL].n + R->ver[PRA_DELTA].n) * kw);
        u64 entries = R->ver[PRA_FULL].n + R->ver[PRA_DELTA].n;
        u64 cap = pow2_at_least(std::max<u64>(1024, 2 * (entries + incoming)));
        R->mt.reserve(cap * stride * 8, c->st);
        pra_launch_table_clear(c->st, R->mt.as<u64>(), cap, stride);
        for (int v : {PRA_FULL, PRA_DELTA}) {
            Rows& r = R->ver[v];
            pra_launch_table_insert_rows(c->st, r.d(), r.n, R->w, kw, R->mt.as<u64>(), cap, c->dflags());
            check_launch("table_insert_rows");
        }
        for (int v : {PRA_FULL, PRA_DELTA}) {
            Rows& r = R->ver[v];
            pra_launch_table_verify_rows(c->st, r.d(), r.n, R->w, kw, R->mt.as<u64>(), cap, c->dflags());
            check_launch("table_verify_rows");
        }
        sync_check(c);
        R->mt_cap = cap;
        R->mt_entries = entries;
        R->mt_kw = kw;
        R->mt_valid = true;
        return;
    }
    if (2 * (R->mt_entries + incoming) <= R->mt_cap) return;
    Prof p(c, K_TABLE_BUILD, 8.0 * R->mt_cap * stride);
    // grow 4x while the table is small, 2x beyond 2^28 slots
    u64 cap = pow2_at_least(2 * (R->mt_entries + incoming));
    if (cap < (1ULL << 28)) cap *= 2;
    DBuf nt;
    nt.reserve(cap * stride * 8, c->st);
    pra_launch_table_clear(c->st, nt.as<u64>(), cap, stride);
    pra_launch_table_rehash(c->st, R->mt.as<u64>(), R->mt_cap, nt.as<u64>(), cap, kw, c->dflags());
    check_launch("table_rehash");
    sync_check(c);
    R->mt.swap(nt);
    R->mt_cap = cap;
}

// Dense membership of interned output R keyed on kw columns: the bitmap of
// full u delta for the current dictionary (rebuilt when the dictionary or the
// relation changed since it was built) and a claim table of at least 2C
// slots.  False when the output does not allow it: more than 4 key columns, a
// rank space over 2^35 bits, 2^28 or more candidates, or a value of full u
// delta without a rank.
static bool ensure_dense(pra_ctx* c, pra_rel* R, int kw, u64 C, u64 slots_per_candidate)
{
    Group& G = grp(c, R->group);
    if (!c->dense_on || G.dict_state != 1 || kw < 1 || kw > 4 || G.dict_n == 0 || C >= (1ULL << PRA_DENSE_PBITS))
        return false;
    unsigned __int128 bits = 1;
    for (int i = 0; i < kw; i++) {
        bits *= G.dict_n;
        if (bits > ((unsigned __int128)1 << 35)) return false;
    }
    if (!R->bm_valid || R->bm_gen != G.dict_gen || R->bm_kw != kw) {
        const u64 words = ((u64)bits + 63) / 64;
        Prof p(c, K_TABLE_BUILD, 8.0 * words + 8.0 * kw * (R->ver[PRA_FULL].n + R->ver[PRA_DELTA].n));
        R->bm.reserve(words * 8, c->st);
        pra_launch_fill_u64(c->st, R->bm.as<u64>(), words, 0);
        c->orand.reserve(32 * 8, c->st);
        pra_launch_fill_u64(c->st, c->orand.as<u64>(), 1, 0);
        for (int v : {PRA_FULL, PRA_DELTA}) {
            Rows& r = R->ver[v];
            pra_launch_dense_build(c->st, r.d(), r.n, R->w, kw, G.dict.as<u64>(), G.dict_cap, G.dict_n,
                                   R->bm.as<u32>(), c->orand.as<u64>());
            check_launch("dense_build");
        }
        if (readback(c, c->orand.p, 1)[0]) {
            R->bm_valid = false;
            return false;
        }
        R->bm_valid = true;
        R->bm_gen = G.dict_gen;
        R->bm_kw = kw;
    }
    // slots in use: a power of two of at least slots_per_candidate x C; the
    // table is allocated for the largest use so far, and every slot is empty
    // between iterations, so a smaller use is a prefix of it
    const u64 cap = pow2_at_least(std::max<u64>(1024, slots_per_candidate * C));
    // the bits x + 1 <= D^kw needs leave the rest of its 36 bits to an epoch
    // tag (at most 16, or the context's setting): each use of the table takes
    // the next epoch, so the slots of earlier ones count as empty, and the
    // table is filled with zeros only when the epochs wrap
    int kbits = 0;
    while (kbits < 36 && ((unsigned __int128)1 << kbits) <= bits) kbits++;
    const int eb = std::min(c->dense_epoch_bits, 36 - kbits);
    bool clear = false;
    if (cap > R->dtab_cap) {
        R->dtab.reserve(cap * 8, c->st);
        R->dtab_cap = cap;
        clear = true;
    }
    if (eb != R->dtab_eb) {
        R->dtab_eb = eb;
        clear = true;
    }
    if (eb > 0 && !clear && R->dtab_epoch + 1 >= (1u << eb)) clear = true;
    if (clear) {
        pra_launch_dense_table_clear(c->st, R->dtab.as<u64>(), R->dtab_cap);
        check_launch("dense_table_clear");
        R->dtab_epoch = 0;
    }
    if (eb > 0) R->dtab_epoch++;
    R->dtab_use = cap;
    return true;
}

static DBuf& seq_counters(pra_ctx* c, int gi, u32 tag)
{
    Group& G = grp(c, gi);
    auto it = G.seq.find(tag);
    if (it != G.seq.end()) return *it->second;
    std::unique_ptr<DBuf> b(new DBuf());
    b->reserve((size_t)c->nb * 8, c->st);
    pra_launch_fill_u64(c->st, b->as<u64>(), (u64)c->nb, 0);
    DBuf& ref = *b;
    G.seq[tag] = std::move(b);
    return ref;
}

// append src rows (already on device) to a version; with `jobs`, the copy is
// left to the caller's batched copy (the version grows at once)
static void append_rows(pra_ctx* c, pra_rel* R, int v, const u64* src, u64 n, std::vector<CopyJob>* jobs = nullptr)
{
    if (!n) return;
    Rows& r = R->ver[v];
    size_t used = (size_t)r.n * R->w * 8;
    r.buf.reserve(used + (size_t)n * R->w * 8, c->st, used);
    if (jobs)
        jobs->push_back(CopyJob{src, r.d() + r.n * R->w, n * (u64)R->w});
    else
        HIPCHK(hipMemcpyAsync(r.d() + r.n * R->w, src, (size_t)n * R->w * 8, hipMemcpyDeviceToDevice, c->st));
    r.n += n;
    r.sorted = false;
    R->touched(v);
}

// full <- full u delta keeping full sorted: both are sorted by (bucket,
// columns) and disjoint, so each delta row's place is a binary search inside
// its bucket and every row moves once (falls back to an append when either
// side is unsorted)
static void merge_into_full(pra_ctx* c, pra_rel* R)
{
    Rows& D = R->ver[PRA_DELTA];
    Rows& F = R->ver[PRA_FULL];
    if (D.n == 0) return;
    if (F.n == 0 || !F.sorted || !D.sorted) {
        bool sorted = F.n == 0 && D.sorted;
        append_rows(c, R, PRA_FULL, D.d(), D.n);
        F.sorted = sorted;
        return;
    }
    const u64* fb = bucket_bounds(c, R, PRA_FULL);
    const u64* db = bucket_bounds(c, R, PRA_DELTA);
    u64 n = F.n + D.n;
    c->tmp64a.reserve(D.n * 8, c->st);
    c->sort_tmp.reserve(n * R->w * 8, c->st);
    {
        Prof p(c, K_MERGE, 16.0 * R->w * n);
        // the merged bucket s holds full's and delta's rows of bucket s: the
        // merge adds delta's bounds to full's
        pra_launch_merge(c->st, F.d(), F.n, fb, D.d(), D.n, db, (u64)c->nb, R->w, c->tmp64a.as<u64>(),
                         c->sort_tmp.as<u64>(), R->bnd[PRA_FULL].as<u64>());
        check_launch("merge");
    }
    F.buf.swap(c->sort_tmp);
    F.n = n;
    F.sorted = true;
    R->touched(PRA_FULL);
    R->bnd_valid[PRA_FULL] = true;
}

// The SCCs of one lock-step iteration seen as one: the rules of every member
// in member order, then rule order, and every member's relations.  Members
// come from different instances of the context and share no relation.
struct RAView {
    std::vector<RADesc*> v;
    RADesc& operator[](size_t i) const { return *v[i]; }
    size_t size() const { return v.size(); }
    struct It {
        RADesc* const* p;
        RADesc& operator*() const { return **p; }
        It& operator++()
        {
            ++p;
            return *this;
        }
        bool operator!=(const It& o) const { return p != o.p; }
    };
    It begin() const { return It{v.data()}; }
    It end() const { return It{v.data() + v.size()}; }
};
struct Merged {
    pra_ctx* ctx = nullptr;
    RAView ras;
    std::vector<pra_rel*> rels;
    explicit Merged(pra_ctx* c, const std::vector<pra_scc*>& members) : ctx(c)
    {
        for (pra_scc* m : members) {
            for (auto& ra : m->ras) ras.v.push_back(&ra);
            rels.insert(rels.end(), m->rels.begin(), m->rels.end());
        }
    }
};

// One threshold round of an output relation: the rules emitting (RA index
// order) and, per listed rule, the outer rows [lo, hi) of every source.
struct RoundSpec {
    std::vector<int> ras;
    std::vector<u64> lo, hi;   // per listed rule, nb each
    DBuf* bounds = nullptr;    // the same on the device: per listed rule lo[nb] then hi[nb]
    u64 nt_base = 0;           // newt rows written by earlier rounds
    u64 cap = 0;               // candidates of this round
};

static u64 read_word(pra_ctx* c, const u64* p) { return readback(c, p, 1)[0]; }

// Where the reference's local_join call over outer rows [start, end) of rule r
// stops (src/RA/parallel_join.cpp:157-230): its outputs are deduplicated per
// call, and the call returns after the first row through which it emitted
// more than the threshold.  Returns that row relative to start, or end-start.
static u64 find_split(pra_ctx* c, Merged* S, int r, u64 start, u64 end, int cw)
{
    RADesc& ra = S->ras[r];
    const int nb = c->nb;
    const u64 n = end - start;
    const u64 base = read_word(c, ra.off.as<u64>() + start);
    const u64 raw = read_word(c, ra.off.as<u64>() + end) - base;
    c->rctl.reserve((size_t)(nb + 4) * 8, c->st);
    u64* tctl = c->rctl.as<u64>();        // [0] candidates, [1] overflow
    u64* tout = tctl + 2;                  // split row
    u64* tposadj = tctl + 4;               // per source: -base
    pra_launch_fill_u64(c->st, tctl, 1, raw);
    pra_launch_fill_u64(c->st, tctl + 1, 1, 0);
    pra_launch_fill_u64(c->st, tposadj, (u64)nb, (u64)0 - base);
    c->cand.reserve(raw * cw * 8, c->st);
    c->dest.reserve(raw * 4, c->st);
    c->slot_of.reserve(raw * 8, c->st);
    c->keep.reserve(raw * 4, c->st);
    {
        JoinExpandArgs a;
        memset(&a, 0, sizeof(a));
        a.ctl = tctl;
        a.outer = ra.outer->ver[ra.outer_v].d() + start * ra.outer->w;
        a.inner = ra.inner->ver[ra.inner_v].d();
        a.off = ra.off.as<u64>() + start;
        a.istart = ra.istart.as<u64>() + start;
        a.posadj = tposadj;
        a.cand = c->cand.as<u64>();
        a.dest = c->dest.as<u32>();
        a.n_outer = n;
        a.nb = (u64)nb;
        a.wo = ra.outer->w;
        a.wi = ra.inner->w;
        a.jcc = ra.in0->jcc;
        a.left = ra.left;
        a.cw = cw;
        a.out_jcc = ra.out->jcc;
        for (int i = 0; i < cw; i++) a.map[i] = ra.map[i];
        pra_launch_join_expand(c->st, a, raw);
        check_launch("split_expand");
    }
    // first occurrences inside the call: a fresh table keyed on whole outputs
    const u64 tcap = pow2_at_least(std::max<u64>(1024, 2 * raw));
    c->ttab.reserve(tcap * (2 + cw) * 8, c->st);
    pra_launch_table_clear(c->st, c->ttab.as<u64>(), tcap, 2 + cw);
    pra_launch_dedup_claim2(c->st, c->cand.as<u64>(), tctl, raw, cw, cw, c->ttab.as<u64>(), tcap,
                            c->slot_of.as<u64>(), 1, c->dflags());
    const u32 nblocks = pra_tiles(raw);
    c->hist.reserve((size_t)nb * nblocks * 4, c->st);
    pra_launch_dedup_decide_hist(c->st, c->cand.as<u64>(), tctl, cw, cw, c->ttab.as<u64>(), c->slot_of.as<u64>(), 1,
                                 c->keep.as<u32>(), c->dest.as<u32>(), (u32)nb, c->hist.as<u32>(), nblocks,
                                 c->dflags());
    check_launch("split_dedup");
    c->kp.reserve((raw + 1) * 4, c->st);
    scan_u32(c, c->keep.as<u32>(), c->kp.as<u32>(), raw, c->kp.as<u32>() + raw);
    pra_launch_find_split(c->st, ra.off.as<u64>() + start, n, c->kp.as<u32>(), c->threshold, tout);
    check_launch("find_split");
    return read_word(c, tout);
}

// The round loop of RAM::execute_in_batches for one output relation whose
// join calls exceed the local-join threshold (src/RAM/RA_tasks.cpp:1121-1269):
// round k emits, per source and rule, the rows from where that call stopped
// in round k-1 to where it stops now; copies and finished calls would only
// re-emit tuples that already arrived, so they emit in the first round only.
// Interning and first-arrival dedup run per round, rounds in order.  Returns
// the new tuples of all rounds.
template <class F>
static u64 run_rounds(Merged* S, pra_rel* R, F&& run_round, u64* st2o_dev)
{
    pra_ctx* c = S->ctx;
    const int nb = c->nb;
    const int cw = R->arity;   // rounds happen for join outputs, which are interned
    struct Call {
        int r;
        std::vector<u64> bst, start;
        std::vector<char> done;
    };
    std::vector<Call> joins;
    std::vector<int> copies;
    for (int r = 0; r < (int)S->ras.size(); r++) {
        RADesc& ra = S->ras[r];
        if (ra.out != R || !ra.active) continue;
        if (ra.type != RA_JOIN) {
            copies.push_back(r);
            continue;
        }
        Call cl;
        cl.r = r;
        u64* hb = readback(c, ra.outer->bnd[ra.outer_v].p, (size_t)nb + 1);
        cl.bst.assign(hb, hb + nb + 1);
        cl.start.assign(cl.bst.begin(), cl.bst.begin() + nb);
        cl.done.resize(nb);
        for (int s = 0; s < nb; s++) cl.done[s] = cl.start[s] == cl.bst[s + 1];
        joins.push_back(std::move(cl));
    }
    DBuf bounds;
    u64 total = 0;
    for (int round = 0;; round++) {
        // (rule, lo, hi) of this round, in rule order
        std::vector<std::tuple<int, std::vector<u64>, std::vector<u64>>> rules;
        u64 cap = 0;
        for (auto& cl : joins) {
            RADesc& ra = S->ras[cl.r];
            std::vector<u64> lo(cl.start), hi(cl.start);
            bool any = false;
            for (int s = 0; s < nb; s++) {
                if (cl.done[s]) continue;
                u64 end = cl.bst[s + 1];
                u64 o0 = read_word(c, ra.off.as<u64>() + cl.start[s]);
                u64 stop = end;
                if ((long long)(read_word(c, ra.off.as<u64>() + end) - o0) > c->threshold)
                    stop = std::min(end, cl.start[s] + find_split(c, S, cl.r, cl.start[s], end, cw) + 1);
                hi[s] = stop;
                cap += read_word(c, ra.off.as<u64>() + stop) - o0;
                any = true;
            }
            if (any) rules.emplace_back(cl.r, lo, hi);
        }
        if (round == 0)
            for (int r : copies) {
                RADesc& ra = S->ras[r];
                u64* hb = readback(c, ra.outer->bnd[ra.outer_v].p, (size_t)nb + 1);
                std::vector<u64> lo(hb, hb + nb), hi(hb + 1, hb + nb + 1);
                cap += read_word(c, ra.off.as<u64>() + ra.n_outer);
                rules.emplace_back(r, lo, hi);
            }
        if (rules.empty()) break;
        std::sort(rules.begin(), rules.end(),
                  [](const auto& a, const auto& b) { return std::get<0>(a) < std::get<0>(b); });
        RoundSpec rs;
        rs.bounds = &bounds;
        rs.nt_base = total;
        rs.cap = cap;
        std::vector<u64> dev;
        for (auto& t : rules) {
            rs.ras.push_back(std::get<0>(t));
            rs.lo.insert(rs.lo.end(), std::get<1>(t).begin(), std::get<1>(t).end());
            rs.hi.insert(rs.hi.end(), std::get<2>(t).begin(), std::get<2>(t).end());
            dev.insert(dev.end(), std::get<1>(t).begin(), std::get<1>(t).end());
            dev.insert(dev.end(), std::get<2>(t).begin(), std::get<2>(t).end());
        }
        bounds.reserve(dev.size() * 8, c->st);
        HIPCHK(hipMemcpyAsync(bounds.p, dev.data(), dev.size() * 8, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (cap) {
            run_round(&rs);
            u64 cnt = publish_readback(c, st2o_dev, 8)[0];
            total += cnt;
            R->mt_entries += cnt;
        }
        bool all = true;
        for (auto& cl : joins) {
            for (auto& t : rules)
                if (std::get<0>(t) == cl.r)
                    for (int s = 0; s < nb; s++)
                        if (!cl.done[s]) {
                            cl.start[s] = std::get<2>(t)[s];
                            cl.done[s] = cl.start[s] == cl.bst[s + 1];
                        }
            for (int s = 0; s < nb; s++) all = all && cl.done[s];
        }
        if (all) break;
    }
    return total;
}

// ------------------------------------------------------------ iteration ---

static bool is_intern_type(int type) { return type != RA_ACOPY; }

// One semi-naive iteration (src/RAM/RA_tasks.cpp:1121-1269) of one SCC, or
// of the current SCCs of several instances in lock step: their rules, outputs
// and per-phase launches are taken together, while dedup, intern ids and
// dictionaries stay per instance.
// A join whose inputs have different join column counts matches on input 0's
// count (src/RAM/RA_tasks.cpp:448 passes input0->get_join_column_count() to
// local_join) and reads input 1 through a view keyed on that count.
// local_join probes input 1 in the bucket of the jcc0-column prefix
// (src/RA/parallel_join.cpp:180-186), and a rank holds only the rows of its own
// buckets, so a row of input 1 takes part only when its bucket under its first
// jcc0 columns equals its own bucket; for a RIGHT join it is the outer row
// whose bucket is checked the same way.  The view holds exactly those rows.
// Input 1's version is sorted by (bucket, columns), so the kept rows are in
// the view's order already.
static void refresh_view(pra_ctx* c, pra_rel* V, int v)
{
    pra_rel* src = V->view_of;
    ensure_sorted(c, src, v);
    if (V->view_gen[v] == src->gen[v]) return;
    Rows& S = src->ver[v];
    Rows& D = V->ver[v];
    const u64 n = S.n;
    if (n >= (1ULL << 32)) throw PraError(PRA_E_UNSUPPORTED, "join view of 2^32 rows or more");
    u64 kept = 0;
    if (n) {
        c->keep.reserve(n * 4, c->st);
        c->kp.reserve(n * 4, c->st);
        c->tmp64a.reserve(8, c->st);
        pra_launch_fill_u64(c->st, c->tmp64a.as<u64>(), 1, 0);
        pra_launch_view_keep(c->st, S.d(), n, src->w, V->jcc, src->jcc, (u64)c->nb, c->keep.as<u32>());
        check_launch("view_keep");
        scan_u32(c, c->keep.as<u32>(), c->kp.as<u32>(), n, c->tmp64a.as<u32>());
        kept = readback(c, c->tmp64a.p, 1)[0] & 0xFFFFFFFFULL;
        D.buf.reserve(std::max<u64>(1, kept) * src->w * 8, c->st);
        pra_launch_view_compact(c->st, S.d(), n, src->w, c->keep.as<u32>(), c->kp.as<u32>(), D.d());
        check_launch("view_compact");
    }
    D.n = kept;
    D.sorted = true;
    V->touched(v);
    V->view_gen[v] = src->gen[v];
}

static void run_iteration(pra_ctx* c, const std::vector<pra_scc*>& members)
{
    for (pra_scc* m : members)
        for (auto& ra : m->ras)
            if (ra.type == RA_JOIN && ra.in1->view_of) refresh_view(c, ra.in1, ra.v1);
    Merged merged(c, members);
    Merged* S = &merged;
    const int nb = c->nb;
    const int nra = (int)S->ras.size();

    // 0. inputs ready: sorted versions, join indexes.  Rules whose outer (or
    //    inner) version is empty emit nothing and launch nothing.
    for (auto& ra : S->ras) {
        if (ra.type == RA_JOIN) {
            if (ra.v0 == PRA_DELTA) {
                ra.outer = ra.in0; ra.outer_v = PRA_DELTA; ra.inner = ra.in1; ra.inner_v = ra.v1; ra.left = 1;
            } else {
                ra.outer = ra.in1; ra.outer_v = ra.v1; ra.inner = ra.in0; ra.inner_v = PRA_FULL; ra.left = 0;
            }
            ra.active = ra.outer->ver[ra.outer_v].n > 0 && ra.inner->ver[ra.inner_v].n > 0;
            ra.sorted_probe = ra.inner->scc_changing;
            if (!ra.active) continue;
            ensure_sorted(c, ra.outer, ra.outer_v);
            if (ra.sorted_probe) {
                ensure_sorted(c, ra.inner, ra.inner_v);
                bucket_bounds(c, ra.inner, ra.inner_v);
            } else {
                ensure_index(c, ra.inner, ra.inner_v);
            }
        } else {
            ra.outer = ra.in0; ra.outer_v = ra.v0; ra.inner = nullptr; ra.left = 1;
            ra.active = ra.outer->ver[ra.outer_v].n > 0;
            if (ra.active) ensure_sorted(c, ra.outer, ra.outer_v);
        }
    }

    // 1. per-RA output counts and their exclusive offsets (the arrival plan
    //    reads them at the source-bucket bounds)
    c->stage.reserve((size_t)std::max(1, nra) * 2 * nb * 8, c->st);
    CountScanBatch csb;
    memset(&csb, 0, sizeof(csb));
    double cs_bytes = 0;
    auto flush_count_scans = [&]() {
        if (!csb.njobs) return;
        prepare_lookback(c, (u64)csb.tile_base[csb.njobs] * PRA_TILE);
        const u32 ep = next_sort_epoch(c);
        {
            Prof p(c, K_JOIN_COUNT, cs_bytes);
            pra_launch_count_scan_many(c->st, csb, c->lb_status.as<u64>(), c->lb_ctr.as<u32>(), ep, c->dflags());
            check_launch("count_scan_many");
        }
        memset(&csb, 0, sizeof(csb));
        cs_bytes = 0;
    };
    for (int r = 0; r < nra; r++) {
        RADesc& ra = S->ras[r];
        Rows& O = ra.outer->ver[ra.outer_v];
        u64 n = O.n;
        ra.n_outer = ra.active ? n : 0;
        if (!ra.active) continue;
        ra.off.reserve((n + 1) * 8, c->st);
        ra.istart.reserve(std::max<u64>(1, n) * 8, c->st);
        if (n == 0) {
            pra_launch_fill_u64(c->st, ra.off.as<u64>(), 1, 0);
        } else {
            // the counts and their offsets in one pass, batched over rules
            CountScanJob& j = csb.job[csb.njobs];
            memset(&j, 0, sizeof(j));
            j.outer = O.d();
            j.n = n;
            j.wo = ra.outer->w;
            j.off = ra.off.as<u64>();
            const bool join = ra.type == RA_JOIN;
            j.kind = !join ? PRA_CS_FILTER : ra.sorted_probe ? PRA_CS_SORTED : PRA_CS_INDEX;
            if (join) {
                j.inner = ra.inner->ver[ra.inner_v].d();
                j.wi = ra.inner->w;
                j.jcc = ra.in0->jcc;
                j.istart = ra.istart.as<u64>();
                if (ra.sorted_probe) {
                    j.ibnd = ra.inner->bnd[ra.inner_v].as<u64>();
                    j.nb = (u64)nb;
                } else {
                    j.table = ra.inner->idx[ra.inner_v].as<u64>();
                    j.cap = ra.inner->idx_cap[ra.inner_v];
                }
            } else {
                j.op = ra.type == RA_COPY_FILTER ? ra.op : PRA_F_TRUE;
            }
            csb.tile_base[csb.njobs + 1] = csb.tile_base[csb.njobs] + pra_tiles(n);
            cs_bytes += 8.0 * n * ra.outer->w + (join ? 24.0 : 8.0) * n;
            if (++csb.njobs == PRA_CS_JOBS) flush_count_scans();
        }
        bucket_bounds(c, ra.outer, ra.outer_v);
    }
    flush_count_scans();

    // 2. output relations, their rule kinds, and candidate capacities: the
    //    previous iteration's count with headroom, else an upper bound from the
    //    outer sizes and the inner index's longest key run.  No host
    //    synchronization: the arrival plan is computed on the device and an
    //    output whose candidates exceed its capacity is re-run.
    std::vector<pra_rel*> outs;
    for (auto& ra : S->ras)
        if (std::find(outs.begin(), outs.end(), ra.out) == outs.end()) outs.push_back(ra.out);
    const size_t no = outs.size();
    std::vector<int> out_intern(no, 1);
    std::vector<u64> ub(no, 0), cap(no, 0);
    for (size_t oi = 0; oi < no; oi++) {
        int type0 = -1, nr = 0;
        for (auto& ra : S->ras)
            if (ra.out == outs[oi]) {
                nr++;
                if (type0 < 0) type0 = is_intern_type(ra.type);
                else if (type0 != (int)is_intern_type(ra.type))
                    throw PraError(PRA_E_UNSUPPORTED, "relation is the target of both acopy and join/copy rules");
                if (!ra.active) continue;
                if (ra.type != RA_JOIN)
                    ub[oi] += ra.n_outer;
                else if (ra.sorted_probe)   // no run-length bound without an index: outer x inner
                    ub[oi] += ra.inner->ver[ra.inner_v].n > (~0ULL >> 8) / ra.n_outer
                                  ? (~0ULL >> 8)
                                  : ra.n_outer * ra.inner->ver[ra.inner_v].n;
                else
                    ub[oi] += ra.n_outer * ra.inner->idx_maxlen[ra.inner_v];
            }
        if (nr > PRA_MAXRA) throw PraError(PRA_E_UNSUPPORTED, "too many rules produce one relation");
        out_intern[oi] = type0;
        pra_rel* R = outs[oi];
        cap[oi] = R->jhint ? std::max<u64>(2 * R->jhint, 65536) : std::min<u64>(ub[oi], 1ULL << 22);
        cap[oi] = std::min(cap[oi], ub[oi]);
    }
    for (size_t i = 0; i < no; i++)
        for (size_t j = i + 1; j < no; j++)
            if (out_intern[i] && out_intern[j] && outs[i]->tag == outs[j]->tag && outs[i]->group == outs[j]->group)
                throw PraError(PRA_E_UNSUPPORTED, "two interned relations of one SCC share an intern tag");

    // per output relation: [0] new tuples, [1] packed-key OR, [2] AND, [3] dictionary miss,
    // [4] candidates, [5] capacity overflow, [6] largest top-byte bin of the packed keys,
    // [8..23] key-column OR, [24..39] AND
    const size_t SW = 40;
    c->stage2.reserve(std::max<size_t>(1, no) * SW * 8, c->st);
    c->posadj_dev.reserve((size_t)std::max(1, nra) * nb * 8, c->st);
    for (size_t oi = 0; oi < no; oi++) {
        Group& G = grp(c, outs[oi]->group);
        if (!G.joined.p) {
            G.joined.reserve(8, c->st);
            pra_launch_fill_u64(c->st, G.joined.as<u64>(), 1, 0);
        }
    }
    // packed sort keys need the instance's value dictionary (rebuilt after a
    // miss, given up after repeated misses)
    for (size_t oi = 0; oi < no; oi++)
        if (out_intern[oi] && ub[oi] && grp(c, outs[oi]->group).dict_state == 0)
            build_dictionary(c, outs[oi]->group);
    std::vector<char> packed(no, 0);
    std::vector<u64> pk_mul(no, 0);
    std::vector<int> pk_ls(no, 0);
    for (size_t oi = 0; oi < no; oi++) {
        const Group& G = grp(c, outs[oi]->group);
        packed[oi] = G.dict_state == 1 && out_intern[oi] && ub[oi] &&
                     packed_layout((u64)nb, G.dict_n, outs[oi]->arity, &pk_mul[oi], &pk_ls[oi]);
    }

    // 3+4. per output: arrival plan, emit candidates, deduplicate, partition
    //      and intern, summaries for the sort
    // rs == nullptr: the iteration's whole emission; else one threshold round
    // the last output of the main pass leaves its fold to the publication
    size_t last_oi = no;
    for (size_t oi = 0; oi < no; oi++)
        if (cap[oi]) last_oi = oi;
    NewtFoldArgs last_fold;
    bool have_last_fold = false;
    // batched partitions (unpacked, packed) and the folds that wait for them
    std::vector<NewtScatterArgs> scatter_jobs[2];
    std::vector<size_t> scatter_outs[2];
    std::vector<NewtFoldArgs> later_folds;
    // batched expansions: join (arguments, blocks, per-candidate bytes and
    // output) and copy (arguments, blocks, bytes)
    std::vector<JoinExpandArgs> jexp_jobs;
    std::vector<u32> jexp_blocks;
    std::vector<std::pair<double, size_t>> jexp_per;
    std::vector<CopyExpandArgs> cexp_jobs;
    std::vector<u32> cexp_blocks;
    double cexp_bytes = 0;
    // planned: the arrival plan was made by the iteration's batched plan
    // per output: its rows (dest, slot_of, keep), candidate words and
    // histogram words in the shared buffers when its claim and decision are
    // batched with other outputs
    struct OutRegion {
        u64 rb = 0, cb = 0, hb = 0;
    };
    std::vector<OutRegion> reg(no);
    std::vector<ClaimJob> claim_jobs(no);
    std::vector<DecideJob> decide_jobs(no);
    // dense membership: taken per output, switched off for the rest of the
    // iteration by a dictionary miss; the dense claims of a batched iteration
    std::vector<char> use_dense(no, 0), dense_off(no, 0);
    std::vector<ScanJob> scan_jobs;
    std::vector<std::pair<DenseArgs, size_t>> dense_jobs;
    // phase 0: one output start to end; 1: up to its candidates, 2: its
    // histogram scan, 3: from its partition on
    // kept counts on (the claims and decisions in between batched over the
    // outputs, each in its own region of the candidate buffers)
    auto run_output = [&](size_t oi, const RoundSpec* rs, bool planned = false, int phase = 0) {
        const OutRegion rg = phase ? reg[oi] : OutRegion{};
        pra_rel* R = outs[oi];
        const u64 C = rs ? rs->cap : cap[oi];
        const bool pk = packed[oi] && !rs;
        const u64 nt_base = rs ? rs->nt_base : 0;
        const bool intern = out_intern[oi] == 1;
        const int cw = intern ? R->arity : R->arity + 1;
        const int kw = cw;
        Rows& NT = R->ver[PRA_NEWT];
        u64* st2o = c->stage2.as<u64>() + oi * SW;
        u64* ctl = st2o + 4;
        // dense claim table size: tuned (timed) on single-output iterations
        // of at least kTuneMinCandidates candidates of capacity, the settled
        // choice (or 2 slots per candidate) elsewhere
        int dvar = 0;
        bool dtime = false;
        if (c->tune_on && !rs && C >= kTuneMinCandidates) {
            if (phase == 0) {
                dvar = g_tune.pick(c->device, TUNE_DENSE_LOAD, band_of(C), kTuneVariants[TUNE_DENSE_LOAD]);
                dtime = g_tune.settled(c->device, TUNE_DENSE_LOAD, band_of(C)) < 0;
            } else {
                dvar = std::max(0, g_tune.settled(c->device, TUNE_DENSE_LOAD, band_of(C)));
            }
        }
        // dense membership unless the output was redone the general way in
        // this iteration, or this is a threshold round
        const bool dn = phase < 2 ? (!rs && !dense_off[oi] && packed[oi] && ensure_dense(c, R, kw, C, 2ULL << dvar))
                                   : use_dense[oi] != 0;
        use_dense[oi] = dn;
        if (dn)
            R->mt_valid = false;   // the general table is not kept current meanwhile
        else
            R->bm_valid = false, ensure_member_table(c, R, kw, C);
        c->cand.reserve(C * cw * 8, c->st);
        c->dest.reserve(C * 4, c->st);
        c->slot_of.reserve(C * 8, c->st);
        c->keep.reserve(C * 4, c->st);
        NT.buf.reserve((nt_base + C) * R->w * 8, c->st, nt_base * R->w * 8);
        const u32 nblocks = pra_tiles(C);
        c->hist.reserve((size_t)nb * nblocks * 4, c->st);
        c->offs.reserve((size_t)nb * nblocks * 4 + 8, c->st);
        u64* cand0 = c->cand.as<u64>() + rg.cb;
        u32* dest0 = c->dest.as<u32>() + rg.rb;
        u64* slot0 = c->slot_of.as<u64>() + rg.rb;
        u32* keep0 = c->keep.as<u32>() + rg.rb;
        u32* hist0 = c->hist.as<u32>() + rg.hb;
        u32* offs0 = c->offs.as<u32>() + rg.hb;
        if (phase < 2) {
            if (!planned) {
                PlanArgs pa;
                memset(&pa, 0, sizeof(pa));
                pa.srcoff = c->stage.as<u64>();
                pa.posadj = c->posadj_dev.as<u64>();
                pa.ctl = ctl;
                pa.joined = grp(c, R->group).joined.as<u64>();
                pa.flags = c->dflags();
                pa.cap = C;
                pa.thresh = c->threshold;
                pa.nb = nb;
                pa.rounds_check = rs ? 0 : 1;
                pa.count_joined = rs ? 0 : 1;
                if (rs) {
                    for (size_t k = 0; k < rs->ras.size(); k++) {
                        RADesc& ra = S->ras[rs->ras[k]];
                        pa.ras[pa.nr] = rs->ras[k];
                        pa.is_join[pa.nr] = ra.type == RA_JOIN;
                        pa.off[pa.nr] = ra.off.as<u64>();
                        pa.lo[pa.nr] = rs->bounds->as<u64>() + k * 2 * nb;
                        pa.hi[pa.nr] = rs->bounds->as<u64>() + k * 2 * nb + nb;
                        pa.nr++;
                    }
                } else {
                    for (int r = 0; r < nra; r++) {
                        RADesc& ra = S->ras[r];
                        if (ra.out != R || !ra.active) continue;
                        pa.ras[pa.nr] = r;
                        pa.is_join[pa.nr] = ra.type == RA_JOIN;
                        pa.off[pa.nr] = ra.off.as<u64>();
                        pa.lo[pa.nr] = ra.outer->bnd[ra.outer_v].as<u64>();
                        pa.hi[pa.nr] = ra.outer->bnd[ra.outer_v].as<u64>() + 1;
                        pa.nr++;
                    }
                }
                pra_launch_plan(c->st, pa);
                check_launch("plan");
            }
            for (int r = 0; r < nra; r++) {
                RADesc& ra = S->ras[r];
                if (ra.out != R || !ra.active) continue;
                size_t k = 0;   // position in the round's rule list
                if (rs) {
                    while (k < rs->ras.size() && rs->ras[k] != r) k++;
                    if (k == rs->ras.size()) continue;
                }
                Rows& O = ra.outer->ver[ra.outer_v];
                if (ra.type == RA_JOIN) {
                    // a round emits one window of rows per source; otherwise all rows at once
                    const int nwin = rs ? nb : 1;
                    for (int s = 0; s < nwin; s++) {
                    u64 lo = rs ? rs->lo[k * nb + s] : 0, hi = rs ? rs->hi[k * nb + s] : ra.n_outer;
                    if (hi <= lo) continue;
                    JoinExpandArgs a;
                    memset(&a, 0, sizeof(a));
                    a.ctl = ctl;
                    a.outer = O.d() + lo * ra.outer->w;
                    a.inner = ra.inner->ver[ra.inner_v].d();
                    a.off = ra.off.as<u64>() + lo;
                    a.istart = ra.istart.as<u64>() + lo;
                    a.posadj = c->posadj_dev.as<u64>() + (size_t)r * nb;
                    a.cand = cand0;
                    a.dest = dest0;
                    a.n_outer = hi - lo;
                    a.nb = (u64)nb;
                    a.wo = ra.outer->w;
                    a.wi = ra.inner->w;
                    a.jcc = ra.in0->jcc;
                    a.left = ra.left;
                    a.cw = cw;
                    a.out_jcc = R->jcc;
                    for (int i = 0; i < cw; i++) a.map[i] = ra.map[i];
                    if (phase == 1) {
                        jexp_jobs.push_back(a);
                        jexp_blocks.push_back(pra_join_expand_blocks(std::min<u64>(C, ub[oi])));
                        jexp_per.emplace_back(8.0 * ra.inner->w + 8.0 * cw, oi);
                        continue;
                    }
                    ProfPerCand p(c, K_JOIN_EXPAND, 8.0 * ra.inner->w + 8.0 * cw, oi);
                    pra_launch_join_expand(c->st, a, std::min<u64>(C, ub[oi]));
                    check_launch("join_expand");
                    }
                } else {
                    CopyExpandArgs a;
                    memset(&a, 0, sizeof(a));
                    a.ctl = ctl;
                    a.rows = O.d();
                    a.off = ra.off.as<u64>();
                    a.posadj = c->posadj_dev.as<u64>() + (size_t)r * nb;
                    a.cand = cand0;
                    a.dest = dest0;
                    a.n = ra.n_outer;
                    a.nb = (u64)nb;
                    a.w = ra.outer->w;
                    a.in_jcc = ra.outer->jcc;
                    a.cw = cw;
                    a.out_jcc = R->jcc;
                    for (int i = 0; i < cw; i++) a.map[i] = ra.map[i];
                    if (phase == 1) {
                        if (a.n) {
                            cexp_jobs.push_back(a);
                            cexp_blocks.push_back((u32)((a.n + PRA_BLOCK - 1) / PRA_BLOCK));
                            cexp_bytes += (double)ra.n_outer * 8.0 * (ra.outer->w + cw);
                        }
                        continue;
                    }
                    Prof p(c, K_COPY_EXPAND, (double)ra.n_outer * 8.0 * (ra.outer->w + cw));
                    pra_launch_copy_expand(c->st, a);
                    check_launch("copy_expand");
                }
            }
            if (dn) {
                DenseArgs da;
                memset(&da, 0, sizeof(da));
                da.cand = cand0;
                da.ctl = ctl;
                da.slot_of = slot0;
                da.keep = keep0;
                da.dest = dest0;
                da.hist = hist0;
                da.offs1 = nblocks == 1 ? offs0 : nullptr;
                Group& G = grp(c, R->group);
                da.dict = G.dict.as<u64>();
                da.dict_cap = G.dict_cap;
                da.D = G.dict_n;
                da.bitmap = R->bm.as<u32>();
                da.table = R->dtab.as<u64>();
                da.cap = R->dtab_use;
                if (R->dtab_eb > 0) {
                    da.emask = ~0ULL << (64 - R->dtab_eb);
                    da.stamp = (u64)R->dtab_epoch << (64 - R->dtab_eb);
                }
                da.nblocks = nblocks;
                da.nb = (u32)nb;
                da.cw = cw;
                da.kw = kw;
                if (phase == 1) {
                    // launched after the batched expansions
                    dense_jobs.emplace_back(da, oi);
                    return;
                }
                hipEvent_t ta = c->tune_begin(dtime);
                {
                    ProfPerCand p(c, K_DEDUP_CLAIM, 8.0 * (kw + 1), oi);
                    pra_launch_dense_claim(c->st, da, C, c->dflags());
                    check_launch("dense_claim");
                }
                {
                    ProfPerCand p(c, K_DEDUP_VERIFY, 16.0, oi);
                    pra_launch_dense_decide(c->st, da, c->dflags());
                    check_launch("dense_decide");
                }
                c->tune_end(ta, TUNE_DENSE_LOAD, band_of(C), dvar, oi, C);
            }
        }
        if (phase < 2 && !dn) {
            u64* tab = R->mt.as<u64>();
            const u64 epoch = ++c->dedup_epoch;
            if (phase == 1) {
                // the batched claim and decision take it from here
                ClaimJob& cj = claim_jobs[oi];
                cj = ClaimJob{cand0, ctl, tab, slot0, R->mt_cap, epoch, cw, kw};
                DecideJob& dj = decide_jobs[oi];
                dj = DecideJob{cand0, ctl, tab, slot0, keep0, dest0, hist0, nblocks == 1 ? offs0 : nullptr, epoch,
                               nblocks, cw, kw};
                return;
            }
            {
                ProfPerCand p(c, K_DEDUP_CLAIM, 8.0 * (kw + 1), oi);
                pra_launch_dedup_claim2(c->st, cand0, ctl, C, cw, kw, tab, R->mt_cap, slot0, epoch, c->dflags());
                check_launch("dedup_claim");
            }
            {
                ProfPerCand p(c, K_DEDUP_VERIFY, 8.0 * kw + 16.0, oi);
                // one tile: the decision writes the offsets itself
                pra_launch_dedup_decide_hist(c->st, cand0, ctl, cw, kw, tab, slot0, epoch, keep0, dest0, (u32)nb, hist0,
                                             nblocks, c->dflags(), nblocks == 1 ? offs0 : nullptr);
                check_launch("dedup_decide_hist");
            }
        }
        u32* total_dev = offs0 + (size_t)nb * nblocks;
        if (phase == 2) {
            // the small histogram scans of a batched iteration go out together
            if (nblocks > 1) {
                if ((u64)nb * nblocks <= PRA_SMALL_SCAN)
                    scan_jobs.push_back(ScanJob{hist0, offs0, total_dev, (u64)nb * nblocks});
                else
                    scan_u32(c, hist0, offs0, (u64)nb * nblocks, total_dev);
            }
            return;
        }
        if (phase == 0 && nblocks > 1) scan_u32(c, hist0, offs0, (u64)nb * nblocks, total_dev);
        DBuf* sq = intern ? &seq_counters(c, R->group, R->tag) : nullptr;
        R->kgofs.reserve(PRA_RADIX * 4, c->st);
        NewtFoldArgs f;   // bin_cur: set with the scatter's bins
        memset(&f, 0, sizeof(f));
        // batched partitions: one summary replica set per output
        u64* rep0 = c->rep_words.as<u64>() + (phase ? oi : 0) * (size_t)PRA_NREP * PRA_PART_WORDS;
        f.rep_words = rep0;
        f.ctl = ctl;
        f.stat = st2o;
        f.kgofs = R->kgofs.as<u32>();
        f.seqbase = sq ? sq->as<u64>() : nullptr;
        f.offs = offs0;
        f.total = total_dev;
        f.nblocks = nblocks;
        f.nb = (u32)nb;
        f.packed = pk ? 1 : 0;
        f.intern = intern ? 1 : 0;
        // one tile: the scatter's block folds the summaries itself
        const bool fold_here = nblocks == 1;
        bool batched = false;
        {
            NewtScatterArgs a;
            memset(&a, 0, sizeof(a));
            a.ctl = ctl;
            a.cand = cand0;
            a.keep = keep0;
            a.dest = dest0;
            a.offs = offs0;
            a.seqbase = sq ? sq->as<u64>() : nullptr;
            a.newt = NT.d() + nt_base * R->w;
            a.nblocks = nblocks;
            a.nb = (u32)nb;
            a.cw = cw;
            a.wn = R->w;
            a.intern = intern ? 1 : 0;
            a.tag = R->tag;
            a.rep_words = rep0;
            if (pk) {
                // bins of twice the average top-byte share (a fuller bin sends
                // the sort to its byte-pass fallback)
                const u32 bcap = (u32)std::min<u64>(PRA_BIN_CAP, pow2_at_least(std::max<u64>(256, 2 * C / 256 + 1)));
                if (bcap > R->bin_cap) {
                    R->bin_keys.reserve((size_t)PRA_RADIX * bcap * 8, c->st);
                    R->bin_idx.reserve((size_t)PRA_RADIX * bcap * 4, c->st);
                    R->bin_cap = bcap;
                }
                if (!R->bin_cur.p) {
                    R->bin_cur.reserve(PRA_RADIX * 4, c->st);
                    pra_launch_fill_u64(c->st, R->bin_cur.as<u64>(), PRA_RADIX / 2, 0);
                }
                a.bin_cap = R->bin_cap;
                a.bin_keys = R->bin_keys.as<u64>();
                a.bin_idx = R->bin_idx.as<u32>();
                a.bin_cur = R->bin_cur.as<u32>();
                f.bin_cur = a.bin_cur;
                R->nkeys.reserve(C * 8, c->st);
                Group& G = grp(c, R->group);
                a.dict = G.dict.as<u64>();
                a.dict_cap = G.dict_cap;
                a.keys = R->nkeys.as<u64>();
                a.rank_radix = G.dict_n;
                a.key_lshift = pk_ls[oi];
                if (dn) {
                    // the dense decision leaves each kept candidate's rank
                    // index in its slot_of word
                    u64 scale = 1;
                    for (int i = 0; i < cw; i++) scale *= G.dict_n;
                    a.xk = slot0;
                    a.xk_scale = scale;
                }
            }
            a.fold_here = fold_here ? 1 : 0;
            a.fold = f;
            // column masks only pick the sort passes of outputs too large for
            // the one-block sort, which compares whole key words
            a.summarize = C > PRA_SMALL_SORT || C * (u64)(R->w + 1) > (u64)PRA_SMALL_SORT * PRA_SMALL_SORT_KW;
            if (phase == 3 && nblocks < PRA_SCATTER_WIDE_TILES) {
                // narrow partitions of a batched iteration go out together
                scatter_jobs[pk ? 1 : 0].push_back(a);
                scatter_outs[pk ? 1 : 0].push_back(oi);
                batched = true;
            } else {
                // block width: tuned (timed) on single-output iterations of
                // at least two tiles
                int wide = -1;
                bool stime = false;
                const int sband = band_of(nblocks);
                if (c->tune_on && nblocks >= 2) {
                    if (phase == 0) {
                        wide = g_tune.pick(c->device, TUNE_SCATTER_WIDE, sband, kTuneVariants[TUNE_SCATTER_WIDE]);
                        stime = g_tune.settled(c->device, TUNE_SCATTER_WIDE, sband) < 0;
                    } else {
                        wide = g_tune.settled(c->device, TUNE_SCATTER_WIDE, sband);
                    }
                }
                hipEvent_t ta = c->tune_begin(stime);
                ProfPerCand p(c, K_DEDUP_SCATTER, 8.0 * cw + 8.0, oi);
                pra_launch_dedup_scatter(c->st, a, wide);
                check_launch("dedup_scatter");
                c->tune_end(ta, TUNE_SCATTER_WIDE, sband, wide, oi, C);
            }
        }
        if (!fold_here && planned && oi == last_oi) {
            // the iteration's last output: folded by the summary's publication
            last_fold = f;
            have_last_fold = true;
        } else if (!fold_here && batched) {
            later_folds.push_back(f);
        } else if (!fold_here) {
            pra_launch_newt_fold(c->st, f);
            check_launch("newt_fold");
        }
    };
    // every output's arrival plan in one launch: its per-rule tables as kernel
    // arguments, or, past their size, from tables in device memory
    const bool plan_all = true;
    if (nra > PRA_MAXRA || no > PRA_MAXRA) {
        std::vector<PlanRule> rules((size_t)nra);
        std::vector<u32> ids;
        std::vector<PlanOut> pouts;
        for (int r = 0; r < nra; r++) {
            RADesc& ra = S->ras[r];
            if (ra.active)
                rules[(size_t)r] = PlanRule{ra.off.as<u64>(), ra.outer->bnd[ra.outer_v].as<u64>(),
                                            ra.type == RA_JOIN ? 1u : 0u, 0u};
        }
        for (size_t oi = 0; oi < no; oi++) {
            if (!cap[oi]) continue;
            PlanOut po{cap[oi], grp(c, outs[oi]->group).joined.as<u64>(), (u32)ids.size(), 0, (u32)oi, 0};
            for (int r = 0; r < nra; r++)
                if (S->ras[r].active && S->ras[r].out == outs[oi]) ids.push_back((u32)r);
            po.count = (u32)ids.size() - po.first;
            pouts.push_back(po);
        }
        const size_t rb = (rules.size() * sizeof(PlanRule) + 255) / 256 * 256;
        const size_t ib = (ids.size() * 4 + 255) / 256 * 256;
        std::vector<char> t(rb + ib + pouts.size() * sizeof(PlanOut));
        memcpy(t.data(), rules.data(), rules.size() * sizeof(PlanRule));
        memcpy(t.data() + rb, ids.data(), ids.size() * 4);
        memcpy(t.data() + rb + ib, pouts.data(), pouts.size() * sizeof(PlanOut));
        const char* d = stage_table(c, JOB_PLAN, t.data(), t.size());
        PlanTabArgs pa{c->stage.as<u64>(), c->posadj_dev.as<u64>(), c->stage2.as<u64>(), c->threshold, nb, (int)SW,
                       (const PlanRule*)d, (const u32*)(d + rb), (const PlanOut*)(d + rb + ib)};
        pra_launch_plan_tab(c->st, pa, (int)pouts.size());
        check_launch("plan_tab");
    } else {
        PlanAllArgs pl;
        memset(&pl, 0, sizeof(pl));
        pl.srcoff = c->stage.as<u64>();
        pl.posadj = c->posadj_dev.as<u64>();
        pl.stage2 = c->stage2.as<u64>();
        for (size_t oi = 0; oi < no; oi++) pl.joined[oi] = grp(c, outs[oi]->group).joined.as<u64>();
        pl.thresh = c->threshold;
        pl.nb = nb;
        pl.nra = nra;
        pl.sw = (int)SW;
        int nblk = 0;
        for (size_t oi = 0; oi < no; oi++) {
            pl.cap[oi] = cap[oi];
            if (cap[oi]) pl.outs[nblk++] = (unsigned char)oi;
        }
        for (int r = 0; r < nra; r++) {
            RADesc& ra = S->ras[r];
            pl.rout[r] = 0xFF;
            if (!ra.active) continue;
            const size_t oi = std::find(outs.begin(), outs.end(), ra.out) - outs.begin();
            pl.rout[r] = (unsigned char)oi;
            pl.is_join[r] = ra.type == RA_JOIN;
            pl.off[r] = ra.off.as<u64>();
            pl.bnd[r] = ra.outer->bnd[ra.outer_v].as<u64>();
        }
        pra_launch_plan_all(c->st, pl, nblk);
        check_launch("plan_all");
    }
    size_t nactive = 0;
    for (size_t oi = 0; oi < no; oi++) nactive += cap[oi] ? 1 : 0;
    if (plan_all && nactive >= 2) {
        // every output's candidates in its own region; their claims and
        // decisions batched by key width
        u64 rt = 0, ct = 0, ht = 0;
        for (size_t oi = 0; oi < no; oi++) {
            if (!cap[oi]) continue;
            const int cwo = out_intern[oi] == 1 ? outs[oi]->arity : outs[oi]->arity + 1;
            reg[oi] = OutRegion{rt, ct, ht};
            rt += cap[oi];
            ct += cap[oi] * (u64)cwo;
            ht += (u64)nb * pra_tiles(cap[oi]) + 2;
        }
        c->cand.reserve(ct * 8, c->st);
        c->dest.reserve(rt * 4, c->st);
        c->slot_of.reserve(rt * 8, c->st);
        c->keep.reserve(rt * 4, c->st);
        c->hist.reserve(ht * 4, c->st);
        c->offs.reserve(ht * 4 + 8, c->st);
        for (size_t oi = 0; oi < no; oi++)
            if (cap[oi]) run_output(oi, nullptr, true, 1);
        // the expansions of all rules: job tables through pinned staging into
        // device memory, one launch per kind
        {
            auto table_bytes = [](size_t n, size_t sz) { return n * sz + ((n + 1) * 4 + 255) / 256 * 256; };
            const size_t jb = jexp_jobs.empty() ? 0 : table_bytes(jexp_jobs.size(), sizeof(JoinExpandArgs));
            const size_t cb = cexp_jobs.empty() ? 0 : table_bytes(cexp_jobs.size(), sizeof(CopyExpandArgs));
            if (jb + cb) {
                char* hs = c->job_stage_buf(0, jb + cb);
                c->jobtab[0].reserve(jb + cb, c->st);
                char* dv = c->jobtab[0].as<char>();
                auto put = [&](size_t at, const void* jobs, size_t n, size_t sz, const std::vector<u32>& blocks) {
                    memcpy(hs + at, jobs, n * sz);
                    u32* bb = (u32*)(hs + at + n * sz);
                    bb[0] = 0;
                    for (size_t k = 0; k < n; k++) bb[k + 1] = bb[k] + blocks[k];
                    HIPCHK(hipMemcpyAsync(dv + at, hs + at, table_bytes(n, sz), hipMemcpyHostToDevice, c->st));
                    return bb[n];
                };
                if (jb) {
                    const u32 nbk = put(0, jexp_jobs.data(), jexp_jobs.size(), sizeof(JoinExpandArgs), jexp_blocks);
                    const size_t before = c->pending.size();
                    {
                        Prof p(c, K_JOIN_EXPAND, 0.0);
                        pra_launch_expand_many(c->st, dv, (int)jexp_jobs.size(), nbk, true);
                        check_launch("join_expand_many");
                    }
                    if (c->pending.size() == before + 1)
                        for (auto& pe : jexp_per) c->per_candidate.emplace_back(before, pe.first, pe.second);
                }
                if (cb) {
                    const u32 nbk = put(jb, cexp_jobs.data(), cexp_jobs.size(), sizeof(CopyExpandArgs), cexp_blocks);
                    Prof p(c, K_COPY_EXPAND, cexp_bytes);
                    pra_launch_expand_many(c->st, dv + jb, (int)cexp_jobs.size(), nbk, false);
                    check_launch("copy_expand_many");
                }
            }
        }
        // the dense claims and decisions: one table per key width through
        // pinned staging, one launch each
        if (!dense_jobs.empty()) {
            auto tbytes = [](size_t n) { return (n * sizeof(DenseArgs) + 2 * (n + 1) * 4 + 255) / 256 * 256; };
            size_t total = 0;
            for (int kw = 1; kw <= 4; kw++) {
                size_t n = 0;
                for (auto& dj : dense_jobs) n += dj.first.kw == kw;
                if (n) total += tbytes(n);
            }
            char* hs = c->job_stage_buf(2, total);
            c->jobtab[2].reserve(total, c->st);
            size_t at = 0;
            for (int kw = 1; kw <= 4; kw++) {
                std::vector<size_t> members;
                for (size_t k = 0; k < dense_jobs.size(); k++)
                    if (dense_jobs[k].first.kw == kw) members.push_back(k);
                if (members.empty()) continue;
                const size_t n = members.size();
                DenseArgs* J = (DenseArgs*)(hs + at);
                u32* cb = (u32*)(J + n);
                u32* db = cb + n + 1;
                cb[0] = db[0] = 0;
                for (size_t k = 0; k < n; k++) {
                    const DenseArgs& a = dense_jobs[members[k]].first;
                    const size_t oi = dense_jobs[members[k]].second;
                    J[k] = a;
                    cb[k + 1] = cb[k] + (u32)((cap[oi] + PRA_BLOCK - 1) / PRA_BLOCK);
                    db[k + 1] = db[k] + a.nblocks;
                }
                HIPCHK(hipMemcpyAsync(c->jobtab[2].as<char>() + at, hs + at, tbytes(n), hipMemcpyHostToDevice, c->st));
                const size_t before = c->pending.size();
                {
                    Prof p(c, K_DEDUP_CLAIM, 0.0);
                    pra_launch_dense_many(c->st, (const DenseArgs*)(c->jobtab[2].as<char>() + at), (int)n, kw, cb[n], 0,
                                          c->dflags());
                    check_launch("dense_claim_many");
                }
                if (c->pending.size() == before + 1)
                    for (size_t k : members)
                        c->per_candidate.emplace_back(before, 8.0 * (kw + 1), dense_jobs[k].second);
                const size_t before2 = c->pending.size();
                {
                    Prof p(c, K_DEDUP_VERIFY, 0.0);
                    pra_launch_dense_many(c->st, (const DenseArgs*)(c->jobtab[2].as<char>() + at), (int)n, kw, 0, db[n],
                                          c->dflags());
                    check_launch("dense_decide_many");
                }
                if (c->pending.size() == before2 + 1)
                    for (size_t k : members) c->per_candidate.emplace_back(before2, 16.0, dense_jobs[k].second);
                at += tbytes(n);
            }
        }
        for (int cls = 1; cls <= 7; cls++) {
            ClaimBatch cb;
            DecideBatch db;
            memset(&cb, 0, sizeof(cb));
            memset(&db, 0, sizeof(db));
            db.nb = (u32)nb;
            std::vector<size_t> members;
            auto flush = [&]() {
                if (!cb.njobs) return;
                // timed launches: bytes per candidate of every member output
                size_t before = c->pending.size();
                {
                    Prof p(c, K_DEDUP_CLAIM, 0.0);
                    pra_launch_dedup_claim_many(c->st, cb, cls, c->dflags());
                    check_launch("dedup_claim_many");
                }
                if (c->pending.size() == before + 1)
                    for (size_t oi : members) c->per_candidate.emplace_back(before, 8.0 * (claim_jobs[oi].kw + 1), oi);
                before = c->pending.size();
                {
                    Prof p(c, K_DEDUP_VERIFY, 0.0);
                    pra_launch_dedup_decide_many(c->st, db, cls, c->dflags());
                    check_launch("dedup_decide_many");
                }
                if (c->pending.size() == before + 1)
                    for (size_t oi : members)
                        c->per_candidate.emplace_back(before, 8.0 * decide_jobs[oi].kw + 16.0, oi);
                memset(&cb, 0, sizeof(cb));
                memset(&db, 0, sizeof(db));
                db.nb = (u32)nb;
                members.clear();
            };
            for (size_t oi = 0; oi < no; oi++) {
                if (!cap[oi] || use_dense[oi] || pra_dedup_class(claim_jobs[oi].kw) != cls) continue;
                const int k = cb.njobs;
                cb.job[k] = claim_jobs[oi];
                cb.block_base[k + 1] = cb.block_base[k] + (u32)((cap[oi] + PRA_BLOCK - 1) / PRA_BLOCK);
                cb.njobs++;
                db.job[k] = decide_jobs[oi];
                db.block_base[k + 1] = db.block_base[k] + decide_jobs[oi].nblocks;
                db.njobs++;
                members.push_back(oi);
                if (cb.njobs == PRA_DEDUP_JOBS) flush();
            }
            flush();
        }
        if (c->rep_sets < (int)no) {
            // every set at its identity values (sets at rest always are)
            c->rep_words.reserve(no * (size_t)PRA_NREP * PRA_PART_WORDS * 8, c->st);
            pra_launch_newt_rep_init(c->st, c->rep_words.as<u64>(), (int)no);
            check_launch("newt_rep_init");
            c->rep_sets = (int)no;
        }
        for (size_t oi = 0; oi < no; oi++)
            if (cap[oi]) run_output(oi, nullptr, true, 2);
        if (!scan_jobs.empty()) {
            Prof p(c, K_SCAN, 0.0);
            const char* d = stage_table(c, JOB_SCAN, scan_jobs.data(), scan_jobs.size() * sizeof(ScanJob));
            pra_launch_scan_small_many(c->st, (const ScanJob*)d, (int)scan_jobs.size());
            check_launch("scan_small_many");
        }
        for (size_t oi = 0; oi < no; oi++)
            if (cap[oi]) run_output(oi, nullptr, true, 3);
        // the batched partitions: job tables through pinned staging into
        // device memory, one launch per key kind
        size_t stage_bytes = 0;
        for (int pk = 0; pk < 2; pk++)
            if (!scatter_jobs[pk].empty())
                stage_bytes += scatter_jobs[pk].size() * sizeof(NewtScatterArgs) +
                               ((scatter_jobs[pk].size() + 1) * 4 + 255) / 256 * 256;
        if (stage_bytes) {
            char* hs = c->job_stage_buf(1, stage_bytes);
            c->jobtab[1].reserve(stage_bytes, c->st);
            size_t at = 0;
            for (int pk = 0; pk < 2; pk++) {
                const auto& J = scatter_jobs[pk];
                if (J.empty()) continue;
                const size_t bytes = J.size() * sizeof(NewtScatterArgs) + ((J.size() + 1) * 4 + 255) / 256 * 256;
                memcpy(hs + at, J.data(), J.size() * sizeof(NewtScatterArgs));
                u32* bb = (u32*)(hs + at + J.size() * sizeof(NewtScatterArgs));
                bb[0] = 0;
                for (size_t k = 0; k < J.size(); k++) bb[k + 1] = bb[k] + J[k].nblocks;
                HIPCHK(hipMemcpyAsync(c->jobtab[1].as<char>() + at, hs + at, bytes, hipMemcpyHostToDevice, c->st));
                const size_t before = c->pending.size();
                {
                    Prof p(c, K_DEDUP_SCATTER, 0.0);
                    pra_launch_dedup_scatter_many(c->st, (const NewtScatterArgs*)(c->jobtab[1].as<char>() + at),
                                                  (int)J.size(), bb[J.size()], pk == 1);
                    check_launch("dedup_scatter_many");
                }
                if (c->pending.size() == before + 1)
                    for (size_t k = 0; k < J.size(); k++)
                        c->per_candidate.emplace_back(before, 8.0 * J[k].cw + 8.0, scatter_outs[pk][k]);
                at += bytes;
            }
        }
        if (later_folds.size() == 1) {
            pra_launch_newt_fold(c->st, later_folds[0]);
            check_launch("newt_fold");
        } else if (!later_folds.empty()) {
            const char* d = stage_table(c, JOB_FOLD, later_folds.data(), later_folds.size() * sizeof(NewtFoldArgs));
            pra_launch_newt_fold_many(c->st, (const NewtFoldArgs*)d, (int)later_folds.size());
            check_launch("newt_fold_many");
        }
    } else {
        for (size_t oi = 0; oi < no; oi++)
            if (cap[oi]) run_output(oi, nullptr, plan_all);
    }

    // 5. new-tuple counts, sort-key summaries and error flags: the
    //    per-iteration host readback (repeated once for outputs whose
    //    candidates overflowed their capacity, now sized exactly)
    std::vector<u64> st2(no * SW, 0);
    // deltas of relations whose full version no rule of the SCC reads go to
    // their full logs while the host waits for the first readback
    std::vector<char> early(S->rels.size(), 0);
    // pass 1 redoes outputs whose candidates overflowed their capacity (now
    // sized exactly) or whose dense claims met a value without a dictionary
    // rank (now the general way); pass 2 the few that needed both
    for (int pass = 0; pass < 3; pass++) {
        u64* h = publish_begin(c, c->stage2.as<u64>(), no * SW, pass == 0 && have_last_fold ? &last_fold : nullptr);
        if (pass == 0) {
            std::vector<CopyJob> appends;
            for (size_t i = 0; i < S->rels.size(); i++) {
                pra_rel* R = S->rels[i];
                if (R->full_read || R->merge_full || !R->ver[PRA_DELTA].n) continue;
                append_rows(c, R, PRA_FULL, R->ver[PRA_DELTA].d(), R->ver[PRA_DELTA].n, &appends);
                early[i] = 1;
            }
            copy_ranges(c, (c->early_parity ^= 1) ? JOB_EARLY_APPEND2 : JOB_EARLY_APPEND, appends);
        }
        h = publish_wait(c, h, no * SW);
        std::vector<size_t> redo;
        for (size_t oi = 0; oi < no; oi++) {
            if (!cap[oi]) continue;
            std::copy(h + oi * SW, h + (oi + 1) * SW, st2.begin() + oi * SW);
            if (st2[oi * SW + 5] == 1 || st2[oi * SW + 5] == 3) redo.push_back(oi);
        }
        if (redo.empty()) break;
        if (pass == 2) throw PraError(PRA_E_HIP, "candidate capacity overflow after resizing");
        for (size_t oi : redo) {
            if (st2[oi * SW + 5] == 3) {
                // claims that stopped at the miss left slots behind
                pra_rel* R = outs[oi];
                pra_launch_dense_table_clear(c->st, R->dtab.as<u64>(), R->dtab_cap);
                check_launch("dense_table_clear");
                dense_off[oi] = 1;
            }
            cap[oi] = st2[oi * SW + 4];
            run_output(oi, nullptr);
        }
    }
    // outputs with a join call over the local-join threshold: emitted in
    // rounds (src/RAM/RA_tasks.cpp:1121-1269, src/RA/parallel_join.cpp:157-230)
    std::vector<char> in_rounds(no, 0);
    for (size_t oi = 0; oi < no; oi++)
        if (cap[oi] && st2[oi * SW + 5] == 2) {
            st2[oi * SW] = run_rounds(S, outs[oi], [&](const RoundSpec* rs) { run_output(oi, rs); },
                                      c->stage2.as<u64>() + oi * SW);
            st2[oi * SW + 5] = 0;
            in_rounds[oi] = 1;
        }
    for (auto& pc : c->per_candidate)
        if (std::get<0>(pc) < c->pending.size())
            c->pending[std::get<0>(pc)].bytes += std::get<1>(pc) * (double)st2[std::get<2>(pc) * SW + 4];
    c->per_candidate.clear();
    // tuned launches: cost per candidate; a launch whose candidates passed
    // its capacity did part of the work and is not counted
    for (auto& t : c->tune_pending) {
        const u64 n = st2[t.oi * SW + 4];
        float ms = 0;
        if (n > 0 && n <= t.cap && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess &&
            g_tune.record(c->device, t.knob, t.band, t.var, (double)ms / (double)n, kTuneVariants[t.knob]))
            tune_save();
        c->evpool.push_back(t.a);
        c->evpool.push_back(t.b);
    }
    c->tune_pending.clear();
    for (size_t oi = 0; oi < no; oi++) {
        pra_rel* R = outs[oi];
        u64 cnt = cap[oi] ? st2[oi * SW] : 0;
        R->jhint = cap[oi] ? st2[oi * SW + 4] : 0;
        R->ver[PRA_NEWT].n = cnt;
        R->ver[PRA_NEWT].sorted = false;
        if (!in_rounds[oi]) R->mt_entries += cnt;   // rounds count their own
    }

    // 6. delta -> full, sort newt into trie order, newt -> delta
    //    (src/RAM/RA_tasks.cpp:1044-1058).  The old delta is folded into full
    //    first: the newt sort writes the next delta's bucket bounds.
    {
        std::vector<CopyJob> appends;
        for (size_t i = 0; i < S->rels.size(); i++) {
            pra_rel* R = S->rels[i];
            Rows& D = R->ver[PRA_DELTA];
            if (R->merge_full)
                merge_into_full(c, R);
            else if (!early[i])
                append_rows(c, R, PRA_FULL, D.d(), D.n, &appends);
        }
        copy_ranges(c, JOB_APPEND, appends);
        for (pra_rel* R : S->rels) {
            R->ver[PRA_DELTA].n = 0;
            R->touched(PRA_DELTA);
        }
    }
    std::vector<char> bnd_ok(no, 0);
    // one-block sorts of all outputs go out together, into the emptied delta
    // buffers (swapped back with newt's below)
    SmallSortBatch ssb;
    memset(&ssb, 0, sizeof(ssb));
    ssb.nb = (u64)nb;
    std::vector<size_t> ss_out;
    // several outputs sorting by packed key: their bin sorts go out as one
    // launch, each into its emptied delta buffer; one sorts at once
    std::vector<BinSortJob> bin_jobs;
    std::vector<size_t> bin_out;
    u64 bins_max = 0;   // the largest bin of the deferred bin sorts
    size_t n_packed_sorts = 0;
    for (size_t oi = 0; oi < no; oi++) n_packed_sorts += packed[oi] && !in_rounds[oi] && st2[oi * SW + 3] == 0;
    auto flush_small_sorts = [&]() {
        if (ss_out.empty()) return;
        double bytes = 0;
        for (size_t k = 0; k < ss_out.size(); k++) {
            const SmallSortJob& j = ssb.job[k];
            bytes += 8.0 * j.n * (j.kc + j.bucketed) + 16.0 * j.n * j.w;
        }
        {
            Prof p(c, K_SMALL_SORT, bytes);
            pra_launch_small_sort_many(c->st, ssb, (int)ss_out.size());
            check_launch("small_sort_many");
        }
        for (size_t oi2 : ss_out) outs[oi2]->ver[PRA_NEWT].buf.swap(outs[oi2]->ver[PRA_DELTA].buf);
        ss_out.clear();
    };
    for (size_t oi = 0; oi < no; oi++) {
        pra_rel* R = outs[oi];
        Rows& NT = R->ver[PRA_NEWT];
        R->bnd[PRA_DELTA].reserve((size_t)(nb + 1) * 8, c->st);
        u64* bnd = R->bnd[PRA_DELTA].as<u64>();
        if (packed[oi] && !in_rounds[oi] && st2[oi * SW + 3] == 0) {
            BinSortJob job;
            if (sort_by_packed_key(c, R, st2[oi * SW + 1], st2[oi * SW + 2], st2[oi * SW + 6], pk_mul[oi], pk_ls[oi],
                                   bnd, n_packed_sorts > 1 ? &job : nullptr, &R->ver[PRA_DELTA].buf)) {
                bin_jobs.push_back(job);
                bin_out.push_back(oi);
                bins_max = std::max<u64>(bins_max, st2[oi * SW + 6]);
            }
            bnd_ok[oi] = 1;
            continue;
        }
        // packed mode collects no column masks (a dictionary miss sorts every
        // digit), and a round's masks cover that round only
        const u64* masks = (packed[oi] || in_rounds[oi]) ? nullptr : &st2[oi * SW + 8];
        Group& G = grp(c, R->group);
        if (packed[oi] && !in_rounds[oi] && G.dict_state == 1) G.dict_state = ++G.dict_rebuilds > 3 ? 2 : 0;
        SmallSortJob job;
        bnd_ok[oi] = sort_rows_masked(c, NT, R->w, R->jcc, out_intern[oi] ? R->arity : R->w, masks, bnd, true, &job,
                                      &R->ver[PRA_DELTA].buf);
        if (job.n) {
            ssb.job[ss_out.size()] = job;
            ss_out.push_back(oi);
            if (ss_out.size() == PRA_SORT_JOBS) flush_small_sorts();
        }
        if (!bnd_ok[oi] && nb == 1) {
            pra_launch_fill_u64(c->st, bnd, 1, 0);
            pra_launch_fill_u64(c->st, bnd + 1, 1, NT.n);
            bnd_ok[oi] = 1;
        }
    }
    flush_small_sorts();
    // the bin sorts of all packed outputs, into the emptied delta buffers
    if (!bin_jobs.empty()) {
        double bytes = 0;
        for (const BinSortJob& j : bin_jobs) bytes += 12.0 * j.n + 16.0 * j.n * j.w;
        Prof p(c, K_BIN_SORT, bytes);
        if (bin_jobs.size() == 1) {
            const BinSortJob& j = bin_jobs[0];
            pra_launch_bin_sort(c->st, j.bin_keys, j.bin_idx, j.bin_cap, j.n, j.bofs, j.bytes, j.rows, j.w, j.out,
                                j.bmul, j.ls, j.nb, j.bnd, c->dflags(), bins_max);
        } else {
            const char* d = stage_table(c, JOB_SORT, bin_jobs.data(), bin_jobs.size() * sizeof(BinSortJob));
            pra_launch_bin_sort_many(c->st, (const BinSortJob*)d, (int)bin_jobs.size(), c->dflags(), bins_max);
        }
        check_launch("bin_sort");
        for (size_t oi2 : bin_out) outs[oi2]->ver[PRA_NEWT].buf.swap(outs[oi2]->ver[PRA_DELTA].buf);
    }
    for (pra_rel* R : S->rels) {
        Rows& D = R->ver[PRA_DELTA];
        D.buf.swap(R->ver[PRA_NEWT].buf);
        D.n = R->ver[PRA_NEWT].n;
        D.sorted = true;
        R->ver[PRA_NEWT].n = 0;
        R->touched(PRA_DELTA);
        auto it = std::find(outs.begin(), outs.end(), R);
        if (it != outs.end() && bnd_ok[it - outs.begin()]) R->bnd_valid[PRA_DELTA] = true;
    }
    for (pra_scc* m : members) grp(c, m->group).iterations++;
}

static u64 delta_sum(pra_scc* S)
{
    u64 d = 0;
    for (pra_rel* R : S->rels) d += R->ver[PRA_DELTA].n;
    return d;
}

// full -> delta at SCC start (src/relation/balanced_hash_relation.cpp:1152-1179)
static void full_to_delta(pra_ctx* c, pra_rel* R)
{
    Rows& F = R->ver[PRA_FULL];
    Rows& D = R->ver[PRA_DELTA];
    if (F.n == 0) return;
    if (D.n == 0) {
        D.buf.swap(F.buf);
        D.n = F.n;
        D.sorted = F.sorted;
        F.n = 0;
        F.sorted = true;
        R->touched(PRA_FULL);
        R->touched(PRA_DELTA);
        ensure_sorted(c, R, PRA_DELTA);
        return;
    }
    append_rows(c, R, PRA_DELTA, F.d(), F.n);
    F.n = 0;
    F.sorted = true;
    R->touched(PRA_FULL);
    sort_rows(c, D, R->w, R->jcc, R->w);
    unique_rows(c, D, R->w);
    R->touched(PRA_DELTA);
}

// delta -> full (src/relation/balanced_hash_relation.cpp:1120-1148)
static void delta_to_full(pra_ctx* c, pra_rel* R)
{
    Rows& D = R->ver[PRA_DELTA];
    if (D.n == 0) return;
    append_rows(c, R, PRA_FULL, D.d(), D.n);
    D.n = 0;
    R->touched(PRA_DELTA);
}

// SCC prologue (src/lie/lie.cpp:286-302): note which relations the rules
// produce and which full versions they read in order, then move full to delta
// for the relations added with status true (skipped when a restart loaded the
// deltas from a checkpoint instead)
static void scc_begin(pra_scc* S, bool move_status)
{
    pra_ctx* c = S->ctx;
    if (S->done || S->running) throw PraError(PRA_E_STATE, "SCC already started");
    for (pra_rel* R : S->rels) R->scc_changing = R->merge_full = R->full_read = false;
    for (auto& ra : S->ras) ra.out->scc_changing = true;
    for (auto& ra : S->ras) {
        if (ra.v0 == PRA_FULL) ra.in0->full_read = true;
        if (ra.type == RA_JOIN && ra.v1 == PRA_FULL) ra.in1->full_read = true;
    }
    // a view changes with its relation, and reading its full reads the relation's
    for (auto& ra : S->ras)
        if (ra.type == RA_JOIN && ra.in1->view_of) {
            ra.in1->scc_changing = ra.in1->view_of->scc_changing;
            if (ra.v1 == PRA_FULL) ra.in1->view_of->full_read = true;
        }
    for (auto& ra : S->ras)
        for (pra_rel* R : {ra.in0, ra.in1, ra.out})
            if (R) R->merge_full = R->full_read && R->scc_changing;
    for (auto& ra : S->ras)
        if (ra.type == RA_JOIN && ra.in1->view_of) {
            pra_rel* src = ra.in1->view_of;
            src->merge_full = src->full_read && src->scc_changing;
            ra.in1->merge_full = false;
        }
    if (move_status)
        for (size_t i = 0; i < S->rels.size(); i++)
            if (S->status[i]) full_to_delta(c, S->rels[i]);
    S->running = true;
}

// one batch: a single iteration for an SCC that runs once, else batch_size
// iterations; records and returns the global delta after it
// (RAM::execute_in_batches, src/RAM/RA_tasks.cpp:1121-1269)
static u64 scc_step(pra_scc* S)
{
    pra_ctx* c = S->ctx;
    if (!S->running) throw PraError(PRA_E_STATE, "SCC not started");
    u32 iters = S->fixpoint ? c->batch_size : 1;
    for (u32 b = 0; b < iters; b++) run_iteration(c, {S});
    u64 d = delta_sum(S);
    S->hist.push_back(d);
    return d;
}

// SCC epilogue: delta folded into full for every relation of the SCC (then
// the error flags checked, unless the caller checks them for several SCCs)
static void scc_end(pra_scc* S, bool check = true)
{
    pra_ctx* c = S->ctx;
    if (!S->running) throw PraError(PRA_E_STATE, "SCC not started");
    for (pra_rel* R : S->rels) {
        delta_to_full(c, R);
        R->scc_changing = R->merge_full = R->full_read = false;
    }
    if (check) sync_check(c);
    S->running = false;
    S->done = true;
}

static void scc_run(pra_scc* S)
{
    scc_begin(S, true);
    if (!S->fixpoint)
        scc_step(S);
    else
        while (scc_step(S) != 0) {
        }
    scc_end(S);
}

// first runnable SCC (of instance g, or of any instance when g < 0) in
// creation order (src/lie/lie.cpp:27-60)
static pra_scc* one_runnable(pra_ctx* c, int g = -1)
{
    for (pra_scc* s : c->sccs) {
        if (s->done || (g >= 0 && s->group != g)) continue;
        bool blocked = false;
        for (auto& d : c->deps)
            if (d.second == s && !d.first->done) blocked = true;
        if (!blocked) return s;
    }
    return nullptr;
}

// One lock-step iteration over every instance's LIE::execute
// (src/lie/lie.cpp:197-448): an instance without an SCC in progress starts
// its next runnable one; then one merged iteration runs over the SCCs in
// progress, and each SCC records its batch and ends where its own LIE would:
// after its one iteration, or after a batch that leaves its delta at zero.
// Returns the instances with work left.
static u64 lockstep_step(pra_ctx* c)
{
    std::vector<pra_scc*> members;
    for (size_t g = 0; g < c->groups.size(); g++) {
        Group& G = *c->groups[g];
        if (!G.running) {
            pra_scc* s = one_runnable(c, (int)g);
            if (!s) continue;
            scc_begin(s, true);
            G.running = s;
            G.batch_it = 0;
        }
        members.push_back(G.running);
    }
    if (members.empty()) return 0;
    run_iteration(c, members);
    u64 left = 0;
    for (pra_scc* s : members) {
        Group& G = grp(c, s->group);
        if (!s->fixpoint || ++G.batch_it == c->batch_size) {
            G.batch_it = 0;
            const u64 d = delta_sum(s);
            s->hist.push_back(d);
            if (!s->fixpoint || d == 0) {
                scc_end(s, false);
                G.running = nullptr;
            }
        }
    }
    for (size_t g = 0; g < c->groups.size(); g++)
        if (c->groups[g]->running || one_runnable(c, (int)g)) left++;
    return left;
}

static void execute_lockstep(pra_ctx* c)
{
    while (lockstep_step(c)) {
    }
    sync_check(c);
}

static int fail(pra_ctx* c, int code, const std::string& m)
{
    if (c) c->err = m;
    return code;
}

#define API_BEGIN(ctxexpr)                                                                         \
    pra_ctx* ctx_ = (ctxexpr);                                                                     \
    try {
#define API_END                                                                                    \
    }                                                                                              \
    catch (const PraError& e) { return fail(ctx_, e.code, e.what()); }                           \
    catch (const std::bad_alloc&) { return fail(ctx_, PRA_E_HIP, "host out of memory"); }        \
    catch (const std::exception& e) { return fail(ctx_, PRA_E_HIP, e.what()); }                  \
    return PRA_OK;

static bool valid_ver(int v) { return v == PRA_DELTA || v == PRA_FULL || v == PRA_NEWT; }

} // namespace

// ----------------------------------------------------------------- C ABI ---

extern "C" {

int pra_ctx_create(int device, int nbuckets, uint64_t hbm_budget, pra_ctx** out)
{
    if (!out || nbuckets < 1 || nbuckets > 256) return PRA_E_ARG;
    *out = nullptr;
    std::unique_ptr<pra_ctx> c(new pra_ctx());
    try {
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) return PRA_E_ARG;
        HIPCHK(hipSetDevice(device));
        HIPCHK(pra_partition_scratch_init(device));
        // freed stream-ordered allocations stay in the device's pool for reuse
        // (instead of going back to the driver at every synchronization) up to
        // a quarter of the device memory, so that many contexts in a row do
        // not hoard fragments until allocations fail
        hipMemPool_t pool = nullptr;
        size_t free_b = 0, total_b = 0;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool &&
            hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            uint64_t keep = total_b / 4;
            const char* env = getenv("PRA_POOL_RELEASE_GB");   // bytes the pool keeps, in GiB
            if (env && *env) keep = (uint64_t)(atof(env) * (double)(1ULL << 30));
            HIPCHK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
        }
        HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        {
            std::lock_guard<std::mutex> lk(g_acct_mu);
            g_acct[c->st].budget = hbm_budget;
        }
        c->device = device;
        c->nb = nbuckets;
        c->budget = hbm_budget;
        grp(c.get(), 0);
        const char* dense_env = getenv("PRA_DENSE_MEMBERSHIP");   // "0": the general hash table for every output
        c->dense_on = !(dense_env && dense_env[0] == '0');
        const char* tune_env = getenv("PRA_AUTOTUNE");            // "0": the fixed defaults of every knob
        c->tune_on = !(tune_env && tune_env[0] == '0');
        tune_init();
        const char* seg_env = getenv("PRA_SEG_SORT");             // "0": byte passes for newts over the bins
        c->seg_sort_on = !(seg_env && seg_env[0] == '0');
        const char* cap_env = getenv("PRA_SORT_TEST_BIN_CAP");
        c->bin_cap_test = cap_env ? strtoull(cap_env, nullptr, 10) : 0;
        const char* eb_env = getenv("PRA_DENSE_EPOCH_BITS");
        if (eb_env) c->dense_epoch_bits = std::max(0, std::min(16, atoi(eb_env)));
        c->flags_buf.reserve(64, c->st);
        pra_launch_fill_u64(c->st, c->flags_buf.as<u64>(), 1, 0);
        c->rep_words.reserve((size_t)PRA_NREP * PRA_PART_WORDS * 8, c->st);
        pra_launch_newt_rep_init(c->st, c->rep_words.as<u64>());
        c->rep_sets = 1;
        check_launch("init");
        HIPCHK(hipStreamSynchronize(c->st));
    } catch (const PraError& e) {
        return e.code;
    }
    *out = c.release();
    return PRA_OK;
}

int pra_ctx_destroy(pra_ctx* ctx)
{
    if (!ctx) return PRA_E_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    for (auto* s : ctx->sccs) delete s;
    for (auto* r : ctx->rels) delete r;
    ctx->sccs.clear();
    ctx->rels.clear();
    delete ctx;
    return PRA_OK;
}

const char* pra_last_error(pra_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pra_ctx_set_batch_size(pra_ctx* ctx, uint32_t bs)
{
    if (!ctx || bs == 0) return PRA_E_ARG;
    ctx->batch_size = bs;
    return PRA_OK;
}

int pra_ctx_set_join_threshold(pra_ctx* ctx, int64_t t)
{
    if (!ctx || t < 0) return PRA_E_ARG;
    ctx->threshold = t;
    return PRA_OK;
}

int pra_ctx_joined_tuples(pra_ctx* ctx, uint64_t* out)
{
    if (!ctx || !out) return PRA_E_ARG;
    API_BEGIN(ctx)
    u64 t = 0;
    for (auto& g : ctx->groups)
        if (g->joined.p) t += readback(ctx, g->joined.p, 1)[0];
    *out = t;
    API_END
}

int pra_ctx_iterations(pra_ctx* ctx, uint64_t* out)
{
    if (!ctx || !out) return PRA_E_ARG;
    u64 t = 0;
    for (auto& g : ctx->groups) t += g->iterations;
    *out = t;
    return PRA_OK;
}

int pra_ctx_set_profiling(pra_ctx* ctx, int on)
{
    if (!ctx) return PRA_E_ARG;
    API_BEGIN(ctx)
    ctx->resolve_prof();
    ctx->prof_mask = (u32)on;
    if (on)
        for (auto& k : ctx->kst) { k.launches = 0; k.ms = 0; k.bytes = 0; }
    API_END
}

int pra_ctx_kernel_stat(pra_ctx* ctx, int idx, const char** name, uint64_t* launches, double* ms, double* bytes)
{
    if (!ctx || idx < 0 || idx >= K_COUNT) return PRA_E_ARG;
    ctx->resolve_prof();
    if (name) *name = ctx->kst[idx].name;
    if (launches) *launches = ctx->kst[idx].launches;
    if (ms) *ms = ctx->kst[idx].ms;
    if (bytes) *bytes = ctx->kst[idx].bytes;
    return PRA_OK;
}

static int partition_args(pra_ctx* ctx, const uint64_t* rows, uint64_t n, uint32_t w, uint32_t jcc,
                          uint64_t* out_rows, uint64_t* bucket_bounds)
{
    if (!ctx) return PRA_E_ARG;
    if (w < 1 || w > PRA_MAXW || jcc > w) return fail(ctx, PRA_E_ARG, "row width or join-column count out of range");
    if (n >> 32) return fail(ctx, PRA_E_ARG, "more than 2^32 - 1 rows");
    if (!bucket_bounds || (n && (!rows || !out_rows))) return fail(ctx, PRA_E_ARG, "null buffer");
    return PRA_OK;
}

int pra_partition_rows(pra_ctx* ctx, const uint64_t* rows, uint64_t n, uint32_t w, uint32_t jcc, uint64_t* out_rows,
                       uint64_t* bucket_bounds)
{
    const int rc = partition_args(ctx, rows, n, w, jcc, out_rows, bucket_bounds);
    if (rc != PRA_OK) return rc;
    API_BEGIN(ctx)
    HIPCHK(hipSetDevice(ctx->device));
    ctx->part_scratch.reserve(pra_partition_scratch_words(n, (u32)ctx->nb) * 4 + 256, ctx->st);
    pra_launch_partition(ctx->st, rows, n, (int)w, (int)jcc, (u32)ctx->nb, out_rows, bucket_bounds,
                         ctx->part_scratch.as<u32>());
    check_launch("partition");
    API_END
}

int pra_partition_rows_noalloc(pra_ctx* ctx, const uint64_t* rows, uint64_t n, uint32_t w, uint32_t jcc,
                               uint64_t* out_rows, uint64_t* bucket_bounds)
{
    if (ctx && n > PRA_NOALLOC_MAX_ROWS)
        return fail(ctx, PRA_E_ARG, "allocation-free partition takes at most " + std::to_string(PRA_NOALLOC_MAX_ROWS) +
                                        " rows, got " + std::to_string(n));
    const int rc = partition_args(ctx, rows, n, w, jcc, out_rows, bucket_bounds);
    if (rc != PRA_OK) return rc;
    API_BEGIN(ctx)
    HIPCHK(hipSetDevice(ctx->device));
    u32* scratch = pra_partition_scratch(ctx->device);
    if (!scratch) throw PraError(PRA_E_STATE, "allocation-free scratch of this device was not set up");
    pra_launch_partition(ctx->st, rows, n, (int)w, (int)jcc, (u32)ctx->nb, out_rows, bucket_bounds, scratch);
    check_launch("partition_noalloc");
    API_END
}

int pra_ctx_set_group(pra_ctx* ctx, int group)
{
    if (!ctx || group < 0 || group >= PRA_MAX_GROUPS) return PRA_E_ARG;
    API_BEGIN(ctx)
    grp(ctx, group);
    ctx->cur_group = group;
    API_END
}

int pra_ctx_group_joined(pra_ctx* ctx, int group, uint64_t* out)
{
    if (!ctx || !out || group < 0 || group >= (int)ctx->groups.size()) return PRA_E_ARG;
    API_BEGIN(ctx)
    Group& G = *ctx->groups[(size_t)group];
    *out = G.joined.p ? readback(ctx, G.joined.p, 1)[0] : 0;
    API_END
}

int pra_ctx_step(pra_ctx* ctx, uint64_t* active)
{
    if (!ctx || !active) return PRA_E_ARG;
    API_BEGIN(ctx)
    HIPCHK(hipSetDevice(ctx->device));
    *active = lockstep_step(ctx);
    if (!*active) sync_check(ctx);
    API_END
}

int pra_ctx_group_status(pra_ctx* ctx, int group, int* done, uint64_t* iterations, uint64_t* bytes)
{
    if (!ctx || group < 0 || group >= (int)ctx->groups.size()) return PRA_E_ARG;
    Group& G = *ctx->groups[(size_t)group];
    if (done) *done = !G.running && !one_runnable(ctx, group);
    if (iterations) *iterations = G.iterations;
    if (bytes) {
        u64 b = 0;
        for (pra_rel* R : ctx->rels) {
            if (R->group != group) continue;
            for (auto& v : R->ver) b += v.buf.cap;
            for (const DBuf* d : {&R->mt, &R->bm, &R->dtab, &R->idx[0], &R->idx[1], &R->nkeys, &R->kgofs,
                                  &R->bin_keys, &R->bin_idx, &R->bin_cur, &R->bnd[0], &R->bnd[1]})
                b += d->cap;
        }
        *bytes = b;
    }
    return PRA_OK;
}

int pra_ctx_release_group(pra_ctx* ctx, int group)
{
    if (!ctx || group < 0 || group >= (int)ctx->groups.size()) return PRA_E_ARG;
    API_BEGIN(ctx)
    auto& deps = ctx->deps;
    deps.erase(std::remove_if(deps.begin(), deps.end(),
                              [&](const std::pair<pra_scc*, pra_scc*>& d) { return d.first->group == group; }),
               deps.end());
    std::vector<pra_scc*> keep_s;
    for (pra_scc* s : ctx->sccs)
        if (s->group == group)
            delete s;
        else
            keep_s.push_back(s);
    ctx->sccs.swap(keep_s);
    std::vector<pra_rel*> keep_r;
    for (pra_rel* r : ctx->rels)
        if (r->group == group)
            delete r;
        else
            keep_r.push_back(r);
    ctx->rels.swap(keep_r);
    ctx->groups[(size_t)group].reset(new Group());
    API_END
}

int pra_autotune_report(char* buf, uint64_t cap, uint64_t* len)
{
    if (!len) return PRA_E_ARG;
    const std::string r = g_tune.report(kTuneNames, TUNE_KNOBS);
    *len = r.size();
    if (buf && cap) {
        const size_t n = std::min<size_t>(r.size(), cap - 1);
        memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return PRA_OK;
}

int pra_ctx_trim(pra_ctx* ctx)
{
    if (!ctx) return PRA_E_ARG;
    API_BEGIN(ctx)
    HIPCHK(hipSetDevice(ctx->device));
    // written before they are read inside every iteration that uses them;
    // released on the context's stream, so reuse stays ordered behind the
    // work already enqueued
    for (DBuf* b : {&ctx->sort_tmp, &ctx->stage, &ctx->stage2, &ctx->tmp64a, &ctx->tmp64b, &ctx->tmp64c, &ctx->cand,
                    &ctx->dest, &ctx->slot_of, &ctx->keep, &ctx->hist, &ctx->offs, &ctx->ktmp, &ctx->idxa, &ctx->idxb,
                    &ctx->posadj_dev, &ctx->ttab, &ctx->kp, &ctx->lb_status})
        b->release();
    ctx->lb_zeroed = 0;   // look-back words are zeroed again on next use
    API_END
}

int pra_ctx_device_bytes(pra_ctx* ctx, uint64_t* live, uint64_t* peak)
{
    if (!ctx) return PRA_E_ARG;
    std::lock_guard<std::mutex> lk(g_acct_mu);
    auto it = g_acct.find(ctx->st);
    if (live) *live = it == g_acct.end() ? 0 : (uint64_t)it->second.live;
    if (peak) *peak = it == g_acct.end() ? 0 : (uint64_t)it->second.peak;
    return PRA_OK;
}

int pra_ctx_synchronize(pra_ctx* ctx)
{
    if (!ctx) return PRA_E_ARG;
    API_BEGIN(ctx)
    HIPCHK(hipStreamSynchronize(ctx->st));
    stream_idle(ctx->st);
    API_END
}

int pra_rel_create(pra_ctx* ctx, uint32_t arity, uint32_t jcc, uint32_t tag, int canonical, pra_rel** out)
{
    if (!ctx || !out) return PRA_E_ARG;
    if (arity < 1 || arity + 1 > PRA_MAXW || jcc > arity || tag >= (1u << 18))
        return fail(ctx, PRA_E_ARG, "relation shape out of range");
    pra_rel* r = new pra_rel();
    r->ctx = ctx;
    r->arity = (int)arity;
    r->jcc = (int)jcc;
    r->tag = tag;
    r->canonical = canonical;
    r->w = (int)arity + 1;
    r->group = ctx->cur_group;
    ctx->rels.push_back(r);
    *out = r;
    return PRA_OK;
}

int pra_rel_destroy(pra_rel* rel)
{
    if (!rel) return PRA_E_ARG;
    pra_ctx* c = rel->ctx;
    (void)hipStreamSynchronize(c->st);
    std::vector<pra_rel*> keep;
    for (pra_rel* r : c->rels)
        if (r == rel || r->view_of == rel) {
            if (r != rel) delete r;
        } else {
            keep.push_back(r);
        }
    c->rels.swap(keep);
    delete rel;
    return PRA_OK;
}

int pra_rel_load_rows(pra_rel* rel, const uint64_t* rows, uint64_t nrows, int version)
{
    if (!rel || (nrows && !rows) || (version != PRA_FULL && version != PRA_DELTA)) return PRA_E_ARG;
    API_BEGIN(rel->ctx)
    pra_ctx* c = rel->ctx;
    if (nrows) {
        DBuf tmp;
        tmp.reserve(nrows * rel->w * 8, c->st);
        HIPCHK(hipMemcpyAsync(tmp.p, rows, nrows * rel->w * 8, hipMemcpyHostToDevice, c->st));
        append_rows(c, rel, version, tmp.as<u64>(), nrows);
        Rows& r = rel->ver[version];
        sort_rows(c, r, rel->w, rel->jcc, rel->w);
        unique_rows(c, r, rel->w);
        HIPCHK(hipStreamSynchronize(c->st));
        rel->touched(version);
        rel->mt_valid = false;
        rel->bm_valid = false;
    }
    API_END
}

int pra_rel_load_bin(pra_rel* rel, const char* path, int version)
{
    if (!rel || !path) return PRA_E_ARG;
    pra_ctx* c = rel->ctx;
    std::string sp = std::string(path) + ".size";
    FILE* f = fopen(sp.c_str(), "r");
    if (!f) return fail(c, PRA_E_IO, "cannot open " + sp);
    long long rows = -1, cols = -1;
    int got = fscanf(f, "%lld\n%lld", &rows, &cols);
    fclose(f);
    if (got != 2 || rows < 0) return fail(c, PRA_E_IO, "malformed " + sp);
    if (cols != rel->w) return fail(c, PRA_E_IO, "column count in " + sp + " does not match arity+1");
    std::vector<u64> buf((size_t)rows * cols);
    if (rows) {
        FILE* d = fopen(path, "rb");
        if (!d) return fail(c, PRA_E_IO, std::string("cannot open ") + path);
        size_t n = fread(buf.data(), 8, buf.size(), d);
        fclose(d);
        if (n != buf.size()) return fail(c, PRA_E_IO, std::string("short read of ") + path);
    }
    return pra_rel_load_rows(rel, buf.data(), (uint64_t)rows, version);
}

int pra_rel_count(pra_rel* rel, int version, uint64_t* out)
{
    if (!rel || !out || !valid_ver(version)) return PRA_E_ARG;
    *out = rel->ver[version].n;
    return PRA_OK;
}

int pra_rel_device_rows(pra_rel* rel, int version, const uint64_t** rows, uint64_t* nrows)
{
    if (!rel || !rows || !nrows || !valid_ver(version)) return PRA_E_ARG;
    API_BEGIN(rel->ctx)
    pra_ctx* c = rel->ctx;
    if (version != PRA_NEWT) ensure_sorted(c, rel, version);
    HIPCHK(hipStreamSynchronize(c->st));
    *rows = rel->ver[version].d();
    *nrows = rel->ver[version].n;
    API_END
}

int pra_rel_checksum(pra_rel* rel, int version, uint64_t* out)
{
    if (!rel || !out || !valid_ver(version)) return PRA_E_ARG;
    API_BEGIN(rel->ctx)
    pra_ctx* c = rel->ctx;
    if (version != PRA_NEWT) ensure_sorted(c, rel, version);
    Rows& r = rel->ver[version];
    c->orand.reserve(32 * 8, c->st);
    pra_launch_fill_u64(c->st, c->orand.as<u64>(), 1, 0);
    pra_launch_rows_checksum(c->st, r.d(), r.n, rel->w, c->orand.as<u64>());
    check_launch("rows_checksum");
    *out = readback(c, c->orand.p, 1)[0];
    API_END
}

int pra_rel_download(pra_rel* rel, int version, uint64_t* rows, uint64_t cap, uint64_t* nrows)
{
    if (!rel || !nrows || !valid_ver(version)) return PRA_E_ARG;
    API_BEGIN(rel->ctx)
    pra_ctx* c = rel->ctx;
    if (version != PRA_NEWT) ensure_sorted(c, rel, version);
    Rows& r = rel->ver[version];
    *nrows = r.n;
    u64 m = std::min<u64>(cap, r.n);
    if (m && rows) HIPCHK(hipMemcpyAsync(rows, r.d(), m * rel->w * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    API_END
}

int pra_rel_write_bin(pra_rel* rel, const char* path, int version)
{
    if (!rel || !path || !valid_ver(version)) return PRA_E_ARG;
    pra_ctx* c = rel->ctx;
    uint64_t n = 0, got = 0;
    int rc = pra_rel_count(rel, version, &n);
    if (rc != PRA_OK) return rc;
    std::vector<u64> buf((size_t)n * rel->w);
    rc = pra_rel_download(rel, version, buf.data(), n, &got);
    if (rc != PRA_OK) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(c, PRA_E_IO, std::string("cannot create ") + path);
    size_t wrote = n ? fwrite(buf.data(), 8, buf.size(), f) : 0;
    int closed = fclose(f);
    if (wrote != buf.size() || closed != 0) return fail(c, PRA_E_IO, std::string("short write of ") + path);
    std::string sp = std::string(path) + ".size";
    FILE* m = fopen(sp.c_str(), "w");
    if (!m) return fail(c, PRA_E_IO, "cannot create " + sp);
    fprintf(m, "%llu\n%d", (unsigned long long)n, rel->w);
    if (fclose(m) != 0) return fail(c, PRA_E_IO, "short write of " + sp);
    return PRA_OK;
}

int pra_scc_create(pra_ctx* ctx, int fixpoint, int id, pra_scc** out)
{
    if (!ctx || !out) return PRA_E_ARG;
    pra_scc* s = new pra_scc();
    s->ctx = ctx;
    s->fixpoint = fixpoint != 0;
    s->id = id;
    s->group = ctx->cur_group;
    ctx->sccs.push_back(s);
    *out = s;
    return PRA_OK;
}

int pra_scc_destroy(pra_scc* scc)
{
    if (!scc) return PRA_E_ARG;
    pra_ctx* c = scc->ctx;
    c->sccs.erase(std::remove(c->sccs.begin(), c->sccs.end(), scc), c->sccs.end());
    c->deps.erase(std::remove_if(c->deps.begin(), c->deps.end(),
                                 [&](const std::pair<pra_scc*, pra_scc*>& d) {
                                     return d.first == scc || d.second == scc;
                                 }),
                  c->deps.end());
    delete scc;
    return PRA_OK;
}

int pra_scc_add_relation(pra_scc* scc, pra_rel* rel, int move_full_to_delta)
{
    if (!scc || !rel || rel->ctx != scc->ctx || rel->group != scc->group) return PRA_E_ARG;
    if (scc->rels.size() >= 64) return fail(scc->ctx, PRA_E_ARG, "at most 64 relations per SCC");
    scc->rels.push_back(rel);
    scc->status.push_back(move_full_to_delta != 0);
    return PRA_OK;
}

// the relations belong to the SCC's context and instance
static bool same_ctx(pra_scc* s, pra_rel* a, pra_rel* b = nullptr, pra_rel* d = nullptr)
{
    auto ok = [&](pra_rel* r) { return r->ctx == s->ctx && r->group == s->group; };
    return a && ok(a) && (!b || ok(b)) && (!d || ok(d));
}

int pra_scc_add_join(pra_scc* scc, pra_rel* out, pra_rel* in0, int ver0, pra_rel* in1, int ver1, const int* proj,
                     int nproj)
{
    if (!scc || !same_ctx(scc, out, in0, in1) || (nproj && !proj)) return PRA_E_ARG;
    pra_ctx* c = scc->ctx;
    if ((ver0 != PRA_DELTA && ver0 != PRA_FULL) || (ver1 != PRA_DELTA && ver1 != PRA_FULL))
        return fail(c, PRA_E_ARG, "join versions must be DELTA or FULL");
    if (nproj != out->arity) return fail(c, PRA_E_ARG, "projection width must equal the output arity");
    if (in0->jcc > in1->arity) return fail(c, PRA_E_ARG, "input 0's join column count exceeds input 1's arity");
    int span = in0->w + in1->w - in0->jcc;
    for (int i = 0; i < nproj; i++)
        if (proj[i] < 0 || proj[i] >= span) return fail(c, PRA_E_ARG, "projection index out of range");
    RADesc ra;
    ra.type = RA_JOIN;
    ra.out = out;
    ra.in0 = in0;
    ra.v0 = ver0;
    ra.in1 = in1;
    ra.v1 = ver1;
    if (in0->jcc != in1->jcc) {
        // input 1 seen through its view keyed on input 0's join column count
        pra_rel* view = nullptr;
        for (pra_rel* r : c->rels)
            if (r->view_of == in1 && r->jcc == in0->jcc) view = r;
        if (!view) {
            view = new pra_rel();
            view->ctx = c;
            view->group = in1->group;
            view->arity = in1->arity;
            view->jcc = in0->jcc;
            view->tag = in1->tag;
            view->canonical = in1->canonical;
            view->w = in1->w;
            view->view_of = in1;
            c->rels.push_back(view);
        }
        ra.in1 = view;
    }
    ra.map.assign(proj, proj + nproj);
    scc->ras.push_back(std::move(ra));
    return PRA_OK;
}

static int add_copylike(pra_scc* scc, int type, pra_rel* out, pra_rel* in, int ver, const int* map, int n, int op)
{
    if (!scc || !same_ctx(scc, out, in) || (n && !map)) return PRA_E_ARG;
    pra_ctx* c = scc->ctx;
    if (ver != PRA_DELTA && ver != PRA_FULL) return fail(c, PRA_E_ARG, "input version must be DELTA or FULL");
    int want = type == RA_ACOPY ? out->w : out->arity;
    if (n != want) return fail(c, PRA_E_ARG, "reorder width does not match the output relation");
    for (int i = 0; i < n; i++)
        if (map[i] < 0 || map[i] >= in->w) return fail(c, PRA_E_ARG, "reorder index out of range");
    if (type == RA_COPY_FILTER && (op < PRA_FILTER_EQ || op > PRA_FILTER_FALSE))
        return fail(c, PRA_E_ARG, "unknown filter predicate");
    if (type == RA_COPY_FILTER && in->w < 2) return fail(c, PRA_E_ARG, "filter reads two input columns");
    RADesc ra;
    ra.type = type;
    ra.out = out;
    ra.in0 = in;
    ra.v0 = ver;
    ra.op = op;
    ra.map.assign(map, map + n);
    scc->ras.push_back(std::move(ra));
    return PRA_OK;
}

int pra_scc_add_copy(pra_scc* scc, pra_rel* out, pra_rel* in, int ver, const int* reorder, int n)
{
    return add_copylike(scc, RA_COPY, out, in, ver, reorder, n, PRA_F_TRUE);
}

int pra_scc_add_acopy(pra_scc* scc, pra_rel* out, pra_rel* in, int ver, const int* reorder, int n)
{
    return add_copylike(scc, RA_ACOPY, out, in, ver, reorder, n, PRA_F_TRUE);
}

int pra_scc_add_copy_filter(pra_scc* scc, pra_rel* out, pra_rel* in, int ver, const int* reorder, int n, int op)
{
    return add_copylike(scc, RA_COPY_FILTER, out, in, ver, reorder, n, op);
}

int pra_scc_add_dependency(pra_scc* before, pra_scc* after)
{
    if (!before || !after || before->ctx != after->ctx || before->group != after->group) return PRA_E_ARG;
    before->ctx->deps.push_back({before, after});
    return PRA_OK;
}

int pra_scc_run(pra_scc* scc, uint64_t* delta_hist, uint64_t cap, uint64_t* nbatches)
{
    if (!scc) return PRA_E_ARG;
    API_BEGIN(scc->ctx)
    HIPCHK(hipSetDevice(scc->ctx->device));
    size_t before = scc->hist.size();
    scc_run(scc);
    size_t nb = scc->hist.size() - before;
    if (nbatches) *nbatches = nb;
    for (size_t i = 0; delta_hist && i < nb && i < cap; i++) delta_hist[i] = scc->hist[before + i];
    API_END
}

int pra_scc_begin(pra_scc* scc, int move_full_to_delta)
{
    if (!scc) return PRA_E_ARG;
    API_BEGIN(scc->ctx)
    HIPCHK(hipSetDevice(scc->ctx->device));
    scc_begin(scc, move_full_to_delta != 0);
    API_END
}

int pra_scc_step(pra_scc* scc, uint64_t* delta)
{
    if (!scc) return PRA_E_ARG;
    API_BEGIN(scc->ctx)
    HIPCHK(hipSetDevice(scc->ctx->device));
    u64 d = scc_step(scc);
    if (delta) *delta = d;
    API_END
}

int pra_scc_end(pra_scc* scc)
{
    if (!scc) return PRA_E_ARG;
    API_BEGIN(scc->ctx)
    HIPCHK(hipSetDevice(scc->ctx->device));
    scc_end(scc);
    API_END
}

int pra_scc_history(pra_scc* scc, uint64_t* out, uint64_t cap, uint64_t* n)
{
    if (!scc || !n) return PRA_E_ARG;
    *n = scc->hist.size();
    for (size_t i = 0; out && i < scc->hist.size() && i < cap; i++) out[i] = scc->hist[i];
    return PRA_OK;
}

int pra_execute(pra_ctx* ctx)
{
    if (!ctx) return PRA_E_ARG;
    API_BEGIN(ctx)
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->groups.size() > 1) {
        execute_lockstep(ctx);
    } else {
        pra_scc* s;
        while ((s = one_runnable(ctx)) != nullptr) scc_run(s);
    }
    API_END
}

int pra_execute_many(pra_ctx** ctxs, int n, int concurrency, int* rc)
{
    if (!ctxs || n < 0 || concurrency < 1) return PRA_E_ARG;
    for (int i = 0; i < n; i++)
        if (!ctxs[i]) return PRA_E_ARG;
    std::vector<int> status(n, PRA_OK);
    std::atomic<int> next(0);
    auto worker = [&]() {
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= n) return;
            status[i] = pra_execute(ctxs[i]);
        }
    };
    int nt = std::min(concurrency, std::max(n, 1));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    int first = PRA_OK;
    for (int i = 0; i < n; i++) {
        if (rc) rc[i] = status[i];
        if (first == PRA_OK && status[i] != PRA_OK) first = status[i];
    }
    return first;
}

} // extern "C"
