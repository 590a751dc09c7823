// Launch interface of the HIP kernels (pra_kernels.hip) used by the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pra_device.hpp"

// copy_filter predicates over columns 0 and 1 of the input row
#define PRA_F_EQ 0
#define PRA_F_NEQ 1
#define PRA_F_LT 2
#define PRA_F_LE 3
#define PRA_F_GT 4
#define PRA_F_GE 5
#define PRA_F_TRUE 6
#define PRA_F_FALSE 7

struct JoinExpandArgs {
    const u64* ctl;      // [0] candidates of the output, [1] capacity overflow
    const u64* outer;
    const u64* inner;
    const u64* off;      // n_outer + 1 exclusive offsets
    const u64* istart;   // first matching inner row per outer row
    const u64* posadj;   // per source bucket: arrival base minus source offset
    u64* cand;
    u32* dest;
    u64 n_outer;
    u64 total;
    u64 nb;
    int wo, wi, jcc, left, cw, out_jcc;
    int map[PRA_MAXW];
};

struct CopyExpandArgs {
    const u64* ctl;
    const u64* rows;
    const u64* off;
    const u64* posadj;
    u64* cand;
    u32* dest;
    u64 n;
    u64 nb;
    int w, in_jcc, cw, out_jcc;
    int map[PRA_MAXW];
};

// One rule's per-row output counts and their exclusive prefix sums
struct CountScanArgs {
    const u64* outer;    // outer rows (wo words each)
    const u64* table;    // JOIN: the static inner's key index, cap slots of [hash, first row, run length]
    const u64* inner;    // JOIN: inner rows (wi words each)
    u64* off;            // n + 1: exclusive prefix of the counts, off[n] = total
    u64* istart;         // JOIN: first inner row of each outer row's key run
    const u64* ibnd;     // sorted inner: its bucket bounds
    u64 n, cap, nb;
    int wo, wi, jcc, op;
};

// Several rules' counts and scans in one launch: tiles of all jobs in ticket
// order, job k owning tiles [tile_base[k], tile_base[k + 1]) of the launch
// and the status words of those tiles
#define PRA_CS_FILTER 0   // copy / copy_filter: 1 per row passing the filter
#define PRA_CS_INDEX 1    // join over a static inner's key index
#define PRA_CS_SORTED 2   // join over an inner kept sorted by (bucket, columns)
struct CountScanJob {
    const u64* outer;
    const u64* table;
    const u64* inner;
    const u64* ibnd;     // PRA_CS_SORTED: the inner's bucket bounds (nb + 1)
    u64* off;
    u64* istart;
    u64 n, cap, nb;
    int wo, wi, jcc, op, kind;
};
#define PRA_CS_JOBS 36
struct CountScanBatch {
    CountScanJob job[PRA_CS_JOBS];
    u32 tile_base[PRA_CS_JOBS + 1];
    int njobs;
};

#define PRA_MAXRA 64

struct PlanArgs {
    u64* srcoff;         // per RA: nb first offsets and nb counts per source (written here)
    u64* posadj;         // per RA: nb arrival adjustments
    u64* ctl;            // [0] candidates, [1] overflow
    u64* joined;
    u32* flags;
    u64 cap;
    long long thresh;
    int nb, nr;
    int ras[PRA_MAXRA];
    int is_join[PRA_MAXRA];
    const u64* off[PRA_MAXRA];   // per listed RA: exclusive output offsets over its outer rows
    const u64* lo[PRA_MAXRA];    // per listed RA: first outer row emitted from each source (nb)
    const u64* hi[PRA_MAXRA];    // per listed RA: end of the rows emitted from each source (nb)
    int rounds_check;            // mark threshold rounds instead of emitting
    int count_joined;            // add the join pairs to *joined
};

// Arrival plans of every output of one iteration (one block per output;
// PlanArgs's fields per rule and per output, with the threshold-round check
// and the join-pair count on)
struct PlanAllArgs {
    u64* srcoff;
    u64* posadj;
    u64* stage2;         // per output sw words; ctl = stage2 + output * sw + 4
    long long thresh;
    int nb, nra, sw;
    const u64* off[PRA_MAXRA];       // per rule: exclusive output offsets over its outer rows
    const u64* bnd[PRA_MAXRA];       // per rule: first outer row of each source bucket (nb + 1)
    u64 cap[PRA_MAXRA];              // per output: candidate capacity
    u64* joined[PRA_MAXRA];          // per output: its instance's join-pair count
    unsigned char rout[PRA_MAXRA];   // per rule: its output (0xFF: inactive)
    unsigned char is_join[PRA_MAXRA];
    unsigned char outs[PRA_MAXRA];   // per block: the output it plans
};

// k_plan_all for any number of rules and outputs, from tables in device
// memory: block b plans outs[b] over rules rule_ids[first .. first + count)
// (ascending), with srcoff / posadj indexed by the rule ids.
struct PlanRule {
    const u64* off;
    const u64* bnd;
    u32 is_join, pad;
};
struct PlanOut {
    u64 cap;
    u64* joined;
    u32 first, count, oi, pad;
};
struct PlanTabArgs {
    u64* srcoff;
    u64* posadj;
    u64* stage2;
    long long thresh;
    int nb, sw;
    const PlanRule* rules;
    const u32* rule_ids;
    const PlanOut* outs;
};
void pra_launch_plan_tab(hipStream_t st, const PlanTabArgs& a, int nouts);

// Membership claims and decisions of several output relations in one launch
// each (one job per output; blocks [block_base[k], block_base[k + 1]) serve
// job k)
struct ClaimJob {
    const u64* cand;
    const u64* ctl;
    u64* table;
    u64* slot_of;
    u64 cap, epoch;
    int cw, kw;
};
struct DecideJob {
    const u64* cand;
    const u64* ctl;
    const u64* table;
    const u64* slot_of;
    u32* keep;
    const u32* dest;
    u32* hist;
    u32* offs1;          // one tile: bucket offsets and total written directly
    u64 epoch;
    u32 nblocks;
    int cw, kw;
};
// Dense membership of an interned output whose key columns all have
// dictionary ranks: the key's rank index x = ((r0 * D) + r1) * D + ..
// (D = dictionary size) selects one bit of a bitmap over D^kw bits that holds
// the keys of full u delta (and of newt, set as tuples are kept).  Within an
// iteration, candidates that are not members claim a slot of a small table
// keyed exactly by x + 1, one word per slot: epoch | (x + 1) <<
// PRA_DENSE_PBITS | earliest arrival; the earliest arrival keeps the tuple and
// sets its bit.  A slot is live in this iteration when its top bits (emask)
// hold this iteration's epoch (stamp) and it is not 0; any other slot counts
// as empty, so a new epoch empties the table without touching it.  When x + 1
// leaves no bits for an epoch (emask 0), the winner writes its slot back to
// 0 instead.  Needs D^kw <= 2^35 (x + 1 in 36 bits) and fewer than 2^28
// candidates.  A value missing from the dictionary sets ctl[1] = 3 and the
// output is redone the general way.
#define PRA_DENSE_PBITS 28
#define PRA_DENSE_PMASK ((1ULL << PRA_DENSE_PBITS) - 1)
struct DenseArgs {
    const u64* cand;
    u64* ctl;
    u64* slot_of;        // claim: the slot (~0: dropped); decision: the kept candidates' x
    u32* keep;
    const u32* dest;
    u32* hist;
    u32* offs1;          // one tile: bucket offsets and total written directly
    const u64* dict;
    u64 dict_cap, D;
    u32* bitmap;
    u64* table;
    u64 cap;
    u64 emask, stamp;    // the slots' epoch bits and this iteration's epoch in them
    u32 nblocks, nb;
    int cw, kw;
};

// job tables (device memory) of the batched launches of a lock-step iteration
struct BinSortJob {
    const u64* bin_keys;
    const u32* bin_idx;
    const u32* bofs;
    const u64* rows;
    u64* out;
    u64* bnd;
    u64 n, bmul, nb;
    u32 bin_cap, bytes;
    int w, ls;
};
struct ScanJob {
    const u32* in;
    u32* out;
    u32* total;
    u64 n;
};
struct CopyJob {
    const u64* src;
    u64* dst;
    u64 words;
};

#define PRA_DEDUP_JOBS 32
struct ClaimBatch {
    ClaimJob job[PRA_DEDUP_JOBS];
    u32 block_base[PRA_DEDUP_JOBS + 1];
    int njobs;
};
struct DecideBatch {
    DecideJob job[PRA_DEDUP_JOBS];
    u32 block_base[PRA_DEDUP_JOBS + 1];
    int njobs;
    u32 nb;
};

// One small sort of k_small_sort (rows -> out, bucket bounds -> bnd)
struct SmallSortJob {
    const u64* rows;
    u64* out;
    u64* bnd;
    u64 n;
    int w, jcc, kc, bucketed;
};
#define PRA_SORT_JOBS 64
struct SmallSortBatch {
    SmallSortJob job[PRA_SORT_JOBS];
    u64 nb;
};

// Summaries of one newt scatter (k_newt_fold)
struct NewtFoldArgs {
    u32* bin_cur;      // top-byte bin counters of the packed keys (reset here)
    u64* rep_words;
    const u64* ctl;
    u64* stat;         // per-output summary words read back by the host
    u32* kgofs;        // top-byte digit offsets of the packed keys
    u64* seqbase;      // per-bucket seq counters (interned outputs)
    const u32* offs;   // scanned per-tile destination histograms
    const u32* total;  // kept candidates
    u32 nblocks, nb;
    int packed, intern;
};

struct NewtScatterArgs {
    const u64* ctl;
    const u64* dict;   // value -> rank dictionary (null: no packed sort key)
    u64 dict_cap;
    u64* keys;         // packed sort key per written row
    u64* rep_words;    // PRA_NREP x PRA_PART_WORDS summary words (OR / AND)
    u64 rank_radix;    // packed key = bucket, then each column's rank, in this radix
    int key_lshift;    // packed key shifted left so that its top bit is bit 63
    u32 bin_cap;       // (key, row) pairs per top-byte bin
    u64* bin_keys;     // 256 bins of bin_cap packed keys
    u32* bin_idx;      // 256 bins of bin_cap newt positions
    u32* bin_cur;      // 256 bin fill counters (all rows, also those past bin_cap)
    // packed keys of a dense-membership iteration: each kept candidate's rank
    // index (sum of rank_c * D^(cw-1-c), left by the dense decision), so the
    // packed key is bucket * xk_scale + index with xk_scale = D^cw; null:
    // the ranks are looked up in the dictionary
    const u64* xk;
    u64 xk_scale;
    const u64* cand;
    const u32* keep;
    const u32* dest;
    const u32* offs;
    const u64* seqbase;
    u64* newt;
    u64 n;
    u32 nblocks, nb;
    int cw, wn, intern;
    u32 tag;
    int fold_here;       // one tile (PRA_BLOCK threads): the block also does k_newt_fold's work
    int summarize;       // unpacked: collect the key columns' OR / AND (else reported as "varies")
    NewtFoldArgs fold;
};

void pra_launch_fill_u64(hipStream_t st, u64* p, u64 n, u64 v);
void pra_launch_table_clear(hipStream_t st, u64* t, u64 cap, int stride);
// exclusive scans; inputs above PRA_SMALL_SCAN elements take a one-pass
// look-back scan over `status` (PRA_RADIX words per PRA_TILE elements, as the
// sort's) with its own self-resetting tile counter
void pra_scan_u64(hipStream_t st, const u64* in, u64* out, u64 n, u64* status, u32* tile_ctr, u32 epoch, u32* flags,
                  u64* total);
void pra_scan_u32(hipStream_t st, const u32* in, u32* out, u64 n, u64* status, u32* tile_ctr, u32 epoch, u32* flags,
                  u32* total);
u32 pra_tiles(u64 n);
void pra_launch_bucket_bounds(hipStream_t st, const u64* rows, u64 n, int w, int jcc, u64 nb, u64* bstart);
void pra_launch_group_flags(hipStream_t st, const u64* rows, u64 n, int w, int jcc, u64* flags);
void pra_launch_group_compact(hipStream_t st, const u64* flags, const u64* gpos, u64 n, u64* gstart);
void pra_launch_compact_rows(hipStream_t st, const u64* in, u64 n, int w, const u64* flags, const u64* pos, u64* out);
void pra_launch_index_insert(hipStream_t st, const u64* rows, int w, int jcc, const u64* gstart, u64 ng, u64* table,
                             u64 cap, u32* flags);
void pra_launch_join_expand(hipStream_t st, const JoinExpandArgs& a, u64 cap);
u32 pra_join_expand_blocks(u64 cap);   // grid of a join expansion of up to cap outputs
// jobs: device memory holding njobs JoinExpandArgs (join) or CopyExpandArgs,
// then njobs + 1 u32 block bases
void pra_launch_expand_many(hipStream_t st, const void* jobs, int njobs, u32 nblocks, bool join);
void pra_launch_plan(hipStream_t st, const PlanArgs& a);
void pra_launch_plan_all(hipStream_t st, const PlanAllArgs& a, int nblocks);
// host[0..words) = src, host[words] = *flags, then host[-1] = seq once they
// are visible to the host
void pra_launch_publish(hipStream_t st, const u64* src, u64 words, const u32* flags, u64* host, u64 seq);
void pra_launch_find_split(hipStream_t st, const u64* off, u64 n, const u32* kp, long long thresh, u64* out);
void pra_launch_group_maxlen(hipStream_t st, const u64* flags, const u64* gpos, const u64* gstart, u64 n, u64* maxlen);
void pra_launch_count_scan_many(hipStream_t st, const CountScanBatch& b, u64* status, u32* tile_ctr, u32 epoch,
                                u32* flags);
void pra_launch_copy_expand(hipStream_t st, const CopyExpandArgs& a);
// wide: 1 = 1,024-thread blocks, 0 = 256-thread blocks, -1 = wide from
// PRA_SCATTER_WIDE_TILES tiles on
void pra_launch_dedup_scatter(hipStream_t st, const NewtScatterArgs& a, int wide = -1);
// jobs: device memory holding njobs arguments, then njobs + 1 u32 block bases
void pra_launch_dedup_scatter_many(hipStream_t st, const NewtScatterArgs* jobs, int njobs, u32 nblocks, bool packed);
void pra_launch_table_insert_rows(hipStream_t st, const u64* rows, u64 n, int w, int kw, u64* table, u64 cap,
                                  u32* flags);
void pra_launch_table_verify_rows(hipStream_t st, const u64* rows, u64 n, int w, int kw, const u64* table, u64 cap,
                                  u32* flags);
void pra_launch_table_rehash(hipStream_t st, const u64* old, u64 oldcap, u64* nt, u64 cap, int kw, u32* flags);
void pra_launch_sort_hist(hipStream_t st, const u64* rows, u64 n, int w, const PassList& pl, u32* ghist);
void pra_launch_sort_scan(hipStream_t st, const u32* ghist, int np, u32* gofs, u64* bnd, u32 nb, u64 n);
void pra_launch_sort_scatter(hipStream_t st, const u64* in, u64* out, u64 n, int w, DigitSpec ds, const u32* gofs,
                             u64* status, u32* tile_ctr, u32 epoch, u32* flags, const u32* iin, u32* iout);
void pra_launch_collect_cols(hipStream_t st, const u64* rows, u64 n, int w, int ncols, u64* out);
// rows of a join input seen under another join column count: keep[i] = 1 when
// row i's bucket under its first ja columns equals its bucket under its first
// jb columns; then the kept rows compacted in order (pos = exclusive scan of keep)
void pra_launch_view_keep(hipStream_t st, const u64* rows, u64 n, int w, int ja, int jb, u64 nb, u32* keep);
void pra_launch_view_compact(hipStream_t st, const u64* rows, u64 n, int w, const u32* keep, const u32* pos, u64* out);
void pra_launch_dict_insert(hipStream_t st, const u64* vals, u64 n, u64* tab, u64 cap, u32* flags);
void pra_launch_gather_rows(hipStream_t st, const u64* in, const u32* idx, u64 n, int w, u64* out);
void pra_launch_key_bounds(hipStream_t st, const u64* keys, u64 n, u64 bmul, int ls, u64 nb, u64* bnd);
void pra_launch_small_sort(hipStream_t st, const u64* rows, u64 n, int w, int jcc, int kc, u64 nb, int bucketed,
                           u64* out, u64* bnd);
void pra_launch_small_sort_many(hipStream_t st, const SmallSortBatch& b, int njobs);
void pra_launch_block_sort(hipStream_t st, const u64* rows, u64 n, int w, const PassList& pl, u64* out, u64* bnd,
                           int jcc, u64 nb);
// Segmented packed-key sort of a newt too large for the top-byte bins: keys
// (and their newt positions idx) already ordered by their top 16 bits;
// segment v = the keys whose bits from `top` up equal v (nseg segments).
// pra_launch_seg_bounds writes seg[0..nseg] and the largest segment into
// *maxseg (zeroed by the caller); pra_launch_seg_sort sorts every segment in
// LDS by its varying bytes (bit p of `bytes`), gathers the rows into out and
// writes the bucket bounds as pra_launch_bin_sort does.  Segments of at most
// PRA_SEG_SMALL pairs are sorted by 256-thread blocks; the nlarge larger ones
// (their ids in `large`, listed by pra_launch_seg_bounds) by a second launch
// of the narrowest blocks, 512 or PRA_BIN_BLOCK threads, that hold maxseg
// (the largest); a segment over PRA_BIN_CAP pairs raises
// PRA_FLAG_BIN_OVERFLOW.
#define PRA_SEG_SMALL 2048
void pra_launch_seg_bounds(hipStream_t st, const u64* keys, u64 n, int top, u32 nseg, u64* seg, u64* stat, u32* large);
void pra_launch_seg_sort(hipStream_t st, const u64* keys, const u32* idx, const u64* seg, int top, u32 nseg, u64 n,
                         u32 bytes, const u64* rows, int w, u64* out, u64 bmul, int ls, u64 nb, u64* bnd, u32* flags,
                         u64 maxseg, const u32* large, u64 nlarge);
// maxbin: the largest bin (of every job, for the batched form), which picks
// the narrowest block of 256, 512 or PRA_BIN_BLOCK threads that holds it
void pra_launch_bin_sort(hipStream_t st, const u64* bin_keys, const u32* bin_idx, u32 bin_cap, u64 n, const u32* bofs,
                         u32 bytes, const u64* rows, int w, u64* out, u64 bmul, int ls, u64 nb, u64* bnd, u32* flags,
                         u64 maxbin);
void pra_launch_col_orand2(hipStream_t st, const u64* rows, u64 n, int w, int kc, u64* out);
void pra_launch_dedup_claim2(hipStream_t st, const u64* cand, const u64* ctl, u64 capn, int cw, int kw, u64* table,
                             u64 cap, u64* slot_of, u64 epoch, u32* flags);
void pra_launch_dedup_decide_hist(hipStream_t st, const u64* cand, const u64* ctl, int cw, int kw, const u64* table,
                                  const u64* slot_of, u64 epoch, u32* keep, const u32* dest, u32 nb, u32* hist,
                                  u32 nblocks, u32* flags, u32* offs1 = nullptr);
void pra_launch_newt_fold(hipStream_t st, const NewtFoldArgs& a);
// key-width class of a claim / decision launch (jobs of one batch share it)
int pra_dedup_class(int kw);
// dense membership: per-iteration claims and decisions of one output, and the
// bitmap of a version's rows (misses counted into *miss)
void pra_launch_dense_claim(hipStream_t st, const DenseArgs& a, u64 capn, u32* flags);
void pra_launch_dense_decide(hipStream_t st, const DenseArgs& a, u32* flags);
// dense claims, then decisions, of several outputs of key width kw: `jobs`
// (device) holds njobs DenseArgs, then njobs + 1 first claim blocks, then
// njobs + 1 first decision blocks (one per tile)
void pra_launch_dense_many(hipStream_t st, const DenseArgs* jobs, int njobs, int kw, u32 claim_blocks,
                           u32 decide_blocks, u32* flags);
void pra_launch_dense_build(hipStream_t st, const u64* rows, u64 n, int w, int kw, const u64* dict, u64 dict_cap, u64 D,
                            u32* bitmap, u64* miss);
void pra_launch_dense_table_clear(hipStream_t st, u64* table, u64 cap);
// batched forms of k_bin_sort (256 blocks per job), k_scan_small (u32, one
// block per job of at most PRA_SMALL_SCAN elements), D2D copies (job table
// followed by njobs + 1 first blocks; 2,048 words per block) and k_newt_fold
void pra_launch_bin_sort_many(hipStream_t st, const BinSortJob* jobs, int njobs, u32* flags, u64 maxbin);
void pra_launch_scan_small_many(hipStream_t st, const ScanJob* jobs, int njobs);
void pra_launch_copy_many(hipStream_t st, const CopyJob* jobs, int njobs, u32 nblocks);
void pra_launch_newt_fold_many(hipStream_t st, const NewtFoldArgs* jobs, int njobs);
void pra_launch_dedup_claim_many(hipStream_t st, const ClaimBatch& b, int cls, u32* flags);
void pra_launch_dedup_decide_many(hipStream_t st, const DecideBatch& b, int cls, u32* flags);
void pra_launch_newt_fold_publish(hipStream_t st, const NewtFoldArgs& a, const u64* src, u64 words,
                                  const u32* flags, u64* host, u64 seq);
// `sets` consecutive replica sets (one per output of a batched partition)
void pra_launch_newt_rep_init(hipStream_t st, u64* rep_words, int sets = 1);
void pra_launch_rows_checksum(hipStream_t st, const u64* rows, u64 n, int w, u64* out);
// merge of sorted delta into sorted full (out); add_to (nullable): full's
// bucket bounds, which become the merged version's
void pra_launch_merge(hipStream_t st, const u64* F, u64 nf, const u64* fbnd, const u64* D, u64 nd, const u64* dbnd,
                      u64 nb, int w, u64* rank, u64* out, u64* add_to = nullptr);
