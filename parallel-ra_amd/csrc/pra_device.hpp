// Device-side building blocks shared by the semi-naive RA kernels (gfx950).
//
// Tuples are rows of w = arity+1 little-endian u64 words, the last word being
// the intern id.  Every sorted row array is ordered by (bucket, col0, col1, ..),
// bucket = tuple_hash(row[0:jcc]) % nbuckets, which is the order in which the
// reference walks its per-rank nested B-tree tries
// (src/relation/shmap_relation_exp.cpp:77-112) concatenated over ranks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;

#define PRA_BLOCK 256
#define PRA_WAVES (PRA_BLOCK / 64)
#define PRA_ITEMS 8
#define PRA_TILE (PRA_BLOCK * PRA_ITEMS)
// packed-key sort bins: rows per bin sorted in LDS by one block
#define PRA_BIN_BLOCK 1024
// one-block LDS sort of small inputs: rows, key words per row at that many rows
// (n x key words must fit PRA_SMALL_SORT x PRA_SMALL_SORT_KW), threads
#define PRA_SMALL_SORT 2048
#define PRA_SMALL_SORT_KW 8
#define PRA_SMALL_SORT_BLOCK 1024
// one-block LSD sort (all passes in one launch) of up to this many rows
#define PRA_BLOCK_SORT 16384
// threads of the dedup decision and of the newt scatter per PRA_TILE candidates
#define PRA_DECIDE_BLOCK 1024
#define PRA_DECIDE_GRID 512
#define PRA_COUNT_SCAN_BLOCK 1024
#define PRA_SCATTER_BLOCK 1024
#define PRA_SCATTER_WIDE_TILES 64   // tiles from which the newt scatter uses PRA_SCATTER_BLOCK
#define PRA_BIN_ITEMS 8
#define PRA_BIN_CAP (PRA_BIN_BLOCK * PRA_BIN_ITEMS)
#define PRA_RADIX_BITS 8
#define PRA_RADIX (1 << PRA_RADIX_BITS)
#define PRA_MAXW 32

// flags word bits reported back to the host
#define PRA_FLAG_COLLISION 1u
#define PRA_FLAG_TABLE_FULL 2u
#define PRA_FLAG_INDEX_COLLISION 4u
#define PRA_FLAG_SORT_STALL 8u
#define PRA_FLAG_THRESHOLD 16u
#define PRA_FLAG_BIN_OVERFLOW 32u
#define PRA_FLAG_SCAN_OVERFLOW 64u

#define PRA_MAXPASS 40
#define PRA_PART_WORDS 40   // summary words of the newt scatter
#define PRA_SMALL_SCAN 8192  // scans up to this many elements run in one block
#define PRA_NREP 16         // replicas the scatter blocks fold their summaries into
#define PRA_LOOKBACK_SPIN_LIMIT (1u << 22)
#define PRA_LOOKBACK 8       // predecessor tiles read per look-back step

// decoupled look-back status words: [63:48] epoch, [47:46] flag, [45:0] count
#define PRA_ST_AGG 1ULL
#define PRA_ST_INC 2ULL

// summary word j of the newt scatter combines by AND (packed key AND, key
// column ANDs) or by OR (everything else)
__host__ __device__ inline bool pra_summary_is_and(int j) { return j == 1 || (j >= 24 && j < 40); }

// FNV-1a-64 over the 8 little-endian bytes of each word, exactly the
// reference's tuple_hash (src/hash/hash.h:11-30).
#define PRA_FNV_OFFSET 14695981039346656037ULL

// one u64 word (8 little-endian bytes) of FNV-1a-64
__device__ inline u64 pra_fnv_word(u64 h, u64 w)
{
#pragma unroll
    for (int b = 0; b < 8; b++) {
        h ^= (w >> (8 * b)) & 255ULL;
        h *= 1099511628211ULL;
    }
    return h;
}

__device__ inline u64 pra_fnv(const u64* t, int n)
{
    u64 h = PRA_FNV_OFFSET;
    for (int i = 0; i < n; i++) h = pra_fnv_word(h, t[i]);
    return h;
}

__host__ __device__ inline u64 pra_mix64(u64 x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// 64-bit table hash of a key of n words; never 0 (0 marks an empty slot).
__host__ __device__ inline u64 pra_keyhash(const u64* k, int n)
{
    u64 h = 0x9E3779B97F4A7C15ULL ^ (u64)n;
    for (int i = 0; i < n; i++) h = pra_mix64(h ^ (k[i] + 0x632BE59BD9B4E019ULL * (u64)(i + 1)));
    return h ? h : 1;
}

// Copy a row of w words between distinct buffers with every load of a chunk
// of 8 words issued before its stores.
__device__ inline void pra_copy_row(u64* __restrict__ dst, const u64* __restrict__ src, int w)
{
    int c = 0;
    for (; c + 8 <= w; c += 8) {
        u64 t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = src[c + j];
#pragma unroll
        for (int j = 0; j < 8; j++) dst[c + j] = t[j];
    }
    u64 t[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (c + j < w) t[j] = src[c + j];
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (c + j < w) dst[c + j] = t[j];
}

// One radix digit of a row: either bits of a column or bits of the bucket.
struct DigitSpec {
    int col;       // column index, or -1 for the bucket digit
    int shift;     // bit offset of the digit
    int jcc;       // bucket: number of leading columns hashed
    int pad;
    u64 nb;        // bucket count
};

struct PassList {
    int np;
    int pad;
    DigitSpec p[PRA_MAXPASS];
};

__device__ inline u32 pra_digit(const u64* row, const DigitSpec& ds)
{
    u64 v;
    if (ds.col >= 0)
        v = row[ds.col];
    else
        v = pra_fnv(row, ds.jcc) % ds.nb;
    return (u32)((v >> ds.shift) & (PRA_RADIX - 1));
}

__device__ inline u64 pra_lanemask_lt()
{
    u32 lane = __lane_id();
    return lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
}

// Mask of the active lanes holding the same `bits`-bit value as this lane.
template <int BITS>
__device__ inline u64 pra_match(u32 v, u64 active)
{
    u64 peers = active;
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        u64 m = __ballot((v >> b) & 1u);
        peers &= ((v >> b) & 1u) ? m : ~m;
    }
    return peers;
}
