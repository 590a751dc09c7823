// HIP kernels of the semi-naive RA hot path for gfx950 (MI355X).
//
// Work is integer hashing, scanning, partitioning and sorting over rows of
// u64 words; it is HBM-bandwidth bound, so kernels stream rows in tiles of
// PRA_TILE rows per 256-thread workgroup (4 wave64s) and keep per-tile
// histograms in LDS.  Launch wrappers at the bottom are called by the host
// engine (pra_engine.cpp) on the context's stream.
#include "pra_device.hpp"
#include "pra_kernels.hpp"

// ------------------------------------------------------------------ utils --

__global__ void k_fill_u64(u64* p, u64 n, u64 v)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Table slots of `stride` words: [hash, minpos, key...]; empty = hash 0,
// minpos = all ones so that atomicMin by the first claimant wins.
__global__ void k_table_clear(u64* t, u64 cap, int stride)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 step = (u64)gridDim.x * blockDim.x;
    for (; i < cap; i += step) {
        u64* s = t + i * stride;
        s[0] = 0;
        s[1] = ~0ULL;
        for (int c = 2; c < stride; c++) s[c] = 0;
    }
}

template <typename T, int NW = PRA_WAVES>
__device__ inline T block_exclusive_scan(T v, T* lds, T* total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int w = 0; w < NW; w++) {
            T t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[NW] = acc;
    }
    __syncthreads();
    T res = x - v + lds[wave];
    *total = lds[NW];
    __syncthreads();
    return res;
}

// ---------------------------------------------------- one-sweep LSD sort ---
//
// k_sort_hist counts the digits of every pass in one read of the rows;
// k_sort_scan turns them into per-pass digit offsets; one k_sort_scatter per
// pass then finds its tile's offset inside each digit by decoupled look-back
// over the preceding tiles' published counts (tiles are numbered in dispatch
// order, so a tile only ever waits on tiles that are already running).

__global__ void k_sort_hist(const u64* rows, u64 n, int w, PassList pl, u32* ghist)
{
    extern __shared__ u32 sh[];
    const int nbins = pl.np * PRA_RADIX;
    for (int i = threadIdx.x; i < nbins; i += PRA_BLOCK) sh[i] = 0;
    __syncthreads();
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 step = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += step) {
        const u64* r = rows + i * w;
        for (int p = 0; p < pl.np; p++) atomicAdd(&sh[p * PRA_RADIX + pra_digit(r, pl.p[p])], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nbins; j += PRA_BLOCK)
        if (sh[j]) atomicAdd(&ghist[j], sh[j]);
}

// exclusive scan of each pass's digit counts; when `bnd` is given the last
// pass is the bucket digit and its offsets are the bucket bounds
__global__ void k_sort_scan(const u32* ghist, int np, u32* gofs, u64* bnd, u32 nb, u64 n)
{
    __shared__ u32 lds[PRA_WAVES + 1];
    for (int p = 0; p < np; p++) {
        u32 v = ghist[p * PRA_RADIX + threadIdx.x];
        u32 tot;
        u32 ex = block_exclusive_scan<u32>(v, lds, &tot);
        gofs[p * PRA_RADIX + threadIdx.x] = ex;
        if (bnd && p == np - 1) {
            if (threadIdx.x < nb) bnd[threadIdx.x] = ex;
            if (threadIdx.x == 0) bnd[nb] = n;
        }
    }
}

static_assert(PRA_BLOCK == PRA_RADIX, "look-back assigns one digit per thread");

__device__ inline u64 st_pack(u32 epoch, u64 flag, u64 cnt) { return ((u64)epoch << 48) | (flag << 46) | cnt; }

// Inclusive prefix of digit d over the tiles before `tile`: walks back over
// the predecessors' published counts, PRA_LOOKBACK predecessors per step with
// all their loads in flight, until an inclusive prefix is found.
__device__ inline u64 pra_lookback(const u64* status, u32 tile, int d, u32 epoch, u32* flags)
{
    u64 excl = 0;
    long long t = (long long)tile - 1;
    u32 spins = 0;
    while (t >= 0) {
        u64 v[PRA_LOOKBACK];
#pragma unroll
        for (int j = 0; j < PRA_LOOKBACK; j++)
            v[j] = t - j >= 0 ? __hip_atomic_load(status + (u64)(t - j) * PRA_RADIX + d, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : 0;
        int used = 0;
        bool done = false, stall = false;
#pragma unroll
        for (int j = 0; j < PRA_LOOKBACK; j++) {
            if (done || stall || t - j < 0) continue;
            u64 fl = (v[j] >> 46) & 3ULL;
            if ((u32)(v[j] >> 48) != epoch || fl == 0) {
                stall = true;
                continue;
            }
            excl += v[j] & ((1ULL << 46) - 1);
            used = j + 1;
            if (fl == PRA_ST_INC) done = true;
        }
        if (done) break;
        t -= used;
        if (stall && ++spins > PRA_LOOKBACK_SPIN_LIMIT) {
            atomicOr(flags, PRA_FLAG_SORT_STALL);
            break;
        }
    }
    return excl;
}

template <bool IDX, int ITEMS>
__global__ void k_sort_scatter(const u64* in, u64* out, u64 n, int w, DigitSpec ds, const u32* gofs, u64* status,
                               u32* tile_ctr, u32 epoch, u32* flags, const u32* iin, u32* iout)
{
    constexpr u64 TILE = (u64)PRA_BLOCK * ITEMS;
    __shared__ u32 wh[PRA_WAVES][PRA_RADIX];
    __shared__ u32 s_base[PRA_RADIX];
    __shared__ u32 s_tile;
    if (threadIdx.x == 0) {
        // tiles are handed out in launch order; the block taking the last one
        // returns the counter to zero for the next launch
        u32 t = atomicAdd(tile_ctr, 1u);
        if (t == gridDim.x - 1) atomicExch(tile_ctr, 0u);
        s_tile = t;
    }
    for (int i = threadIdx.x; i < PRA_WAVES * PRA_RADIX; i += PRA_BLOCK) (&wh[0][0])[i] = 0;
    __syncthreads();
    const u32 tile = s_tile;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 lt = pra_lanemask_lt();
    u64 base = (u64)tile * TILE + (u64)wave * (TILE / PRA_WAVES);
    u32 dg[ITEMS], lc[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        u64 idx = base + (u64)k * 64 + lane;
        bool valid = idx < n;
        u32 d = valid ? pra_digit(in + idx * w, ds) : 0u;
        u64 act = __ballot(valid);
        u64 peers = pra_match<PRA_RADIX_BITS>(d, act);
        dg[k] = d;
        lc[k] = 0;
        if (valid) {
            u32 before = wh[wave][d];
            lc[k] = before + (u32)__popcll(peers & lt);
            if ((peers >> lane) == 1ULL) wh[wave][d] = before + (u32)__popcll(peers);
        }
    }
    __syncthreads();
    {
        const int d = threadIdx.x;   // PRA_BLOCK == PRA_RADIX: one digit per thread
        u32 acc = 0;
        for (int v = 0; v < PRA_WAVES; v++) {
            u32 t = wh[v][d];
            wh[v][d] = acc;
            acc += t;
        }
        u64* my = status + (u64)tile * PRA_RADIX + d;
        u64 excl = 0;
        if (tile == 0) {
            __hip_atomic_store(my, st_pack(epoch, PRA_ST_INC, acc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(my, st_pack(epoch, PRA_ST_AGG, acc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            excl = pra_lookback(status, tile, d, epoch, flags);
            __hip_atomic_store(my, st_pack(epoch, PRA_ST_INC, excl + acc), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_base[d] = gofs[d] + (u32)excl;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        u64 idx = base + (u64)k * 64 + lane;
        if (idx < n) {
            u64 pos = (u64)s_base[dg[k]] + wh[wave][dg[k]] + lc[k];
            const u64* src = in + idx * w;
            u64* dst = out + pos * w;
            pra_copy_row(dst, src, w);
            if (IDX) iout[pos] = iin ? iin[idx] : (u32)idx;
        }
    }
}

// One block sorts n <= PRA_SMALL_SORT rows, whose key words (n x kw) fit the
// PRA_SMALL_SORT x PRA_SMALL_SORT_KW words of LDS, by (bucket, col0 .. col[kc-1])
// (the bucket FNV(row[0:jcc]) % nb leads when `bucketed`; columns compare as
// unsigned words) with a bitonic network over row indices in LDS, then
// gathers the rows into `out`.  Unstable: rows with equal keys are equal.
// With bnd (bucketed), bnd[s] = first row of bucket s, bnd[nb] = n.
__device__ inline void pra_small_sort_block(const u64* rows, u64 n, int w, int jcc, int kc, u64 nb, int bucketed,
                                            u64* out, u64* bnd)
{
    __shared__ u64 key[PRA_SMALL_SORT * PRA_SMALL_SORT_KW];
    __shared__ u32 idx[PRA_SMALL_SORT];
    const int kw = kc + (bucketed ? 1 : 0);
    const u32 m = (u32)n;
    for (u32 i = threadIdx.x; i < m; i += PRA_SMALL_SORT_BLOCK) {
        const u64* r = rows + (u64)i * w;
        u64* k = key + (u64)i * kw;
        int o = 0;
        if (bucketed) k[o++] = pra_fnv(r, jcc) % nb;
        for (int c = 0; c < kc; c++) k[o++] = r[c];
    }
    u32 N = 2;
    while (N < m) N <<= 1;
    for (u32 i = threadIdx.x; i < N; i += PRA_SMALL_SORT_BLOCK) idx[i] = i;   // i >= m: past every row
    __syncthreads();
    for (u32 k = 2; k <= N; k <<= 1)
        for (u32 j = k >> 1; j > 0; j >>= 1) {
            for (u32 i = threadIdx.x; i < N; i += PRA_SMALL_SORT_BLOCK) {
                const u32 l = i ^ j;
                if (l <= i) continue;
                const u32 a = idx[i], b = idx[l];
                bool gt;
                if (a >= m || b >= m) {
                    gt = a >= m && (b < m || a > b);
                } else {
                    const u64* ka = key + (u64)a * kw;
                    const u64* kb = key + (u64)b * kw;
                    int c = 0;
                    while (c < kw && ka[c] == kb[c]) c++;
                    gt = c < kw && ka[c] > kb[c];
                }
                if (((i & k) == 0) == gt) {
                    idx[i] = b;
                    idx[l] = a;
                }
            }
            __syncthreads();
        }
    for (u32 i = threadIdx.x; i < m; i += PRA_SMALL_SORT_BLOCK) {
        const u64* src = rows + (u64)idx[i] * w;
        u64* dst = out + (u64)i * w;
        pra_copy_row(dst, src, w);
    }
    if (bnd && bucketed) {
        for (u32 i = threadIdx.x; i <= m; i += PRA_SMALL_SORT_BLOCK) {
            // buckets (prev, cur] start at row i
            const u64 prev = i == 0 ? 0 : key[(u64)idx[i - 1] * kw] + 1;
            const u64 cur = i == m ? nb : key[(u64)idx[i] * kw];
            for (u64 s = prev; s <= cur && s <= nb; s++) bnd[s] = i;
        }
    }
}

__global__ __launch_bounds__(PRA_SMALL_SORT_BLOCK) void k_small_sort(const u64* rows, u64 n, int w, int jcc, int kc,
                                                                     u64 nb, int bucketed, u64* out, u64* bnd)
{
    pra_small_sort_block(rows, n, w, jcc, kc, nb, bucketed, out, bnd);
}

// Independent small sorts (one per output relation of an iteration), one
// block each.
__global__ __launch_bounds__(PRA_SMALL_SORT_BLOCK) void k_small_sort_many(SmallSortBatch b)
{
    const SmallSortJob& j = b.job[blockIdx.x];
    pra_small_sort_block(j.rows, j.n, j.w, j.jcc, j.kc, b.nb, j.bucketed, j.out, j.bnd);
}

// One block sorts n <= PRA_BLOCK_SORT rows with all the LSD passes of `pl`
// (least significant first; digits from the rows, as k_sort_scatter's): the
// row indices stay in registers and LDS, every pass ranks them stably by
// wave-local counts and a scan over waves and digits, and the rows are
// gathered into `out` at the end.  With bnd, the last pass is the bucket
// digit and bnd[s] = first row of bucket s (FNV(row[0:jcc]) % nb), bnd[nb] = n.
__global__ __launch_bounds__(PRA_BIN_BLOCK) void k_block_sort(const u64* rows, u64 n, int w, PassList pl, u64* out,
                                                              u64* bnd, int jcc, u64 nb)
{
    constexpr int IT = PRA_BLOCK_SORT / PRA_BIN_BLOCK;
    constexpr int WAVES = PRA_BIN_BLOCK / 64;
    constexpr u32 PER_WAVE = PRA_BLOCK_SORT / WAVES;
    __shared__ u32 sidx[PRA_BLOCK_SORT];
    __shared__ u32 wh[WAVES][PRA_RADIX];
    __shared__ u32 wsum[PRA_RADIX / 64];
    const u32 m = (u32)n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 lt = pra_lanemask_lt();
    u32 id[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) id[k] = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
#pragma unroll 1
    for (int p = 0; p < pl.np; p++) {
        const DigitSpec ds = pl.p[p];
        for (int i = threadIdx.x; i < WAVES * PRA_RADIX; i += PRA_BIN_BLOCK) (&wh[0][0])[i] = 0;
        __syncthreads();
        u32 dg[IT], lc[IT];
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            bool valid = i < m;
            dg[k] = 0;
            lc[k] = 0;
            if ((u32)wave * PER_WAVE + (u32)k * 64 >= m) continue;   // item row wholly past the rows
            u32 d = valid ? pra_digit(rows + (u64)id[k] * w, ds) : 0u;
            u64 peers = pra_match<PRA_RADIX_BITS>(d, __ballot(valid));
            dg[k] = d;
            if (valid) {
                u32 before = wh[wave][d];
                lc[k] = before + (u32)__popcll(peers & lt);
                if ((peers >> lane) == 1ULL) wh[wave][d] = before + (u32)__popcll(peers);
            }
        }
        __syncthreads();
        const int d = threadIdx.x;   // digit of this thread in the scan (threads < PRA_RADIX)
        u32 acc = 0;
        if (d < PRA_RADIX)
            for (int v = 0; v < WAVES; v++) {
                u32 t = wh[v][d];
                wh[v][d] = acc;
                acc += t;
            }
        u32 x = acc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            u32 y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (d < PRA_RADIX && lane == 63) wsum[wave] = x;
        __syncthreads();
        if (d < PRA_RADIX) {
            u32 base = x - acc;
            for (int v = 0; v < wave; v++) base += wsum[v];
            for (int v = 0; v < WAVES; v++) wh[v][d] += base;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            if (i < m) sidx[wh[wave][dg[k]] + lc[k]] = id[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            if (i < m) id[k] = sidx[i];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < IT; k++) {
        u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
        if (i < m) {
            const u64* src = rows + (u64)id[k] * w;
            u64* dst = out + (u64)i * w;
            pra_copy_row(dst, src, w);
        }
    }
    if (bnd) {
        // buckets (bucket of row i-1, bucket of row i] start at row i
        for (u32 i = threadIdx.x; i <= m; i += PRA_BIN_BLOCK) {
            const u64 prev = i == 0 ? 0 : pra_fnv(rows + (u64)sidx[i - 1] * w, jcc) % nb + 1;
            const u64 cur = i == m ? nb : pra_fnv(rows + (u64)sidx[i] * w, jcc) % nb;
            for (u64 s = prev; s <= cur && s <= nb; s++) bnd[s] = i;
        }
    }
}

// bitwise OR and AND of column blockIdx.y over all rows, used to skip radix
// passes whose digit is the same for every row
__global__ void k_col_orand2(const u64* rows, u64 n, int w, u64* out /* [0..15] or, [16..31] and */)
{
    const int c = blockIdx.y;
    u64 o = 0, a = ~0ULL;
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 step = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += step) {
        u64 v = rows[i * w + c];
        o |= v;
        a &= v;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        o |= __shfl_xor(o, d, 64);
        a &= __shfl_xor(a, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicOr((unsigned long long*)&out[c], (unsigned long long)o);
        atomicAnd((unsigned long long*)&out[16 + c], (unsigned long long)a);
    }
}

// ------------------------------------------------------ value dictionary ---
//
// An order-preserving dictionary of the values in the relations' key columns
// (rank(a) < rank(b) iff a < b) lets a newt row's sort key
// (bucket, col0, .., col[a-1]) be packed into one u64: the bucket and the
// column ranks as digits of radix D (the dictionary size), which the sort
// then orders in a few byte passes.
// Slots: [value, rank + 1]; rank + 1 == 0 marks an empty slot.

__device__ inline u64 pra_dict_slot(u64 v, u64 cap) { return pra_mix64(v ^ 0x5851F42D4C957F2DULL) & (cap - 1); }

__device__ inline bool pra_dict_rank(const u64* tab, u64 cap, u64 v, u64* rank)
{
    u64 m = cap - 1, s = pra_dict_slot(v, cap);
    for (u64 q = 0; q < cap; q++) {
        const u64* sl = tab + 2 * s;
        u64 r = sl[1];
        if (r == 0) return false;
        if (sl[0] == v) {
            *rank = r - 1;
            return true;
        }
        s = (s + 1) & m;
    }
    return false;
}

__global__ void k_collect_cols(const u64* rows, u64 n, int w, int ncols, u64* out)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < ncols; c++) out[i * ncols + c] = rows[i * w + c];
}

// keep[i] = 1 when row i falls in the same bucket under the hash of its first
// ja columns as under the hash of its first jb columns
__global__ void k_view_keep(const u64* rows, u64 n, int w, int ja, int jb, u64 nb, u32* keep)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64* r = rows + i * w;
    keep[i] = (pra_fnv(r, ja) % nb) == (pra_fnv(r, jb) % nb) ? 1u : 0u;
}

// the kept rows, in order, at their exclusive-prefix positions
__global__ void k_view_compact(const u64* rows, u64 n, int w, const u32* keep, const u32* pos, u64* out)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const u64* r = rows + i * w;
    u64* o = out + (u64)pos[i] * w;
    for (int k = 0; k < w; k++) o[k] = r[k];
}

// values sorted and distinct: rank = position
__global__ void k_dict_insert(const u64* vals, u64 n, u64* tab, u64 cap, u32* flags)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 v = vals[i];
    u64 m = cap - 1, s = pra_dict_slot(v, cap);
    for (u64 q = 0; q < cap; q++) {
        u64* sl = tab + 2 * s;
        if (atomicCAS((unsigned long long*)(sl + 1), 0ULL, (unsigned long long)(i + 1)) == 0) {
            sl[0] = v;
            return;
        }
        s = (s + 1) & m;
    }
    atomicOr(flags, PRA_FLAG_TABLE_FULL);
}

// out[i] = in[idx[i]] for rows of w words
__global__ void k_gather_rows(const u64* in, const u32* idx, u64 n, int w, u64* out)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64* src = in + (u64)idx[i] * w;
    u64* dst = out + i * w;
    pra_copy_row(dst, src, w);
}

// Packed-key sort of the newt rows.  The newt scatter has put every row's
// (key, newt position) pair into the bin of the key's top byte (bin b holds
// its pairs, in no particular order, at bin_keys / bin_idx + b * bin_cap, and
// its rows go to [bofs[b], bofs[b + 1]) of the sorted order); block b orders
// its bin by the key's lower varying bytes (bit p of `bytes` = byte p varies,
// p < 7) with LSD passes in LDS, then gathers the rows into `out` in that
// order.  Keys are distinct, so the result is the total order.  A bin holds
// at most min(bin_cap, PRA_BIN_CAP) pairs; the host checks the largest bin
// before it launches this kernel (a larger one raises PRA_FLAG_BIN_OVERFLOW
// and is left unsorted).  Bucket s starts at the first key >= (s * bmul) <<
// ls (the key is bucket * bmul + column ranks, shifted left by ls); the block
// whose bin holds that value counts its keys below it.
// The sort of one bin (or segment) b by a block of BLOCK threads (at most
// BLOCK * PRA_BIN_ITEMS pairs): its pairs are keys[0..end-start) /
// idx[0..end-start), all with the same key bits from bit `top` up (the bin's
// value b), and its rows go to [start, end) of the sorted order.
template <int BLOCK>
__device__ __forceinline__ void pra_bin_sort_range(const u64* keys, const u32* idx, u64 start, u64 end, u32 cap, u64 n,
                                                   u32 bytes, const u64* rows, int w, u64* out, u64 bmul, int ls,
                                                   u64 nb, u64* bnd, u32* flags, const u32 b, const int top)
{
    constexpr int IT = PRA_BIN_ITEMS;
    constexpr u32 CAP = BLOCK * PRA_BIN_ITEMS;
    constexpr int WAVES = BLOCK / 64;
    constexpr u32 PER_WAVE = CAP / WAVES;
    static_assert(BLOCK >= PRA_RADIX, "the digit scan takes a thread per digit");
    __shared__ u64 skey[CAP];
    __shared__ u32 sidx[CAP];
    __shared__ u32 wh[WAVES][PRA_RADIX];
    __shared__ u32 wsum[PRA_RADIX / 64];
    __shared__ u32 below;
    if (b == 0 && threadIdx.x == 0) {
        bnd[0] = 0;
        bnd[nb] = n;
    }
    if (end <= start) {
        for (u64 s = 1 + threadIdx.x; s < nb; s += BLOCK)
            if ((((s * bmul) << ls) >> top) == b) bnd[s] = start;
        return;
    }
    if (end - start > CAP || end - start > cap) {
        if (threadIdx.x == 0) atomicOr(flags, PRA_FLAG_BIN_OVERFLOW);
        return;
    }
    const u32 m = (u32)(end - start);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 lt = pra_lanemask_lt();
    u64 key[IT];
    u32 id[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
        u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
        key[k] = 0;
        id[k] = 0;
        if (i < m) {
            key[k] = keys[i];
            id[k] = idx[i];
        }
    }
#pragma unroll 1
    for (int p = 0; p < 8; p++) {
        if (!((bytes >> p) & 1u)) continue;
        for (int i = threadIdx.x; i < WAVES * PRA_RADIX; i += BLOCK) (&wh[0][0])[i] = 0;
        __syncthreads();
        // stable rank of every pair among the pairs of its digit: wave-local
        // counts in element order, then a scan over waves and digits
        u32 dg[IT], lc[IT];
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            bool valid = i < m;
            dg[k] = 0;
            lc[k] = 0;
            // a wave skips the ranking of item rows wholly past the bin
            if ((u32)wave * PER_WAVE + (u32)k * 64 >= m) continue;
            u32 d = valid ? (u32)(key[k] >> (8 * p)) & 255u : 0u;
            u64 peers = pra_match<PRA_RADIX_BITS>(d, __ballot(valid));
            dg[k] = d;
            if (valid) {
                u32 before = wh[wave][d];
                lc[k] = before + (u32)__popcll(peers & lt);
                if ((peers >> lane) == 1ULL) wh[wave][d] = before + (u32)__popcll(peers);
            }
        }
        __syncthreads();
        const int d = threadIdx.x;   // digit of this thread in the scan (threads < PRA_RADIX)
        u32 acc = 0;
        if (d < PRA_RADIX)
            for (int v = 0; v < WAVES; v++) {
                u32 t = wh[v][d];
                wh[v][d] = acc;
                acc += t;
            }
        u32 x = acc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            u32 y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (d < PRA_RADIX && lane == 63) wsum[wave] = x;
        __syncthreads();
        if (d < PRA_RADIX) {
            u32 base = x - acc;
            for (int v = 0; v < wave; v++) base += wsum[v];
            for (int v = 0; v < WAVES; v++) wh[v][d] += base;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            if (i < m) {
                u32 pos = wh[wave][dg[k]] + lc[k];
                skey[pos] = key[k];
                sidx[pos] = id[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            if (i < m) {
                key[k] = skey[i];
                id[k] = sidx[i];
            }
        }
        __syncthreads();
    }
    for (u64 s = 1; s < nb; s++) {
        const u64 T = (s * bmul) << ls;
        if ((T >> top) != b) continue;
        if (threadIdx.x == 0) below = 0;
        __syncthreads();
        u32 cnt = 0;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
            if (i < m && key[k] < T) cnt++;
        }
        if (cnt) atomicAdd(&below, cnt);
        __syncthreads();
        if (threadIdx.x == 0) bnd[s] = start + below;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < IT; k++) {
        u32 i = (u32)wave * PER_WAVE + (u32)k * 64 + lane;
        if (i < m) {
            const u64* src = rows + (u64)id[k] * w;
            u64* dst = out + (start + i) * w;
            pra_copy_row(dst, src, w);
        }
    }
}

// bin b of the newt scatter's top-byte bins (bofs = the bins' first sorted
// positions) by a block of BLOCK threads
template <int BLOCK>
__device__ __forceinline__ void pra_bin_sort_block(const u64* bin_keys, const u32* bin_idx, u32 bin_cap, u64 n,
                                                   const u32* bofs, u32 bytes, const u64* rows, int w, u64* out,
                                                   u64 bmul, int ls, u64 nb, u64* bnd, u32* flags, const u32 b)
{
    const u64 start = bofs[b], end = b + 1 < PRA_RADIX ? bofs[b + 1] : n;
    pra_bin_sort_range<BLOCK>(bin_keys + (u64)b * bin_cap, bin_idx + (u64)b * bin_cap, start, end, bin_cap, n,
                              bytes & 0x7Fu, rows, w, out, bmul, ls, nb, bnd, flags, b, 56);
}

// BLOCK threads per bin: the narrowest of 256, 512 and PRA_BIN_BLOCK whose
// BLOCK * PRA_BIN_ITEMS pairs hold the largest bin (a wave holds 512 of a
// bin's pairs, so small bins leave most waves of a wide block without items,
// and every wave still takes part in each pass's barriers and digit scan)
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bin_sort(const u64* bin_keys, const u32* bin_idx, u32 bin_cap, u64 n,
                                                    const u32* bofs, u32 bytes, const u64* rows, int w, u64* out,
                                                    u64 bmul, int ls, u64 nb, u64* bnd, u32* flags)
{
    pra_bin_sort_block<BLOCK>(bin_keys, bin_idx, bin_cap, n, bofs, bytes, rows, w, out, bmul, ls, nb, bnd, flags,
                              blockIdx.x);
}

// The bin sorts of several outputs in one launch: block b sorts bin b % 256
// of job b / 256 (job table in device memory).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bin_sort_many(const BinSortJob* jobs, u32* flags)
{
    const BinSortJob& j = jobs[blockIdx.x / PRA_RADIX];
    pra_bin_sort_block<BLOCK>(j.bin_keys, j.bin_idx, j.bin_cap, j.n, j.bofs, j.bytes, j.rows, j.w, j.out, j.bmul, j.ls,
                              j.nb, j.bnd, flags, blockIdx.x % PRA_RADIX);
}

// Segmented form for newts too large for the top-byte bins: the pairs are
// already ordered by the keys' top 16 bits (two counting passes), segment b
// is the pairs whose key bits from `top` up equal b ([seg[b], seg[b + 1])),
// and block b sorts it by the varying bytes below (bit p of `bytes`).
// Block i sorts segment list[i] (list null: segment i, skipping a segment
// of more than skip_above pairs, which a launch over the list sorts).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_seg_sort(const u64* keys, const u32* idx, const u64* seg, int top, u64 n,
                                                    u32 bytes, const u64* rows, int w, u64* out, u64 bmul, int ls,
                                                    u64 nb, u64* bnd, u32* flags, const u32* list, u32 skip_above)
{
    const u32 b = list ? list[blockIdx.x] : blockIdx.x;
    const u64 start = seg[b], end = seg[b + 1];
    if (end - start > skip_above) return;
    pra_bin_sort_range<BLOCK>(keys + start, idx + start, start, end, BLOCK * PRA_BIN_ITEMS, n, bytes, rows, w, out,
                              bmul, ls, nb, bnd, flags, b, top);
}

// seg[v] = first position of a sorted key array whose bits from `top` up are
// >= v, for v in [0, nseg] (seg[nseg] = n); stat[0] = the largest segment,
// and the stat[1] segments of more than PRA_SEG_SMALL pairs are listed in
// large (in no particular order)
__global__ void k_seg_bounds(const u64* keys, u64 n, int top, u32 nseg, u64* seg, u64* stat, u32* large)
{
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nseg) return;
    u64 b2[2];
    for (int k = 0; k < 2; k++) {
        const u64 t = v + (u64)k;
        if (t >= nseg) {
            b2[k] = n;
            continue;
        }
        u64 lo = 0, hi = n;
        while (lo < hi) {
            const u64 mid = (lo + hi) >> 1;
            if ((keys[mid] >> top) < t) lo = mid + 1; else hi = mid;
        }
        b2[k] = lo;
    }
    seg[v] = b2[0];
    if (v + 1 == nseg) seg[nseg] = n;
    const u64 size = b2[1] - b2[0];
    atomicMax((unsigned long long*)stat, (unsigned long long)size);
    if (size > PRA_SEG_SMALL) large[atomicAdd((unsigned long long*)(stat + 1), 1ULL)] = (u32)v;
}

// bnd[s] = first position whose key is >= (s * bmul) << ls (the key is
// bucket * bmul + column ranks, shifted left by ls)
__global__ void k_key_bounds(const u64* keys, u64 n, u64 bmul, int ls, u64 nb, u64* bnd)
{
    u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nb) return;
    if (s == nb) {
        bnd[s] = n;
        return;
    }
    const u64 T = (s * bmul) << ls;
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if (keys[mid] < T) lo = mid + 1; else hi = mid;
    }
    bnd[s] = lo;
}

// order-sensitive checksum of rows: sum over i of hash(row i, i)
__global__ void k_rows_checksum(const u64* rows, u64 n, int w, u64* out)
{
    __shared__ u64 red[PRA_WAVES];
    u64 acc = 0;
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 step = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += step) acc += pra_mix64(pra_keyhash(rows + i * w, w) ^ pra_mix64(i + 0x2545F4914F6CDD1DULL));
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t = 0;
        for (int v = 0; v < PRA_WAVES; v++) t += red[v];
        atomicAdd((unsigned long long*)out, (unsigned long long)t);
    }
}

// bstart[s] = first row whose bucket is >= s, for s in [0, nb]
__global__ void k_bucket_bounds(const u64* rows, u64 n, int w, int jcc, u64 nb, u64* bstart)
{
    u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nb) return;
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        u64 b = pra_fnv(rows + mid * w, jcc) % nb;
        if (b < s) lo = mid + 1; else hi = mid;
    }
    bstart[s] = lo;
}

// keep rows whose flag is set, in order
__global__ void k_compact_rows(const u64* in, u64 n, int w, const u64* flags, const u64* pos, u64* out)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    const u64* src = in + i * w;
    u64* dst = out + pos[i] * w;
    pra_copy_row(dst, src, w);
}

// ------------------------------------------------------- join index --------

__global__ void k_group_flags(const u64* rows, u64 n, int w, int jcc, u64* flags)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 f = 1;
    if (i > 0) {
        f = 0;
        const u64* a = rows + i * w;
        const u64* b = a - w;
        for (int c = 0; c < jcc; c++)
            if (a[c] != b[c]) { f = 1; break; }
    }
    flags[i] = f;
}

__global__ void k_group_compact(const u64* flags, const u64* gpos, u64 n, u64* gstart)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        gstart[n == 0 ? 0 : gpos[n - 1] + flags[n - 1]] = n;
        return;
    }
    if (flags[i]) gstart[gpos[i]] = i;
}

// longest group (largest number of rows sharing a key) of a sorted version
__global__ void k_group_maxlen(const u64* flags, const u64* gpos, const u64* gstart, u64 n, u64* maxlen)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    u64 len = gstart[gpos[i] + 1] - i;
    atomicMax((unsigned long long*)maxlen, (unsigned long long)len);
}

// index slots: [hash, start row, row count]
__global__ void k_index_insert(const u64* rows, int w, int jcc, const u64* gstart, u64 ng, u64* table, u64 cap,
                               u32* flags)
{
    u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    u64 i0 = gstart[g], cnt = gstart[g + 1] - i0;
    u64 h = pra_keyhash(rows + i0 * w, jcc);
    u64 m = cap - 1, s = h & m;
    for (u64 p = 0; p < cap; p++) {
        u64* slot = table + s * 3;
        u64 prev = atomicCAS((unsigned long long*)slot, 0ULL, (unsigned long long)h);
        if (prev == 0) {
            slot[1] = i0;
            slot[2] = cnt;
            return;
        }
        if (prev == h) {
            atomicOr(flags, PRA_FLAG_INDEX_COLLISION);
            return;
        }
        s = (s + 1) & m;
    }
    atomicOr(flags, PRA_FLAG_TABLE_FULL);
}

// --------------------------------------------------------------- join ------

// lexicographic comparison of the first n words of two rows
__device__ inline int pra_row_cmp(const u64* a, const u64* b, int n)
{
    for (int x = 0; x < n; x++)
        if (a[x] != b[x]) return a[x] < b[x] ? -1 : 1;
    return 0;
}

// Run of inner rows matching an outer key, in an inner version kept sorted by
// (bucket, columns): one run inside the key's bucket, found by a binary
// search and a gallop (used for inner relations that change every iteration,
// where a hash index would be rebuilt each time).  Returns its length.
__device__ inline u64 pra_sorted_run(const u64* key, int jcc, const u64* inner, int wi, const u64* ibnd, u64 nb,
                                     u64* first_out)
{
    u64 b = nb == 1 ? 0 : pra_fnv(key, jcc) % nb;
    u64 lo = ibnd[b], hi = ibnd[b + 1];
    u64 a = lo, z = hi;   // first row >= key
    while (a < z) {
        u64 m = (a + z) >> 1;
        if (pra_row_cmp(inner + m * wi, key, jcc) < 0) a = m + 1; else z = m;
    }
    u64 first = a;
    // first row > key: gallop from the first match (key runs are short), then
    // a binary search inside the last step
    u64 step = 1;
    z = first;
    while (z < hi && pra_row_cmp(inner + z * wi, key, jcc) <= 0) {
        a = z + 1;
        z = first + step < hi ? first + step : hi;
        step <<= 1;
    }
    while (a < z) {
        u64 m = (a + z) >> 1;
        if (pra_row_cmp(inner + m * wi, key, jcc) <= 0) a = m + 1; else z = m;
    }
    *first_out = first;
    return a - first;
}

// Merge of a sorted delta into a sorted full (both ordered by (bucket, all w
// columns), disjoint).  rank[j] = number of full rows before delta row j,
// searched inside delta row j's bucket.
__global__ void k_merge_rank(const u64* D, u64 nd, const u64* dbnd, const u64* F, const u64* fbnd, u64 nb, int w,
                             u64* rank)
{
    u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nd) return;
    u64 b = 0, bh = nb;   // bucket of row j: last b with dbnd[b] <= j
    while (b + 1 < bh) {
        u64 m = (b + bh) >> 1;
        if (dbnd[m] <= j) b = m; else bh = m;
    }
    const u64* d = D + j * w;
    u64 a = fbnd[b], z = fbnd[b + 1];
    while (a < z) {
        u64 m = (a + z) >> 1;
        if (pra_row_cmp(F + m * w, d, w) < 0) a = m + 1; else z = m;
    }
    rank[j] = a;
}

// Place every row of the merge: delta row j at j + rank[j], full row i after
// the delta rows ranked at or before it (rank is non-decreasing).
// (with add_to: block 0 also adds delta's bucket bounds to full's, the
// bounds of the merged version; k_merge_rank has read full's before)
__global__ void k_merge_scatter(const u64* F, u64 nf, const u64* D, u64 nd, const u64* rank, int w, u64* out,
                                u64* add_to, const u64* dbnd, u64 nbb)
{
    if (add_to && blockIdx.x == 0)
        for (u64 i = threadIdx.x; i < nbb; i += blockDim.x) add_to[i] += dbnd[i];
    u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf + nd) return;
    const u64* src;
    u64 pos;
    if (t < nf) {
        u64 a = 0, z = nd;   // delta rows with rank <= t
        while (a < z) {
            u64 m = (a + z) >> 1;
            if (rank[m] <= t) a = m + 1; else z = m;
        }
        src = F + t * w;
        pos = t + a;
    } else {
        u64 j = t - nf;
        src = D + j * w;
        pos = j + rank[j];
    }
    u64* dst = out + pos * w;
    pra_copy_row(dst, src, w);
}

// One thread per join pair: locate the outer row by binary search over the
// exclusive offsets, build the projection of (outer ++ inner[jcc:]) (LEFT) or
// (inner ++ outer[jcc:]) (RIGHT), and write it at its arrival position.
// largest i in [lo, hi] with off[i] <= x (off[lo] <= x holds)
__device__ inline u64 pra_row_of(const u64* off, u64 lo, u64 hi, u64 x)
{
    while (lo < hi) {
        u64 mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Row of output o0 for the whole block: the largest i in [0, n] with off[i]
// <= o0 (off[0] <= o0 holds), found with PRA_BLOCK probes per step, so the
// range shrinks PRA_BLOCK-fold per dependent load.
__device__ inline u64 pra_block_row_of(const u64* off, u64 n, u64 o0, u32* s_cnt)
{
    u64 lo = 0, hi = n + 1;   // answer in [lo, hi)
    while (hi - lo > 1) {
        const u64 span = hi - lo;
        const u64 idx = lo + (span * threadIdx.x) / PRA_BLOCK;
        const bool pred = off[idx] <= o0;   // true for a prefix of the threads
        if (threadIdx.x == 0) *s_cnt = 0;
        __syncthreads();
        const u64 b = __ballot(pred);
        if ((threadIdx.x & 63) == 0 && b) atomicAdd(s_cnt, (u32)__popcll(b));
        __syncthreads();
        const u32 cnt = *s_cnt;
        __syncthreads();
        const u64 nlo = lo + (span * (cnt - 1)) / PRA_BLOCK;
        hi = cnt < PRA_BLOCK ? lo + (span * cnt) / PRA_BLOCK : hi;
        lo = nlo;
    }
    return lo;
}

// block `bid` of `gdim` of one rule's join expansion
__device__ __forceinline__ void pra_join_expand_block(const JoinExpandArgs& a, u32 bid, u32 gdim)
{
    if (a.ctl[1]) return;
    const u64 first = a.off[0], total = a.off[a.n_outer];   // first > 0 for a window of rows
    __shared__ u64 s_off[PRA_BLOCK + 1];
    __shared__ u32 s_cnt;
    const u64 step = (u64)gdim * blockDim.x;
    for (u64 o0 = first + (u64)bid * blockDim.x; o0 < total; o0 += step) {
        __syncthreads();
        // the block's first row, then the offsets of the next PRA_BLOCK rows
        // in LDS, where each output finds its row
        const u64 r0 = pra_block_row_of(a.off, a.n_outer, o0, &s_cnt);
        s_off[threadIdx.x] = r0 + threadIdx.x <= a.n_outer ? a.off[r0 + threadIdx.x] : ~0ULL;
        if (threadIdx.x == 0) s_off[PRA_BLOCK] = r0 + PRA_BLOCK <= a.n_outer ? a.off[r0 + PRA_BLOCK] : ~0ULL;
        __syncthreads();
        u64 o = o0 + threadIdx.x;
        if (o >= total) continue;
        u64 i;
        if (o < s_off[PRA_BLOCK]) {
            u32 lo = 0, hi = PRA_BLOCK;   // largest j in [lo, hi] with s_off[j] <= o
            while (lo < hi) {
                u32 mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= o) lo = mid; else hi = mid - 1;
            }
            i = r0 + lo;
        } else {
            i = pra_row_of(a.off, r0 + PRA_BLOCK, a.n_outer, o);
        }
        const u64* d = a.outer + i * a.wo;
        const u64* f = a.inner + (a.istart[i] + (o - a.off[i])) * a.wi;
        u64 s = a.nb == 1 ? 0 : pra_fnv(d, a.jcc) % a.nb;
        u64 pos = a.posadj[s] + o;
        u64* dst = a.cand + pos * a.cw;
        // the destination bucket hashes the output's first out_jcc columns
        // as they are written
        u64 h = PRA_FNV_OFFSET;
        for (int c = 0; c < a.cw; c++) {
            int m = a.map[c];
            u64 v;
            if (a.left)
                v = m < a.wo ? d[m] : f[a.jcc + (m - a.wo)];
            else
                v = m < a.wi ? f[m] : d[a.jcc + (m - a.wi)];
            dst[c] = v;
            if (c < a.out_jcc) h = pra_fnv_word(h, v);
        }
        a.dest[pos] = a.nb == 1 ? 0u : (u32)(h % a.nb);
    }
}

__global__ void k_join_expand(JoinExpandArgs a)
{
    pra_join_expand_block(a, blockIdx.x, gridDim.x);
}

// job of block `bid` in a device job table (block bases after the njobs
// arguments): the number of later jobs' bases at or below it, 64 bases per
// load round
__device__ inline int pra_job_of(const u32* block_base, int njobs, u32 bid)
{
    int k = 0;
    for (int j0 = 1; j0 < njobs; j0 += 64) {
        const int j = j0 + (int)(threadIdx.x & 63);
        k += __popcll(__ballot(j < njobs && block_base[j] <= bid));
    }
    return k;
}

// Join expansions of several rules in one launch (job table in device
// memory); a block whose job has no outputs left for it leaves before it
// reads the job's arguments.
__global__ void k_join_expand_many(const JoinExpandArgs* jobs, int njobs)
{
    const u32* block_base = (const u32*)(jobs + njobs);
    const int k = pra_job_of(block_base, njobs, blockIdx.x);
    const u32 bid = blockIdx.x - block_base[k];
    const JoinExpandArgs* jp = jobs + k;
    if (jp->ctl[1] || jp->off[0] + (u64)bid * PRA_BLOCK >= jp->off[jp->n_outer]) return;
    const JoinExpandArgs a = *jp;
    pra_join_expand_block(a, bid, block_base[k + 1] - block_base[k]);
}

__device__ inline bool pra_filter(int op, u64 x, u64 y)
{
    switch (op) {
    case PRA_F_EQ: return x == y;
    case PRA_F_NEQ: return x != y;
    case PRA_F_LT: return x < y;
    case PRA_F_LE: return x <= y;
    case PRA_F_GT: return x > y;
    case PRA_F_GE: return x >= y;
    case PRA_F_TRUE: return true;
    default: return false;
    }
}

// outer row i of one rule's copy expansion
__device__ __forceinline__ void pra_copy_expand_row(const CopyExpandArgs& a, u64 i)
{
    if (a.ctl[1]) return;
    if (i >= a.n) return;
    if (a.off[i + 1] == a.off[i]) return;
    const u64* r = a.rows + i * a.w;
    u64 s = a.nb == 1 ? 0 : pra_fnv(r, a.in_jcc) % a.nb;
    u64 pos = a.posadj[s] + a.off[i];
    u64* dst = a.cand + pos * a.cw;
    u64 h = PRA_FNV_OFFSET;
    for (int c = 0; c < a.cw; c++) {
        u64 v = r[a.map[c]];
        dst[c] = v;
        if (c < a.out_jcc) h = pra_fnv_word(h, v);
    }
    a.dest[pos] = a.nb == 1 ? 0u : (u32)(h % a.nb);
}

__global__ void k_copy_expand(CopyExpandArgs a)
{
    pra_copy_expand_row(a, (u64)blockIdx.x * blockDim.x + threadIdx.x);
}

// Copy expansions of several rules in one launch (job table as
// k_join_expand_many's; every block has rows)
__global__ void k_copy_expand_many(const CopyExpandArgs* jobs, int njobs)
{
    const u32* block_base = (const u32*)(jobs + njobs);
    const int k = pra_job_of(block_base, njobs, blockIdx.x);
    const CopyExpandArgs a = jobs[k];
    pra_copy_expand_row(a, (u64)(blockIdx.x - block_base[k]) * blockDim.x + threadIdx.x);
}

// Arrival plan of one output relation, computed on the device: for source
// bucket s and the RAs producing the relation (in RA order), the position of
// their outputs in the arrival order source-major, RA-minor
// (src/comm/all_to_allv_comm.cpp:64-142), the candidate count, whether it
// exceeds the allocated capacity, and the join pairs.  RA j emits the outer
// rows [lo[j][s], hi[j][s]) of source s (whole buckets, or one round's rows).
// With rounds_check, a join call with more than `thresh` outputs from one
// source (src/RA/parallel_join.cpp:196-201, src/RAM/RA_tasks.cpp:312) marks
// the output for threshold rounds (ctl[1] = 2) instead of emitting.
__global__ void k_plan(PlanArgs a)
{
    const int nb = a.nb;
    for (int e = threadIdx.x; e < a.nr * nb; e += blockDim.x) {
        int j = e / nb, s = e % nb;
        u64 lo = a.off[j][a.lo[j][s]], hi = a.off[j][a.hi[j][s]];
        a.srcoff[(u64)a.ras[j] * 2 * nb + s] = lo;
        a.srcoff[(u64)a.ras[j] * 2 * nb + nb + s] = hi - lo;
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x != 0) return;
    u64 tot = 0, jtot = 0;
    bool rounds = false;
    for (int s = 0; s < nb; s++)
        for (int j = 0; j < a.nr; j++) {
            const u64* so = a.srcoff + (u64)a.ras[j] * 2 * nb;
            a.posadj[(u64)a.ras[j] * nb + s] = tot - so[s];
            u64 cnt = so[nb + s];
            if (a.is_join[j]) {
                if (a.rounds_check && (long long)cnt > a.thresh) rounds = true;
                jtot += cnt;
            }
            tot += cnt;
        }
    a.ctl[0] = tot;
    a.ctl[1] = rounds ? 2 : tot > a.cap ? 1 : 0;
    if (a.count_joined && (rounds || tot <= a.cap)) *a.joined += jtot;
}

// k_plan for every output of an iteration at once: block b plans output
// outs[b] over its active rules in rule order (source-major, rule-minor
// arrival, as k_plan).
__global__ void k_plan_all(PlanAllArgs a)
{
    const int oi = a.outs[blockIdx.x];
    const int nb = a.nb;
    __shared__ unsigned char lr[PRA_MAXRA];
    __shared__ int nlr;
    if (threadIdx.x == 0) {
        int k = 0;
        for (int r = 0; r < a.nra; r++)
            if (a.rout[r] == oi) lr[k++] = (unsigned char)r;
        nlr = k;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nlr * nb; e += blockDim.x) {
        const int r = lr[e / nb], s = e % nb;
        const u64 lo = a.off[r][a.bnd[r][s]], hi = a.off[r][a.bnd[r][s + 1]];
        a.srcoff[(u64)r * 2 * nb + s] = lo;
        a.srcoff[(u64)r * 2 * nb + nb + s] = hi - lo;
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x != 0) return;
    u64 tot = 0, jtot = 0;
    bool rounds = false;
    for (int s = 0; s < nb; s++)
        for (int k = 0; k < nlr; k++) {
            const int r = lr[k];
            const u64* so = a.srcoff + (u64)r * 2 * nb;
            a.posadj[(u64)r * nb + s] = tot - so[s];
            const u64 cnt = so[nb + s];
            if (a.is_join[r]) {
                if ((long long)cnt > a.thresh) rounds = true;
                jtot += cnt;
            }
            tot += cnt;
        }
    u64* ctl = a.stage2 + (u64)oi * a.sw + 4;
    ctl[0] = tot;
    ctl[1] = rounds ? 2 : tot > a.cap[oi] ? 1 : 0;
    if ((rounds || tot <= a.cap[oi]) && jtot) atomicAdd((unsigned long long*)a.joined[oi], (unsigned long long)jtot);
}

__global__ void k_plan_tab(PlanTabArgs a)
{
    const PlanOut o = a.outs[blockIdx.x];
    const int nb = a.nb;
    for (u32 e = threadIdx.x; e < o.count * (u32)nb; e += blockDim.x) {
        const u32 r = a.rule_ids[o.first + e / nb];
        const int s = (int)(e % nb);
        const PlanRule& pr = a.rules[r];
        const u64 lo = pr.off[pr.bnd[s]], hi = pr.off[pr.bnd[s + 1]];
        a.srcoff[(u64)r * 2 * nb + s] = lo;
        a.srcoff[(u64)r * 2 * nb + nb + s] = hi - lo;
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x != 0) return;
    u64 tot = 0, jtot = 0;
    bool rounds = false;
    for (int s = 0; s < nb; s++)
        for (u32 k = 0; k < o.count; k++) {
            const u32 r = a.rule_ids[o.first + k];
            const u64* so = a.srcoff + (u64)r * 2 * nb;
            a.posadj[(u64)r * nb + s] = tot - so[s];
            const u64 cnt = so[nb + s];
            if (a.rules[r].is_join) {
                if ((long long)cnt > a.thresh) rounds = true;
                jtot += cnt;
            }
            tot += cnt;
        }
    u64* ctl = a.stage2 + (u64)o.oi * a.sw + 4;
    ctl[0] = tot;
    ctl[1] = rounds ? 2 : tot > o.cap ? 1 : 0;
    if ((rounds || tot <= o.cap) && jtot) atomicAdd((unsigned long long*)o.joined, (unsigned long long)jtot);
}

// Split of one threshold round: the first outer row i (of n, offsets off
// relative to base = off[0]) such that the call's distinct outputs through
// row i, kp[off[i + 1] - base] (exclusive prefix of first-occurrence flags),
// exceed thresh; n when there is none.
__global__ void k_find_split(const u64* off, u64 n, const u32* kp, long long thresh, u64* out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const u64 base = off[0];
    u64 a = 0, z = n;
    while (a < z) {
        u64 m = (a + z) >> 1;
        if ((long long)kp[off[m + 1] - base] > thresh) z = m; else a = m + 1;
    }
    *out = a;
}

// Copy `words` summary words and then the error flags into pinned host
// memory; the host reads them once the kernel has completed.
// The summary words and the error flags into pinned host memory, then
// host[-1] = seq once every word is visible to the host (each thread's
// system-scope release before the barrier, then one system-scope store), so
// the host can start on the summary before the kernel's completion is
// signalled.  host[-1] is written by publications only, so a sequence number
// never meets a stale summary word there.
__device__ __forceinline__ void pra_publish_words(const u64* src, u64 words, const u32* flags, u64* host, u64 seq)
{
    for (u64 i = threadIdx.x; i < words; i += blockDim.x) host[i] = src[i];
    if (threadIdx.x == 0) host[words] = *flags;
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(host - 1, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_publish(const u64* src, u64 words, const u32* flags, u64* host, u64 seq)
{
    pra_publish_words(src, words, flags, host, seq);
}

// ---------------------------------------------------- dedup / newt ---------

// Count this lane in LDS counter h[d] when `on`, adding once per group of
// lanes of the wave with the same d (d < 256): a tile's candidates fall into
// few destination buckets, and one LDS atomic per lane would serialise on
// them.  Every lane of the wave must call it.
__device__ __forceinline__ void pra_wave_count(u32* h, u32 d, bool on)
{
    const u64 act = __ballot(on);
    if (!act) return;
    const u64 peers = pra_match<PRA_RADIX_BITS>(on ? d : 0u, act);
    if (on && __lane_id() == (u32)(__ffsll((long long)peers) - 1)) atomicAdd(&h[d], (u32)__popcll(peers));
}


// Claim (or find) the slot of candidate p's key (p < the candidate count).
// The first claimant stores the key; every candidate then lowers the slot's
// arrival mark to (epoch << 40 | p+1), so the earliest arrival of this
// iteration wins, while a mark from an earlier iteration (key already in full
// or delta) always wins.  MAXKW > 0: keys of at most MAXKW words (exactly
// MAXKW with EXACT), loaded once into registers with every word in flight
// together; MAXKW == 0: any key width, read from the row as needed.
template <int MAXKW, bool EXACT>
__device__ __forceinline__ void pra_claim_one(const u64* cand, int cw, int kw_, u64* table, u64 cap, u64* slot_of,
                                              u64 epoch, u32* flags, u64 p)
{
    const int kw = EXACT ? MAXKW : kw_;
    const u64* key = cand + p * cw;
    u64 kr[MAXKW > 0 ? MAXKW : 1];
    u64 h;
    if (MAXKW > 0) {
#pragma unroll
        for (int c = 0; c < MAXKW; c++) kr[c] = c < kw ? key[c] : 0;
        // pra_keyhash over the register copy
        h = 0x9E3779B97F4A7C15ULL ^ (u64)kw;
#pragma unroll
        for (int c = 0; c < MAXKW; c++)
            if (c < kw) h = pra_mix64(h ^ (kr[c] + 0x632BE59BD9B4E019ULL * (u64)(c + 1)));
        h = h ? h : 1;
    } else {
        h = pra_keyhash(key, kw);
    }
    const int stride = 2 + kw;
    const u64 mark = (epoch << 40) | (p + 1);
    u64 m = cap - 1, s = h & m;
    for (u64 q = 0; q < cap; q++) {
        u64* slot = table + s * stride;
        u64 cur = slot[0];
        bool mine = false;
        if (cur == 0) {
            cur = atomicCAS((unsigned long long*)slot, 0ULL, (unsigned long long)h);
            if (cur == 0) {
                mine = true;
                if (MAXKW > 0) {
#pragma unroll
                    for (int c = 0; c < MAXKW; c++)
                        if (c < kw) slot[2 + c] = kr[c];
                } else {
                    for (int c = 0; c < kw; c++) slot[2 + c] = key[c];
                }
            }
        }
        if (mine || cur == h) {
            // the claimant lowers the mark at once; a later candidate first
            // reads it: a mark from an earlier iteration (key already present)
            // or an earlier arrival of this one already beats ours
            if (mine) {
                atomicMin((unsigned long long*)(slot + 1), (unsigned long long)mark);
            } else {
                u64 old = slot[1];
                if (old > mark && old >= (epoch << 40))
                    atomicMin((unsigned long long*)(slot + 1), (unsigned long long)mark);
            }
            slot_of[p] = s;
            return;
        }
        s = (s + 1) & m;
    }
    slot_of[p] = ~0ULL;
    atomicOr(flags, PRA_FLAG_TABLE_FULL);
}

// One output's claims: one thread per candidate.
template <int MAXKW, bool EXACT>
__global__ void k_dedup_claim2(const u64* cand, const u64* ctl, int cw, int kw_, u64* table, u64 cap, u64* slot_of,
                               u64 epoch, u32* flags)
{
    if (ctl[1]) return;
    const u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ctl[0]) return;
    pra_claim_one<MAXKW, EXACT>(cand, cw, kw_, table, cap, slot_of, epoch, flags, p);
}

// Claims of several outputs in one launch: a block serves the job whose
// block range holds it, as k_dedup_claim2 serves its one output.
template <int MAXKW, bool EXACT>
__global__ void k_dedup_claim_many(ClaimBatch b, u32* flags)
{
    int k = 0;
    while (k + 1 < b.njobs && b.block_base[k + 1] <= blockIdx.x) k++;
    const ClaimJob& j = b.job[k];
    if (j.ctl[1]) return;
    const u64 p = (u64)(blockIdx.x - b.block_base[k]) * blockDim.x + threadIdx.x;
    if (p >= j.ctl[0]) return;
    pra_claim_one<MAXKW, EXACT>(j.cand, j.cw, j.kw, j.table, j.cap, j.slot_of, j.epoch, flags, p);
}

// Keep the winners, verify every key against its slot (a 64-bit hash
// collision between distinct keys is reported, never merged) and count kept
// candidates per destination bucket per tile.
// MAXKW > 0: keys of at most MAXKW words (exactly MAXKW with EXACT; the
// slot's words and the candidate's are loaded together); MAXKW == 0: any key
// width.
// Tile `tile` of PRA_TILE candidates (h, sl: the block's LDS counters and
// scan words).
template <int MAXKW, bool EXACT>
__device__ __forceinline__ void pra_decide_tile(const u64* cand, u64 n, int cw, int kw_, const u64* table,
                                                const u64* slot_of, u64 epoch, u32* keep, const u32* dest, u32 nb,
                                                u32* hist, u32 nblocks, u32* flags, u32* offs1, u32 tile, u32* h,
                                                u32* sl)
{
    const int kw = EXACT ? MAXKW : kw_;
    u64 base = (u64)tile * PRA_TILE;
    if (base >= n) {
        // a tile past the candidates keeps nothing
        for (u32 i = threadIdx.x; i < nb; i += PRA_DECIDE_BLOCK) hist[(u64)i * nblocks + tile] = 0;
        if (offs1)
            for (u32 i = threadIdx.x; i <= nb; i += PRA_DECIDE_BLOCK) offs1[i] = 0;
        return;
    }
    for (int i = threadIdx.x; i < PRA_RADIX; i += PRA_DECIDE_BLOCK) h[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PRA_TILE / PRA_DECIDE_BLOCK; k++) {
        u64 p = base + (u64)k * PRA_DECIDE_BLOCK + threadIdx.x;
        u32 kp = 0;
        if (p < n) {
            u64 s = slot_of[p];
            if (s != ~0ULL) {
                const u64* slot = table + s * (2 + kw);
                const u64* key = cand + p * cw;
                bool eq = true;
                if (MAXKW > 0) {
                    u64 x[MAXKW > 0 ? MAXKW : 1], y[MAXKW > 0 ? MAXKW : 1];
#pragma unroll
                    for (int c = 0; c < MAXKW; c++) {
                        x[c] = c < kw ? slot[2 + c] : 0;
                        y[c] = c < kw ? key[c] : 0;
                    }
#pragma unroll
                    for (int c = 0; c < MAXKW; c++) eq = eq && x[c] == y[c];
                } else {
                    for (int c = 0; c < kw; c++)
                        if (slot[2 + c] != key[c]) { eq = false; break; }
                }
                if (!eq) atomicOr(flags, PRA_FLAG_COLLISION);
                if (slot[1] == ((epoch << 40) | (p + 1))) kp = 1;
            }
            keep[p] = kp;
        }
        pra_wave_count(h, kp ? dest[p] : 0u, kp != 0);
    }
    __syncthreads();
    if (offs1) {
        // a single tile: its counts are the whole histogram, whose
        // exclusive offsets (and total, offs1[nb]) go out directly
        u32 tot;
        const u32 v = threadIdx.x < nb ? h[threadIdx.x] : 0u;
        const u32 ex = block_exclusive_scan<u32, PRA_DECIDE_BLOCK / 64>(v, sl, &tot);
        if (threadIdx.x < nb) offs1[threadIdx.x] = ex;
        if (threadIdx.x == 0) offs1[nb] = tot;
    } else {
        for (u32 i = threadIdx.x; i < nb; i += PRA_DECIDE_BLOCK) hist[(u64)i * nblocks + tile] = h[i];
    }
    __syncthreads();
}

// One output's decisions: blocks of PRA_DECIDE_BLOCK threads take tiles of
// PRA_TILE candidates in turn (each candidate is a chain of dependent loads,
// so many waves per tile hide them).
template <int MAXKW, bool EXACT>
__global__ __launch_bounds__(PRA_DECIDE_BLOCK) void k_dedup_decide_hist(const u64* cand, const u64* ctl, int cw,
                                                                        int kw_, const u64* table,
                                                                        const u64* slot_of, u64 epoch, u32* keep,
                                                                        const u32* dest, u32 nb, u32* hist,
                                                                        u32 nblocks, u32* flags, u32* offs1)
{
    if (ctl[1]) return;
    const u64 n = ctl[0];
    __shared__ u32 h[PRA_RADIX];
    __shared__ u32 sl[PRA_DECIDE_BLOCK / 64 + 1];
    for (u32 tile = blockIdx.x; tile < nblocks; tile += gridDim.x)
        pra_decide_tile<MAXKW, EXACT>(cand, n, cw, kw_, table, slot_of, epoch, keep, dest, nb, hist, nblocks, flags,
                                      offs1, tile, h, sl);
}

// Decisions of several outputs in one launch: block -> (job, tile) by the
// jobs' block ranges; no stride, every tile has its block.
template <int MAXKW, bool EXACT>
__global__ __launch_bounds__(PRA_DECIDE_BLOCK) void k_dedup_decide_many(DecideBatch b, u32* flags)
{
    __shared__ u32 h[PRA_RADIX];
    __shared__ u32 sl[PRA_DECIDE_BLOCK / 64 + 1];
    int k = 0;
    while (k + 1 < b.njobs && b.block_base[k + 1] <= blockIdx.x) k++;
    const DecideJob& j = b.job[k];
    if (j.ctl[1]) return;
    pra_decide_tile<MAXKW, EXACT>(j.cand, j.ctl[0], j.cw, j.kw, j.table, j.slot_of, j.epoch, j.keep, j.dest, b.nb,
                                  j.hist, j.nblocks, flags, j.offs1, blockIdx.x - b.block_base[k], h, sl);
}

// ----------------------------------------------------- dense membership ---

// rank index of a key of KW words, false when a value has no dictionary rank
template <int KW>
__device__ __forceinline__ bool pra_dense_index(const u64* key, const u64* dict, u64 dcap, u64 D, u64* idx)
{
    u64 v[KW];
#pragma unroll
    for (int c = 0; c < KW; c++) v[c] = key[c];
    u64 x = 0;
#pragma unroll
    for (int c = 0; c < KW; c++) {
        u64 r;
        if (!pra_dict_rank(dict, dcap, v[c], &r)) return false;
        x = x * D + r;
    }
    *idx = x;
    return true;
}

// a slot written in this iteration (the others count as empty)
__device__ __forceinline__ bool pra_dense_live(u64 v, const DenseArgs& a) { return v != 0 && (v & a.emask) == a.stamp; }

// A candidate whose key is a member (bit set) is dropped at once
// (slot_of = ~0); any other claims or finds the slot holding x + 1 and lowers
// the slot's arrival to its own position, so the earliest arrival wins.  A
// slot is one word, epoch | (x + 1) << PRA_DENSE_PBITS | arrival: the first
// candidate of a key claims an empty slot with one compare-and-swap, a later
// one reads it and swaps only while its own arrival is earlier.
template <int KW>
__device__ __forceinline__ void pra_dense_claim_one(const DenseArgs& a, u64 p, u32* flags)
{
    u64 x;
    if (!pra_dense_index<KW>(a.cand + p * a.cw, a.dict, a.dict_cap, a.D, &x)) {
        atomicCAS((unsigned long long*)(a.ctl + 1), 0ULL, 3ULL);
        return;
    }
    if ((a.bitmap[x >> 5] >> (x & 31)) & 1u) {
        a.slot_of[p] = ~0ULL;
        return;
    }
    const u64 key = x + 1, m = a.cap - 1, mine = a.stamp | (key << PRA_DENSE_PBITS) | p;
    u64 s = pra_mix64(key) & m;
    for (u64 q = 0; q < a.cap; q++) {
        unsigned long long* slot = (unsigned long long*)(a.table + s);
        // a plain load first: a slot held by another key then costs no
        // atomic (swapping straight away measured slower)
        u64 cur = *slot;
        if (!pra_dense_live(cur, a)) {
            // empty: every other write in this launch leaves a live slot, so
            // a failed swap returns one
            const u64 prev = atomicCAS(slot, (unsigned long long)cur, (unsigned long long)mine);
            if (prev == cur) {
                a.slot_of[p] = s;
                return;
            }
            cur = prev;
        }
        if (((cur & ~a.emask) >> PRA_DENSE_PBITS) == key) {
            // the key's slot: lower its arrival while ours is earlier (the
            // key bits never change once set); a candidate that never got in
            // has lost to an earlier arrival for good, and the decision skips it
            bool in = false;
            while ((cur & PRA_DENSE_PMASK) > p) {
                const u64 prev = atomicCAS(slot, (unsigned long long)cur, (unsigned long long)mine);
                if (prev == cur) {
                    in = true;
                    break;
                }
                cur = prev;
            }
            a.slot_of[p] = in ? s : ~0ULL;
            return;
        }
        s = (s + 1) & m;
    }
    a.slot_of[p] = ~0ULL;
    atomicOr(flags, PRA_FLAG_TABLE_FULL);
}

template <int KW>
__global__ void k_dense_claim(DenseArgs a, u32* flags)
{
    if (a.ctl[1]) return;
    const u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.ctl[0]) return;
    pra_dense_claim_one<KW>(a, p, flags);
}

// Dense claims of several outputs (of one key width) in one launch; the job
// table in device memory is followed by the claim blocks' and the decision
// blocks' first block per job (njobs + 1 words each).
template <int KW>
__global__ void k_dense_claim_many(const DenseArgs* jobs, int njobs, u32* flags)
{
    const u32* cbase = (const u32*)(jobs + njobs);
    const int k = pra_job_of(cbase, njobs, blockIdx.x);
    const DenseArgs* jp = jobs + k;
    if (jp->ctl[1]) return;
    const u64 p = (u64)(blockIdx.x - cbase[k]) * blockDim.x + threadIdx.x;
    if (p >= jp->ctl[0]) return;
    pra_dense_claim_one<KW>(*jp, p, flags);
}

// Tile `tile` of the dense decision: the earliest arrival of each claimed key
// keeps it, sets its membership bit and returns its slot to empty; kept
// candidates are counted per destination bucket as in pra_decide_tile.
__device__ __forceinline__ void pra_dense_decide_tile(const DenseArgs& a, u64 n, u32 tile, u32* h, u32* sl)
{
    const u64 base = (u64)tile * PRA_TILE;
    if (base >= n) {
        for (u32 i = threadIdx.x; i < a.nb; i += PRA_DECIDE_BLOCK) a.hist[(u64)i * a.nblocks + tile] = 0;
        if (a.offs1)
            for (u32 i = threadIdx.x; i <= a.nb; i += PRA_DECIDE_BLOCK) a.offs1[i] = 0;
        return;
    }
    for (int i = threadIdx.x; i < PRA_RADIX; i += PRA_DECIDE_BLOCK) h[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PRA_TILE / PRA_DECIDE_BLOCK; k++) {
        const u64 p = base + (u64)k * PRA_DECIDE_BLOCK + threadIdx.x;
        u32 kp = 0, d = 0;
        if (p < n) {
            const u64 s = a.slot_of[p];
            if (s != ~0ULL) {
                u64* slot = a.table + s;
                const u64 v = *slot;
                // the earliest arrival of the key keeps it; without epochs
                // the slot returns to empty, which no later arrival's
                // position matches
                if (pra_dense_live(v, a) && (v & PRA_DENSE_PMASK) == p) {
                    kp = 1;
                    d = a.dest[p];
                    const u64 x = ((v & ~a.emask) >> PRA_DENSE_PBITS) - 1;
                    atomicOr(&a.bitmap[x >> 5], 1u << (x & 31));
                    if (!a.emask) *slot = 0;
                    // the rank index, for the newt scatter's packed key
                    a.slot_of[p] = x;
                }
            }
            a.keep[p] = kp;
        }
        pra_wave_count(h, d, kp != 0);
    }
    __syncthreads();
    if (a.offs1) {
        u32 tot;
        const u32 v = threadIdx.x < a.nb ? h[threadIdx.x] : 0u;
        const u32 ex = block_exclusive_scan<u32, PRA_DECIDE_BLOCK / 64>(v, sl, &tot);
        if (threadIdx.x < a.nb) a.offs1[threadIdx.x] = ex;
        if (threadIdx.x == 0) a.offs1[a.nb] = tot;
    } else {
        for (u32 i = threadIdx.x; i < a.nb; i += PRA_DECIDE_BLOCK) a.hist[(u64)i * a.nblocks + tile] = h[i];
    }
    __syncthreads();
}

__global__ __launch_bounds__(PRA_DECIDE_BLOCK) void k_dense_decide(DenseArgs a)
{
    if (a.ctl[1]) return;
    const u64 n = a.ctl[0];
    __shared__ u32 h[PRA_RADIX];
    __shared__ u32 sl[PRA_DECIDE_BLOCK / 64 + 1];
    for (u32 tile = blockIdx.x; tile < a.nblocks; tile += gridDim.x) pra_dense_decide_tile(a, n, tile, h, sl);
}

// Dense decisions of several outputs in one launch: one block per tile of
// each job (table as k_dense_claim_many).
__global__ __launch_bounds__(PRA_DECIDE_BLOCK) void k_dense_decide_many(const DenseArgs* jobs, int njobs)
{
    __shared__ u32 h[PRA_RADIX];
    __shared__ u32 sl[PRA_DECIDE_BLOCK / 64 + 1];
    const u32* dbase = (const u32*)(jobs + njobs) + njobs + 1;
    const int k = pra_job_of(dbase, njobs, blockIdx.x);
    const DenseArgs a = jobs[k];
    if (a.ctl[1]) return;
    pra_dense_decide_tile(a, a.ctl[0], blockIdx.x - dbase[k], h, sl);
}

// the membership bits of n rows (key = their first kw columns)
template <int KW>
__global__ void k_dense_build(const u64* rows, u64 n, int w, const u64* dict, u64 dcap, u64 D, u32* bitmap, u64* miss)
{
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 x;
    if (!pra_dense_index<KW>(rows + i * w, dict, dcap, D, &x)) {
        atomicAdd((unsigned long long*)miss, 1ULL);
        return;
    }
    atomicOr(&bitmap[x >> 5], 1u << (x & 31));
}

// Combine the PRA_NREP replicas of the scatter summaries and reset them to
// their identity values for the next scatter.  Part 0 (packed keys): the
// top-byte bin counters, whose exclusive offsets go to kgofs and whose
// largest count goes to stat[6]; the counters return to zero.  Part 1:
// summary words -> stat[1..3] (packed key OR / AND / dictionary miss) and
// stat[8..39] (key column OR / AND).  Part 2: the kept total -> stat[0], and
// the seq counters advance by the new tuples of each bucket (bucket d's rows
// start at offs[d * nblocks]).  PRA_BLOCK threads; counters and replicas are
// read by atomic exchange, so the scatter's own block may run it too.
__device__ inline void pra_newt_fold_part(const NewtFoldArgs& a, u32 y, u32 t, u32* lds, u32* mx)
{
    if (y == 0) {
        if (!a.packed) return;
        const u32 v = atomicExch(&a.bin_cur[t], 0u);
        u32 tot;
        a.kgofs[t] = block_exclusive_scan<u32>(v, lds, &tot);
        if (t == 0) *mx = 0;
        __syncthreads();
        atomicMax(mx, v);
        __syncthreads();
        if (t == 0) a.stat[6] = *mx;
    } else if (y == 1) {
        if (t < PRA_PART_WORDS && !(t >= 3 && t < 8)) {
            const bool is_and = pra_summary_is_and(t);
            const u64 id = is_and ? ~0ULL : 0;
            u64 v = id;
            for (int r = 0; r < PRA_NREP; r++) {
                unsigned long long* p = (unsigned long long*)(a.rep_words + (u64)r * PRA_PART_WORDS + t);
                const u64 x = atomicExch(p, (unsigned long long)id);
                v = is_and ? (v & x) : (v | x);
            }
            a.stat[t < 3 ? t + 1 : t] = v;
        }
    } else {
        if (t == 0) a.stat[0] = *a.total;
        if (t < a.nb && a.intern) {
            const u32 start = a.offs[(u64)t * a.nblocks];
            const u32 end = t + 1 < a.nb ? a.offs[(u64)(t + 1) * a.nblocks] : *a.total;
            a.seqbase[t] += (u64)(end - start);
        }
    }
}

// k_newt_fold's three parts in one block after the scatter of a single tile
__device__ inline void pra_scatter_fold_tail(const NewtFoldArgs& f)
{
    __shared__ u32 flds[PRA_WAVES + 1];
    __shared__ u32 fmx;
    __syncthreads();
    for (u32 y = 0; y < 3; y++) {
        pra_newt_fold_part(f, y, threadIdx.x, flds, &fmx);
        __syncthreads();
    }
}

// Stable partition of the kept candidates by destination bucket; JOIN-type
// rows get their intern id tag<<46 | bucket<<28 | seq, seq counted per
// (tag, bucket) in arrival order (src/RAM/RA_tasks.cpp:858-875).
template <bool PACKED, int SB>
__device__ __forceinline__ void pra_scatter_block(NewtScatterArgs& a, const u32 bid)
{
    // PRA_TILE candidates per block: SB threads of IT items, SW waves
    constexpr int SW = SB / 64, IT = PRA_TILE / SB;
    if (a.ctl[1]) return;
    a.n = a.ctl[0];
    // blocks past the candidates hold nothing; the summary reduce skips them
    if ((u64)bid * PRA_TILE >= a.n) {
        if (SB == PRA_BLOCK && a.fold_here) pra_scatter_fold_tail(a.fold);
        return;
    }
    __shared__ u32 wh[SW][PRA_RADIX];
    for (int i = threadIdx.x; i < SW * PRA_RADIX; i += SB) (&wh[0][0])[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 lt = pra_lanemask_lt();
    u64 base = (u64)bid * PRA_TILE + (u64)wave * (PRA_TILE / SW);
    u32 dg[IT], lc[IT], kept = 0;
#pragma unroll
    for (int k = 0; k < IT; k++) {
        u64 p = base + (u64)k * 64 + lane;
        bool valid = p < a.n && a.keep[p];
        kept |= (valid ? 1u : 0u) << k;
        u32 d = valid ? a.dest[p] : 0u;
        u64 act = __ballot(valid);
        u64 peers = pra_match<PRA_RADIX_BITS>(d, act);
        dg[k] = d;
        lc[k] = 0;
        if (valid) {
            u32 before = wh[wave][d];
            lc[k] = before + (u32)__popcll(peers & lt);
            if ((peers >> lane) == 1ULL) wh[wave][d] = before + (u32)__popcll(peers);
        }
    }
    // the wide form loads the kept rows' words (up to PW) and rank indexes
    // now, so that they arrive while the block forms its offsets and bin
    // counts (loads are not moved across the barriers in between)
    constexpr int PW = SB == PRA_SCATTER_BLOCK ? 2 : 0;
    const bool pre_rows = PW > 0 && a.cw <= PW;
    u64 prow[IT][PW > 0 ? PW : 1], pxk[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const u64 p = base + (u64)k * 64 + lane;
        pxk[k] = 0;
        if ((kept >> k) & 1u) {
            if (PACKED && PW > 0 && a.xk) pxk[k] = a.xk[p];
            if (pre_rows)
#pragma unroll
                for (int c = 0; c < (PW > 0 ? PW : 1); c++)
                    if (c < a.cw) prow[k][c] = a.cand[p * a.cw + c];
        }
    }
    __syncthreads();
    for (u32 d = threadIdx.x; d < a.nb; d += SB) {
        u32 acc = a.offs[(u64)d * a.nblocks + bid];
        for (int v = 0; v < SW; v++) {
            u32 t = wh[v][d];
            wh[v][d] = acc;
            acc += t;
        }
    }
    __syncthreads();
    // per-block summaries: the packed key's OR / AND / dictionary miss and the
    // block's share of each top-byte bin (packed mode), or the key columns'
    // OR / AND
    __shared__ u32 khist[PRA_RADIX];
    __shared__ u64 red[SW][PRA_PART_WORDS];
    constexpr bool packed = PACKED;
    if (packed)
        for (int i = threadIdx.x; i < PRA_RADIX; i += SB) khist[i] = 0;
    __syncthreads();
    // key columns summarised (unpacked, when asked): the first 16; later ones
    // report the identity values (OR 0, AND all ones), which the host reads as
    // "varies"
    constexpr int NC = PACKED ? 1 : 16;
    const int kc = !a.summarize ? 0 : a.cw < NC ? a.cw : NC;
    u64 kor = 0, kand = ~0ULL, kmiss = 0;
    u64 o[NC], an[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) { o[c] = 0; an[c] = ~0ULL; }
    // packed keys: from the rank index the dense decision left, or else the
    // column ranks of all items are looked up column by column with every
    // item's first probe in flight at once (a dictionary slot resolves at its
    // first probe unless its chain is longer)
    u64 K[PACKED ? IT : 1];
    u32 lr[PACKED ? IT : 1];
    if (packed) {
        if (a.xk) {
#pragma unroll
            for (int k = 0; k < IT; k++) {
                K[k] = 0;
                if ((kept >> k) & 1u)
                    K[k] = (u64)dg[k] * a.xk_scale + (PW > 0 ? pxk[k] : a.xk[base + (u64)k * 64 + lane]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < IT; k++) K[k] = dg[k];
            for (int c = 0; c < a.cw; c++) {
                u64 v[IT], e[IT], r1[IT];
#pragma unroll
                for (int k = 0; k < IT; k++) {
                    v[k] = e[k] = r1[k] = 0;
                    if ((kept >> k) & 1u) v[k] = a.cand[(base + (u64)k * 64 + lane) * a.cw + c];
                }
#pragma unroll
                for (int k = 0; k < IT; k++)
                    if ((kept >> k) & 1u) {
                        const u64* sl = a.dict + 2 * pra_dict_slot(v[k], a.dict_cap);
                        e[k] = sl[0];
                        r1[k] = sl[1];
                    }
#pragma unroll
                for (int k = 0; k < IT; k++)
                    if ((kept >> k) & 1u) {
                        u64 r = 0;
                        if (r1[k] != 0 && e[k] == v[k])
                            r = r1[k] - 1;
                        else if (!pra_dict_rank(a.dict, a.dict_cap, v[k], &r))
                            kmiss = 1;
                        K[k] = K[k] * a.rank_radix + r;
                    }
            }
        }
        // every row's (key, newt position) goes to the bin of the key's top
        // byte: a rank inside the block's share of the bin, then one counter
        // reservation per bin and block
#pragma unroll
        for (int k = 0; k < IT; k++) {
            K[k] <<= a.key_lshift;
            if ((kept >> k) & 1u) lr[k] = atomicAdd(&khist[K[k] >> 56], 1u);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < PRA_RADIX; i += SB)
            if (khist[i]) khist[i] = atomicAdd(a.bin_cur + i, khist[i]);
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < IT; k++) {
        u64 p = base + (u64)k * 64 + lane;
        if ((kept >> k) & 1u) {
            u32 d = dg[k];
            u64 pos = (u64)wh[wave][d] + lc[k];
            const u64* src = a.cand + p * a.cw;
            u64* dst = a.newt + pos * a.wn;
            if (pre_rows) {
#pragma unroll
                for (int c = 0; c < (PW > 0 ? PW : 1); c++)
                    if (c < a.cw) dst[c] = prow[k][c];
            } else {
                pra_copy_row(dst, src, a.cw);
            }
            if (a.intern) {
                u64 rank = pos - a.offs[(u64)d * a.nblocks];
                u64 key = ((u64)a.tag << 46) | ((u64)d << 28);
                dst[a.cw] = key | (a.seqbase[d] + rank);
            }
            if (packed) {
                const u64 Kk = K[PACKED ? k : 0];
                a.keys[pos] = Kk;
                kor |= Kk;
                kand &= Kk;
                const u32 bin = (u32)(Kk >> 56), slot = khist[bin] + lr[PACKED ? k : 0];
                if (slot < a.bin_cap) {
                    a.bin_keys[(u64)bin * a.bin_cap + slot] = Kk;
                    a.bin_idx[(u64)bin * a.bin_cap + slot] = (u32)pos;
                }
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++)
                    if (c < kc) { o[c] |= src[c]; an[c] &= src[c]; }
            }
        }
    }
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) {
        kor |= __shfl_xor(kor, dd, 64);
        kand &= __shfl_xor(kand, dd, 64);
        kmiss |= __shfl_xor(kmiss, dd, 64);
    }
    if (!packed) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (c < kc) {
#pragma unroll
                for (int dd = 32; dd >= 1; dd >>= 1) {
                    o[c] |= __shfl_xor(o[c], dd, 64);
                    an[c] &= __shfl_xor(an[c], dd, 64);
                }
            }
        }
    }
    if (lane == 0) {
        red[wave][0] = kor;
        red[wave][1] = kand;
        red[wave][2] = kmiss;
#pragma unroll
        for (int c = 0; c < 16; c++) {
            red[wave][8 + c] = (!packed && c < kc) ? o[c < NC ? c : 0] : 0;
            red[wave][24 + c] = (!packed && c < kc) ? an[c < NC ? c : 0] : ~0ULL;
        }
    }
    __syncthreads();
    // fold into replica bid % PRA_NREP of the summaries (identity
    // values are not sent); k_newt_fold combines and resets the replicas
    const u32 rep = bid % PRA_NREP;
    if (threadIdx.x < PRA_PART_WORDS) {
        const int j = threadIdx.x;
        if (j < 3 || j >= 8) {
            u64 v = red[0][j];
            bool is_and = pra_summary_is_and(j);
            for (int w2 = 1; w2 < SW; w2++) v = is_and ? (v & red[w2][j]) : (v | red[w2][j]);
            unsigned long long* dst = (unsigned long long*)(a.rep_words + (u64)rep * PRA_PART_WORDS + j);
            if (is_and && v != ~0ULL) atomicAnd(dst, (unsigned long long)v);
            if (!is_and && v != 0) atomicOr(dst, (unsigned long long)v);
        }
    }
    if (SB == PRA_BLOCK && a.fold_here) pra_scatter_fold_tail(a.fold);
}

template <bool PACKED, int SB>
__global__ __launch_bounds__(SB) void k_dedup_scatter(NewtScatterArgs a)
{
    pra_scatter_block<PACKED, SB>(a, blockIdx.x);
}

// Partitions of several outputs' candidates in one launch (PRA_BLOCK-thread
// blocks): the jobs' arguments and block ranges are read from device memory
// (njobs arguments, then njobs + 1 block bases), a block serving the job whose
// range holds it as k_dedup_scatter serves its one output.
template <bool PACKED>
__global__ __launch_bounds__(PRA_BLOCK) void k_dedup_scatter_many(const NewtScatterArgs* jobs, int njobs)
{
    const u32* block_base = (const u32*)(jobs + njobs);
    // the job: the number of later jobs' block bases at or below this block,
    // 64 bases per load round
    int k = 0;
    for (int j0 = 1; j0 < njobs; j0 += 64) {
        const int j = j0 + (int)(threadIdx.x & 63);
        k += __popcll(__ballot(j < njobs && block_base[j] <= blockIdx.x));
    }
    NewtScatterArgs a = jobs[k];
    pra_scatter_block<PACKED, PRA_BLOCK>(a, blockIdx.x - block_base[k]);
}

// The folds of several outputs in one launch: block b runs part b % 3 of job
// b / 3 (job table in device memory).
__global__ void k_newt_fold_many(const NewtFoldArgs* jobs)
{
    const NewtFoldArgs& a = jobs[blockIdx.x / 3];
    if (a.ctl[1]) return;
    __shared__ u32 lds[PRA_WAVES + 1];
    __shared__ u32 mx;
    pra_newt_fold_part(a, blockIdx.x % 3, threadIdx.x, lds, &mx);
}

// One block per part of pra_newt_fold_part.
__global__ void k_newt_fold(NewtFoldArgs a)
{
    if (a.ctl[1]) return;
    __shared__ u32 lds[PRA_WAVES + 1];
    __shared__ u32 mx;
    pra_newt_fold_part(a, blockIdx.x, threadIdx.x, lds, &mx);
}

// The last output's fold and then the iteration's summary publication
// (k_publish) in one block: the fold's parts run in turn (skipped, as in
// k_newt_fold, when the output overflowed its candidate capacity), then
// `words` summary words and the error flags go to pinned host memory.
__global__ void k_newt_fold_publish(NewtFoldArgs a, const u64* src, u64 words, const u32* flags, u64* host, u64 seq)
{
    __shared__ u32 lds[PRA_WAVES + 1];
    __shared__ u32 mx;
    if (!a.ctl[1])
        for (u32 y = 0; y < 3; y++) {
            pra_newt_fold_part(a, y, threadIdx.x, lds, &mx);
            __syncthreads();
        }
    pra_publish_words(src, words, flags, host, seq);
}

// identity values of the summary replicas (all counts 0, OR words 0, AND
// words all ones)
__global__ void k_newt_rep_init(u64* rep_words)
{
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < PRA_NREP * PRA_PART_WORDS) rep_words[i] = pra_summary_is_and(i % PRA_PART_WORDS) ? ~0ULL : 0;
}

// ----------------------------------------- membership table maintenance -----

// Insert keys of already-present rows (minpos 0).  Duplicate keys (rows that
// differ only after the key) land on the same slot and are skipped.
__global__ void k_table_insert_rows(const u64* rows, u64 n, int w, int kw, u64* table, u64 cap, u32* flags)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64* key = rows + i * w;
    const int stride = 2 + kw;
    u64 h = pra_keyhash(key, kw);
    u64 m = cap - 1, s = h & m;
    for (u64 q = 0; q < cap; q++) {
        u64* slot = table + s * stride;
        u64 cur = atomicCAS((unsigned long long*)slot, 0ULL, (unsigned long long)h);
        if (cur == 0) {
            slot[1] = 0;
            for (int c = 0; c < kw; c++) slot[2 + c] = key[c];
            return;
        }
        if (cur == h) return;
        s = (s + 1) & m;
    }
    atomicOr(flags, PRA_FLAG_TABLE_FULL);
}

__global__ void k_table_verify_rows(const u64* rows, u64 n, int w, int kw, const u64* table, u64 cap, u32* flags)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64* key = rows + i * w;
    const int stride = 2 + kw;
    u64 h = pra_keyhash(key, kw);
    u64 m = cap - 1, s = h & m;
    for (u64 q = 0; q < cap; q++) {
        const u64* slot = table + s * stride;
        if (slot[0] == h) {
            for (int c = 0; c < kw; c++)
                if (slot[2 + c] != key[c]) { atomicOr(flags, PRA_FLAG_COLLISION); break; }
            return;
        }
        if (slot[0] == 0) { atomicOr(flags, PRA_FLAG_COLLISION); return; }
        s = (s + 1) & m;
    }
}

__global__ void k_table_rehash(const u64* old, u64 oldcap, u64* nt, u64 cap, int kw, u32* flags)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= oldcap) return;
    const int stride = 2 + kw;
    const u64* os = old + i * stride;
    u64 h = os[0];
    if (h == 0) return;
    u64 m = cap - 1, s = h & m;
    for (u64 q = 0; q < cap; q++) {
        u64* slot = nt + s * stride;
        u64 cur = atomicCAS((unsigned long long*)slot, 0ULL, (unsigned long long)h);
        if (cur == 0) {
            slot[1] = 0;
            for (int c = 0; c < kw; c++) slot[2 + c] = os[2 + c];
            return;
        }
        s = (s + 1) & m;
    }
    atomicOr(flags, PRA_FLAG_TABLE_FULL);
}

// ------------------------------------------------------- launch wrappers ---

static inline unsigned grid_for(u64 n, unsigned bs = PRA_BLOCK)
{
    u64 g = (n + bs - 1) / bs;
    return (unsigned)(g ? g : 1);
}

void pra_launch_fill_u64(hipStream_t st, u64* p, u64 n, u64 v)
{
    if (!n) return;
    u64 g = (n + PRA_BLOCK - 1) / PRA_BLOCK;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)g), dim3(PRA_BLOCK), 0, st, p, n, v);
}

void pra_launch_table_clear(hipStream_t st, u64* t, u64 cap, int stride)
{
    u64 g = (cap + PRA_BLOCK - 1) / PRA_BLOCK;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_table_clear, dim3((unsigned)(g ? g : 1)), dim3(PRA_BLOCK), 0, st, t, cap, stride);
}


// one block scans inputs of up to PRA_SMALL_SCAN elements, every thread
// holding a run of consecutive elements in registers
template <typename T>
__global__ void k_scan_small(const T* in, T* out, u64 n, T* total)
{
    constexpr int PER = PRA_SMALL_SCAN / PRA_BLOCK;
    __shared__ T lds[PRA_WAVES + 1];
    const u64 base = (u64)threadIdx.x * PER;
    T v[PER];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        sum += v[k];
    }
    T tot;
    T ex = block_exclusive_scan<T>(sum, lds, &tot);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

// Exclusive scans of several u32 arrays of at most PRA_SMALL_SCAN elements in
// one launch, one block each (job table in device memory).
__global__ void k_scan_small_many(const ScanJob* jobs)
{
    const ScanJob& j = jobs[blockIdx.x];
    constexpr int PER = PRA_SMALL_SCAN / PRA_BLOCK;
    __shared__ u32 lds[PRA_WAVES + 1];
    const u64 base = (u64)threadIdx.x * PER;
    u32 v[PER];
    u32 sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        v[k] = base + k < j.n ? j.in[base + k] : 0;
        sum += v[k];
    }
    u32 tot;
    u32 ex = block_exclusive_scan<u32>(sum, lds, &tot);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        if (base + k < j.n) j.out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0 && j.total) *j.total = tot;
}

// Word copies of several ranges in one launch: job k takes blocks
// [block_base[k], block_base[k + 1]), 8 words per thread (block_base follows
// the job table in device memory).
__global__ void k_copy_many(const CopyJob* jobs, int njobs)
{
    const u32* block_base = (const u32*)(jobs + njobs);
    const int k = pra_job_of(block_base, njobs, blockIdx.x);
    const CopyJob j = jobs[k];
    const u64 i0 = ((u64)(blockIdx.x - block_base[k]) * PRA_BLOCK + threadIdx.x) * 8;
    u64 t[8];
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (i0 + q < j.words) t[q] = j.src[i0 + q];
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (i0 + q < j.words) j.dst[i0 + q] = t[q];
}

// Exclusive prefix of `tile` over its predecessors' published values (digit
// 0 of the status words), walked by the 64 lanes of one wave: each step reads
// the 64 nearest unconsumed predecessors at once and adds the ready ones up to
// the nearest inclusive prefix.  Every lane returns the same value.
__device__ inline u64 pra_lookback_wave(const u64* status, u32 tile, u32 epoch, u32* flags)
{
    constexpr u64 VAL = (1ULL << 46) - 1;
    const int lane = __lane_id();
    u64 excl = 0;
    long long t = (long long)tile - 1;
    u32 spins = 0;
    while (t >= 0) {
        long long j = t - lane;
        u64 v = j >= 0 ? __hip_atomic_load(status + (u64)j * PRA_RADIX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                       : st_pack(epoch, PRA_ST_INC, 0);
        u64 fl = (v >> 46) & 3ULL;
        bool ready = (u32)(v >> 48) == epoch && fl != 0;
        u64 notready = __ballot(!ready);
        u64 inc = __ballot(ready && fl == PRA_ST_INC);
        int first_nr = notready ? __ffsll((unsigned long long)notready) - 1 : 64;
        int first_inc = inc ? __ffsll((unsigned long long)inc) - 1 : 64;
        bool stop = first_inc < first_nr;
        int take = stop ? first_inc + 1 : first_nr;
        u64 x = lane < take ? (v & VAL) : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        excl += x;
        if (stop) break;
        t -= take;
        if (take == 0 && ++spins > PRA_LOOKBACK_SPIN_LIMIT) {
            if (lane == 0) atomicOr(flags, PRA_FLAG_SORT_STALL);
            break;
        }
    }
    return excl;
}

// One-pass exclusive scan: tiles of PRA_TILE elements are taken in ticket
// order; each publishes its sum, finds the sum of the tiles before it by
// decoupled look-back (digit 0 of the sort's status words), publishes its
// inclusive prefix and writes its outputs.  Prefixes must stay below 2^46
// (PRA_FLAG_SCAN_OVERFLOW otherwise).
template <typename T>
__global__ void k_scan_lookback(const T* in, T* out, u64 n, u64* status, u32* tile_ctr, u32 epoch, u32* flags,
                                T* total)
{
    __shared__ T lds[PRA_WAVES + 1];
    __shared__ u32 s_tile;
    __shared__ u64 s_excl;
    if (threadIdx.x == 0) {
        u32 t = atomicAdd(tile_ctr, 1u);
        if (t == gridDim.x - 1) atomicExch(tile_ctr, 0u);
        s_tile = t;
    }
    __syncthreads();
    const u32 tile = s_tile;
    const u64 base = (u64)tile * PRA_TILE + (u64)threadIdx.x * PRA_ITEMS;
    T v[PRA_ITEMS];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < PRA_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        sum += v[k];
    }
    T tot;
    T ex = block_exclusive_scan<T>(sum, lds, &tot);
    if (threadIdx.x < 64) {
        constexpr u64 LIM = 1ULL << 46;
        u64* my = status + (u64)tile * PRA_RADIX;
        const u64 agg = (u64)tot;
        u64 excl = 0;
        if (threadIdx.x == 0) {
            if (agg >= LIM) atomicOr(flags, PRA_FLAG_SCAN_OVERFLOW);
            __hip_atomic_store(my, st_pack(epoch, tile == 0 ? PRA_ST_INC : PRA_ST_AGG, agg & (LIM - 1)),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (tile != 0) {
            excl = pra_lookback_wave(status, tile, epoch, flags);
            if (threadIdx.x == 0) {
                if (excl + agg >= LIM) atomicOr(flags, PRA_FLAG_SCAN_OVERFLOW);
                __hip_atomic_store(my, st_pack(epoch, PRA_ST_INC, (excl + agg) & (LIM - 1)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (threadIdx.x == 0) {
            s_excl = excl;
            if (tile == gridDim.x - 1 && total) *total = (T)(excl + agg);
        }
    }
    __syncthreads();
    T run = (T)s_excl + ex;
#pragma unroll
    for (int k = 0; k < PRA_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// One tile (`tile` of `ntiles`, status words from `status`) of one rule's
// per-row output counts and their exclusive prefix sums.  PRA_CS_INDEX
// (join): a row's count is the length of its key's run of inner rows, found
// in the static inner's key index (istart = the run's first row);
// PRA_CS_SORTED (join): the same run searched in the inner's sorted rows;
// PRA_CS_FILTER (copy, copy_filter): 1 for rows passing the filter `op`.  A tile holds PRA_TILE rows (PRA_TILE /
// PRA_COUNT_SCAN_BLOCK consecutive rows per thread) and finds its offset by
// decoupled look-back over digit 0 of the status words, as k_scan_lookback
// does; the last tile writes off[n] = total.
template <int KIND>
__device__ inline void pra_count_scan_tile(const CountScanArgs& a, u64* status, u32 tile, u32 ntiles, u32 epoch,
                                           u32* flags, u64* lds, u64* s_excl)
{
    constexpr int IT = PRA_TILE / PRA_COUNT_SCAN_BLOCK;
    constexpr int NW = PRA_COUNT_SCAN_BLOCK / 64;
    const u64 base = (u64)tile * PRA_TILE + (u64)threadIdx.x * IT;
    u64 c[IT];
    if (KIND == PRA_CS_SORTED) {
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const u64 i = base + k;
            c[k] = 0;
            if (i < a.n) {
                u64 first;
                c[k] = pra_sorted_run(a.outer + i * a.wo, a.jcc, a.inner, a.wi, a.ibnd, a.nb, &first);
                a.istart[i] = first;
            }
        }
    } else if (KIND == PRA_CS_INDEX) {
        // every item's key hash and first index slot are loaded before any
        // item walks on (a probe mostly ends at its first slot)
        const u64 m = a.cap - 1;
        u64 h[IT], f0[IT], f1[IT], f2[IT];
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const u64 i = base + k;
            h[k] = i < a.n && a.cap ? pra_keyhash(a.outer + i * a.wo, a.jcc) : 0;
        }
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const u64* sl = a.table + (h[k] & m) * 3;
            f0[k] = h[k] ? sl[0] : 0;
            f1[k] = h[k] ? sl[1] : 0;
            f2[k] = h[k] ? sl[2] : 0;
        }
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const u64 i = base + k;
            u64 cnt = 0, st = 0;
            if (h[k]) {
                const u64* key = a.outer + i * a.wo;
                u64 s = h[k] & m, sh = f0[k], ss = f1[k], sc = f2[k];
                for (u64 q = 0; q < a.cap && sh != 0; q++) {
                    if (sh == h[k]) {
                        const u64* f = a.inner + ss * a.wi;
                        bool eq = true;
                        for (int x = 0; x < a.jcc; x++)
                            if (f[x] != key[x]) { eq = false; break; }
                        if (eq) { cnt = sc; st = ss; }
                        break;
                    }
                    s = (s + 1) & m;
                    const u64* sl = a.table + s * 3;
                    sh = sl[0];
                    ss = sl[1];
                    sc = sl[2];
                }
            }
            c[k] = cnt;
            if (i < a.n) a.istart[i] = st;
        }
    } else {
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const u64 i = base + k;
            const u64* r = a.outer + i * a.wo;
            c[k] = i < a.n && pra_filter(a.op, r[0], r[1]) ? 1 : 0;
        }
    }
    u64 sum = 0;
#pragma unroll
    for (int k = 0; k < IT; k++) sum += c[k];
    u64 tot;
    const u64 ex = block_exclusive_scan<u64, NW>(sum, lds, &tot);
    if (threadIdx.x < 64) {
        constexpr u64 LIM = 1ULL << 46;
        u64* my = status + (u64)tile * PRA_RADIX;
        u64 excl = 0;
        if (threadIdx.x == 0) {
            if (tot >= LIM) atomicOr(flags, PRA_FLAG_SCAN_OVERFLOW);
            __hip_atomic_store(my, st_pack(epoch, tile == 0 ? PRA_ST_INC : PRA_ST_AGG, tot & (LIM - 1)),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (tile != 0) {
            excl = pra_lookback_wave(status, tile, epoch, flags);
            if (threadIdx.x == 0) {
                if (excl + tot >= LIM) atomicOr(flags, PRA_FLAG_SCAN_OVERFLOW);
                __hip_atomic_store(my, st_pack(epoch, PRA_ST_INC, (excl + tot) & (LIM - 1)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (threadIdx.x == 0) {
            *s_excl = excl;
            if (tile == ntiles - 1) a.off[a.n] = excl + tot;
        }
    }
    __syncthreads();
    u64 run = *s_excl + ex;
#pragma unroll
    for (int k = 0; k < IT; k++) {
        if (base + k < a.n) a.off[base + k] = run;
        run += c[k];
    }
}

// Counts and scans of several rules in one launch: tiles are taken in ticket
// order; a block's ticket names its rule (the job whose tile range holds it)
// and its tile within that rule, whose look-back covers that rule's tiles.
__global__ __launch_bounds__(PRA_COUNT_SCAN_BLOCK) void k_count_scan_many(CountScanBatch b, u64* status,
                                                                          u32* tile_ctr, u32 epoch, u32* flags)
{
    __shared__ u64 lds[PRA_COUNT_SCAN_BLOCK / 64 + 1];
    __shared__ u32 s_tile;
    __shared__ int s_job;
    __shared__ u64 s_excl;
    if (threadIdx.x == 0) {
        u32 t = atomicAdd(tile_ctr, 1u);
        if (t == gridDim.x - 1) atomicExch(tile_ctr, 0u);
        int k = 0;
        while (k + 1 < b.njobs && b.tile_base[k + 1] <= t) k++;
        s_job = k;
        s_tile = t;
    }
    __syncthreads();
    const int k = s_job;
    const CountScanJob& j = b.job[k];
    CountScanArgs a;
    a.outer = j.outer;
    a.table = j.table;
    a.inner = j.inner;
    a.off = j.off;
    a.istart = j.istart;
    a.n = j.n;
    a.cap = j.cap;
    a.wo = j.wo;
    a.wi = j.wi;
    a.jcc = j.jcc;
    a.op = j.op;
    a.ibnd = j.ibnd;
    a.nb = j.nb;
    const u32 t0 = b.tile_base[k], nt = b.tile_base[k + 1] - t0;
    u64* st = status + (u64)t0 * PRA_RADIX;
    if (j.kind == PRA_CS_INDEX)
        pra_count_scan_tile<PRA_CS_INDEX>(a, st, s_tile - t0, nt, epoch, flags, lds, &s_excl);
    else if (j.kind == PRA_CS_SORTED)
        pra_count_scan_tile<PRA_CS_SORTED>(a, st, s_tile - t0, nt, epoch, flags, lds, &s_excl);
    else
        pra_count_scan_tile<PRA_CS_FILTER>(a, st, s_tile - t0, nt, epoch, flags, lds, &s_excl);
}

// kernel arguments of one launch are limited to 4 KiB
static_assert(sizeof(CountScanBatch) + 4 * sizeof(void*) <= 4096, "count-scan batch exceeds the kernel arguments");
static_assert(sizeof(SmallSortBatch) <= 4096, "small-sort batch exceeds the kernel arguments");
static_assert(sizeof(PlanAllArgs) <= 4096, "plan batch exceeds the kernel arguments");

void pra_launch_count_scan_many(hipStream_t st, const CountScanBatch& b, u64* status, u32* tile_ctr, u32 epoch,
                                u32* flags)
{
    if (b.njobs <= 0 || !b.tile_base[b.njobs]) return;
    hipLaunchKernelGGL(k_count_scan_many, dim3(b.tile_base[b.njobs]), dim3(PRA_COUNT_SCAN_BLOCK), 0, st, b, status,
                       tile_ctr, epoch, flags);
}

template <typename T>
static void scan_impl(hipStream_t st, const T* in, T* out, u64 n, u64* status, u32* tile_ctr, u32 epoch, u32* flags,
                      T* total)
{
    if (n <= PRA_SMALL_SCAN) {
        hipLaunchKernelGGL(k_scan_small<T>, dim3(1), dim3(PRA_BLOCK), 0, st, in, out, n, total);
        return;
    }
    hipLaunchKernelGGL(k_scan_lookback<T>, dim3(pra_tiles(n)), dim3(PRA_BLOCK), 0, st, in, out, n, status, tile_ctr,
                       epoch, flags, total);
}

void pra_scan_u64(hipStream_t st, const u64* in, u64* out, u64 n, u64* status, u32* tile_ctr, u32 epoch, u32* flags,
                  u64* total)
{
    scan_impl<u64>(st, in, out, n, status, tile_ctr, epoch, flags, total);
}

void pra_scan_u32(hipStream_t st, const u32* in, u32* out, u64 n, u64* status, u32* tile_ctr, u32 epoch, u32* flags,
                  u32* total)
{
    scan_impl<u32>(st, in, out, n, status, tile_ctr, epoch, flags, total);
}

u32 pra_tiles(u64 n) { return (u32)((n + PRA_TILE - 1) / PRA_TILE); }

void pra_launch_bucket_bounds(hipStream_t st, const u64* rows, u64 n, int w, int jcc, u64 nb, u64* bstart)
{
    hipLaunchKernelGGL(k_bucket_bounds, dim3(grid_for(nb + 1)), dim3(PRA_BLOCK), 0, st, rows, n, w, jcc, nb, bstart);
}

void pra_launch_group_flags(hipStream_t st, const u64* rows, u64 n, int w, int jcc, u64* flags)
{
    hipLaunchKernelGGL(k_group_flags, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, rows, n, w, jcc, flags);
}

void pra_launch_group_compact(hipStream_t st, const u64* flags, const u64* gpos, u64 n, u64* gstart)
{
    hipLaunchKernelGGL(k_group_compact, dim3(grid_for(n + 1)), dim3(PRA_BLOCK), 0, st, flags, gpos, n, gstart);
}

void pra_launch_compact_rows(hipStream_t st, const u64* in, u64 n, int w, const u64* flags, const u64* pos, u64* out)
{
    if (!n) return;
    hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, in, n, w, flags, pos, out);
}

void pra_launch_index_insert(hipStream_t st, const u64* rows, int w, int jcc, const u64* gstart, u64 ng, u64* table,
                             u64 cap, u32* flags)
{
    if (!ng) return;
    hipLaunchKernelGGL(k_index_insert, dim3(grid_for(ng)), dim3(PRA_BLOCK), 0, st, rows, w, jcc, gstart, ng, table,
                       cap, flags);
}

u32 pra_join_expand_blocks(u64 cap)
{
    u64 g = (cap + PRA_BLOCK - 1) / PRA_BLOCK;
    return (u32)(g > 16384 ? 16384 : g);
}

void pra_launch_join_expand(hipStream_t st, const JoinExpandArgs& a, u64 cap)
{
    if (!cap) return;
    hipLaunchKernelGGL(k_join_expand, dim3(pra_join_expand_blocks(cap)), dim3(PRA_BLOCK), 0, st, a);
}

void pra_launch_expand_many(hipStream_t st, const void* jobs, int njobs, u32 nblocks, bool join)
{
    if (njobs <= 0 || !nblocks) return;
    if (join)
        hipLaunchKernelGGL(k_join_expand_many, dim3(nblocks), dim3(PRA_BLOCK), 0, st, (const JoinExpandArgs*)jobs,
                           njobs);
    else
        hipLaunchKernelGGL(k_copy_expand_many, dim3(nblocks), dim3(PRA_BLOCK), 0, st, (const CopyExpandArgs*)jobs,
                           njobs);
}

void pra_launch_plan(hipStream_t st, const PlanArgs& a)
{
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(PRA_BLOCK), 0, st, a);
}

void pra_launch_plan_all(hipStream_t st, const PlanAllArgs& a, int nblocks)
{
    if (nblocks > 0) hipLaunchKernelGGL(k_plan_all, dim3(nblocks), dim3(PRA_BLOCK), 0, st, a);
}

void pra_launch_plan_tab(hipStream_t st, const PlanTabArgs& a, int nouts)
{
    if (nouts > 0) hipLaunchKernelGGL(k_plan_tab, dim3(nouts), dim3(PRA_BLOCK), 0, st, a);
}

void pra_launch_find_split(hipStream_t st, const u64* off, u64 n, const u32* kp, long long thresh, u64* out)
{
    hipLaunchKernelGGL(k_find_split, dim3(1), dim3(1), 0, st, off, n, kp, thresh, out);
}

void pra_launch_newt_fold_publish(hipStream_t st, const NewtFoldArgs& a, const u64* src, u64 words,
                                  const u32* flags, u64* host, u64 seq)
{
    hipLaunchKernelGGL(k_newt_fold_publish, dim3(1), dim3(PRA_BLOCK), 0, st, a, src, words, flags, host, seq);
}

void pra_launch_publish(hipStream_t st, const u64* src, u64 words, const u32* flags, u64* host, u64 seq)
{
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(PRA_BLOCK), 0, st, src, words, flags, host, seq);
}

void pra_launch_copy_expand(hipStream_t st, const CopyExpandArgs& a)
{
    if (!a.n) return;
    hipLaunchKernelGGL(k_copy_expand, dim3(grid_for(a.n)), dim3(PRA_BLOCK), 0, st, a);
}


void pra_launch_dedup_scatter(hipStream_t st, const NewtScatterArgs& a, int wide_choice)
{
    // many tiles: wide blocks hide the per-candidate loads; few tiles: narrow
    // blocks keep the per-block scans short (the autotuner may choose instead)
    const bool wide = wide_choice < 0 ? a.nblocks >= PRA_SCATTER_WIDE_TILES : wide_choice != 0;
    if (a.dict && wide)
        hipLaunchKernelGGL((k_dedup_scatter<true, PRA_SCATTER_BLOCK>), dim3(a.nblocks), dim3(PRA_SCATTER_BLOCK), 0, st,
                           a);
    else if (a.dict)
        hipLaunchKernelGGL((k_dedup_scatter<true, PRA_BLOCK>), dim3(a.nblocks), dim3(PRA_BLOCK), 0, st, a);
    else if (wide)
        hipLaunchKernelGGL((k_dedup_scatter<false, PRA_SCATTER_BLOCK>), dim3(a.nblocks), dim3(PRA_SCATTER_BLOCK), 0,
                           st, a);
    else
        hipLaunchKernelGGL((k_dedup_scatter<false, PRA_BLOCK>), dim3(a.nblocks), dim3(PRA_BLOCK), 0, st, a);
}

void pra_launch_dedup_scatter_many(hipStream_t st, const NewtScatterArgs* jobs, int njobs, u32 nblocks, bool packed)
{
    if (njobs <= 0 || !nblocks) return;
    if (packed)
        hipLaunchKernelGGL(k_dedup_scatter_many<true>, dim3(nblocks), dim3(PRA_BLOCK), 0, st, jobs, njobs);
    else
        hipLaunchKernelGGL(k_dedup_scatter_many<false>, dim3(nblocks), dim3(PRA_BLOCK), 0, st, jobs, njobs);
}

void pra_launch_table_insert_rows(hipStream_t st, const u64* rows, u64 n, int w, int kw, u64* table, u64 cap,
                                  u32* flags)
{
    if (!n) return;
    hipLaunchKernelGGL(k_table_insert_rows, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, rows, n, w, kw, table, cap,
                       flags);
}

void pra_launch_table_verify_rows(hipStream_t st, const u64* rows, u64 n, int w, int kw, const u64* table, u64 cap,
                                  u32* flags)
{
    if (!n) return;
    hipLaunchKernelGGL(k_table_verify_rows, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, rows, n, w, kw, table, cap,
                       flags);
}

void pra_launch_table_rehash(hipStream_t st, const u64* old, u64 oldcap, u64* nt, u64 cap, int kw, u32* flags)
{
    if (!oldcap) return;
    hipLaunchKernelGGL(k_table_rehash, dim3(grid_for(oldcap)), dim3(PRA_BLOCK), 0, st, old, oldcap, nt, cap, kw,
                       flags);
}

void pra_launch_sort_hist(hipStream_t st, const u64* rows, u64 n, int w, const PassList& pl, u32* ghist)
{
    u64 g = (n + PRA_BLOCK * 4 - 1) / (PRA_BLOCK * 4);
    if (g > 2048) g = 2048;
    if (g == 0) g = 1;
    size_t lds = (size_t)pl.np * PRA_RADIX * sizeof(u32);
    hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)g), dim3(PRA_BLOCK), lds, st, rows, n, w, pl, ghist);
}

void pra_launch_sort_scan(hipStream_t st, const u32* ghist, int np, u32* gofs, u64* bnd, u32 nb, u64 n)
{
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(PRA_BLOCK), 0, st, ghist, np, gofs, bnd, nb, n);
}

// rows are sorted in tiles of PRA_TILE; (key, index) pairs in tiles of 2 x PRA_TILE
void pra_launch_sort_scatter(hipStream_t st, const u64* in, u64* out, u64 n, int w, DigitSpec ds, const u32* gofs,
                             u64* status, u32* tile_ctr, u32 epoch, u32* flags, const u32* iin, u32* iout)
{
    if (iout) {
        u32 nt = pra_tiles(n);
        hipLaunchKernelGGL((k_sort_scatter<true, PRA_ITEMS>), dim3(nt ? nt : 1), dim3(PRA_BLOCK), 0, st, in, out, n, w, ds,
                           gofs, status, tile_ctr, epoch, flags, iin, iout);
    } else {
        u32 nt = pra_tiles(n);
        hipLaunchKernelGGL((k_sort_scatter<false, PRA_ITEMS>), dim3(nt ? nt : 1), dim3(PRA_BLOCK), 0, st, in, out, n, w, ds,
                           gofs, status, tile_ctr, epoch, flags, iin, iout);
    }
}

void pra_launch_seg_bounds(hipStream_t st, const u64* keys, u64 n, int top, u32 nseg, u64* seg, u64* stat, u32* large)
{
    hipLaunchKernelGGL(k_seg_bounds, dim3(grid_for(nseg)), dim3(PRA_BLOCK), 0, st, keys, n, top, nseg, seg, stat,
                       large);
}

void pra_launch_seg_sort(hipStream_t st, const u64* keys, const u32* idx, const u64* seg, int top, u32 nseg, u64 n,
                         u32 bytes, const u64* rows, int w, u64* out, u64 bmul, int ls, u64 nb, u64* bnd, u32* flags,
                         u64 maxseg, const u32* large, u64 nlarge)
{
#define PRA_SEG_SORT(B, G, L, S)                                                                                     \
    hipLaunchKernelGGL(k_seg_sort<B>, dim3(G), dim3(B), 0, st, keys, idx, seg, top, n, bytes, rows, w, out, bmul, ls, \
                       nb, bnd, flags, L, S)
    static_assert(PRA_SEG_SMALL == 256 * PRA_BIN_ITEMS, "the small segments fit the 256-thread blocks");
    // every segment that fits the 256-thread blocks, then the larger ones
    PRA_SEG_SORT(256, nseg, (const u32*)nullptr, maxseg <= PRA_SEG_SMALL ? ~0u : (u32)PRA_SEG_SMALL);
    if (maxseg > PRA_SEG_SMALL && nlarge > 0) {
        if (maxseg <= 512 * PRA_BIN_ITEMS)
            PRA_SEG_SORT(512, nlarge, large, ~0u);
        else
            PRA_SEG_SORT(PRA_BIN_BLOCK, nlarge, large, ~0u);
    }
#undef PRA_SEG_SORT
}

void pra_launch_bin_sort(hipStream_t st, const u64* bin_keys, const u32* bin_idx, u32 bin_cap, u64 n, const u32* bofs,
                         u32 bytes, const u64* rows, int w, u64* out, u64 bmul, int ls, u64 nb, u64* bnd, u32* flags,
                         u64 maxbin)
{
#define PRA_BIN_SORT(B)                                                                                              \
    hipLaunchKernelGGL(k_bin_sort<B>, dim3(PRA_RADIX), dim3(B), 0, st, bin_keys, bin_idx, bin_cap, n, bofs, bytes, rows, \
                       w, out, bmul, ls, nb, bnd, flags)
    if (maxbin <= 256 * PRA_BIN_ITEMS)
        PRA_BIN_SORT(256);
    else if (maxbin <= 512 * PRA_BIN_ITEMS)
        PRA_BIN_SORT(512);
    else
        PRA_BIN_SORT(PRA_BIN_BLOCK);
#undef PRA_BIN_SORT
}

void pra_launch_small_sort(hipStream_t st, const u64* rows, u64 n, int w, int jcc, int kc, u64 nb, int bucketed,
                           u64* out, u64* bnd)
{
    hipLaunchKernelGGL(k_small_sort, dim3(1), dim3(PRA_SMALL_SORT_BLOCK), 0, st, rows, n, w, jcc, kc, nb, bucketed, out,
                       bnd);
}

void pra_launch_small_sort_many(hipStream_t st, const SmallSortBatch& b, int njobs)
{
    if (njobs > 0) hipLaunchKernelGGL(k_small_sort_many, dim3(njobs), dim3(PRA_SMALL_SORT_BLOCK), 0, st, b);
}

void pra_launch_block_sort(hipStream_t st, const u64* rows, u64 n, int w, const PassList& pl, u64* out, u64* bnd,
                           int jcc, u64 nb)
{
    hipLaunchKernelGGL(k_block_sort, dim3(1), dim3(PRA_BIN_BLOCK), 0, st, rows, n, w, pl, out, bnd, jcc, nb);
}

void pra_launch_collect_cols(hipStream_t st, const u64* rows, u64 n, int w, int ncols, u64* out)
{
    if (!n) return;
    hipLaunchKernelGGL(k_collect_cols, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, rows, n, w, ncols, out);
}

void pra_launch_view_keep(hipStream_t st, const u64* rows, u64 n, int w, int ja, int jb, u64 nb, u32* keep)
{
    if (!n) return;
    hipLaunchKernelGGL(k_view_keep, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, rows, n, w, ja, jb, nb, keep);
}

void pra_launch_view_compact(hipStream_t st, const u64* rows, u64 n, int w, const u32* keep, const u32* pos, u64* out)
{
    if (!n) return;
    hipLaunchKernelGGL(k_view_compact, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, rows, n, w, keep, pos, out);
}

void pra_launch_dict_insert(hipStream_t st, const u64* vals, u64 n, u64* tab, u64 cap, u32* flags)
{
    if (!n) return;
    hipLaunchKernelGGL(k_dict_insert, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, vals, n, tab, cap, flags);
}

void pra_launch_gather_rows(hipStream_t st, const u64* in, const u32* idx, u64 n, int w, u64* out)
{
    if (!n) return;
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, in, idx, n, w, out);
}

void pra_launch_key_bounds(hipStream_t st, const u64* keys, u64 n, u64 bmul, int ls, u64 nb, u64* bnd)
{
    hipLaunchKernelGGL(k_key_bounds, dim3(grid_for(nb + 1)), dim3(PRA_BLOCK), 0, st, keys, n, bmul, ls, nb, bnd);
}

void pra_launch_col_orand2(hipStream_t st, const u64* rows, u64 n, int w, int kc, u64* out)
{
    u64 g = (n + PRA_BLOCK - 1) / PRA_BLOCK;
    if (g > 512) g = 512;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_col_orand2, dim3((unsigned)g, (unsigned)kc), dim3(PRA_BLOCK), 0, st, rows, n, w, out);
}

void pra_launch_dedup_claim2(hipStream_t st, const u64* cand, const u64* ctl, u64 capn, int cw, int kw, u64* table,
                             u64 cap, u64* slot_of, u64 epoch, u32* flags)
{
    if (!capn) return;
    const dim3 g(grid_for(capn)), b(PRA_BLOCK);
#define PRA_CLAIM(K, E)                                                                                                \
    hipLaunchKernelGGL((k_dedup_claim2<K, E>), g, b, 0, st, cand, ctl, cw, kw, table, cap, slot_of, epoch, flags)
    switch (kw) {
    case 1: PRA_CLAIM(1, true); break;
    case 2: PRA_CLAIM(2, true); break;
    case 3: PRA_CLAIM(3, true); break;
    case 4: PRA_CLAIM(4, true); break;
    default:
        if (kw <= 8)
            PRA_CLAIM(8, false);
        else if (kw <= 16)
            PRA_CLAIM(16, false);
        else
            PRA_CLAIM(0, false);
    }
#undef PRA_CLAIM
}

int pra_dedup_class(int kw)
{
    return kw >= 1 && kw <= 4 ? kw : kw <= 8 ? 5 : kw <= 16 ? 6 : 7;
}

#define PRA_DEDUP_DISPATCH(CLS, LAUNCH)                                                                                \
    switch (CLS) {                                                                                                     \
    case 1: LAUNCH(1, true); break;                                                                                    \
    case 2: LAUNCH(2, true); break;                                                                                    \
    case 3: LAUNCH(3, true); break;                                                                                    \
    case 4: LAUNCH(4, true); break;                                                                                    \
    case 5: LAUNCH(8, false); break;                                                                                   \
    case 6: LAUNCH(16, false); break;                                                                                  \
    default: LAUNCH(0, false);                                                                                         \
    }

void pra_launch_dense_claim(hipStream_t st, const DenseArgs& a, u64 capn, u32* flags)
{
    if (!capn) return;
    const dim3 g(grid_for(capn)), b(PRA_BLOCK);
    switch (a.kw) {
    case 1: hipLaunchKernelGGL(k_dense_claim<1>, g, b, 0, st, a, flags); break;
    case 2: hipLaunchKernelGGL(k_dense_claim<2>, g, b, 0, st, a, flags); break;
    case 3: hipLaunchKernelGGL(k_dense_claim<3>, g, b, 0, st, a, flags); break;
    default: hipLaunchKernelGGL(k_dense_claim<4>, g, b, 0, st, a, flags); break;
    }
}

void pra_launch_dense_many(hipStream_t st, const DenseArgs* jobs, int njobs, int kw, u32 claim_blocks,
                           u32 decide_blocks, u32* flags)
{
    if (njobs <= 0) return;
    if (claim_blocks) {
        const dim3 g(claim_blocks), b(PRA_BLOCK);
        switch (kw) {
        case 1: hipLaunchKernelGGL(k_dense_claim_many<1>, g, b, 0, st, jobs, njobs, flags); break;
        case 2: hipLaunchKernelGGL(k_dense_claim_many<2>, g, b, 0, st, jobs, njobs, flags); break;
        case 3: hipLaunchKernelGGL(k_dense_claim_many<3>, g, b, 0, st, jobs, njobs, flags); break;
        default: hipLaunchKernelGGL(k_dense_claim_many<4>, g, b, 0, st, jobs, njobs, flags); break;
        }
    }
    if (decide_blocks)
        hipLaunchKernelGGL(k_dense_decide_many, dim3(decide_blocks), dim3(PRA_DECIDE_BLOCK), 0, st, jobs, njobs);
}

void pra_launch_dense_decide(hipStream_t st, const DenseArgs& a, u32*)
{
    const dim3 g(std::min<u32>(a.nblocks, PRA_DECIDE_GRID)), b(PRA_DECIDE_BLOCK);
    hipLaunchKernelGGL(k_dense_decide, g, b, 0, st, a);
}

void pra_launch_dense_build(hipStream_t st, const u64* rows, u64 n, int w, int kw, const u64* dict, u64 dict_cap, u64 D,
                            u32* bitmap, u64* miss)
{
    if (!n) return;
    const dim3 g(grid_for(n)), b(PRA_BLOCK);
    switch (kw) {
    case 1: hipLaunchKernelGGL(k_dense_build<1>, g, b, 0, st, rows, n, w, dict, dict_cap, D, bitmap, miss); break;
    case 2: hipLaunchKernelGGL(k_dense_build<2>, g, b, 0, st, rows, n, w, dict, dict_cap, D, bitmap, miss); break;
    case 3: hipLaunchKernelGGL(k_dense_build<3>, g, b, 0, st, rows, n, w, dict, dict_cap, D, bitmap, miss); break;
    default: hipLaunchKernelGGL(k_dense_build<4>, g, b, 0, st, rows, n, w, dict, dict_cap, D, bitmap, miss); break;
    }
}

void pra_launch_dense_table_clear(hipStream_t st, u64* table, u64 cap) { pra_launch_fill_u64(st, table, cap, 0); }

void pra_launch_dedup_claim_many(hipStream_t st, const ClaimBatch& b, int cls, u32* flags)
{
    if (b.njobs <= 0 || !b.block_base[b.njobs]) return;
    const dim3 g(b.block_base[b.njobs]), bl(PRA_BLOCK);
#define PRA_CLAIM_MANY(K, E) hipLaunchKernelGGL((k_dedup_claim_many<K, E>), g, bl, 0, st, b, flags)
    PRA_DEDUP_DISPATCH(cls, PRA_CLAIM_MANY)
#undef PRA_CLAIM_MANY
}

void pra_launch_dedup_decide_many(hipStream_t st, const DecideBatch& b, int cls, u32* flags)
{
    if (b.njobs <= 0 || !b.block_base[b.njobs]) return;
    const dim3 g(b.block_base[b.njobs]), bl(PRA_DECIDE_BLOCK);
#define PRA_DECIDE_MANY(K, E) hipLaunchKernelGGL((k_dedup_decide_many<K, E>), g, bl, 0, st, b, flags)
    PRA_DEDUP_DISPATCH(cls, PRA_DECIDE_MANY)
#undef PRA_DECIDE_MANY
}
#undef PRA_DEDUP_DISPATCH

static_assert(sizeof(ClaimBatch) + sizeof(void*) <= 4096, "claim batch exceeds the kernel arguments");
static_assert(sizeof(DecideBatch) + sizeof(void*) <= 4096, "decision batch exceeds the kernel arguments");

void pra_launch_dedup_decide_hist(hipStream_t st, const u64* cand, const u64* ctl, int cw, int kw, const u64* table,
                                  const u64* slot_of, u64 epoch, u32* keep, const u32* dest, u32 nb, u32* hist,
                                  u32 nblocks, u32* flags, u32* offs1)
{
    // tiles past the candidate count (known only on the device) cost a block
    // each, and a block fills a compute unit: a bounded grid strides over them
    const dim3 g(std::min<u32>(nblocks, PRA_DECIDE_GRID)), b(PRA_DECIDE_BLOCK);
#define PRA_DECIDE(K, E)                                                                                               \
    hipLaunchKernelGGL((k_dedup_decide_hist<K, E>), g, b, 0, st, cand, ctl, cw, kw, table, slot_of, epoch, keep, dest, \
                       nb, hist, nblocks, flags, offs1)
    switch (kw) {
    case 1: PRA_DECIDE(1, true); break;
    case 2: PRA_DECIDE(2, true); break;
    case 3: PRA_DECIDE(3, true); break;
    case 4: PRA_DECIDE(4, true); break;
    default:
        if (kw <= 8)
            PRA_DECIDE(8, false);
        else if (kw <= 16)
            PRA_DECIDE(16, false);
        else
            PRA_DECIDE(0, false);
    }
#undef PRA_DECIDE
}

void pra_launch_bin_sort_many(hipStream_t st, const BinSortJob* jobs, int njobs, u32* flags, u64 maxbin)
{
    if (njobs <= 0) return;
#define PRA_BIN_SORT(B) hipLaunchKernelGGL(k_bin_sort_many<B>, dim3(PRA_RADIX * njobs), dim3(B), 0, st, jobs, flags)
    if (maxbin <= 256 * PRA_BIN_ITEMS)
        PRA_BIN_SORT(256);
    else if (maxbin <= 512 * PRA_BIN_ITEMS)
        PRA_BIN_SORT(512);
    else
        PRA_BIN_SORT(PRA_BIN_BLOCK);
#undef PRA_BIN_SORT
}

void pra_launch_scan_small_many(hipStream_t st, const ScanJob* jobs, int njobs)
{
    if (njobs > 0) hipLaunchKernelGGL(k_scan_small_many, dim3(njobs), dim3(PRA_BLOCK), 0, st, jobs);
}

void pra_launch_copy_many(hipStream_t st, const CopyJob* jobs, int njobs, u32 nblocks)
{
    if (njobs > 0 && nblocks) hipLaunchKernelGGL(k_copy_many, dim3(nblocks), dim3(PRA_BLOCK), 0, st, jobs, njobs);
}

void pra_launch_newt_fold_many(hipStream_t st, const NewtFoldArgs* jobs, int njobs)
{
    if (njobs > 0) hipLaunchKernelGGL(k_newt_fold_many, dim3(3 * njobs), dim3(PRA_BLOCK), 0, st, jobs);
}

void pra_launch_newt_fold(hipStream_t st, const NewtFoldArgs& a)
{
    static_assert(PRA_RADIX == PRA_BLOCK && PRA_PART_WORDS <= PRA_BLOCK, "one thread per bin / word");
    hipLaunchKernelGGL(k_newt_fold, dim3(3), dim3(PRA_BLOCK), 0, st, a);
}

void pra_launch_newt_rep_init(hipStream_t st, u64* rep_words, int sets)
{
    for (int k = 0; k < sets; k++)
        hipLaunchKernelGGL(k_newt_rep_init, dim3((PRA_NREP * PRA_PART_WORDS + PRA_BLOCK - 1) / PRA_BLOCK),
                           dim3(PRA_BLOCK), 0, st, rep_words + (size_t)k * PRA_NREP * PRA_PART_WORDS);
}

void pra_launch_merge(hipStream_t st, const u64* F, u64 nf, const u64* fbnd, const u64* D, u64 nd, const u64* dbnd,
                      u64 nb, int w, u64* rank, u64* out, u64* add_to)
{
    if (!nd) return;
    hipLaunchKernelGGL(k_merge_rank, dim3(grid_for(nd)), dim3(PRA_BLOCK), 0, st, D, nd, dbnd, F, fbnd, nb, w, rank);
    hipLaunchKernelGGL(k_merge_scatter, dim3(grid_for(nf + nd)), dim3(PRA_BLOCK), 0, st, F, nf, D, nd, rank, w, out,
                       add_to, dbnd, nb + 1);
}

void pra_launch_rows_checksum(hipStream_t st, const u64* rows, u64 n, int w, u64* out)
{
    u64 g = (n + PRA_BLOCK - 1) / PRA_BLOCK;
    if (g > 1024) g = 1024;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_rows_checksum, dim3((unsigned)g), dim3(PRA_BLOCK), 0, st, rows, n, w, out);
}

void pra_launch_group_maxlen(hipStream_t st, const u64* flags, const u64* gpos, const u64* gstart, u64 n, u64* maxlen)
{
    if (!n) return;
    hipLaunchKernelGGL(k_group_maxlen, dim3(grid_for(n)), dim3(PRA_BLOCK), 0, st, flags, gpos, gstart, n, maxlen);
}
