// Stable partition of device rows by bucket = tuple_hash(row[0:jcc]) % nb:
// the routing of input rows to their owner rank
// (buffer_data_to_hash_buffer_col, src/IO/parallel_io.cpp:239-297) and of the
// all-to-all tuple exchange (src/comm/all_to_allv_comm.cpp:64-142), as one
// on-device step: the output holds bucket 0's rows, then bucket 1's, ..., each
// bucket's rows in input order, and the bucket bounds.
//
// Three launches over tiles of PRA_PART_TILE rows:
//   k_part_hist     per tile, the rows of each bucket          -> H[b][tile]
//   k_part_scan     per bucket, exclusive prefix over tiles    -> H, T[b]
//   k_part_scatter  per tile, bucket starts (scan of T) + H, then every row to
//                   its place: rows are taken in input order, 256 at a time,
//                   ranked inside each wave by matching bucket ballots and
//                   across the block's waves by their per-bucket counts.
// Scratch: nb * tiles + nb u32 words (H then T); every word a call reads was
// written earlier in the same call, so the scratch carries no state between
// calls and needs no clearing.
//
// Algorithmic bytes per row: 8 * jcc (histogram) + 8 * (w + jcc) + 8 * w
// (scatter: key words again, the row read and written).
#include "pra_device.hpp"
#include "pra_partition.hpp"

#include <mutex>

namespace {

constexpr int kPartBlock = 256;
constexpr int kPartRounds = PRA_PART_TILE / kPartBlock;
constexpr int kScanBlock = 1024;

__device__ inline u32 part_bucket(const u64* __restrict__ rows, u64 i, int w, int jcc, u32 nb)
{
    return (u32)(pra_fnv(rows + i * (u64)w, jcc) % nb);
}

__global__ __launch_bounds__(kPartBlock) void k_part_hist(const u64* __restrict__ rows, u64 n, int w, int jcc,
                                                           u32 nb, u32* __restrict__ H, u32 ntiles)
{
    __shared__ u32 h[PRA_RADIX];
    const u32 tid = threadIdx.x, tile = blockIdx.x;
    h[tid] = 0;
    __syncthreads();
    const u64 base = (u64)tile * PRA_PART_TILE;
#pragma unroll
    for (int r = 0; r < kPartRounds; r++) {
        const u64 i = base + (u64)r * kPartBlock + tid;
        if (i < n) atomicAdd(&h[part_bucket(rows, i, w, jcc, nb)], 1u);
    }
    __syncthreads();
    if (tid < nb) H[(u64)tid * ntiles + tile] = h[tid];
}

// block b: exclusive prefix of H[b][0..ntiles) in place, the total into T[b]
__global__ __launch_bounds__(kScanBlock) void k_part_scan(u32* __restrict__ H, u32 ntiles, u32* __restrict__ T)
{
    __shared__ u32 part[kScanBlock];
    const u32 tid = threadIdx.x;
    u32* h = H + (u64)blockIdx.x * ntiles;
    const u32 per = (ntiles + kScanBlock - 1) / kScanBlock;
    const u32 lo = tid * per, hi = min(ntiles, lo + per);
    u32 s = 0;
    for (u32 t = lo; t < hi; t++) s += h[t];
    part[tid] = s;
    __syncthreads();
    for (u32 d = 1; d < kScanBlock; d <<= 1) {   // inclusive Hillis-Steele scan of the chunk sums
        const u32 v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    u32 run = part[tid] - s;
    for (u32 t = lo; t < hi; t++) {
        const u32 v = h[t];
        h[t] = run;
        run += v;
    }
    if (tid == kScanBlock - 1) T[blockIdx.x] = part[tid];
}

__global__ __launch_bounds__(kPartBlock) void k_part_scatter(const u64* __restrict__ rows, u64 n, int w, int jcc,
                                                              u32 nb, const u32* __restrict__ H,
                                                              const u32* __restrict__ T, u32 ntiles,
                                                              u64* __restrict__ out, u64* __restrict__ bounds)
{
    __shared__ u32 start[PRA_RADIX];
    __shared__ u32 base[PRA_RADIX];
    __shared__ u32 wc[PRA_WAVES][PRA_RADIX];
    const u32 tid = threadIdx.x, tile = blockIdx.x;
    const u32 wave = tid / 64, lane = __lane_id();
    // bucket starts: exclusive scan of the bucket totals
    const u32 tot = tid < nb ? T[tid] : 0;
    start[tid] = tot;
    __syncthreads();
    for (u32 d = 1; d < PRA_RADIX; d <<= 1) {
        const u32 v = tid >= d ? start[tid - d] : 0;
        __syncthreads();
        start[tid] += v;
        __syncthreads();
    }
    const u32 st = start[tid] - tot;
    if (tile == 0) {
        if (tid < nb) bounds[tid] = st;
        if (tid == 0) bounds[nb] = n;
    }
    base[tid] = tid < nb ? st + H[(u64)tid * ntiles + tile] : 0;
#pragma unroll
    for (int q = 0; q < PRA_WAVES; q++) wc[q][tid] = 0;
    __syncthreads();
    const u64 tbase = (u64)tile * PRA_PART_TILE;
    for (int r = 0; r < kPartRounds; r++) {
        const u64 i = tbase + (u64)r * kPartBlock + tid;
        const bool valid = i < n;
        const u32 b = valid ? part_bucket(rows, i, w, jcc, nb) : 0;
        const u64 active = __ballot(valid);
        const u64 peers = pra_match<PRA_RADIX_BITS>(b, active);
        const u32 rank = (u32)__popcll(peers & pra_lanemask_lt());
        if (valid && lane == (u32)(__ffsll((long long)peers) - 1)) wc[wave][b] = (u32)__popcll(peers);
        __syncthreads();
        if (valid) {
            u32 off = base[b] + rank;
            for (u32 q = 0; q < wave; q++) off += wc[q][b];
            pra_copy_row(out + (u64)off * w, rows + i * (u64)w, w);
        }
        __syncthreads();
        if (tid < nb) {
            u32 add = 0;
#pragma unroll
            for (int q = 0; q < PRA_WAVES; q++) {
                add += wc[q][tid];
                wc[q][tid] = 0;
            }
            base[tid] += add;
        }
        __syncthreads();
    }
}

__global__ void k_part_empty(u64* bounds, u32 nb)
{
    const u32 t = threadIdx.x;
    if (t <= nb) bounds[t] = 0;
}

// The allocation-free entry point's scratch: one array per device, for the
// whole process, holding H and T of a partition of at most
// PRA_NOALLOC_MAX_ROWS rows into at most 256 buckets.
static __device__ __align__(256) u32 scratch[PRA_NOALLOC_SCRATCH_WORDS];

std::mutex g_scratch_mu;
u32* g_scratch_dev[PRA_MAX_DEVICES] = {};

} // namespace

u32 pra_partition_tiles(u64 n) { return (u32)((n + PRA_PART_TILE - 1) / PRA_PART_TILE); }

size_t pra_partition_scratch_words(u64 n, u32 nb) { return (size_t)nb * pra_partition_tiles(n) + nb; }

void pra_launch_partition(hipStream_t st, const u64* rows, u64 n, int w, int jcc, u32 nb, u64* out, u64* bounds,
                          u32* scratch_words)
{
    if (n == 0) {
        hipLaunchKernelGGL(k_part_empty, dim3(1), dim3(PRA_RADIX + 1), 0, st, bounds, nb);
        return;
    }
    const u32 ntiles = pra_partition_tiles(n);
    u32* H = scratch_words;
    u32* T = scratch_words + (size_t)nb * ntiles;
    hipLaunchKernelGGL(k_part_hist, dim3(ntiles), dim3(kPartBlock), 0, st, rows, n, w, jcc, nb, H, ntiles);
    hipLaunchKernelGGL(k_part_scan, dim3(nb), dim3(kScanBlock), 0, st, H, ntiles, T);
    hipLaunchKernelGGL(k_part_scatter, dim3(ntiles), dim3(kPartBlock), 0, st, rows, n, w, jcc, nb, H, T, ntiles, out,
                       bounds);
}

// Once per device, with that device current: the device address of this
// device's copy of `scratch`.  Host code uses only this address, never the
// array's name (which names a host-side stand-in).
hipError_t pra_partition_scratch_init(int device)
{
    if (device < 0 || device >= PRA_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return hipSuccess;
    u32* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = ptr;
    return hipSuccess;
}

u32* pra_partition_scratch(int device)
{
    if (device < 0 || device >= PRA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[device];
}
