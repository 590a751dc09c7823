// Host interface of the stable bucket partition (pra_partition.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define PRA_PART_TILE 2048
// largest input of the allocation-free partition, and its scratch: per-tile
// bucket counts of 256 buckets over PRA_NOALLOC_MAX_ROWS / PRA_PART_TILE
// tiles, plus 256 bucket totals (4,195,328 bytes per device)
#define PRA_NOALLOC_MAX_ROWS (1u << 23)
#define PRA_NOALLOC_SCRATCH_WORDS (256u * (PRA_NOALLOC_MAX_ROWS / PRA_PART_TILE) + 256u)
#define PRA_MAX_DEVICES 64

uint32_t pra_partition_tiles(uint64_t n);
// u32 scratch words a partition of n rows into nb buckets needs
size_t pra_partition_scratch_words(uint64_t n, uint32_t nb);
// rows: n rows of w u64 words (device); out: n rows (device); bounds: nb + 1
// u64 words (device); scratch: pra_partition_scratch_words(n, nb) u32 words
void pra_launch_partition(hipStream_t st, const uint64_t* rows, uint64_t n, int w, int jcc, uint32_t nb,
                          uint64_t* out, uint64_t* bounds, uint32_t* scratch);

// Allocation-free partition: the device address of this device's copy of the
// static scratch array, looked up once per device by
// pra_partition_scratch_init (called with the device current, from
// pra_ctx_create); null before that
hipError_t pra_partition_scratch_init(int device);
uint32_t* pra_partition_scratch(int device);
