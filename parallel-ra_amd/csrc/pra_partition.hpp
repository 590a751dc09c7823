// Host interface of the stable bucket partition (pra_partition.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define PRA_PART_TILE 2048

uint32_t pra_partition_tiles(uint64_t n);
// u32 scratch words a partition of n rows into nb buckets needs
size_t pra_partition_scratch_words(uint64_t n, uint32_t nb);
// rows: n rows of w u64 words (device); out: n rows (device); bounds: nb + 1
// u64 words (device); scratch: pra_partition_scratch_words(n, nb) u32 words
void pra_launch_partition(hipStream_t st, const uint64_t* rows, uint64_t n, int w, int jcc, uint32_t nb,
                          uint64_t* out, uint64_t* bounds, uint32_t* scratch);
