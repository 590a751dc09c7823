// Exposes the reference's own tuple_hash (src/hash/hash.h, header-only) as a
// C symbol so the oracle's FNV-1a restatement can be checked against it.
// Built only where the reference tree is present; output goes to oracle/_ref/.
#include <cstdint>
typedef uint64_t u64;
#include "hash.h"

extern "C" u64 fnv_ref_tuple_hash(const u64* t, u64 n) { return tuple_hash(t, n); }
