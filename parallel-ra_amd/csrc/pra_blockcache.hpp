// Process-wide cache of large device blocks, with its ownership rule kept free
// of any HIP type so that the rule can be exercised on the CPU with fake
// stream handles and a fake allocator.
//
// Returning tens of GB to the driver, and first use of freshly mapped device
// memory, cost far more than the work large blocks hold, so released blocks of
// at least `min_block` bytes are kept and handed to later requests of the same
// size class or of the two classes below it.  Classes are quarter octaves
// (min_block x 2^k x 1, 1.25, 1.5, 1.75), so a block is at most 25 % larger
// than what was asked for when fresh, and at most about 50 % when reused.
//
// Ownership rule (no events, no cross-stream ordering):
//   * put(stream, block): the block was released while work enqueued on
//     `stream` may still read or write it.  It is OWNED by that stream: only a
//     request on the same stream may take it, and that reuse is ordered behind
//     the earlier work by the stream itself.
//   * idle(stream): the host has observed that every piece of work enqueued on
//     `stream` so far has completed (a stream synchronization returned).  The
//     stream's owned blocks become IDLE: any stream may take them, and flush()
//     may free them.
//   * flush(device) frees idle blocks only; an owned block is never freed or
//     handed to another stream before its owner is declared idle.
// Out of memory, the idle blocks released longest ago are freed first, and
// only until the failed request fits: freeing tens of GB and touching fresh
// memory again costs far more than the request.
// The caller must call idle() only after a synchronization that happened after
// the put() calls it covers (the cache records no time, so it relies on the
// host order of put, synchronize, idle).
#ifndef PRA_BLOCKCACHE_HPP
#define PRA_BLOCKCACHE_HPP

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <vector>

namespace pra {

struct BlockAllocator {
    // returns nullptr when the device has no room
    void* (*alloc)(size_t bytes, void* user);
    void (*release)(void* p, void* user);
    void* user;
};

class BlockCache {
  public:
    explicit BlockCache(size_t min_block) : min_(min_block) {}

    size_t min_block() const { return min_; }

    // quarter-octave size class of a request: the smallest
    // min_block x 2^k x q / 4 (q = 4..7) that holds it
    size_t size_class(size_t bytes) const
    {
        size_t oct = min_;
        while (oct * 2 <= bytes) oct <<= 1;
        for (size_t q = 4; q <= 8; q++)
            if (oct / 4 * q >= bytes) return oct / 4 * q;
        return oct * 2;
    }

    // the class above class `cls`
    size_t next_class(size_t cls) const { return size_class(cls + 1); }

    // A cached block for `bytes` on `stream`: idle or owned by `stream`, of
    // the request's class or one of the next two, smallest first.  nullptr on
    // a miss.
    void* take(int device, const void* stream, size_t bytes, size_t* got)
    {
        const size_t cls = size_class(bytes), top = next_class(next_class(cls));
        std::lock_guard<std::mutex> lk(mu_);
        long best = -1;
        for (size_t i = 0; i < free_.size(); i++) {
            const Entry& b = free_[i];
            if (b.device == device && (!b.owner || b.owner == stream) && b.size >= cls && b.size <= top &&
                (best < 0 || b.size < free_[(size_t)best].size))
                best = (long)i;
        }
        if (best < 0) return nullptr;
        Entry e = free_[(size_t)best];
        free_.erase(free_.begin() + best);
        *got = e.size;
        return e.p;
    }

    // take(), else a fresh block of the request's class from the allocator;
    // when the allocator fails, idle blocks of the device are freed, oldest
    // first, until the request's size is freed, and the allocation retried;
    // then once more after freeing every idle block.  nullptr when that fails
    // too.
    void* get(int device, const void* stream, size_t bytes, size_t* got, const BlockAllocator& a)
    {
        if (void* p = take(device, stream, bytes, got)) return p;
        const size_t cls = size_class(bytes);
        void* p = a.alloc(cls, a.user);
        if (!p && flush(device, a, cls)) p = a.alloc(cls, a.user);
        if (!p && flush(device, a)) p = a.alloc(cls, a.user);
        if (p) *got = cls;
        return p;
    }

    void put(int device, const void* stream, void* p, size_t size)
    {
        std::lock_guard<std::mutex> lk(mu_);
        free_.push_back(Entry{p, size, device, stream, ++clock_});
    }

    // every piece of work enqueued on `stream` so far has completed
    void idle(const void* stream)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& b : free_)
            if (b.owner == stream) b.owner = nullptr;
    }

    // free idle blocks of a device, released longest ago first, until at
    // least `want` bytes are freed (all of them by default); returns the
    // bytes freed
    size_t flush(int device, const BlockAllocator& a, size_t want = ~(size_t)0)
    {
        std::vector<Entry> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            std::vector<size_t> idle;
            for (size_t i = 0; i < free_.size(); i++)
                if (free_[i].device == device && !free_[i].owner) idle.push_back(i);
            std::sort(idle.begin(), idle.end(), [&](size_t x, size_t y) { return free_[x].stamp < free_[y].stamp; });
            size_t freed = 0, k = 0;
            for (; k < idle.size() && freed < want; k++) freed += free_[idle[k]].size;
            idle.resize(k);
            std::sort(idle.begin(), idle.end());
            for (size_t j = idle.size(); j-- > 0;) {
                drop.push_back(free_[idle[j]]);
                free_.erase(free_.begin() + (long)idle[j]);
            }
        }
        size_t bytes = 0;
        for (auto& b : drop) {
            a.release(b.p, a.user);
            bytes += b.size;
        }
        return bytes;
    }

    // bytes cached for a device: owned by `stream` (any owner when stream is
    // null), and idle
    size_t owned_bytes(int device, const void* stream) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        size_t s = 0;
        for (auto& b : free_)
            if (b.device == device && b.owner && (!stream || b.owner == stream)) s += b.size;
        return s;
    }
    size_t idle_bytes(int device) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        size_t s = 0;
        for (auto& b : free_)
            if (b.device == device && !b.owner) s += b.size;
        return s;
    }

  private:
    struct Entry {
        void* p;
        size_t size;
        int device;
        const void* owner;   // stream that may still use the block, or null when idle
        unsigned long long stamp;   // order of release
    };
    size_t min_;
    unsigned long long clock_ = 0;
    mutable std::mutex mu_;
    std::vector<Entry> free_;
};

} // namespace pra

#endif
