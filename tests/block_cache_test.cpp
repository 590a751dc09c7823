// CPU check of the large-block cache's ownership rule
// (parallel-ra_amd/csrc/pra_blockcache.hpp) with fake stream handles and a
// fake allocator that records every allocation and release.  Prints "ok" and
// exits 0 when every check holds; otherwise names the first failed check.
#include "../parallel-ra_amd/csrc/pra_blockcache.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>

namespace {

struct FakeDevice {
    size_t capacity = 0, used = 0;
    std::set<void*> live;
    std::map<void*, size_t> sizes;
    int allocs = 0, frees = 0;
    char* next = reinterpret_cast<char*>(0x100000);
};

void* fake_alloc(size_t bytes, void* user)
{
    auto* d = static_cast<FakeDevice*>(user);
    if (d->used + bytes > d->capacity) return nullptr;
    void* p = d->next;
    d->next += bytes + 4096;
    d->used += bytes;
    d->live.insert(p);
    d->sizes[p] = bytes;
    d->allocs++;
    return p;
}

void fake_release(void* p, void* user)
{
    auto* d = static_cast<FakeDevice*>(user);
    if (!d->live.count(p)) {
        std::printf("FAIL: release of a block that is not live\n");
        std::exit(1);
    }
    d->live.erase(p);
    d->frees++;
    d->used -= d->sizes[p];
}

int failures = 0;
void expect(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        failures++;
    }
}

} // namespace

int main()
{
    const size_t MB = 1 << 20;
    FakeDevice dev;
    dev.capacity = 1024 * MB;
    pra::BlockAllocator a{fake_alloc, fake_release, &dev};
    pra::BlockCache cache(64 * MB);
    int sa_tag = 0, sb_tag = 0;
    const void* sA = &sa_tag;   // two fake streams
    const void* sB = &sb_tag;
    size_t got = 0;

    // size classes are quarter octaves from the minimum block up
    expect(cache.size_class(1) == 64 * MB, "class of 1 byte is the minimum block");
    expect(cache.size_class(64 * MB + 1) == 80 * MB, "class rounds up to the next quarter octave");
    expect(cache.size_class(113 * MB) == 128 * MB && cache.size_class(129 * MB) == 160 * MB &&
               cache.size_class(400 * MB) == 448 * MB,
           "classes 64, 80, 96, 112, 128, 160, ... MiB");

    // a block released while stream A may still use it is owned by A
    void* b = cache.get(0, sA, 100 * MB, &got, a);
    expect(b != nullptr && got == 112 * MB, "fresh block of the request's class");
    cache.put(0, sA, b, got);
    expect(cache.owned_bytes(0, sA) == 112 * MB, "released block is owned by its stream");

    // another stream never gets it while A owns it
    void* c = cache.get(0, sB, 100 * MB, &got, a);
    expect(c != nullptr && c != b, "an owned block is not handed to another stream");
    cache.put(0, sB, c, got);

    // a flush never frees an owned block
    size_t freed = cache.flush(0, a);
    expect(freed == 0 && dev.live.count(b) && dev.live.count(c), "flush frees no owned block");

    // the owning stream reuses it without any synchronization
    void* b2 = cache.get(0, sA, 90 * MB, &got, a);
    expect(b2 == b, "the owning stream reuses its own block");
    cache.put(0, sA, b2, got);

    // after A is declared idle its blocks are free for every stream
    int sc_tag = 0;
    const void* sC = &sc_tag;   // a stream that owns nothing
    cache.idle(sA);
    expect(cache.owned_bytes(0, sA) == 0 && cache.idle_bytes(0) == 112 * MB, "idle moves A's blocks to the pool");
    expect(cache.owned_bytes(0, sB) == 112 * MB, "idle of A leaves B's blocks owned");
    void* b3 = cache.get(0, sC, 70 * MB, &got, a);
    expect(b3 == b, "an idle block goes to another stream");
    cache.put(0, sC, b3, got);

    // classes: a request takes its class or one of the next two, never larger
    cache.idle(sB);
    cache.idle(sC);   // pool: b, c (112 MiB each), all idle
    void* small = cache.get(0, sA, 90 * MB, &got, a);
    expect((small == b || small == c) && got == 112 * MB, "a 96 MiB request may take a block two classes up");
    cache.put(0, sA, small, got);
    cache.idle(sA);
    void* big = cache.get(0, sA, 300 * MB, &got, a);
    expect(big != b && big != c && got == 320 * MB, "a 320 MiB request gets a fresh block");
    cache.put(0, sA, big, got);
    void* tiny = cache.get(0, sB, 1 * MB, &got, a);
    expect(tiny != big && tiny != b && tiny != c && got == 64 * MB,
           "a 64 MiB request takes no block three classes up (112 MiB) or more");
    cache.put(0, sB, tiny, got);
    cache.idle(sB);

    // out of room: 608 MiB are held (b, c and tiny idle, `big` owned by A)
    // and C needs 448 MiB more; the idle blocks are freed to make space, the
    // owned one is not
    expect(dev.used == 608 * MB, "bookkeeping before the pressure case");
    void* d0 = cache.get(0, sC, 400 * MB, &got, a);
    expect(d0 != nullptr && got == 448 * MB, "allocation succeeds after freeing idle blocks");
    expect(!dev.live.count(b) && !dev.live.count(c), "idle blocks were freed");
    expect(dev.live.count(big) == 1, "a block owned by A survives C's out-of-memory flush");
    if (d0) cache.put(0, sC, d0, got);
    // still no room and nothing idle: B's request fails instead of taking the
    // blocks A and C own
    void* d1 = cache.get(0, sB, 400 * MB, &got, a);
    expect(d1 == nullptr && dev.live.count(big) == 1 && dev.live.count(d0) == 1,
           "a request that cannot fit fails; owned blocks stay");

    // flush after everything is idle returns every block
    cache.idle(sA);
    cache.idle(sB);
    cache.idle(sC);
    cache.flush(0, a);
    expect(dev.live.empty(), "flush of an all-idle cache frees every block");
    expect(dev.allocs == dev.frees, "allocations and releases balance");

    // out of room with several idle blocks: only the ones released longest
    // ago are freed, and only as many as the request needs
    {
        FakeDevice small;
        small.capacity = 256 * MB;
        pra::BlockAllocator sa{fake_alloc, fake_release, &small};
        pra::BlockCache lru(64 * MB);
        void* x[4];
        for (void*& p : x) p = lru.get(0, sA, 64 * MB, &got, sa);
        expect(small.used == 256 * MB, "four 64 MiB blocks fill the small device");
        lru.put(0, sA, x[2], 64 * MB);   // released first
        lru.put(0, sA, x[0], 64 * MB);
        lru.put(0, sA, x[1], 64 * MB);   // released last
        lru.idle(sA);
        void* y = lru.get(0, sB, 100 * MB, &got, sa);   // 112 MiB: two blocks must go
        expect(y != nullptr && got == 112 * MB, "the request fits after a partial flush");
        expect(!small.live.count(x[2]) && !small.live.count(x[0]) && small.live.count(x[1]),
               "the two blocks released longest ago were freed, the newest kept");
        lru.put(0, sB, y, got);
        lru.put(0, sA, x[3], 64 * MB);
        lru.idle(sA);
        lru.idle(sB);
        lru.flush(0, sa);
        expect(small.live.empty(), "a full flush still frees every idle block");
    }

    // devices are separate
    void* e = cache.get(1, sA, 64 * MB, &got, a);
    cache.put(1, sA, e, got);
    cache.idle(sA);
    size_t took = 0;
    expect(cache.take(0, sA, 64 * MB, &took) == nullptr, "a block of device 1 is not handed out on device 0");
    cache.flush(1, a);

    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
