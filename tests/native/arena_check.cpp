// Host check of the arena bookkeeping (dn_arena_pool.cpp): random
// allocation / release sequences against a brute-force model of the arena.
// Exit status 0 when every invariant holds.
#include <algorithm>
#include <cstdio>
#include <iterator>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "dn_arena.h"

static int fails = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            ++fails;                                                  \
        }                                                             \
    } while (0)

int main()
{
    const size_t SIZE = 1 << 24;
    dn_arena a;
    a.base = (char *)0x100000;  // bookkeeping only: never dereferenced
    a.init(SIZE);
    std::mt19937_64 rng(7);
    std::map<char *, size_t> live;  // pointer -> rounded size
    for (int it = 0; it < 20000; ++it) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            const size_t n = 1 + rng() % (rng() % 4 == 0 ? SIZE / 8 : 20000);
            const size_t r = dn_arena::round(n);
            // best fit exists iff some free block is large enough
            const bool can = a.largest() >= r;
            const bool fit = a.fits({n});
            CHECK(fit == can);
            char *p = (char *)a.alloc(n);
            CHECK((p != nullptr) == can);
            if (!p) continue;
            CHECK(((size_t)(p - a.base)) % dn_arena::ALIGN == 0);
            CHECK(p >= a.base && p + r <= a.base + SIZE);
            // no overlap with live buffers
            auto nx = live.lower_bound(p);
            if (nx != live.end()) CHECK(p + r <= nx->first);
            if (nx != live.begin()) {
                auto pv = std::prev(nx);
                CHECK(pv->first + pv->second <= p);
            }
            live[p] = r;
        } else {
            auto it2 = live.begin();
            std::advance(it2, rng() % live.size());
            CHECK(a.release(it2->first));
            live.erase(it2);
        }
        size_t used = 0;
        for (auto &kv : live) used += kv.second;
        CHECK(a.free_bytes + used == SIZE);
        // free blocks are coalesced: no two free blocks touch
        size_t prev_end = (size_t)-1;
        for (auto &kv : a.free_off) {
            CHECK(kv.first != prev_end);
            prev_end = kv.first + kv.second;
        }
        if (fails) break;
    }
    CHECK(!a.release(a.base + 3));  // not a live buffer
    for (auto &kv : live) CHECK(a.release(kv.first));
    CHECK(a.free_off.size() == 1 && a.free_off.begin()->second == SIZE && a.free_bytes == SIZE);
    // fits() of a set of sizes agrees with allocating them in that order
    for (int t = 0; t < 200 && !fails; ++t) {
        std::vector<size_t> sz;
        for (int k = 0, m = 1 + rng() % 12; k < m; ++k) sz.push_back(1 + rng() % (SIZE / 4));
        std::sort(sz.rbegin(), sz.rend());
        const bool f = a.fits(sz);
        std::vector<void *> got;
        bool ok = true;
        for (size_t s : sz) {
            void *p = a.alloc(s);
            if (!p) {
                ok = false;
                break;
            }
            got.push_back(p);
        }
        CHECK(f == ok);
        for (void *p : got) CHECK(a.release(p));
    }
    std::printf("%s\n", fails ? "FAIL" : "arena bookkeeping ok");
    return fails ? 1 : 0;
}
