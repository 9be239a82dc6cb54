// Bookkeeping of the device-memory arena (plain C++, no device calls):
// best-fit allocation over the free blocks, split on allocation, coalesced
// with free neighbours on release.
#include <algorithm>
#include <cstdint>
#include <set>

#include "dn_arena.h"

void dn_arena::init(size_t n)
{
    size = n;
    free_bytes = n;
    insert_free(0, n);
}

void dn_arena::insert_free(size_t off, size_t len)
{
    free_off[off] = len;
    free_len.emplace(len, off);
}

void dn_arena::erase_free(std::map<size_t, size_t>::iterator it)
{
    auto range = free_len.equal_range(it->second);
    for (auto r = range.first; r != range.second; ++r)
        if (r->second == it->first) {
            free_len.erase(r);
            break;
        }
    free_off.erase(it);
}

void *dn_arena::alloc(size_t n)
{
    n = round(std::max<size_t>(n, 1));
    std::lock_guard<std::mutex> lk(mu);
    auto best = free_len.lower_bound(n);  // smallest free block that fits
    if (best == free_len.end()) return nullptr;
    const size_t off = best->second, len = best->first;
    erase_free(free_off.find(off));
    if (len > n) insert_free(off + n, len - n);
    used[off] = n;
    free_bytes -= n;
    return base + off;
}

bool dn_arena::release(void *p)
{
    std::lock_guard<std::mutex> lk(mu);
    const std::uintptr_t a = reinterpret_cast<std::uintptr_t>(p), b = reinterpret_cast<std::uintptr_t>(base);
    if (a < b || a >= b + size) return false;
    const size_t off = (size_t)(a - b);
    auto u = used.find(off);
    if (u == used.end()) return false;
    size_t start = off, len = u->second;
    used.erase(u);
    free_bytes += len;
    auto next = free_off.find(start + len);
    if (next != free_off.end()) {
        len += next->second;
        erase_free(next);
    }
    auto prev = free_off.lower_bound(start);
    if (prev != free_off.begin()) {
        --prev;
        if (prev->first + prev->second == start) {
            start = prev->first;
            len += prev->second;
            erase_free(prev);
        }
    }
    insert_free(start, len);
    return true;
}

bool dn_arena::fits(const std::vector<size_t> &sizes)
{
    std::lock_guard<std::mutex> lk(mu);
    std::multiset<size_t> blocks;
    for (const auto &kv : free_len) blocks.insert(kv.first);
    for (size_t s : sizes) {
        const size_t n = round(std::max<size_t>(s, 1));
        auto b = blocks.lower_bound(n);
        if (b == blocks.end()) return false;
        const size_t rest = *b - n;
        blocks.erase(b);
        if (rest) blocks.insert(rest);
    }
    return true;
}

size_t dn_arena::largest()
{
    std::lock_guard<std::mutex> lk(mu);
    return free_len.empty() ? 0 : free_len.rbegin()->first;
}

