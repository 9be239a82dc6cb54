// Device-memory arena: one hipMalloc carved into the buffers of many
// contexts (best fit, 256-byte aligned, free neighbours coalesced).  Creating
// and destroying contexts in an arena costs host bookkeeping only: no
// hipMalloc / hipFree, whose cost grows with the size (the driver maps and
// clears the pages) and which serialise with other HIP calls of the process.
#pragma once
#include <cstddef>
#include <map>
#include <mutex>
#include <vector>

struct dn_arena {
    int device = 0;
    char *base = nullptr;
    size_t size = 0;
    std::mutex mu;
    std::map<size_t, size_t> free_off;        // offset -> length of free blocks
    std::multimap<size_t, size_t> free_len;   // length -> offset of free blocks
    std::map<size_t, size_t> used;            // offset -> length of live buffers
    size_t free_bytes = 0;

    static constexpr size_t ALIGN = 256;
    static size_t round(size_t n) { return (n + ALIGN - 1) / ALIGN * ALIGN; }

    void init(size_t n);
    void *alloc(size_t n);      // nullptr when no free block fits
    bool release(void *p);      // false when p is not a live buffer
    // would buffers of these sizes (in this order) fit now
    bool fits(const std::vector<size_t> &sizes);
    size_t largest();

  private:
    void insert_free(size_t off, size_t len);
    void erase_free(std::map<size_t, size_t>::iterator it);
};
