// Host-side check of the serving arena's first-fit bookkeeping
// (csrc/first_fit.hpp), no GPU: random allocations and releases against a
// model of the live blocks.  After every step: live blocks are 256-byte
// aligned, inside the region and disjoint; free blocks are coalesced (no two
// touch) and, with the live ones, tile the region exactly; used() is the sum
// of the live blocks and peak() its maximum; each allocation returns the
// lowest free block that fits, and fails only when no free block does.
// Built plainly and with AddressSanitizer by tests/test_first_fit.py.
#include <cstdint>
#include <cstdio>
#include <map>
#include <vector>

#include "first_fit.hpp"

using dspmm::FirstFit;

static uint64_t g_s = 1;
static uint64_t rnd() {
    g_s = g_s * 6364136223846793005ull + 1442695040888963407ull;
    return g_s >> 33;
}

static int run(const char *name, size_t region, size_t max_req, int steps, uint64_t seed) {
    g_s = seed;
    FirstFit ff;
    ff.reset(region);
    std::map<size_t, size_t> live;     // offset -> granule bytes
    size_t used = 0, peak = 0;
    int fails = 0;
    for (int step = 0; step < steps; step++) {
        if (live.empty() || rnd() % 100 < 55) {
            const size_t n = 1 + rnd() % max_req, g = FirstFit::granule(n);
            size_t want = SIZE_MAX;
            for (const auto &fb : ff.free_blocks())
                if (fb.second >= g) {
                    want = fb.first;
                    break;
                }
            size_t off = 0;
            const bool ok = ff.alloc(n, &off);
            if (ok != (want != SIZE_MAX) || (ok && off != want)) {
                std::printf("%s: step %d: alloc of %zu gave %d at %zu, first fit is %zu\n", name, step, n,
                            (int)ok, off, want);
                return 1;
            }
            if (!ok) {
                fails++;
                continue;
            }
            live[off] = g;
            used += g;
            peak = std::max(peak, used);
        } else {
            auto it = live.begin();
            std::advance(it, (long)(rnd() % live.size()));
            // release with the caller's unrounded size when it rounds to the same granule
            ff.release(it->first, it->second - (rnd() % 2 ? 255 : 0));
            used -= it->second;
            live.erase(it);
        }
        // invariants
        size_t pos = 0, prev_free_end = SIZE_MAX;
        auto li = live.begin();
        auto fi = ff.free_blocks().begin();
        while (pos < region) {
            if (li != live.end() && li->first == pos) {
                pos += li->second;
                ++li;
            } else if (fi != ff.free_blocks().end() && fi->first == pos) {
                if (prev_free_end == pos) {
                    std::printf("%s: step %d: free blocks touch at %zu\n", name, step, pos);
                    return 1;
                }
                pos += fi->second;
                prev_free_end = pos;
                ++fi;
                continue;
            } else {
                std::printf("%s: step %d: nothing at %zu (a gap or an overlap)\n", name, step, pos);
                return 1;
            }
            prev_free_end = SIZE_MAX;
        }
        if (pos != region || li != live.end() || fi != ff.free_blocks().end() || ff.used() != used ||
            ff.peak() != peak) {
            std::printf("%s: step %d: region not tiled exactly, or used/peak off\n", name, step);
            return 1;
        }
        for (const auto &b : live)
            if (b.first % 256 || b.second % 256) {
                std::printf("%s: step %d: unaligned block\n", name, step);
                return 1;
            }
    }
    std::printf("%s: ok, %d steps, %d refused, peak %zu of %zu\n", name, steps, fails, peak, region);
    return 0;
}

int main() {
    int rc = 0, ok = 0;
    const struct { const char *name; size_t region, max_req; int steps; uint64_t seed; } cases[] = {
        {"small region, near-full", 64 * 256, 2000, 20000, 3},
        {"large region, mixed sizes", 1 << 24, 1 << 20, 20000, 5},
        {"one-granule requests", 4096, 200, 5000, 7},
        {"requests as large as the region", 8 * 256, 8 * 256, 5000, 9},
    };
    for (const auto &c : cases) {
        const int e = run(c.name, c.region, c.max_req, c.steps, c.seed);
        rc |= e;
        ok += !e;
    }
    std::printf("%d of %d checks ok\n", ok, (int)(sizeof(cases) / sizeof(cases[0])));
    return rc;
}
