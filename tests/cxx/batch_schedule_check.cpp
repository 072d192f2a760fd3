// Host-side check of the batched launch's schedule (assemble_schedule in
// csrc/batch_schedule.cpp), no GPU: over random batches (empty items, one
// dominant item, many equal ones) every item's work units [0, units) are
// covered exactly once across the eight XCD lists; each list's "units
// before" column is its running count and the XCD offsets and unit counts
// agree; an item heavier than 1/32 of the batch (8 units or more) is cut
// into contiguous ranges in XCD order; carry offsets and split counts are
// prefix sums of the fragments'; operands and strides are bound from the
// items.  Built plainly and with the host code under AddressSanitizer by
// tests/test_batch_schedule.py.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "batch_schedule.hpp"
#include "common.hpp"

using namespace dspmm;

static uint64_t g_s = 1;
static uint64_t rnd() {
    g_s = g_s * 6364136223846793005ull + 1442695040888963407ull;
    return g_s >> 33;
}

static int check(const char *name, const std::vector<ItemFragment> &fr) {
    const int n = (int)fr.size();
    std::vector<dspmm_batch_item> items((size_t)n);
    std::vector<const ItemFragment *> fp((size_t)n);
    for (int i = 0; i < n; i++) {
        items[(size_t)i] = dspmm_batch_item{nullptr, 8, reinterpret_cast<const void *>((uintptr_t)(4096 * (i + 1))),
                                            8 + i, reinterpret_cast<float *>((uintptr_t)(8192 * (i + 1))), 16 + i};
        fp[(size_t)i] = &fr[(size_t)i];
    }
    BatchSchedule s;
    if (assemble_schedule(fp.data(), items.data(), n, s) != DSPMM_STATUS_SUCCESS) {
        std::printf("%s: refused\n", name);
        return 1;
    }
    double total = 0;
    int64_t units = 0;
    for (const auto &f : fr) {
        total += f.cost;
        units += f.units;
    }
    std::vector<std::vector<int>> cover((size_t)n);
    for (int i = 0; i < n; i++) cover[(size_t)i].assign((size_t)fr[(size_t)i].units, 0);
    if ((int)s.xoff.size() != 2 * kNumXcd + 1 || s.xoff[0] != 0 || s.xoff[kNumXcd] != (int)s.segs.size()) {
        std::printf("%s: xoff shape\n", name);
        return 1;
    }
    int64_t nunits = 0;
    int maxb = 0;
    std::vector<int> last_x((size_t)n, -1), next_begin((size_t)n, 0);
    for (int x = 0; x < kNumXcd; x++) {
        int before = 0;
        for (int k = s.xoff[(size_t)x]; k < s.xoff[(size_t)x + 1]; k++) {
            const int4 g = s.segs[(size_t)k];
            if (g.x < 0 || g.x >= n || g.z <= 0 || g.y < 0 || g.y + g.z > fr[(size_t)g.x].units || g.w != before) {
                std::printf("%s: segment %d of XCD %d\n", name, k, x);
                return 1;
            }
            for (int u = g.y; u < g.y + g.z; u++) cover[(size_t)g.x][(size_t)u]++;
            // a cut item: its ranges follow one another in XCD order
            if (g.z != fr[(size_t)g.x].units) {
                if (last_x[(size_t)g.x] >= x || g.y != next_begin[(size_t)g.x]) {
                    std::printf("%s: item %d cut out of order\n", name, g.x);
                    return 1;
                }
                last_x[(size_t)g.x] = x;
                next_begin[(size_t)g.x] = g.y + g.z;
                if (!(fr[(size_t)g.x].cost > total / 32.0 && fr[(size_t)g.x].units >= kNumXcd)) {
                    std::printf("%s: a light item %d cut\n", name, g.x);
                    return 1;
                }
            }
            before += g.z;
        }
        if (s.xoff[(size_t)kNumXcd + 1 + x] != before || s.units_x[x] != before) {
            std::printf("%s: unit count of XCD %d\n", name, x);
            return 1;
        }
        nunits += before;
        maxb = std::max(maxb, div_up(before, kBlock / kWave));
    }
    for (int i = 0; i < n; i++)
        for (int c : cover[(size_t)i])
            if (c != 1) {
                std::printf("%s: item %d units covered %d times\n", name, i, c);
                return 1;
            }
    if (nunits != units || s.nunits != units || s.max_blocks_x != maxb) {
        std::printf("%s: totals\n", name);
        return 1;
    }
    int64_t carry = 0;
    int splits = 0;
    for (int i = 0; i < n; i++) {
        const ItemFragment &f = fr[(size_t)i];
        const ItemDesc &d = s.items[(size_t)i];
        if (s.carry_off[(size_t)i] != carry || s.split_prefix[(size_t)i] != splits ||
            s.carry_bytes[(size_t)i] != sizeof(float) * (size_t)f.nslots * (size_t)f.d.ldcarry ||
            d.B != items[(size_t)i].B || d.C != items[(size_t)i].C || d.ldb != items[(size_t)i].ldb ||
            d.ldc != items[(size_t)i].ldc || d.carry != nullptr || d.ntiles != f.d.ntiles) {
            std::printf("%s: item %d bindings or prefix sums\n", name, i);
            return 1;
        }
        carry += (int64_t)f.nslots * f.d.ldcarry;
        splits += f.nsplits;
    }
    if (s.carry_floats != carry || s.nsplits != splits || s.split_prefix[(size_t)n] != splits) {
        std::printf("%s: carry or split totals\n", name);
        return 1;
    }
    std::printf("%s: ok, %d items, %lld units, %d segments\n", name, n, (long long)units, (int)s.segs.size());
    return 0;
}

static ItemFragment frag(int units, double cost) {
    ItemFragment f;
    f.units = units;
    f.cost = cost;
    f.nslots = (int)(rnd() % 5);
    f.nsplits = f.nslots ? 1 + (int)(rnd() % 2) : 0;
    f.d.ldcarry = 8 * (1 + (int)(rnd() % 4));
    f.d.ntiles = units * 2;
    f.lds_ints = 1 + (int)(rnd() % 100);
    return f;
}

int main() {
    int rc = 0, ok = 0, n = 0;
    {
        g_s = 3;
        std::vector<ItemFragment> fr;
        for (int i = 0; i < 512; i++) {
            const int u = (int)(rnd() % 300);
            fr.push_back(frag(u, u * (1.0 + (double)(rnd() % 100) / 10.0)));
        }
        rc |= check("512 random items", fr) ? 1 : (ok++, 0);
        n++;
    }
    {
        g_s = 5;
        std::vector<ItemFragment> fr;
        fr.push_back(frag(100000, 1e9));            // dominates: cut over the XCDs
        for (int i = 0; i < 40; i++) fr.push_back(frag(1 + (int)(rnd() % 50), 1e3));
        fr.push_back(frag(0, 0));                   // an empty matrix
        fr.push_back(frag(5, 1e9));                 // heavy but under 8 units: whole
        rc |= check("one dominant item", fr) ? 1 : (ok++, 0);
        n++;
    }
    {
        g_s = 7;
        std::vector<ItemFragment> fr;
        for (int i = 0; i < 3; i++) fr.push_back(frag(0, 0));
        rc |= check("only empty items", fr) ? 1 : (ok++, 0);
        n++;
    }
    {
        g_s = 9;
        std::vector<ItemFragment> fr;
        for (int i = 0; i < 20; i++) fr.push_back(frag(64, 64.0));   // 20 equal: each over 1/32
        rc |= check("twenty equal items", fr) ? 1 : (ok++, 0);
        n++;
    }
    {
        g_s = 11;
        std::vector<ItemFragment> fr;
        fr.push_back(frag(7, 1.0));
        rc |= check("one item", fr) ? 1 : (ok++, 0);
        n++;
    }
    std::printf("%d of %d checks ok\n", ok, n);
    return rc;
}
