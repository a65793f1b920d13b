// Randomised check of zkx::Placement (the commit server's placement of
// requests in its device region) against a brute-force model: one flag per
// unit of the region.  Prints "ok <operations>" or the first mismatch.
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <utility>
#include <vector>

#include "placement.h"

int main(int argc, char **argv)
{
    const size_t units = argc > 1 ? strtoull(argv[1], 0, 10) : 512;
    const int ops = argc > 2 ? atoi(argv[2]) : 200000;
    std::mt19937_64 rng(argc > 3 ? strtoull(argv[3], 0, 10) : 1);
    zkx::Placement p;
    p.reset(units);
    std::vector<char> used(units, 0);
    std::vector<std::pair<size_t, size_t>> held;
    auto fail = [&](const char *what, int i) {
        printf("mismatch: %s at operation %d\n", what, i);
        exit(1);
    };
    for (int i = 0; i < ops; i++) {
        const int kind = (int)(rng() % 3);
        if (kind < 2) {
            const size_t n = 1 + rng() % (rng() % 4 == 0 ? units / 2 : units / 16);
            // brute force: lowest and highest offsets of n free units
            long lo = -1, hi = -1;
            for (size_t o = 0; o + n <= units; o++) {
                bool ok = true;
                for (size_t k = 0; k < n && ok; k++) ok = !used[o + k];
                if (ok) {
                    if (lo < 0) lo = (long)o;
                    hi = (long)o;
                }
            }
            size_t off = 0;
            bool placed;
            if (kind == 0) {
                const size_t limit = rng() % 2 ? units : rng() % (units + 1);
                const bool exp = lo >= 0 && (size_t)lo + n <= limit;
                placed = p.take_low(n, limit, off);
                if (placed != exp || (placed && off != (size_t)lo)) fail("take_low", i);
            } else {
                placed = p.take_high(n, off);
                if (placed != (hi >= 0) || (placed && off != (size_t)hi)) fail("take_high", i);
            }
            if (placed) {
                for (size_t k = 0; k < n; k++) {
                    if (used[off + k]) fail("placed over a used unit", i);
                    used[off + k] = 1;
                }
                held.push_back({off, n});
            }
        } else if (!held.empty()) {
            const size_t j = rng() % held.size();
            p.give(held[j].first, held[j].second);
            for (size_t k = 0; k < held[j].second; k++) used[held[j].first + k] = 0;
            held[j] = held.back();
            held.pop_back();
        }
        // free ranges == maximal runs of free units (coalesced)
        std::vector<std::pair<size_t, size_t>> runs;
        for (size_t o = 0; o < units;) {
            if (used[o]) {
                o++;
                continue;
            }
            size_t e = o;
            while (e < units && !used[e]) e++;
            runs.push_back({o, e - o});
            o = e;
        }
        const auto &fr = p.free_ranges();
        if (fr.size() != runs.size()) fail("number of free ranges", i);
        size_t k = 0;
        for (const auto &r : fr)
            if (r.first != runs[k].first || r.second != runs[k++].second) fail("free range", i);
    }
    printf("ok %d\n", ops);
    return 0;
}
