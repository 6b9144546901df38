// Host-side exhaustive check of the update kernel's work plans
// (decompose() and the partial-buffer bounds in h4_kernels.hpp): for every
// active count of the tested system sizes and every lane-group width, the
// chunks of a plan cover the j tiles exactly and the plan's j-split
// partials fit the bound the library allocates; random ensembles check the
// summed bound; partial_at is one to one in both layouts.  Built and run
// by tests/test_plan_bounds.py; exit status 0 means every case held.
#include "h4_kernels.hpp"

#include <cstdio>
#include <random>
#include <vector>

using namespace h4;

static int failures = 0;

static void check_single(uint32_t n, uint32_t grid, uint32_t max_split) {
    const uint32_t n_jt = (n + JT - 1) / JT;
    const uint32_t widths[4] = {4, 16, 64, 256};
    const unsigned long long bound = partial_bound(n, grid * QUEUE_FACTOR);
    for (uint32_t nact = 1; nact <= n; ++nact) {
        for (uint32_t iw : widths) {
            const Decomp d = decompose(nact, iw, n_jt, grid, max_split);
            const bool covers = d.S >= 1 && (unsigned long long)d.S * d.tiles_per_chunk >= n_jt &&
                                (unsigned long long)(d.S - 1) * d.tiles_per_chunk < n_jt;
            const unsigned long long part = d.S > 1 ? 6ull * d.S * d.rows : 0ull;
            if (!covers || part > bound || d.n_items != d.n_it * d.S ||
                d.rows != d.n_it * iw || (unsigned long long)d.n_it * iw < nact) {
                if (failures++ < 10)
                    std::printf("FAIL n=%u grid=%u max_split=%u nact=%u iw=%u: S=%u tpc=%u "
                                "rows=%u partials=%llu bound=%llu\n",
                                n, grid, max_split, nact, iw, d.S, d.tiles_per_chunk, d.rows,
                                part, bound);
            }
        }
    }
}

static void check_ensembles(uint32_t trials, uint32_t grid) {
    const uint32_t target = grid * QUEUE_FACTOR;
    std::mt19937_64 rng(20261202);
    for (uint32_t t = 0; t < trials; ++t) {
        const uint32_t K = 2 + rng() % 80;
        std::vector<uint32_t> n(K), nact(K);
        unsigned long long n_total = 0, per_cluster = 0;
        double W = 0.0;
        for (uint32_t c = 0; c < K; ++c) {
            n[c] = 1 + rng() % 40000;
            nact[c] = (rng() % 4 == 0) ? 0 : 1 + rng() % n[c];
            n_total += n[c];
            per_cluster += partial_bound(n[c], target);
            W += (double)nact[c] * (double)n[c];
        }
        unsigned long long used = 0;
        for (uint32_t c = 0; c < K; ++c) {
            if (!nact[c]) continue;
            // the share rule of the device plan (plan_cluster)
            const double share = W > 0.0 ? (double)nact[c] * (double)n[c] / W : 0.0;
            uint32_t gc = (uint32_t)(share * grid);
            if (gc < 1) gc = 1;
            const Decomp d = decompose(nact[c], pick_iw(nact[c], DEFAULT_IW_FILL),
                                       (n[c] + JT - 1) / JT, gc, 1024);
            if (d.S > 1) used += 6ull * d.S * d.rows;
        }
        const unsigned long long bound =
            per_cluster < ensemble_partial_bound(K, n_total, target)
                ? per_cluster
                : ensemble_partial_bound(K, n_total, target);
        if (used > bound && failures++ < 10)
            std::printf("FAIL ensemble trial %u K=%u: partials %llu > bound %llu\n", t, K, used,
                        bound);
    }
}

// partial_at maps (component, chunk, particle slot) one to one onto the
// plan's 6 S rows doubles in both layouts.
static void check_layouts() {
    const uint32_t shapes[][2] = {{2, 4}, {3, 64}, {32, 320}, {342, 320}, {1024, 8}, {17, 256}};
    for (const auto &sh : shapes)
        for (uint32_t cm = 0; cm < 2; ++cm) {
            ClPlan pl{};
            pl.S = sh[0];
            pl.rows = sh[1];
            pl.cmaj = cm;
            const size_t total = 6ull * pl.S * pl.rows;
            std::vector<unsigned char> seen(total, 0);
            bool ok = true;
            for (uint32_t q = 0; q < 6; ++q)
                for (uint32_t sc = 0; sc < pl.S; ++sc)
                    for (uint32_t k = 0; k < pl.rows; ++k) {
                        const size_t at = partial_at(pl, q, sc, k);
                        if (at >= total || seen[at]++) ok = false;
                    }
            if (!ok && failures++ < 10)
                std::printf("FAIL partial_at S=%u rows=%u cmaj=%u\n", pl.S, pl.rows, cm);
        }
}

int main() {
    check_layouts();
    const uint32_t sizes[] = {1, 2, 3, 255, 256, 257, 1000, 1024, 4097, 65536};
    const uint32_t grids[] = {1, 7, 256, 2048, 4096};
    const uint32_t splits[] = {1, 16, 1024, 1u << 30};
    for (uint32_t n : sizes)
        for (uint32_t gr : grids)
            for (uint32_t ms : splits) check_single(n, gr, ms);
    check_single(262144, 2048, 1024);
    check_single(262144, 128, 1024);
    check_single(300000, 2048, 1024);
    check_ensembles(2000, 4096);
    check_ensembles(2000, 256);
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("plan bounds ok\n");
    return 0;
}
