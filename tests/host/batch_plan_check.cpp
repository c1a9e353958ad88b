// CPU checks of the batched engine's host planning (mhd_gpu_solver_amd/csrc/mhd_batch_plan.hpp):
// every interior cell of every instance is marched exactly once, each instance has exactly one
// lead tile (its step bookkeeping), the per-wave offsets are consistent, the capacity fill
// balances the waves, the timed choice's shortlist ranks by the load model, and the replay
// accounting / progress check behave on fake step counts.
// Built and run by tests/test_batch_plan.py; prints "ok" and exits 0 when every check holds.
#include "mhd_batch_plan.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace mhdhip;

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                     \
            std::printf("\n");                            \
            ++failures;                                   \
        }                                                 \
    } while (0)

// each (row, strip) of every instance covered once; one lead tile per instance
static void check_cover(const std::vector<PlanInst>& insts, const WavePlan& p, int waves,
                        const char* what)
{
    CHECK(static_cast<int>(p.first.size()) == waves + 1, "%s: offsets size", what);
    CHECK(p.first.front() == 0 && p.first.back() == static_cast<int>(p.tiles.size()),
          "%s: offsets ends", what);
    for (int w = 0; w < waves; ++w) CHECK(p.first[w] <= p.first[w + 1], "%s: offsets order", what);
    for (const PlanInst& in : insts) {
        const StripShape s = strip_shape(in.nx, in.ny);
        const int nstrips = s.wide + (s.half ? 1 : 0);
        std::vector<int> hits(static_cast<size_t>(nstrips) * in.ny, 0);
        int leads = 0;
        for (const BatchTile& t : p.tiles) {
            if (t.inst != in.slot) continue;
            CHECK(t.strip >= 0 && t.strip < nstrips, "%s: strip index", what);
            CHECK(t.r0 >= kPlanGhosts && t.r0 < t.r1 && t.r1 <= kPlanGhosts + in.ny,
                  "%s: rows %d..%d of %d", what, t.r0, t.r1, in.ny);
            if (t.strip == 0 && t.r0 == kPlanGhosts) ++leads;
            const bool half = s.half && t.strip == s.wide;
            CHECK(half == (t.r0b >= 0), "%s: tile kind", what);
            for (int r = t.r0; r < t.r1; ++r) {
                if (half && r - kPlanGhosts >= s.half_rows) continue;
                ++hits[static_cast<size_t>(t.strip) * in.ny + (r - kPlanGhosts)];
            }
            if (half) {
                CHECK(t.r0b - t.r0 == s.half_rows, "%s: half offset", what);
                for (int r = t.r0b; r < t.r1b; ++r) {
                    ++hits[static_cast<size_t>(t.strip) * in.ny + (r - kPlanGhosts)];
                }
            }
        }
        CHECK(leads == 1, "%s: %d lead tiles for slot %d", what, leads, in.slot);
        for (size_t k = 0; k < hits.size(); ++k) {
            if (hits[k] != 1) {
                CHECK(false, "%s: slot %d (%dx%d) strip %zu row %zu marched %d times", what,
                      in.slot, in.nx, in.ny, k / in.ny, k % in.ny, hits[k]);
                break;
            }
        }
    }
}

int main()
{
    std::mt19937 rng(12345);
    // edge shapes: one-cell grids, narrow tails of 1..28 columns, 1..3 rows, wide grids
    std::vector<std::vector<PlanInst>> sets;
    sets.push_back({{0, 1, 1}});
    sets.push_back({{3, 61, 1}, {5, 88, 2}, {6, 28, 3}, {9, 121, 9}, {1, 5, 40}});
    sets.push_back({{0, 4096, 4096}});
    sets.push_back({{0, 64, 64}, {1, 96, 96}, {2, 128, 128}, {3, 192, 192}, {4, 256, 256}});
    {
        // the mixed-size ensemble's shapes
        std::vector<PlanInst> ens;
        const int sizes[5] = {64, 96, 128, 192, 256};
        for (int k = 0; k < 512; ++k) ens.push_back({k, sizes[rng() % 5], 0});
        for (auto& e : ens) e.ny = e.nx;
        sets.push_back(ens);
    }
    {
        std::vector<PlanInst> rnd;
        for (int k = 0; k < 40; ++k) rnd.push_back({k, 1 + static_cast<int>(rng() % 300),
                                                    1 + static_cast<int>(rng() % 300)});
        sets.push_back(rnd);
    }
    const int wave_counts[] = {1, 3, 64, 2048, 4096};
    for (const auto& set : sets) {
        for (const int waves : wave_counts) {
            for (const int rows : {1, 5, 32, 200}) {
                check_cover(set, plan_fixed(set, rows, waves), waves, "fixed");
            }
            for (const int min_rows : {1, 2, 4}) {
                const WavePlan p = plan_fill(set, waves, min_rows);
                check_cover(set, p, waves, "fill");
                // balance: the busiest wave carries at most the mean plus one piece's overhead
                // and one minimum piece, unless one strip alone is longer (it is cut anyway)
                const long long mean = (p.total_cost + waves - 1) / waves;
                CHECK(p.max_cost <= mean + tile_cost(min_rows) + 5 + 2 * min_rows ||
                          p.max_cost <= tile_cost(min_rows) + 5 + 2 * min_rows,
                      "fill balance: max %lld mean %lld (waves %d, min_rows %d)", p.max_cost,
                      mean, waves, min_rows);
            }
        }
    }
    // the ensemble: the fill beats round-robin fixed segments on the busiest wave
    {
        const auto& ens = sets[4];
        const WavePlan f = plan_fill(ens, 4096, 2);
        long long best_fixed = -1;
        for (const int rows : {8, 16, 24, 32, 48, 64}) {
            const long long c = plan_fixed(ens, rows, 4096).max_cost;
            best_fixed = best_fixed < 0 ? c : std::min(best_fixed, c);
        }
        std::printf("ensemble 4096 waves: fill max %lld (mean %lld), best fixed max %lld\n",
                    f.max_cost, f.total_cost / 4096, best_fixed);
        CHECK(f.max_cost < best_fixed, "fill %lld not below fixed %lld", f.max_cost, best_fixed);
    }
    // the load model: its cost is plan_fixed's, and it picks the cheapest candidate
    for (const auto& set : sets) {
        for (const int waves : {64, 4096}) {
            const std::vector<int> cands = {8, 16, 32, 64};
            const int pick = model_rows(set, waves, cands);
            long long best = -1;
            for (const int r : cands) {
                const long long c = plan_fixed(set, r, waves).max_cost;
                CHECK(c == fixed_max_cost(set, r, waves), "fixed_max_cost %d", r);
                best = best < 0 ? c : std::min(best, c);
            }
            CHECK(fixed_max_cost(set, pick, waves) == best, "model pick %d", pick);
        }
    }
    // the timed choice's shortlist: at most 3 lengths, cheapest first, all within 12 % of the
    // cheapest, which is the load model's pick; summed over launches; empty sets cost nothing
    for (const auto& set : sets) {
        const std::vector<int> cands = {8, 12, 16, 20, 24, 28, 32, 40, 48, 64, 96, 128};
        const std::vector<PlanInst> none;
        for (const int waves : {64, 2048, 4096}) {
            const std::vector<int> sl =
                shortlist_rows({PlanLaunch{&set, waves}, PlanLaunch{&none, waves}}, cands);
            CHECK(!sl.empty() && sl.size() <= 3, "shortlist size %zu", sl.size());
            if (sl.empty()) continue;
            const long long c0 = fixed_max_cost(set, sl[0], waves);
            CHECK(sl[0] == model_rows(set, waves, cands), "shortlist head %d", sl[0]);
            for (size_t i = 0; i < sl.size(); ++i) {
                const long long c = fixed_max_cost(set, sl[i], waves);
                CHECK(c * 100 <= c0 * 112, "shortlist %d beyond 12 %%", sl[i]);
                if (i > 0) CHECK(c >= fixed_max_cost(set, sl[i - 1], waves), "shortlist order");
            }
            for (const int r : cands) {  // nothing cheaper left out
                const bool in = std::find(sl.begin(), sl.end(), r) != sl.end();
                if (!in && sl.size() < 3) {
                    CHECK(fixed_max_cost(set, r, waves) * 100 > c0 * 112, "left out %d", r);
                }
            }
            // two launches of the same set: the same ranking
            const std::vector<int> two =
                shortlist_rows({PlanLaunch{&set, waves}, PlanLaunch{&set, waves}}, cands);
            CHECK(two == sl, "two equal launches rank alike");
        }
    }
    // even segments: a strip of n rows in ceil(n / rows) tiles whose lengths differ by at most
    // one, every row in exactly one of them
    for (int n = 1; n <= 300; ++n) {
        for (int rows = 1; rows <= 40; ++rows) {
            const int k = segment_count(n, rows);
            int lo = n, hi = 0, covered = 0;
            for (int j = 0; j < k; ++j) {
                const int len = segment_start(n, k, j + 1) - segment_start(n, k, j);
                lo = std::min(lo, len);
                hi = std::max(hi, len);
                covered += len;
            }
            CHECK(segment_start(n, k, 0) == 0 && covered == n && hi <= rows && hi - lo <= 1 &&
                      lo >= 1,
                  "segments of %d rows at %d: lengths %d..%d", n, rows, lo, hi);
        }
    }
    // the untimed choice on the queued hand-out (4096 waves): 12 rows for a set that fills
    // the waves less than twice, longer for more cells, never below kPickMinRows
    {
        std::vector<int> cands;
        for (const int r : {8, 12, 16, 20, 24, 28, 32, 40, 48, 64, 96, 128}) {
            if (r >= kPickMinRows) cands.push_back(r);
        }
        auto pick = [&](const std::vector<PlanInst>& set) {
            return model_rows_for(set, 4096, cands, HandOut::kQueued, kPickTileOverhead);
        };
        CHECK(pick({PlanInst{0, 1024, 1024}}) == 12, "one 1024^2: %d", pick({PlanInst{0, 1024, 1024}}));
        std::vector<PlanInst> big;
        for (int i = 0; i < 8; ++i) big.push_back(PlanInst{i, 1000, 1000});
        const int r8 = pick(big);
        CHECK(r8 >= 16 && r8 <= 32, "eight 1000^2: %d", r8);
    }
    // the XCD remap is a permutation of the workgroups, each XCD's logical ids contiguous
    for (int nb = 1; nb <= 3000; nb += (nb < 100 ? 1 : 37)) {
        for (const bool rev : {false, true}) {
            std::vector<int> seen(static_cast<size_t>(nb), 0);
            std::vector<int> lo(kPlanXcds, nb), hi(kPlanXcds, -1), cnt(kPlanXcds, 0);
            for (int p = 0; p < nb; ++p) {
                const int l = xcd_logical_block(p, nb, rev);
                CHECK(l >= 0 && l < nb, "remap of %d of %d: %d", p, nb, l);
                if (l < 0 || l >= nb) continue;
                ++seen[static_cast<size_t>(l)];
                const int x = p % kPlanXcds;
                lo[x] = std::min(lo[x], l);
                hi[x] = std::max(hi[x], l);
                ++cnt[x];
            }
            for (int v : seen) CHECK(v == 1, "remap of %d workgroups not a permutation", nb);
            for (int x = 0; x < kPlanXcds; ++x) {
                CHECK(cnt[x] == 0 || hi[x] - lo[x] + 1 == cnt[x], "XCD %d of %d not contiguous",
                      x, nb);
            }
        }
    }
    // the single-run grid (k_stage): one wave per (strip, row segment) in row-band order,
    // four per workgroup; every tile marched by exactly one wave
    for (int strips = 1; strips <= 70; strips += 3) {
        for (int segs = 1; segs <= 90; segs += 7) {
            const long long tiles = static_cast<long long>(strips) * segs;
            const int nb = static_cast<int>((tiles + 3) / 4);
            std::vector<int> hit(static_cast<size_t>(tiles), 0);
            for (const bool rev : {false, true}) {
                std::fill(hit.begin(), hit.end(), 0);
                for (int p = 0; p < nb; ++p) {
                    for (int w = 0; w < 4; ++w) {
                        const unsigned wl =
                            static_cast<unsigned>(xcd_logical_block(p, nb, rev)) * 4 + w;
                        const int strip = static_cast<int>(wl % static_cast<unsigned>(strips));
                        const int seg = static_cast<int>(wl / static_cast<unsigned>(strips));
                        if (seg >= segs) continue;
                        ++hit[static_cast<size_t>(seg) * strips + strip];
                    }
                }
                for (int h : hit) CHECK(h == 1, "k_stage grid %d x %d: a tile marched %d times",
                                        strips, segs, h);
            }
        }
    }
    // replay accounting
    {
        const std::vector<int> before = {0, 16, 5, 9};
        const std::vector<int> after = {8, 20, 5, 9};
        const std::vector<char> stepping = {1, 1, 0, 1};
        CHECK(replay_active_launches(before, after, stepping, 8) == 8, "active launches 8");
        CHECK(replay_active_launches({0, 16}, {3, 20}, {1, 1}, 8) == 4, "active launches 4");
        CHECK(replay_active_launches({4}, {4}, {1}, 8) == 0, "active launches 0");
        CHECK(replay_active_launches({0}, {30}, {1}, 8) == 8, "clamped to the replay");
        CHECK(replay_progressed(before, after, stepping, {0, 0, 0, 0}), "progress");
        CHECK(!replay_progressed({3, 7}, {3, 7}, {1, 1}, {0, 0}), "no progress");
        CHECK(replay_progressed({3, 7}, {3, 7}, {1, 1}, {0, 1}), "finished counts");
        CHECK(replay_progressed({3, 7}, {3, 7}, {0, 0}, {0, 0}), "nothing stepping");
        CHECK(!replay_progressed({3, 7}, {3, 9}, {1, 0}, {0, 0}), "only a paused one moved");
    }
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
