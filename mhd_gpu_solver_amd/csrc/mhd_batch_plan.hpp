// Host-side planning of the batched engine, free of HIP so that it compiles and is tested on
// the CPU as well:
//   * the work partition: the tiles (instance, strip, row range) of the persistent stage
//     kernel, either as per-wave lists (CSR offsets into a tile array) or as one queue per
//     XCD from which the waves take tiles as they finish (plan_queues);
//   * the accounting of a replay of timed steps: how many of its launches stepped anything;
//   * the progress check that turns a replay that advanced no stepping instance into an error.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <vector>

namespace mhdhip {

constexpr int kPlanStripCells = 60;  // cells per wave strip (64 lanes, 2 halo lanes each side)
constexpr int kPlanHalfCells = 28;   // cells per half-wave strip (narrow tail strips)
constexpr int kPlanGhosts = 2;

struct BatchTile {
    int inst;
    int strip;
    int r0, r1;   // global rows [r0, r1) marched by the wave (by its lower half when r0b >= 0)
    int r0b, r1b; // half-wave tile of a narrow strip: the upper half's rows start at r0b and
                  // are updated below r1b; r0b < 0 for a whole-wave tile
    int flip = 0; // march the rows in the other direction (the odd bands of an instance)
};

// One instance as the planner sees it: its slot and interior size.
struct PlanInst {
    int slot;
    int nx, ny;
};

// The strips of an nx x ny grid: `wide` strips of up to 60 columns marched over all ny rows,
// and, when the last strip is at most 28 columns wide, that strip as a half-wave strip whose
// wave pairs a chunk of its lower ceil(ny/2) rows (lanes 0..31) with the same chunk of its
// upper rows (lanes 32..63); it then has `half_rows` = ceil(ny/2) rows to march.
struct StripShape {
    int wide;       // whole-wave strips
    bool half;      // a half-wave tail strip follows them (strip index `wide`)
    int half_rows;  // rows the half-wave strip marches
};

inline StripShape strip_shape(int nx, int ny)
{
    const int full = nx / kPlanStripCells;
    const int tail = nx - full * kPlanStripCells;
    StripShape s{};
    s.half = tail > 0 && tail <= kPlanHalfCells;
    s.wide = s.half ? full : (nx + kPlanStripCells - 1) / kPlanStripCells;
    s.half_rows = s.half ? (ny + 1) / 2 : 0;
    return s;
}

// Tile marching rows [a0, a1) (counted from the first interior row) of strip `st`.
inline BatchTile make_tile(int slot, int nx, int ny, int st, int a0, int a1)
{
    const StripShape s = strip_shape(nx, ny);
    const int ng = kPlanGhosts;
    if (s.half && st == s.wide) {
        const int h = s.half_rows;
        const int r0b = ng + h + a0;
        return BatchTile{slot, st, ng + a0, ng + a1, r0b, std::min(r0b + (a1 - a0), ng + ny)};
    }
    return BatchTile{slot, st, ng + a0, ng + a1, -1, 0};
}

// Cost of a tile of L rows in half-rows: it marches L + 2 rows and loads L + 3 (`overhead`
// half-rows for the three extra row loads and the tile's start-up; see kPickTileOverhead).
inline long long tile_cost(int rows, int overhead = 5) { return 2ll * rows + overhead; }

// Per-wave tile lists: the tiles of wave w are tiles[first[w] .. first[w + 1]).
struct WavePlan {
    std::vector<BatchTile> tiles;
    std::vector<int> first;  // waves + 1 offsets
    long long max_cost = 0;  // the busiest wave's cost (half-rows)
    long long total_cost = 0;
};

// Fixed segments of `rows` rows over every strip, tile k handed to wave k mod `waves` (the
// persistent kernel's round-robin over the whole tile list).
inline WavePlan plan_fixed(const std::vector<PlanInst>& insts, int rows, int waves)
{
    rows = std::max(1, rows);
    waves = std::max(1, waves);
    std::vector<BatchTile> all;
    for (const PlanInst& in : insts) {
        const StripShape s = strip_shape(in.nx, in.ny);
        for (int a0 = 0; a0 < in.ny; a0 += rows) {
            for (int st = 0; st < s.wide; ++st) {
                all.push_back(make_tile(in.slot, in.nx, in.ny, st, a0, std::min(a0 + rows, in.ny)));
            }
        }
        for (int a0 = 0; a0 < s.half_rows; a0 += rows) {
            all.push_back(
                make_tile(in.slot, in.nx, in.ny, s.wide, a0, std::min(a0 + rows, s.half_rows)));
        }
    }
    WavePlan p;
    std::vector<int> count(static_cast<size_t>(waves), 0);
    std::vector<long long> load(static_cast<size_t>(waves), 0);
    for (size_t k = 0; k < all.size(); ++k) {
        const size_t w = k % static_cast<size_t>(waves);
        ++count[w];
        load[w] += tile_cost(all[k].r1 - all[k].r0);
        p.total_cost += tile_cost(all[k].r1 - all[k].r0);
    }
    p.first.assign(static_cast<size_t>(waves) + 1, 0);
    for (int w = 0; w < waves; ++w) {
        p.first[w + 1] = p.first[w] + count[w];
    }
    p.tiles.resize(all.size());
    std::vector<int> pos(p.first.begin(), p.first.end() - 1);
    for (size_t k = 0; k < all.size(); ++k) {
        p.tiles[pos[k % static_cast<size_t>(waves)]++] = all[k];
    }
    p.max_cost = load.empty() ? 0 : *std::max_element(load.begin(), load.end());
    return p;
}

// Busiest wave's cost (half-rows) of plan_fixed(insts, rows, waves), without building it.
inline long long fixed_max_cost(const std::vector<PlanInst>& insts, int rows, int waves)
{
    rows = std::max(1, rows);
    waves = std::max(1, waves);
    std::vector<long long> load(static_cast<size_t>(waves), 0);
    long long k = 0;
    auto add = [&](int len) { load[static_cast<size_t>(k++ % waves)] += tile_cost(len); };
    for (const PlanInst& in : insts) {
        const StripShape s = strip_shape(in.nx, in.ny);
        for (int a0 = 0; a0 < in.ny; a0 += rows) {
            for (int st = 0; st < s.wide; ++st) add(std::min(rows, in.ny - a0));
        }
        for (int a0 = 0; a0 < s.half_rows; a0 += rows) add(std::min(rows, s.half_rows - a0));
    }
    return *std::max_element(load.begin(), load.end());
}

// Load model: the fixed row length (of `candidates`) whose round-robin hand-out leaves the
// busiest wave the least work, the longer length on ties (fewer halo re-reads). Round-robin
// hands neighbouring strips of one row band to consecutive waves, which share a workgroup and
// an XCD, so the strips' shared halo lines are fetched once for all of them.
inline int model_rows(const std::vector<PlanInst>& insts, int waves,
                      const std::vector<int>& candidates)
{
    int best = candidates.empty() ? 32 : candidates.front();
    long long best_cost = std::numeric_limits<long long>::max();
    for (const int rows : candidates) {
        const long long c = fixed_max_cost(insts, rows, waves);
        if (c < best_cost || (c == best_cost && rows > best)) {
            best_cost = c;
            best = rows;
        }
    }
    return best;
}

// One persistent launch as the load model sees it: the instances it steps and its wave count.
struct PlanLaunch {
    const std::vector<PlanInst>* insts;
    int waves;
};

// The timed row choice's candidates: the fixed row lengths of least load-model cost (the
// busiest wave's cost, summed over `launches`), at most `count` of them and within `pct` % of
// the least, cheapest first (equal costs: the longer length first).
inline std::vector<int> shortlist_rows(const std::vector<PlanLaunch>& launches,
                                       const std::vector<int>& candidates, int count = 3,
                                       int pct = 12)
{
    std::vector<std::pair<long long, int>> cost;
    for (const int rows : candidates) {
        long long c = 0;
        for (const PlanLaunch& l : launches) {
            if (!l.insts->empty()) c += fixed_max_cost(*l.insts, rows, l.waves);
        }
        cost.emplace_back(c, -rows);
    }
    std::sort(cost.begin(), cost.end());
    std::vector<int> out;
    for (size_t i = 0; i < cost.size() && static_cast<int>(out.size()) < count; ++i) {
        if (cost[i].first * 100 <= cost.front().first * (100 + pct)) out.push_back(-cost[i].second);
    }
    return out;
}

// Capacity fill: the strips of all instances, in instance order, are cut into row pieces and
// poured into the waves in order, each wave filled up to a common capacity C (half-rows,
// piece overheads included), so every wave carries about the same work in one round of the
// persistent grid. Consecutive waves take consecutive pieces (the same instance, neighbouring
// strips), and consecutive logical waves share an XCD, so an instance's halo rows and strip
// edges meet in one L2. C is the least capacity for which the pour fits in `waves` waves.
// Pieces are at least `min_rows` rows unless a strip's remainder is shorter.
inline WavePlan plan_fill(const std::vector<PlanInst>& insts, int waves, int min_rows = 2)
{
    waves = std::max(1, waves);
    min_rows = std::max(1, min_rows);
    struct Strip {
        int inst;  // index into insts
        int st;
        int rows;
    };
    std::vector<Strip> strips;
    for (size_t i = 0; i < insts.size(); ++i) {
        const StripShape s = strip_shape(insts[i].nx, insts[i].ny);
        for (int st = 0; st < s.wide; ++st) {
            strips.push_back(Strip{static_cast<int>(i), st, insts[i].ny});
        }
        if (s.half) {
            strips.push_back(Strip{static_cast<int>(i), s.wide, s.half_rows});
        }
    }
    WavePlan p;
    p.first.assign(static_cast<size_t>(waves) + 1, 0);
    if (strips.empty()) {
        return p;
    }
    // Pour with capacity C. Returns the waves it needs (more than `waves`: C is too small).
    // With `out`, writes the tiles and the offsets; the waves past the last one stay empty.
    auto pour = [&](long long C, std::vector<BatchTile>* out, std::vector<int>* first,
                    long long* max_cost) {
        int w = 0;
        long long used = 0;
        long long worst = 0;
        for (const Strip& s : strips) {
            int a = 0;
            while (a < s.rows) {
                const int left = s.rows - a;
                const int need = std::min(left, min_rows);
                long long fit = (C - used - 5) / 2;  // rows that fit in what this wave has left
                if (fit < need) {
                    if (used > 0 && w + 1 < waves + (out ? 0 : 1)) {
                        // this wave is full: close it and pour into the next one
                        worst = std::max(worst, used);
                        if (first) (*first)[static_cast<size_t>(w) + 1] = static_cast<int>(out->size());
                        ++w;
                        used = 0;
                        if (w >= waves) return w + 1;
                        continue;
                    }
                    fit = need;  // an empty wave (or the last one) takes the piece anyway
                }
                int L = static_cast<int>(std::min<long long>(left, fit));
                if (left - L > 0 && left - L < min_rows) {
                    // leave no stub shorter than min_rows behind
                    L = left - min_rows >= need ? left - min_rows : left;
                }
                if (out) {
                    const PlanInst& in = insts[static_cast<size_t>(s.inst)];
                    out->push_back(make_tile(in.slot, in.nx, in.ny, s.st, a, a + L));
                }
                used += tile_cost(L);
                a += L;
            }
        }
        worst = std::max(worst, used);
        if (first) {
            for (int k = w + 1; k <= waves; ++k) {
                (*first)[static_cast<size_t>(k)] = static_cast<int>(out->size());
            }
        }
        if (max_cost) *max_cost = worst;
        return w + 1;
    };
    long long total = 0;
    for (const Strip& s : strips) total += tile_cost(s.rows);
    long long lo = std::max<long long>(tile_cost(min_rows), (total + waves - 1) / waves);
    long long hi = total + 5;  // one wave takes everything
    while (lo < hi) {
        const long long mid = lo + (hi - lo) / 2;
        if (pour(mid, nullptr, nullptr, nullptr) <= waves) {
            hi = mid;
        } else {
            lo = mid + 1;
        }
    }
    p.tiles.clear();
    pour(lo, &p.tiles, &p.first, &p.max_cost);
    for (const BatchTile& t : p.tiles) p.total_cost += tile_cost(t.r1 - t.r0);
    return p;
}

// ---- dynamic hand-out: one tile queue per XCD ----------------------------------------------
// The stage kernel's waves on XCD x take the tiles of queue x: wave lw of the W_x waves there
// first takes tile lw, then tickets W_x, W_x + 1, ... from the queue's counter (an atomic add),
// until a ticket is past the queue's end. A wave that finishes early takes the next tile, so
// the waves end together without a plan of who takes what; the tiles running at the same time
// on one XCD are neighbours in the list, so their shared halo rows and strip edges meet in one
// L2. Each stage's kernel zeroes the other stage's counters of its variant when it starts
// (the stages alternate on one stream), so no wave has to know that it is the last one.
constexpr int kPlanXcds = 8;

// Workgroups are dispatched round-robin over the XCDs (physical id p runs on XCD p % 8). The
// logical id of workgroup p of nb: each XCD receives a contiguous range of logical ids, in the
// order its workgroups are dispatched (reversed within each XCD with `reverse`). (constexpr:
// the stage kernels use it.)
constexpr int xcd_logical_block(int p, int nb, bool reverse = false)
{
    const int q = nb / kPlanXcds;
    const int r = nb % kPlanXcds;
    const int xcd = p % kPlanXcds;
    const int n = xcd < r ? q + 1 : q;  // workgroups on this XCD
    const int local = reverse ? n - 1 - p / kPlanXcds : p / kPlanXcds;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}
static_assert(xcd_logical_block(0, 20) == 0 && xcd_logical_block(8, 20) == 1 &&
                  xcd_logical_block(1, 20) == 3 && xcd_logical_block(19, 20) == 11 &&
                  xcd_logical_block(15, 20) == 19 &&
                  xcd_logical_block(0, 20, true) == 2,
              "XCD-contiguous logical workgroups");

// Queues of a launch of `blocks` workgroups: one per XCD, or one per workgroup when there are
// fewer workgroups than XCDs (every queue needs a wave). Workgroup b takes queue b mod
// queues, which is its XCD (workgroups are dispatched round-robin over the XCDs).
constexpr int launch_queues(int blocks)
{
    return blocks < kPlanXcds ? (blocks < 1 ? 1 : blocks) : kPlanXcds;
}
// Waves of `wpb` per workgroup that take queue q, and the index of wave `wave` of workgroup
// `block` among the waves of its queue. (constexpr: the stage kernel uses them too.)
constexpr int queue_waves(int blocks, int q, int wpb)
{
    return q < blocks ? ((blocks - 1 - q) / launch_queues(blocks) + 1) * wpb : 0;
}
constexpr int queue_wave_index(int blocks, int block, int wave, int wpb)
{
    return (block / launch_queues(blocks)) * wpb + wave;
}
static_assert(queue_waves(1024, 0, 4) == 512 && queue_waves(1025, 0, 4) == 516 &&
                  queue_waves(1025, 1, 4) == 512 && queue_waves(3, 2, 4) == 4 &&
                  launch_queues(3) == 3 && queue_wave_index(3, 2, 1, 4) == 1,
              "waves per queue");

struct QueuePlan {
    std::vector<BatchTile> tiles;  // queue x holds tiles[begin[x] .. begin[x + 1])
    std::vector<int> begin;        // queues + 1 offsets
    long long total_cost = 0;
    long long max_queue_cost = 0;  // the costliest queue (half-rows)
};

// A strip of n rows cut into segments of at most `rows` rows: ceil(n / rows) segments whose
// lengths differ by at most one, segment j starting at segment_start(n, count, j). (A cut at
// every `rows` rows leaves a short last segment, e.g. 128 rows at 24 end in one of 8; a short
// tile still pays its halo rows and its start-up, and measured much slower per row.)
constexpr int segment_count(int n, int rows) { return (n + rows - 1) / rows; }
constexpr int segment_start(int n, int count, int j)
{
    return static_cast<int>(static_cast<long long>(n) * j / count);
}
static_assert(segment_count(128, 24) == 6 && segment_start(128, 6, 1) == 21 &&
                  segment_start(128, 6, 6) == 128 && segment_count(64, 16) == 4 &&
                  segment_start(64, 4, 3) == 48,
              "even segments");

// Segments of at most `rows` rows, even within each strip, queued largest instance first (the
// short tiles of the small grids come last, where the hand-out ends), each instance band by
// band, a band's strips left to right (its half-wave tile last); the tiles of odd bands march
// the other way (`flip`), so that the rows on each side of a band boundary are loaded by both
// bands' waves at the same end of their march. The list is cut into `queues` contiguous
// queues of about equal load-model cost (launch_queues of the launch's grid).
inline QueuePlan plan_queues(std::vector<PlanInst> insts, int rows, int queues = kPlanXcds,
                             int overhead = 5)
{
    const int xcds = std::max(1, queues);
    rows = std::max(1, rows);
    std::stable_sort(insts.begin(), insts.end(), [](const PlanInst& a, const PlanInst& b) {
        return static_cast<long long>(a.nx) * a.ny > static_cast<long long>(b.nx) * b.ny;
    });
    QueuePlan p;
    for (const PlanInst& in : insts) {
        const StripShape s = strip_shape(in.nx, in.ny);
        const int kw = segment_count(in.ny, rows);
        const int kh = s.half ? segment_count(s.half_rows, rows) : 0;  // kh <= kw
        for (int j = 0; j < kw; ++j) {
            const int a0 = segment_start(in.ny, kw, j);
            const int a1 = segment_start(in.ny, kw, j + 1);
            for (int st = 0; st < s.wide; ++st) {
                p.tiles.push_back(make_tile(in.slot, in.nx, in.ny, st, a0, a1));
                p.tiles.back().flip = j & 1;
            }
            if (j < kh) {
                p.tiles.push_back(make_tile(in.slot, in.nx, in.ny, s.wide,
                                            segment_start(s.half_rows, kh, j),
                                            segment_start(s.half_rows, kh, j + 1)));
                p.tiles.back().flip = j & 1;
            }
        }
    }
    std::vector<long long> prefix(p.tiles.size() + 1, 0);
    for (size_t k = 0; k < p.tiles.size(); ++k) {
        prefix[k + 1] = prefix[k] + tile_cost(p.tiles[k].r1 - p.tiles[k].r0, overhead);
    }
    p.total_cost = prefix.back();
    p.begin.assign(static_cast<size_t>(xcds) + 1, 0);
    size_t k = 0;
    for (int x = 1; x < xcds; ++x) {
        // the first cut at or past x / xcds of the cost, or the one before it when nearer
        const long long target = p.total_cost * x / xcds;
        while (k < p.tiles.size() && prefix[k] < target) ++k;
        size_t cut = k;
        if (cut > 0 && target - prefix[cut - 1] < prefix[cut] - target) --cut;
        cut = std::max(cut, static_cast<size_t>(p.begin[static_cast<size_t>(x) - 1]));
        p.begin[static_cast<size_t>(x)] = static_cast<int>(cut);
    }
    p.begin[static_cast<size_t>(xcds)] = static_cast<int>(p.tiles.size());
    for (int x = 0; x < xcds; ++x) {
        p.max_queue_cost = std::max(p.max_queue_cost,
                                    prefix[static_cast<size_t>(p.begin[x + 1])] -
                                        prefix[static_cast<size_t>(p.begin[x])]);
    }
    return p;
}

// Load model of the queued hand-out on `waves` waves: the costliest queue spread over its
// XCD's waves, plus half a tile for the end, where the last waves finish their last tiles at
// staggered times.
inline long long queue_model_cost(const std::vector<PlanInst>& insts, int rows, int waves,
                                  int wpb = 4, int overhead = 5)
{
    waves = std::max(1, waves);
    const int queues = launch_queues(std::max(1, waves / std::max(1, wpb)));
    const QueuePlan p = plan_queues(insts, rows, queues, overhead);
    const long long per_xcd = std::max(1, waves / queues);
    return (p.max_queue_cost + per_xcd - 1) / per_xcd + tile_cost(std::max(1, rows), overhead) / 2;
}

// The two hand-outs as the load model and the timed choice see them.
enum class HandOut { kStatic, kQueued };

inline long long handout_cost(const std::vector<PlanInst>& insts, int rows, int waves,
                              HandOut how, int overhead = 5)
{
    return how == HandOut::kQueued ? queue_model_cost(insts, rows, waves, 4, overhead)
                                   : fixed_max_cost(insts, rows, waves);
}

// model_rows and shortlist_rows for either hand-out (the static one: the busiest wave of the
// round-robin plan; the queued one: queue_model_cost).
inline int model_rows_for(const std::vector<PlanInst>& insts, int waves,
                          const std::vector<int>& candidates, HandOut how, int overhead = 5)
{
    int best = candidates.empty() ? 32 : candidates.front();
    long long best_cost = std::numeric_limits<long long>::max();
    for (const int rows : candidates) {
        const long long c = handout_cost(insts, rows, waves, how, overhead);
        if (c < best_cost || (c == best_cost && rows > best)) {
            best_cost = c;
            best = rows;
        }
    }
    return best;
}

// The untimed choice (the default partition, re-made at every change of the stepping set):
// the queued model with a tile's fixed cost at 14 half-rows and lengths from 12 rows. Measured
// on FP32 sets of 1-8 M cells (one to thirty instances of 64^2 .. 1024^2), tiles of 8 rows
// were the slowest or next to it every time, and at 5 half-rows the model picked 8-12 rows
// where 16-24 ran 5-7 % faster (the served request stream: 17,930 Mcell-updates/s at its
// picks against 19,000-19,100 at a fixed 16 or 24). At 14 it picks what ran fastest on each
// measured set (12 rows for 1-2 M cells, 16 for three 1024^2 grids, 24 for 8 M cells of
// random sizes). The timed choice's shortlist keeps 5, which ranks the mixed ensemble's
// lengths as measured (16 first).
constexpr int kPickTileOverhead = 14;
constexpr int kPickMinRows = 12;

inline std::vector<int> shortlist_rows_for(const std::vector<PlanLaunch>& launches,
                                           const std::vector<int>& candidates, HandOut how,
                                           int count = 3, int pct = 12)
{
    std::vector<std::pair<long long, int>> cost;
    for (const int rows : candidates) {
        long long c = 0;
        for (const PlanLaunch& l : launches) {
            if (!l.insts->empty()) c += handout_cost(*l.insts, rows, l.waves, how);
        }
        cost.emplace_back(c, -rows);
    }
    std::sort(cost.begin(), cost.end());
    std::vector<int> out;
    for (size_t i = 0; i < cost.size() && static_cast<int>(out.size()) < count; ++i) {
        if (cost[i].first * 100 <= cost.front().first * (100 + pct)) out.push_back(-cost[i].second);
    }
    return out;
}

// ---- replay accounting ---------------------------------------------------------------------
// Launches of a replay of `steps` steps that stepped some instance: the largest step advance
// of the instances stepping in it (the later launches found every instance finished and
// returned at once). before/after: step counts at the start and end of the replay;
// stepping: which entries took part.
inline int replay_active_launches(const std::vector<int>& before, const std::vector<int>& after,
                                  const std::vector<char>& stepping, int steps)
{
    int most = 0;
    for (size_t k = 0; k < before.size() && k < after.size() && k < stepping.size(); ++k) {
        if (stepping[k]) {
            most = std::max(most, after[k] - before[k]);
        }
    }
    return std::min(std::max(most, 0), steps);
}

// True when a replay advanced (or finished) some instance that was stepping in it. A replay
// that leaves every stepping instance where it was can only repeat itself: the host must stop
// with an error instead of replaying forever.
inline bool replay_progressed(const std::vector<int>& before, const std::vector<int>& after,
                              const std::vector<char>& stepping,
                              const std::vector<char>& finished_after)
{
    bool any = false;
    for (size_t k = 0; k < stepping.size(); ++k) {
        if (!stepping[k]) continue;
        any = true;
        if (after[k] > before[k] || finished_after[k]) return true;
    }
    return !any;  // nothing was stepping: nothing to advance
}

}  // namespace mhdhip
