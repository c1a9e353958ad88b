// CPU model of the batched engine's dynamic tile hand-out (mhd_gpu_solver_amd/csrc/
// mhd_batch_plan.hpp, plan_queues; the ticket loop of k_batch_stage in mhd_batch_kernels.hpp).
//
// Workgroups of a launch run in a random interleaving. The waves of a workgroup first take
// the tiles at their own indices among their XCD's waves, then, round by round, one ticket of
// four tiles from their XCD's counter (an atomic add of 4, offset by the XCD's wave count;
// wave w marches the ticket's tile w), until a ticket starts past the queue's end. At its
// start, each stage's kernel zeroes the other stage's counters. Checked over long random
// sequences of launches whose tile lists and grids change between launches: every launch
// marches every tile exactly once, and every workgroup's loop ends.
//
// Also modelled: the protocol that hung in an earlier build (counters zeroed by the last wave
// of each launch, counted on the way out, all tiles from tickets), with a wave that returns
// before it is counted. The counter then stays past the queue's end, the next launch hands out
// no tile, no instance advances, and a host loop without a progress check replays forever.
// Built and run by tests/test_batch_plan.py; prints "ok" and exits 0 when every check holds.
#include "mhd_batch_plan.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

using namespace mhdhip;

static int failures = 0;
#define CHECK(cond, ...)                                     \
    do {                                                     \
        if (!(cond)) {                                       \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                        \
            std::printf("\n");                               \
            ++failures;                                      \
        }                                                    \
    } while (0)

struct Device {
    int next[2][kPlanXcds] = {};  // the counters of one kernel variant, stage 1 and stage 2
                                  // (queue q uses entry q; a launch has at most kPlanXcds)
};

// One launch of stage `stage` (1 or 2) over `plan` with `blocks` workgroups of 4 waves.
// Returns how many times each tile was marched; `steps` counts the tiles marched.
static std::vector<int> launch(Device& d, const QueuePlan& plan, int stage, int blocks,
                               std::mt19937& rng, long long* steps)
{
    const int wpb = 4;
    const int nq = launch_queues(blocks);
    CHECK(static_cast<int>(plan.begin.size()) == nq + 1, "plan for %d queues", nq);
    struct Block {
        int x, round;
        bool done;
    };
    std::vector<Block> bl;
    for (int b = 0; b < blocks; ++b) bl.push_back(Block{b % nq, 0, false});
    std::vector<int> hits(plan.tiles.size(), 0);
    const int other = stage == 1 ? 1 : 0;  // index of the other stage's counters
    const int mine = stage - 1;
    // block 0's first wave zeroes the other stage's counters at some point of the launch
    const long long reset_at = std::uniform_int_distribution<long long>(0, 50)(rng);
    bool reset_done = false;
    long long iter = 0;
    size_t live = bl.size();
    auto march = [&](const Block& k, int u) {
        const int qb = plan.begin[static_cast<size_t>(k.x)];
        const int qn = plan.begin[static_cast<size_t>(k.x) + 1] - qb;
        if (u >= qn) return;  // this wave has no tile in the round
        const int t = qb + (stage == 2 ? qn - 1 - u : u);
        ++hits[static_cast<size_t>(t)];
        ++*steps;
    };
    while (live > 0) {
        if (!reset_done && iter >= reset_at) {
            for (int x = 0; x < kPlanXcds; ++x) d.next[other][x] = 0;
            reset_done = true;
        }
        const size_t bi = std::uniform_int_distribution<size_t>(0, bl.size() - 1)(rng);
        Block& k = bl[bi];
        ++iter;
        if (k.done) continue;
        const int qn = plan.begin[static_cast<size_t>(k.x) + 1] - plan.begin[static_cast<size_t>(k.x)];
        if (k.round == 0) {
            // the waves' own tiles
            for (int w = 0; w < wpb; ++w) {
                march(k, queue_wave_index(blocks, static_cast<int>(bi), w, wpb));
            }
        } else {
            // one ticket of wpb tiles for the workgroup, drawn after its waves met
            const int base = queue_waves(blocks, k.x, wpb) + d.next[mine][k.x];
            d.next[mine][k.x] += wpb;
            if (base >= qn) {
                k.done = true;
                --live;
                continue;
            }
            for (int w = 0; w < wpb; ++w) march(k, base + w);
        }
        ++k.round;
    }
    if (!reset_done) {
        for (int x = 0; x < kPlanXcds; ++x) d.next[other][x] = 0;
    }
    return hits;
}

static std::vector<PlanInst> random_set(std::mt19937& rng)
{
    std::vector<PlanInst> insts;
    const int n = std::uniform_int_distribution<int>(1, 40)(rng);
    for (int i = 0; i < n; ++i) {
        const int nx = std::uniform_int_distribution<int>(1, 300)(rng);
        const int ny = std::uniform_int_distribution<int>(1, 300)(rng);
        insts.push_back(PlanInst{i, nx, ny});
    }
    return insts;
}

// every (row, strip) of every instance in exactly one tile of the queued list, queues in order
static void check_queue_plan(const std::vector<PlanInst>& insts, const QueuePlan& p, int rows,
                             int nq)
{
    CHECK(p.begin.size() == static_cast<size_t>(nq) + 1, "queue offsets size");
    CHECK(p.begin.front() == 0 && p.begin.back() == static_cast<int>(p.tiles.size()),
          "queue offsets ends");
    for (int x = 0; x < nq; ++x) CHECK(p.begin[x] <= p.begin[x + 1], "queue order");
    for (const PlanInst& in : insts) {
        const StripShape s = strip_shape(in.nx, in.ny);
        const int nstrips = s.wide + (s.half ? 1 : 0);
        std::vector<int> hits(static_cast<size_t>(nstrips) * in.ny, 0);
        int leads = 0;
        for (const BatchTile& t : p.tiles) {
            if (t.inst != in.slot) continue;
            CHECK(t.r1 - t.r0 <= rows, "tile longer than %d rows", rows);
            if (t.strip == 0 && t.r0 == kPlanGhosts) ++leads;
            const bool half = s.half && t.strip == s.wide;
            for (int r = t.r0; r < t.r1; ++r) {
                if (half && r - kPlanGhosts >= s.half_rows) continue;
                ++hits[static_cast<size_t>(t.strip) * in.ny + (r - kPlanGhosts)];
            }
            if (half) {
                for (int r = t.r0b; r < t.r1b; ++r) {
                    ++hits[static_cast<size_t>(t.strip) * in.ny + (r - kPlanGhosts)];
                }
            }
        }
        CHECK(leads == 1, "%d lead tiles for slot %d", leads, in.slot);
        for (int h : hits) CHECK(h == 1, "a (row, strip) of slot %d marched %d times", in.slot, h);
    }
    // the queues carry about equal cost: none above the mean by more than its largest tile
    long long big = 0;
    for (const BatchTile& t : p.tiles) big = std::max(big, tile_cost(t.r1 - t.r0));
    CHECK(p.max_queue_cost <= p.total_cost / nq + big,
          "queue balance: max %lld total %lld", p.max_queue_cost, p.total_cost);
}

// The earlier protocol: tickets from 0, the counter zeroed by the wave that finds itself the
// last one counted (done == waves) on its way out. `uncounted` waves return without being
// counted. Waves run in a random interleaving. Returns how often each tile was marched.
static std::vector<int> old_protocol_launch(int& next, int& done, int qn, int waves,
                                            int uncounted, std::mt19937& rng)
{
    std::vector<int> hits(static_cast<size_t>(qn), 0);
    // 0: takes tickets; 1: found its ticket past the end, counts itself next; 2: gone
    std::vector<char> state(static_cast<size_t>(waves), 0);
    int left = waves;
    // every wave issues its first ticket as the launch starts, before any wave is done
    std::vector<int> order(static_cast<size_t>(waves));
    for (int w = 0; w < waves; ++w) order[static_cast<size_t>(w)] = w;
    std::shuffle(order.begin(), order.end(), rng);
    for (const int w : order) {
        const int t = next++;
        if (t < qn) {
            ++hits[static_cast<size_t>(t)];
        } else {
            state[static_cast<size_t>(w)] = 1;
        }
    }
    while (left > 0) {
        const int w = std::uniform_int_distribution<int>(0, waves - 1)(rng);
        char& st = state[static_cast<size_t>(w)];
        if (st == 2) continue;
        if (st == 0) {
            const int t = next++;
            if (t < qn) {
                ++hits[static_cast<size_t>(t)];
            } else {
                st = 1;
            }
            continue;
        }
        st = 2;
        --left;
        if (w < uncounted) continue;
        if (++done == waves) {
            next = 0;
            done = 0;
        }
    }
    return hits;
}

int main()
{
    std::mt19937 rng(12345);
    // plans
    for (int trial = 0; trial < 200; ++trial) {
        const std::vector<PlanInst> insts = random_set(rng);
        const int rows = std::uniform_int_distribution<int>(1, 80)(rng);
        const int nq = std::uniform_int_distribution<int>(1, kPlanXcds)(rng);
        check_queue_plan(insts, plan_queues(insts, rows, nq), rows, nq);
    }
    // long alternating sequences of launches with changing lists and grids
    Device d;
    for (int trial = 0; trial < 300; ++trial) {
        const std::vector<PlanInst> insts = random_set(rng);
        const int rows = std::uniform_int_distribution<int>(1, 64)(rng);
        for (int stage = 1; stage <= 2; ++stage) {
            const int blocks = std::uniform_int_distribution<int>(1, 40)(rng);
            const QueuePlan plan = plan_queues(insts, rows, launch_queues(blocks));
            long long steps = 0;
            const std::vector<int> hits = launch(d, plan, stage, blocks, rng, &steps);
            for (size_t t = 0; t < hits.size(); ++t) {
                CHECK(hits[t] == 1, "trial %d stage %d: tile %zu marched %d times", trial,
                      stage, t, hits[t]);
            }
            CHECK(steps == static_cast<long long>(plan.tiles.size()), "tiles marched");
        }
    }
    // a stage-1 launch with no stage 2 after it (an interrupted launch sequence) leaves the
    // stage-1 counters past their queues: the next stage 1 would march only the first tiles
    // of each queue. The engine zeroes the counters from the host before it launches again.
    {
        Device e;
        const QueuePlan plan =
            plan_queues({PlanInst{0, 256, 256}, PlanInst{1, 64, 64}}, 8, launch_queues(2));
        long long steps = 0;
        (void)launch(e, plan, 1, 2, rng, &steps);
        steps = 0;
        std::vector<int> hits = launch(e, plan, 1, 2, rng, &steps);
        int missed = 0;
        for (int h : hits) missed += h == 0;
        CHECK(missed > 0, "a repeated stage 1 without the reset marched every tile");
        e = Device{};  // the host's zeroing
        steps = 0;
        hits = launch(e, plan, 1, 2, rng, &steps);
        for (int h : hits) CHECK(h == 1, "after the host reset: a tile marched %d times", h);
    }
    // the earlier protocol: with every wave counted it marches every tile once per launch;
    // after one launch with an uncounted wave the done count no longer meets the launches, and
    // later launches march tiles twice or not at all (a launch that marches nothing advances
    // no instance: without a progress check the host replays forever)
    {
        int next = 0, done = 0;
        for (int k = 0; k < 20; ++k) {
            const std::vector<int> h = old_protocol_launch(next, done, 100, 64, 0, rng);
            for (int x : h) CHECK(x == 1, "earlier protocol, all counted: a tile marched %d", x);
        }
        (void)old_protocol_launch(next, done, 100, 64, 1, rng);
        int wrong = 0, empty = 0;
        for (int k = 0; k < 20; ++k) {
            const std::vector<int> h = old_protocol_launch(next, done, 100, 64, 0, rng);
            int marched = 0;
            bool bad = false;
            for (int x : h) {
                marched += x;
                bad = bad || x != 1;
            }
            wrong += bad;
            empty += marched == 0;
        }
        CHECK(wrong > 0 && empty > 0,
              "earlier protocol after an uncounted wave: %d of 20 launches wrong, %d empty",
              wrong, empty);
        std::printf("earlier protocol after one uncounted wave: %d of 20 launches wrong, %d "
                    "marched nothing\n", wrong, empty);
    }
    if (failures == 0) std::printf("ok\n");
    return failures == 0 ? 0 : 1;
}
