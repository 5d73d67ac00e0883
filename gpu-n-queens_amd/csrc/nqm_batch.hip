// nqm_batch.hip -- nqm_solve_batch: many solves in one persistent pass (K2).
//
// Host plan:
//   1. each solve's prefix becomes one item at row k (an empty board of
//      n >= 2 becomes ceil(n/2) row-1 items, the left half weighted x2, as
//      the reference's mirror halving, deffinitions.cuh:26-30);
//   2. k2_expand grows items level by level on the device until each solve
//      reaches its target row (explicit depth) or the queue holds enough
//      items to fill the chip (automatic depth);
//   3. k2_drain empties the queue. With a pass budget, a wave that runs out
//      parks its live stack as a continuation and hands back the items it
//      still holds (right after the queue, in the same buffer), and the
//      drain is relaunched over them (continuations first) until nothing is
//      left (restart);
//   4. per-solve u64 counters come back in one copy.
// With NQM_TUNE_GRAPH, steps 2-4 (and the learned stage, with a model) are
// captured once into a HIP graph (the sizes learned by a first, uncaptured
// run) and launched as one graph; counters are cleared by fill kernels, never
// memset nodes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nqm_internal.h"
#include "nqm_items.h"

namespace nqm {

namespace {

struct HostInit {
    std::vector<Item> items;
    std::vector<uint64_t> sols;   // solutions found while building items (full boards)
    std::vector<uint64_t> nodes;  // weighted placements made on the host
    std::vector<uint8_t> row;     // row of each solve's initial items
    uint64_t exec_nodes = 0;      // unweighted placements made on the host
};

// Relative bitsets of a prefix at row k; false if the prefix is invalid.
bool prefix_state(uint32_t n, uint32_t k, const uint8_t* cols_in, uint32_t* cols, uint32_t* ld,
                  uint32_t* rd) {
    const uint32_t mask = (1u << n) - 1u;
    uint32_t c = 0, l = 0, r = 0;
    for (uint32_t i = 0; i < k; ++i) {
        if (cols_in[i] >= n) return false;
        const uint32_t b = 1u << cols_in[i];
        if (b & (c | l | r)) return false;
        c |= b;
        l = ((l | b) << 1) & mask;
        r = (r | b) >> 1;
    }
    *cols = c;
    *ld = l;
    *rd = r;
    return true;
}

int build_host_items(nqm_ctx* ctx, const nqm_solve* solves, uint32_t ns, bool mirror, HostInit* h) {
    h->sols.assign(ns, 0);
    h->nodes.assign(ns, 0);
    h->row.assign(ns, 0);
    for (uint32_t s = 0; s < ns; ++s) {
        const nqm_solve& q = solves[s];
        if (q.n < 1 || q.n > kMaxN || q.k > q.n || q.k > 32)
            return fail(ctx, NQM_EINVAL, "solve %u: n=%u k=%u out of range", s, q.n, q.k);
        const uint32_t mask = (1u << q.n) - 1u;
        uint32_t cols, ld, rd;
        if (!prefix_state(q.n, q.k, q.prefix, &cols, &ld, &rd))
            return fail(ctx, NQM_EPREFIX, "solve %u: prefix is not a valid placement", s);
        h->row[s] = (uint8_t)q.k;
        if (q.k == q.n) {  // a complete board
            h->sols[s] = 1;
            continue;
        }
        if (q.k == 0 && mirror && q.n >= 2) {
            // row-0 columns < ceil(n/2); x2 weight left of n/2 (integer div)
            h->row[s] = 1;
            for (uint32_t c = 0; c < (q.n + 1) / 2; ++c) {
                const uint32_t ws = c < q.n / 2 ? 1u : 0u;
                const uint32_t b = 1u << c;
                Item it{};
                it.cols = b;
                it.ld = (b << 1) & mask;
                it.rd = b >> 1;
                it.f = ~(it.cols | it.ld | it.rd) & mask;
                it.solve = s;
                it.meta = item_meta(q.n, 1, ws);
                h->nodes[s] += 1ull << ws;
                h->exec_nodes += 1;
                if (it.f) h->items.push_back(it);
            }
            continue;
        }
        Item it{};
        it.cols = cols;
        it.ld = ld;
        it.rd = rd;
        it.f = ~(cols | ld | rd) & mask;
        it.solve = s;
        it.meta = item_meta(q.n, q.k, 0);
        if (it.f) h->items.push_back(it);
    }
    return NQM_OK;
}

// Expand every item below its solve's target row by one row, level by level,
// while the items stay under max_items (k2_expand on the host).
// Items the host expands to before the device takes over: a single solve
// stops early (each device level costs a few launches), a mixed batch goes
// further (its first levels are many small solves).
constexpr size_t kHostItemsOne = 1u << 10, kHostItemsMany = 1u << 14;
void host_expand(const std::vector<uint8_t>& target, size_t max_items, HostInit* h) {
    const uint32_t ns = (uint32_t)h->row.size();
    for (;;) {
        bool any = false;
        for (uint32_t s = 0; s < ns; ++s) any |= h->row[s] < target[s];
        if (!any || h->items.size() >= max_items) return;
        std::vector<Item> next;
        next.reserve(h->items.size() * 8);
        for (const Item& it : h->items) {
            const uint32_t n = meta_n(it.meta), row = meta_row(it.meta);
            if (row >= target[it.solve]) {
                next.push_back(it);
                continue;
            }
            const uint32_t mask = (1u << n) - 1u, ws = meta_wshift(it.meta);
            const uint32_t placed = (uint32_t)__builtin_popcount(it.f);
            h->nodes[it.solve] += (uint64_t)placed << ws;
            h->exec_nodes += placed;
            for (uint32_t f = it.f; f;) {
                const uint32_t b = f & (0u - f);
                f ^= b;
                Item c{};
                c.cols = it.cols | b;
                c.ld = ((it.ld | b) << 1) & mask;
                c.rd = (it.rd | b) >> 1;
                c.f = ~(c.cols | c.ld | c.rd) & mask;
                c.solve = it.solve;
                c.meta = item_meta(n, row + 1, ws);
                if (c.f) next.push_back(c);
            }
        }
        for (uint32_t s = 0; s < ns; ++s)
            if (h->row[s] < target[s]) h->row[s]++;
        h->items.swap(next);
    }
}

// K2 drain grid for a stack depth: as many waves per CU as LDS and the
// requested cap allow.
int drain_shape(nqm_ctx* ctx, uint32_t stack_rows, const nqm_tune& t, uint32_t* blocks,
                uint32_t* threads) {
    int rc;
    uint32_t th = t.threads;
    if (th == 0 && t.waves_per_cu == 0) {
        if ((rc = k2_auto_threads(stack_rows, ctx->lds_per_block, &th)) != NQM_OK)
            return fail(ctx, rc, "occupancy query failed");
    } else if (th == 0) {
        th = 64;  // a wave budget per CU: one-wave workgroups meet it exactly
    }
    if (th != 64 && th != 128 && th != 256) return fail(ctx, NQM_EINVAL, "threads=%u", th);
    while (th > 64 && k2_lds_bytes(stack_rows, th) > ctx->lds_per_block) th >>= 1;
    if (k2_lds_bytes(stack_rows, th) > ctx->lds_per_block)
        return fail(ctx, NQM_EUNSUPPORTED, "stack of %u rows does not fit LDS", stack_rows);
    int per_cu = 0;
    rc = k2_max_blocks_per_cu(th, k2_lds_bytes(stack_rows, th), &per_cu);
    if (rc != NQM_OK) return fail(ctx, rc, "occupancy query failed");
    if (per_cu < 1) per_cu = 1;
    if (t.waves_per_cu) {
        const int cap = (int)std::max(1u, t.waves_per_cu / (th / 64u));
        per_cu = std::min(per_cu, cap);
    }
    *blocks = (uint32_t)per_cu * (uint32_t)ctx->prop.multiProcessorCount;
    *threads = th;
    return NQM_OK;
}

}  // namespace

void drop_plan(GraphPlan& plan) {
    if (plan.exec) (void)hipGraphExecDestroy(plan.exec);
    if (plan.graph) (void)hipGraphDestroy(plan.graph);
    plan = GraphPlan{};
}

namespace {

// Launch the kept graph: clears, expansion levels, drain, commits, all as
// built for these solves; the host adds what it placed when building it.
int replay_plan(nqm_ctx* ctx, uint32_t ns, uint64_t* counts, uint64_t* nodes_out, nqm_stats* stats,
                double t0) {
    BatchWork& b = ctx->batch;
    hipStream_t st = ctx->stream;
    NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
    NQM_HIP(ctx, hipGraphLaunch(b.plan.exec, st));
    NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
    std::vector<uint64_t> res(2 + 2 * (size_t)ns);
    NQM_HIP(ctx, hipMemcpyAsync(res.data(), b.res.ptr, res.size() * 8, hipMemcpyDeviceToHost, st));
    NQM_HIP(ctx, hipStreamSynchronize(st));
    float run_ms = 0.f;
    NQM_HIP(ctx, hipEventElapsedTime(&run_ms, ctx->ev_start, ctx->ev_stop));
    const uint32_t err = (uint32_t)res[1];
    if (err) {
        drop_plan(b.plan);
        return fail(ctx, err & 16u ? NQM_EDEVICE : NQM_ERANGE, "graph replay failed (err=0x%x)", err);
    }
    const uint64_t* const hs = res.data() + 2;
    const uint64_t* const hn = hs + ns;
    uint64_t total = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        counts[s] = hs[s] + b.plan.host_sols[s];
        if (nodes_out) nodes_out[s] = hn[s] + b.plan.host_nodes[s];
        total += counts[s];
    }
    if (stats) {
        stats->total = total;
        stats->nodes = res[0] + b.plan.host_exec;
        stats->kernel_ms = (double)run_ms;
        stats->wall_ms = now_ms() - t0;
        stats->launches = 1;
        stats->blocks = b.plan.blocks;
        stats->search_ms = (double)run_ms;  // the graph as a whole
        stats->rounds = 1;
        stats->lanes = b.plan.lanes;
    }
    return NQM_OK;
}

}  // namespace

int solve_batch_impl(nqm_ctx* ctx, const nqm_solve* solves, uint32_t ns, const nqm_tune* tune_in,
                     uint64_t* counts, uint64_t* nodes_out, nqm_stats* stats) {
    const double t0 = now_ms();
    nqm_tune t{};
    if (tune_in) t = *tune_in;
    if (t.flags & ~(NQM_TUNE_NO_MIRROR | NQM_TUNE_NO_DONATE | NQM_TUNE_GRAPH | NQM_TUNE_NO_ORDER))
        return fail(ctx, NQM_EINVAL, "unknown tune flags 0x%x", t.flags);
    const uint32_t refill = t.refill_idle ? std::min(t.refill_idle, 64u) : 12u;
    const bool mirror = !(t.flags & NQM_TUNE_NO_MIRROR);
    const bool use_graph = (t.flags & NQM_TUNE_GRAPH) && !t.budget;

    // A graph-mode call that repeats the previous call's solves, tuning and
    // model weights launches the graph built then; any other call drops it.
    BatchWork& b = ctx->batch;
    const uint64_t model_version = t.model ? mlp_version(t.model) : 0u;
    std::vector<uint8_t> key(sizeof(nqm_tune) + sizeof(model_version) + (size_t)ns * sizeof(nqm_solve));
    std::memcpy(key.data(), &t, sizeof(nqm_tune));
    std::memcpy(key.data() + sizeof(nqm_tune), &model_version, sizeof(model_version));
    std::memcpy(key.data() + sizeof(nqm_tune) + sizeof(model_version), solves, (size_t)ns * sizeof(nqm_solve));
    if (use_graph && b.plan.valid && b.plan.key == key) return replay_plan(ctx, ns, counts, nodes_out, stats, t0);
    drop_plan(b.plan);

    HostInit h;
    int rc = build_host_items(ctx, solves, ns, mirror, &h);
    if (rc != NQM_OK) return rc;

    uint32_t max_n = 1;
    for (uint32_t s = 0; s < ns; ++s) max_n = std::max(max_n, solves[s].n);

    // Target rows. Explicit depth: min(n-1, max(row, depth)). Automatic:
    // expand toward n-4 but stop once the queue can fill the chip. A
    // stack-row cap raises the target so that n-1-row <= stack_rows.
    std::vector<uint8_t> target(ns);
    const bool auto_depth = t.depth == 0;
    for (uint32_t s = 0; s < ns; ++s) {
        const uint32_t n = solves[s].n, r = h.row[s];
        uint32_t tr = auto_depth ? (n > 4 ? std::max(r, n - 4) : r)
                                 : std::max(r, std::min(t.depth, n - 1));
        if (t.stack_rows && n - 1 > t.stack_rows + tr) tr = n - 1 - t.stack_rows;
        target[s] = (uint8_t)std::min(tr, n - 1);
    }
    // With a model, the plain expansion stops 16x earlier: the learned stage
    // splits only the items it predicts large.
    const uint64_t chip_lanes = (uint64_t)ctx->prop.multiProcessorCount * 16u * 64u;
    const uint64_t want_items = !auto_depth ? UINT64_MAX : t.model ? chip_lanes / 2u : chip_lanes * 8u;

    // The first levels are tiny: place them on the host (same rule and node
    // accounting as k2_expand) and save a device round trip per level.
    host_expand(target, ns == 1 ? kHostItemsOne : kHostItemsMany, &h);

    const uint32_t count0 = (uint32_t)h.items.size();
    if ((rc = ensure_dev(ctx, b.init, std::max<size_t>(count0, 1) * sizeof(Item))) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, b.head, kHeadSegs * kHeadStride * 4)) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, b.spill_count, 64)) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, b.counts, 32 * 4)) != NQM_OK) return rc;  // 32 count slots
    if ((rc = ensure_dev(ctx, b.target, std::max<size_t>(ns, 1))) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, b.res, 16 + (size_t)ns * 16)) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, b.part, kExpandPartBytes)) != NQM_OK) return rc;

    uint32_t* const d_cnt = static_cast<uint32_t*>(b.counts.ptr);  // 32 slots
    uint32_t* const d_head = static_cast<uint32_t*>(b.head.ptr);
    uint32_t* const d_spill_count = static_cast<uint32_t*>(b.spill_count.ptr);
    uint64_t* const d_exec = static_cast<uint64_t*>(b.res.ptr);
    uint32_t* const d_err = reinterpret_cast<uint32_t*>(d_exec + 1);
    uint64_t* const d_sols = d_exec + 2;
    uint64_t* const d_nodes = d_sols + ns;
    const uint8_t* const d_target = static_cast<uint8_t*>(b.target.ptr);
    uint32_t* const d_qcount = d_cnt + 12;  // the drain reads its queue size here
    void* const d_part = b.part.ptr;        // per-block records of the level being expanded
    hipStream_t st = ctx->stream;

    NQM_HIP(ctx, hipMemcpyAsync(b.target.ptr, target.data(), ns, hipMemcpyHostToDevice, st));
    if (count0)
        NQM_HIP(ctx, hipMemcpyAsync(b.init.ptr, h.items.data(), count0 * sizeof(Item),
                                    hipMemcpyHostToDevice, st));

    // ---- plan: expand level by level, learning each level's size ---------
    // Level L reads slot L (slot 0 = initial count) and writes slot L+1;
    // buffers: init -> q[0] -> q[1] -> q[0] ...
    auto level_in = [&](uint32_t L) -> Item* {
        return L == 0 ? static_cast<Item*>(b.init.ptr) : static_cast<Item*>(b.q[(L - 1) & 1].ptr);
    };
    auto level_out = [&](uint32_t L) -> Item* { return static_cast<Item*>(b.q[L & 1].ptr); };
    auto enqueue_clear = [&]() -> hipError_t {
        hipError_t e;
        if ((e = k2_fill_u32(st, d_cnt, 16, 0)) != hipSuccess) return e;
        if ((e = k2_fill_u32(st, d_cnt, 1, count0)) != hipSuccess) return e;
        return k2_fill_u64(st, d_exec, 2 + 2 * ns);  // exec, err, sols, nodes
    };
    NQM_HIP(ctx, enqueue_clear());

    // Level L reads level_in(L) (count in slot L) and writes level_out(L)
    // (count in slot L+1). Automatic depth keeps expanding while the queue is
    // too short to fill the chip, or while the deepest solve would need more
    // than kRowsGoal LDS stack rows (fewer rows, more resident waves), as long
    // as a level stays under kAutoMaxItems. With a model the expansion stops
    // at want_items: the learned stage splits only what it predicts large.
    //
    // Every level is counted on the device before it is written. For a single
    // solve the host waits for that count only when it decides something: a
    // level that the rows goal (or an explicit depth) asks for regardless of
    // its size, and whose upper bound fits kAutoMaxItems, is sized by the
    // bound (an item at row r has at most n - r children) and enqueued
    // without a round trip. (Mixed batches measured slower that way: their
    // loose bounds inflate the next level's grid.)
    constexpr uint32_t kRowsGoal = 9;
    constexpr uint64_t kAutoMaxItems = 1ull << 27;
    std::vector<uint32_t> caps, ins;
    std::vector<uint8_t> final_row(h.row);
    uint32_t count = count0;  // size of the current level (an upper bound unless exact)
    bool exact = true;
    float expand_ms = 0.f;
    NQM_HIP(ctx, hipEventRecord(ctx->ev_x0, st));
    for (uint32_t L = 0; L < 11; ++L) {
        bool any = false;
        uint32_t rows_left = 0, grow = 1;
        for (uint32_t s = 0; s < ns; ++s)
            if (final_row[s] < target[s]) {
                any = true;
                rows_left = std::max(rows_left, solves[s].n - 1u - final_row[s]);
                grow = std::max(grow, solves[s].n - final_row[s]);
            }
        if (!any || count == 0) break;
        const uint64_t bound = (uint64_t)count * grow;
        // with a model, a level whose input (even as an upper bound) is under
        // want_items is needed whatever its size
        const bool blind = ns == 1 && bound <= kAutoMaxItems &&
                           (!auto_depth || (!t.model && rows_left > kRowsGoal) ||
                            (t.model && (uint64_t)count < want_items));
        // the decision below needs the level's true size only where the size
        // decides it (with a model, or within the rows goal)
        if (!blind && !exact && auto_depth && (t.model || rows_left <= kRowsGoal)) {
            NQM_HIP(ctx, hipMemcpyAsync(&count, d_cnt + L, 4, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
            exact = true;
            if (count == 0) break;
        }
        if (auto_depth && !blind && (uint64_t)count >= want_items && (t.model || rows_left <= kRowsGoal)) break;
        const uint32_t grid = std::min<uint32_t>(kExpandMaxGrid, (count + 255u) / 256u);
        NQM_HIP(ctx, k2_expand_count(st, grid, level_in(L), d_cnt + L, d_part, d_cnt + L + 1, d_target));
        uint32_t next = (uint32_t)bound;  // blind: capacity
        if (!blind) {
            NQM_HIP(ctx, hipMemcpyAsync(&next, d_cnt + L + 1, 4, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
            if (auto_depth && next > kAutoMaxItems && (uint64_t)count >= want_items) break;
            if (next > 0x7fffffffu) return fail(ctx, NQM_ERANGE, "expansion exceeds 2^31 items");
        }
        if ((rc = ensure_dev(ctx, b.q[L & 1], (size_t)std::max(next, 1u) * sizeof(Item))) != NQM_OK) return rc;
        NQM_HIP(ctx, k2_expand_write(st, grid, level_in(L), d_cnt + L, d_part, !t.model, level_out(L), next,
                                     d_target, d_nodes, d_exec, d_err));
        caps.push_back(next);
        ins.push_back(count);
        for (uint32_t s = 0; s < ns; ++s)
            if (final_row[s] < target[s]) final_row[s]++;
        count = next;
        exact = !blind;
    }
    NQM_HIP(ctx, hipEventRecord(ctx->ev_x1, st));
    // A single-round drain without a model reads its queue size on the
    // device: nothing below needs the host to wait for the expansion.
    const bool host_count = !exact && (t.model || t.budget || use_graph);
    if (host_count) {
        NQM_HIP(ctx, hipMemcpyAsync(&count, d_cnt + caps.size(), 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        exact = true;
    }
    const uint32_t levels = (uint32_t)caps.size();
    Item* final_q = level_in(levels);
    int spare = levels == 0 ? 0 : (int)(levels & 1);  // q[] not holding final_q
    uint32_t* final_slot = d_cnt + levels;

    // ---- learned splitting: split predicted-large items, order largest first
    //
    // The stage runs live, the host reading the sizes it decides on, and in
    // graph mode once more inside the capture with the sizes of the live run.
    // Features, predictions and the split threshold are formed in a fixed
    // order, so every size is the same on every run over the same items and
    // weights (the plan key holds the weights' version).
    float learn_ms = 0.f;
    uint32_t split_levels = 0;
    std::vector<uint32_t> split_cap, split_out, goal_cap;  // sizes of the live run
    std::vector<uint8_t> goal_rows(ns);                    // rows-goal target rows
    const uint32_t learn_count = count;
    const int learn_spare = spare;
    auto learned_stage = [&](bool live, Item*& q, uint32_t*& slot, uint32_t& cnt, int& sp) -> int {
        int r;
        auto need = [&](DevBuf& d, size_t bytes) -> int {
            if (live) return ensure_dev(ctx, d, bytes);
            return d.cap >= bytes ? NQM_OK : fail(ctx, NQM_EDEVICE, "graph capture: buffer smaller than planned");
        };
        if ((r = need(b.hist, 64 * 4)) != NQM_OK) return r;
        // stat: double sum, u32 over, float threshold; per-block partial sums from byte 64
        if ((r = need(b.stat, 64 + (size_t)kPredSumBlocks * 8)) != NQM_OK) return r;
        double* const d_psum = static_cast<double*>(b.stat.ptr);
        uint32_t* const d_over = reinterpret_cast<uint32_t*>(d_psum + 1);
        float* const d_thresh = reinterpret_cast<float*>(d_over + 1);
        double* const d_bsums = d_psum + 8;
        uint32_t* const d_slot = d_cnt + 16;  // slots 16..31 for this stage
        auto predict = [&]() -> int {
            int r2;
            if ((r2 = need(b.feats, (size_t)cnt * 32)) != NQM_OK) return r2;
            if ((r2 = need(b.pred, (size_t)cnt * 4)) != NQM_OK) return r2;
            NQM_HIP(ctx, mlp_features_items(st, q, slot, b.feats.ptr, cnt));
            return mlp_predict_device(t.model, b.feats.ptr, cnt, static_cast<float*>(b.pred.ptr), slot);
        };
        const uint32_t nsplit = live ? (t.split_rounds ? t.split_rounds : 1u) : (uint32_t)split_cap.size();
        for (uint32_t i = 0;; ++i) {
            if ((r = predict()) != NQM_OK) return r;
            if (i == nsplit) break;
            // Split the heavy tail: items predicted above 8x the level's mean
            // (or the fixed threshold). The largest-first order handles the
            // rest, and the rows-goal expansion that follows divides every
            // item ~10^2-fold anyway; splitting nearly everything (a lane's
            // fair share as the bar) only added a full level of copies.
            const float* const d_pred = static_cast<const float*>(b.pred.ptr);
            NQM_HIP(ctx, k2_fill_u32(st, reinterpret_cast<uint32_t*>(d_psum), 4, 0));
            NQM_HIP(ctx, k2_split_threshold(st, d_pred, slot, cnt, t.split_log2, 8.0, d_bsums, d_psum, d_over,
                                            d_thresh));
            uint32_t cap = 0;
            if (live) {
                // Sized for every item splitting (n children at most) while
                // that stays under 1 GiB, so the host need not wait for the
                // threshold (with nothing over it the level is a copy); else
                // sized by the items over the threshold.
                uint64_t cap64 = (uint64_t)cnt * max_n;
                if (cap64 * sizeof(Item) > (1ull << 30)) {
                    uint32_t over = 0;
                    NQM_HIP(ctx, hipMemcpyAsync(&over, d_over, 4, hipMemcpyDeviceToHost, st));
                    NQM_HIP(ctx, hipStreamSynchronize(st));
                    if (over == 0) break;
                    cap64 = (uint64_t)cnt + (uint64_t)over * max_n;
                }
                if (cap64 > 0x7fffffffull) return fail(ctx, NQM_ERANGE, "split exceeds 2^31 items");
                cap = (uint32_t)cap64;
            } else {
                cap = split_cap[i];
            }
            if ((r = need(b.q[sp], (size_t)cap * sizeof(Item))) != NQM_OK) return r;
            uint32_t* const out_slot = d_slot + (i & 7);
            const uint32_t grid = std::min<uint32_t>(kExpandMaxGrid, (cnt + 255u) / 256u);
            NQM_HIP(ctx, k2_expand_count(st, grid, q, slot, d_part, out_slot, d_target, d_pred, d_thresh));
            NQM_HIP(ctx, k2_expand_write(st, grid, q, slot, d_part, false, static_cast<Item*>(b.q[sp].ptr), cap,
                                         d_target, d_nodes, d_exec, d_err, d_pred, d_thresh));
            if (live) {
                // The split level's size stays on the device (its kernels and
                // the predictor read it there, cap sizes their launches),
                // unless restart rounds need it on the host.
                if (t.budget) {
                    NQM_HIP(ctx, hipMemcpyAsync(&cnt, out_slot, 4, hipMemcpyDeviceToHost, st));
                    NQM_HIP(ctx, hipStreamSynchronize(st));
                } else {
                    cnt = cap;
                    exact = false;
                }
                split_cap.push_back(cap);
                split_out.push_back(cnt);
                ++split_levels;
            } else {
                cnt = split_out[i];
            }
            q = static_cast<Item*>(b.q[sp].ptr);
            slot = out_slot;
            sp ^= 1;
        }
        if (!(t.flags & NQM_TUNE_NO_ORDER)) {
            if ((r = need(b.q[sp], (size_t)std::max(cnt, 1u) * sizeof(Item))) != NQM_OK) return r;
            NQM_HIP(ctx, k2_order_by_prediction(st, q, static_cast<const float*>(b.pred.ptr), slot, cnt,
                                                t.budget ? 1u : kHeadSegs, static_cast<uint32_t*>(b.hist.ptr),
                                                static_cast<Item*>(b.q[sp].ptr)));
            q = static_cast<Item*>(b.q[sp].ptr);
            sp ^= 1;
        }
        // Then the stack-depth goal of the automatic depth: every item is
        // expanded to row n-1-kRowsGoal (split items may be deeper already).
        // Expansion keeps the queue's order, so the largest-predicted-first
        // order carries over to the children. The target rows go to their
        // own buffer: a graph's first levels still read the original ones.
        if (!auto_depth) return NQM_OK;
        if (live) {
            for (uint32_t s = 0; s < ns; ++s) {
                const uint32_t n = solves[s].n;
                const uint32_t g = n > kRowsGoal + 1u ? n - 1u - kRowsGoal : 0u;
                goal_rows[s] = (uint8_t)std::max<uint32_t>(final_row[s], std::min(g, n - 1u));
            }
            if ((r = ensure_dev(ctx, b.goal, std::max<size_t>(ns, 1))) != NQM_OK) return r;
            NQM_HIP(ctx, hipMemcpyAsync(b.goal.ptr, goal_rows.data(), ns, hipMemcpyHostToDevice, st));
        }
        const uint8_t* const d_goal = static_cast<const uint8_t*>(b.goal.ptr);
        for (uint32_t L = 0;; ++L) {
            if (live) {
                bool below = false;
                for (uint32_t s = 0; s < ns; ++s) below |= final_row[s] < goal_rows[s];
                if (!below || cnt == 0) break;
            } else if (L == goal_cap.size()) {
                break;
            }
            uint32_t* const out_slot = d_cnt + 24 + (L & 3);
            const uint32_t grid = std::min<uint32_t>(kExpandMaxGrid, (cnt + 255u) / 256u);
            NQM_HIP(ctx, k2_expand_count(st, grid, q, slot, d_part, out_slot, d_goal));
            uint32_t next = 0;
            if (live) {
                // As in the plain expansion, a single solve's level that fits
                // kAutoMaxItems at its bound (n - row children per item) is
                // sized by the bound, with no round trip; the drain then reads
                // the queue size on the device. Restart rounds need it on the host.
                uint32_t grow = 1;
                for (uint32_t s = 0; s < ns; ++s)
                    if (final_row[s] < goal_rows[s]) grow = std::max(grow, solves[s].n - final_row[s]);
                const uint64_t bound = (uint64_t)cnt * grow;
                if (ns == 1 && !t.budget && bound <= kAutoMaxItems) {
                    next = (uint32_t)bound;
                    exact = false;
                } else {
                    NQM_HIP(ctx, hipMemcpyAsync(&next, out_slot, 4, hipMemcpyDeviceToHost, st));
                    NQM_HIP(ctx, hipStreamSynchronize(st));
                    if (next > kAutoMaxItems) break;
                    exact = true;
                }
                goal_cap.push_back(next);
            } else {
                next = goal_cap[L];
            }
            if ((r = need(b.q[sp], (size_t)std::max(next, 1u) * sizeof(Item))) != NQM_OK) return r;
            NQM_HIP(ctx, k2_expand_write(st, grid, q, slot, d_part, false, static_cast<Item*>(b.q[sp].ptr), next,
                                         d_goal, d_nodes, d_exec, d_err));
            q = static_cast<Item*>(b.q[sp].ptr);
            slot = out_slot;
            sp ^= 1;
            cnt = next;
            if (live)
                for (uint32_t s = 0; s < ns; ++s)
                    if (final_row[s] < goal_rows[s]) final_row[s]++;
        }
        return NQM_OK;
    };
    if (t.model && count) {
        NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
        if ((rc = learned_stage(true, final_q, final_slot, count, spare)) != NQM_OK) return rc;
        NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        NQM_HIP(ctx, hipEventSynchronize(ctx->ev_stop));
        NQM_HIP(ctx, hipEventElapsedTime(&learn_ms, ctx->ev_start, ctx->ev_stop));
    }

    // ---- drain shape -------------------------------------------------------
    uint32_t stack_rows = 1;
    for (uint32_t s = 0; s < ns; ++s) {
        const uint32_t n = solves[s].n;
        if (n > final_row[s] + 1) stack_rows = std::max(stack_rows, n - 1 - final_row[s]);
    }
    uint32_t blocks = 0, threads = 0;
    if ((rc = drain_shape(ctx, stack_rows, t, &blocks, &threads)) != NQM_OK) return rc;

    K2DrainParams p;
    p.solve_sols = d_sols;
    p.solve_nodes = d_nodes;
    p.exec_nodes = d_exec;
    p.err = d_err;
    p.stack_rows = stack_rows;
    p.refill_idle = refill;
    p.donate = (t.flags & NQM_TUNE_NO_DONATE) ? 0u : 1u;
    p.budget = t.budget;
    p.head = d_head;
    p.nseg = t.budget ? 1u : kHeadSegs;  // restart rounds read one head
    if ((rc = ensure_dev(ctx, b.wave_part, (size_t)blocks * (threads / 64u) * 32u)) != NQM_OK) return rc;
    p.wave_part = static_cast<uint64_t*>(b.wave_part.ptr);
    p.spill_count = d_spill_count;
    p.item_count = d_qcount;

    float run_ms = 0.f;
    // expansion levels (the learned stage's included) and drain rounds; a
    // graph counts one launch
    uint32_t launches = levels + split_levels + (uint32_t)goal_cap.size(), rounds = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t round_parked_max = 0, round_handed_max = 0, round_room = 0;

    if (use_graph) {
        // One graph: clears, every expand level, the learned stage, the
        // drain. Replays the whole batch from the initial items with the
        // sizes learned above.
        p.items = final_q;
        p.spill = nullptr;
        p.spill_cap = 0;
        NQM_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_clear();
        for (uint32_t L = 0; e == hipSuccess && L < levels; ++L) {
            const uint32_t grid = std::min<uint32_t>(kExpandMaxGrid, (ins[L] + 255u) / 256u);
            e = k2_expand_count(st, grid, level_in(L), d_cnt + L, d_part, d_cnt + L + 1, d_target);
            if (e == hipSuccess)
                e = k2_expand_write(st, grid, level_in(L), d_cnt + L, d_part, !t.model, level_out(L), caps[L],
                                    d_target, d_nodes, d_exec, d_err);
        }
        int lrc = NQM_OK;
        uint32_t* queue_slot = d_cnt + levels;
        if (e == hipSuccess && t.model && learn_count) {
            Item* q = level_in(levels);
            uint32_t cnt = learn_count;
            int sp = learn_spare;
            lrc = learned_stage(false, q, queue_slot, cnt, sp);
            if (lrc == NQM_OK && q != final_q)
                lrc = fail(ctx, NQM_EDEVICE, "graph capture: the learned stage left another queue than planned");
        }
        if (e == hipSuccess && lrc == NQM_OK)
            e = hipMemcpyAsync(d_qcount, queue_slot, 4, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && lrc == NQM_OK) e = k2_fill_u32(st, d_head, kHeadSegs * kHeadStride, 0);
        if (e == hipSuccess && lrc == NQM_OK) e = k2_drain(st, blocks, threads, p);
        hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (lrc != NQM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return lrc;
        }
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(ctx, NQM_EDEVICE, "graph capture failed: %s / %s", hipGetErrorString(e),
                        hipGetErrorString(e2));
        }
        hipError_t ei = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(ctx, NQM_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
        NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
        NQM_HIP(ctx, hipGraphLaunch(gexec, st));
        NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        ++launches;  // the planning levels ran live, then the graph
        rounds = 1;
        // keep it: the next identical call launches it again
        b.plan.key = key;
        b.plan.graph = graph;
        b.plan.exec = gexec;
        b.plan.host_sols = h.sols;
        b.plan.host_nodes = h.nodes;
        b.plan.host_exec = h.exec_nodes;
        b.plan.blocks = blocks;
        b.plan.lanes = (uint64_t)blocks * threads;
        graph = nullptr;
        gexec = nullptr;
    } else {
        NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
        // Budget rounds run in place: a round's spills are appended right
        // after its queue, and the next round's queue starts where the head
        // stopped. The live range moves to the front of the other buffer only
        // when the end of its buffer is near.
        Item* queue = final_q;
        uint32_t qn = count;
        // b.q[qb] holds the queue (rounds in place), at offset qstart
        int qb = final_q == b.q[0].ptr ? 0 : final_q == b.q[1].ptr ? 1 : -1;
        uint64_t qstart = 0;
        // a round hands back the items its waves hold (the open rows of
        // unfinished pieces are parked as continuations, one per lane)
        const uint64_t lanes = (uint64_t)blocks * threads;
        const uint64_t spill_room = (uint64_t)blocks * (threads / 64u) * kGrabItems;
        const bool rounds_in_place = t.budget != 0;
        uint32_t* d_cc = nullptr;  // in count, head, out count of parked pieces
        uint64_t parked_max = 0, handed_max = 0;  // per round
        uint32_t conts = 0;        // parked pieces waiting for the next round
        int ci = 0;                // b.cont[ci] holds them
        if (t.budget) {
            if ((rc = ensure_dev(ctx, b.cont_counts, 64)) != NQM_OK) return rc;
            for (DevBuf* d : {&b.cont[0], &b.cont[1]})
                if ((rc = ensure_dev(ctx, *d, (size_t)lanes * sizeof(Cont))) != NQM_OK) return rc;
            d_cc = static_cast<uint32_t*>(b.cont_counts.ptr);
        }
        for (;;) {
            NQM_HIP(ctx, k2_fill_u32(st, d_head, kHeadSegs * kHeadStride, 0));
            if (exact)
                NQM_HIP(ctx, k2_fill_u32(st, d_qcount, 1, qn));
            else  // the expansion's last count, left on the device
                NQM_HIP(ctx, hipMemcpyAsync(d_qcount, final_slot, 4, hipMemcpyDeviceToDevice, st));
            if (rounds_in_place) {
                if (qb < 0 || qstart + qn + spill_room > b.q[qb].cap / sizeof(Item)) {
                    const int to = qb < 0 ? spare : qb ^ 1;
                    const uint64_t want = 2 * ((uint64_t)qn + spill_room);
                    if (want > 0x7fffffffull) return fail(ctx, NQM_ERANGE, "spill queue too large");
                    if ((rc = ensure_dev(ctx, b.q[to], (size_t)want * sizeof(Item))) != NQM_OK) return rc;
                    if (qn)
                        NQM_HIP(ctx, hipMemcpyAsync(b.q[to].ptr, queue, (size_t)qn * sizeof(Item),
                                                    hipMemcpyDeviceToDevice, st));
                    queue = static_cast<Item*>(b.q[to].ptr);
                    qb = to;
                    qstart = 0;
                }
                NQM_HIP(ctx, k2_fill_u32(st, d_spill_count, 1, qn));
                p.spill = queue;
                p.spill_cap = (uint32_t)std::min<uint64_t>(b.q[qb].cap / sizeof(Item) - qstart, 0x7fffffffull);
            }
            if (t.budget) {
                // parked pieces: resumed from cont[ci], parked into cont[ci ^ 1]
                if ((rc = ensure_dev(ctx, b.cont[ci ^ 1], (2 * (size_t)lanes + conts) * sizeof(Cont))) != NQM_OK)
                    return rc;
                NQM_HIP(ctx, k2_fill_u32(st, d_cc, 3, 0));
                NQM_HIP(ctx, k2_fill_u32(st, d_cc, 1, conts));
                p.cont_in = static_cast<const Cont*>(b.cont[ci].ptr);
                p.cont_in_count = d_cc;
                p.cont_head = d_cc + 1;
                p.cont_out = static_cast<Cont*>(b.cont[ci ^ 1].ptr);
                p.cont_out_count = d_cc + 2;
                p.cont_cap = (uint32_t)std::min<uint64_t>(b.cont[ci ^ 1].cap / sizeof(Cont), 0x7fffffffull);
            }
            p.items = queue;
            NQM_HIP(ctx, k2_drain(st, blocks, threads, p));
            ++launches;
            ++rounds;
            if (!rounds_in_place) break;
            uint32_t hd = 0, end = 0, cc[3] = {0, 0, 0};
            NQM_HIP(ctx, hipMemcpyAsync(&hd, d_head, 4, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipMemcpyAsync(&end, d_spill_count, 4, hipMemcpyDeviceToHost, st));
            if (t.budget) NQM_HIP(ctx, hipMemcpyAsync(cc, d_cc, 12, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
            const uint32_t h0 = std::min(hd, qn);  // items [h0, qn) were never taken
            parked_max = std::max<uint64_t>(parked_max, cc[2]);
            handed_max = std::max<uint64_t>(handed_max, end > qn ? end - qn : 0u);
            queue += h0;
            qstart += h0;
            qn = end - h0;
            // parked pieces no wave resumed this round follow the new ones
            const uint32_t c0 = std::min(cc[1], conts);
            if (c0 < conts)
                NQM_HIP(ctx, hipMemcpyAsync(static_cast<Cont*>(b.cont[ci ^ 1].ptr) + cc[2],
                                            static_cast<const Cont*>(b.cont[ci].ptr) + c0,
                                            (size_t)(conts - c0) * sizeof(Cont), hipMemcpyDeviceToDevice, st));
            conts = cc[2] + (conts - c0);
            ci ^= 1;
            if (qn == 0 && conts == 0) break;
            if (rounds > 1000000u) return fail(ctx, NQM_ERANGE, "restart rounds do not converge");
        }
        NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        round_parked_max = parked_max;
        round_handed_max = handed_max;
        round_room = spill_room;
    }

    std::vector<uint64_t> res(2 + 2 * (size_t)ns);
    NQM_HIP(ctx, hipMemcpyAsync(res.data(), b.res.ptr, res.size() * 8, hipMemcpyDeviceToHost, st));
    NQM_HIP(ctx, hipStreamSynchronize(st));
    const uint64_t exec = res[0];
    const uint32_t err = (uint32_t)res[1];
    const uint64_t* const hs = res.data() + 2;
    const uint64_t* const hn = hs + ns;
    NQM_HIP(ctx, hipEventElapsedTime(&run_ms, ctx->ev_start, ctx->ev_stop));
    // a graph is timed as a whole (its drain is not timed apart)
    const float search_ms = run_ms;
    if (!use_graph) NQM_HIP(ctx, hipEventElapsedTime(&expand_ms, ctx->ev_x0, ctx->ev_x1));
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (err & 16u) return fail(ctx, NQM_EDEVICE, "drain LDS does not start at address 0");
    if (err) return fail(ctx, NQM_ERANGE, "queue capacity exceeded (err=0x%x)", err);
    if (use_graph) b.plan.valid = true;

    uint64_t total = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        const uint64_t c = hs[s] + h.sols[s];
        counts[s] = c;
        if (nodes_out) nodes_out[s] = hn[s] + h.nodes[s];
        total += c;
    }
    if (stats) {
        stats->total = total;
        stats->nodes = exec + h.exec_nodes;
        stats->kernel_ms = use_graph ? (double)run_ms : (double)expand_ms + (double)learn_ms + (double)run_ms;
        stats->wall_ms = now_ms() - t0;
        stats->launches = launches;
        stats->blocks = blocks;
        stats->search_ms = (double)search_ms;
        stats->rounds = rounds;
        stats->lanes = (uint64_t)blocks * threads;
        stats->parked_max = round_parked_max;
        stats->handed_back_max = round_handed_max;
        stats->hand_back_room = round_room;
    }
    (void)split_levels;
    return NQM_OK;
}

}  // namespace nqm

extern "C" int nqm_solve_batch(nqm_ctx* ctx, const nqm_solve* solves, uint32_t nsolves,
                               const nqm_tune* tune, uint64_t* counts, uint64_t* nodes,
                               nqm_stats* stats) {
    if (!ctx) return NQM_EINVAL;
    if (!solves || nsolves == 0 || !counts) return nqm::fail(ctx, NQM_EINVAL, "empty batch");
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    return nqm::solve_batch_impl(ctx, solves, nsolves, tune, counts, nodes, stats);
}

extern "C" int nqm_autotune(nqm_ctx* ctx, const nqm_solve* solves, uint32_t nsolves,
                            const nqm_tune_space* space, nqm_tune* best, nqm_tune_result* results,
                            uint32_t max_results, uint32_t* nresults) {
    if (!ctx || !solves || !nsolves || !space || !best) return NQM_EINVAL;
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    auto list = [](const uint32_t* v) {
        std::vector<uint32_t> out;
        for (int i = 0; i < 8 && v[i]; ++i) out.push_back(v[i]);
        if (out.empty()) out.push_back(0);  // automatic
        return out;
    };
    const std::vector<uint32_t> depths = list(space->depths);
    const std::vector<uint32_t> waves = list(space->waves_per_cu);
    const std::vector<uint32_t> rows = list(space->stack_rows);
    const uint32_t reps = space->repeats ? space->repeats : 3u;
    std::vector<uint64_t> ref(nsolves), got(nsolves);
    bool have_ref = false, have_best = false;
    double best_ms = 0;
    uint32_t nres = 0;
    int last_rc = NQM_OK;
    for (uint32_t d : depths)
        for (uint32_t w : waves)
            for (uint32_t r : rows) {
                nqm_tune t{};
                t.depth = d;
                t.waves_per_cu = w;
                t.stack_rows = r;
                t.flags = space->base_flags & ~(uint32_t)NQM_TUNE_GRAPH;
                std::vector<double> ms;
                bool feasible = true;
                for (uint32_t k = 0; k < reps && feasible; ++k) {
                    nqm_stats st{};
                    int rc = nqm::solve_batch_impl(ctx, solves, nsolves, &t, got.data(), nullptr, &st);
                    if (rc == NQM_ERANGE || rc == NQM_ENOMEM) {
                        // the queue this combination asks for does not fit
                        // (e.g. a stack-row cap that forces a deep expansion):
                        // it is reported, not tried
                        feasible = false;
                        last_rc = rc;
                        break;
                    }
                    if (rc != NQM_OK) return rc;
                    if (!have_ref) {
                        ref = got;
                        have_ref = true;
                    } else if (got != ref) {
                        return nqm::fail(ctx, NQM_ERANGE, "tune depth=%u waves=%u rows=%u changed a count", d,
                                         w, r);
                    }
                    ms.push_back(st.kernel_ms);
                }
                std::sort(ms.begin(), ms.end());
                const double med = feasible ? ms[ms.size() / 2] : -1.0;
                if (results && nres < max_results) {
                    results[nres].depth = d;
                    results[nres].waves_per_cu = w;
                    results[nres].stack_rows = r;
                    results[nres].pad = 0;
                    results[nres].ms = med;
                }
                ++nres;
                if (feasible && (!have_best || med < best_ms)) {
                    best_ms = med;
                    *best = t;
                    have_best = true;
                }
            }
    if (nresults) *nresults = std::min(nres, max_results);
    if (!have_best) return last_rc != NQM_OK ? last_rc : nqm::fail(ctx, NQM_EINVAL, "no tuning to try");
    return NQM_OK;
}
