// nqm_server.hip -- count requests served by continuous batching.
//
// No reference counterpart (the reference counts one pool per process run).
// Requests of any n and prefix join a device-resident work queue between
// drain rounds; every round is one budgeted K2 drain over the whole queue,
// so work of requests that arrived at different times overlaps on the
// device, and a request's result is ready the round its last item ends.
//
//   submit: admit while a counter slot is free and the request's expanded
//           items fit the queue memory; the request is expanded on the
//           device (staging buffer) and put in front of the queue when
//           there is room before it (new requests start first), else at
//           the tail.
//   step:   one drain round (pass budget) over parked continuations and
//           queued items; waves out of budget park their stacks as new
//           continuations, held items go back right after the queue (in
//           place), and flag kernels over what is left say which requests
//           are done.
//   poll:   finished results (count, nodes, admission and completion clock).
// Counts per request are exactly those of nqm_solve_batch on that request
// alone: items are disjoint subtrees, and counters are per slot.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "nqm_internal.h"
#include "nqm_items.h"

namespace nqm {

// Automatic staging depth: large boards are expanded until a piece needs at
// most this many LDS stack rows (14 one-wave workgroups per CU), as far as
// the queue memory left allows; small boards stop at row 6.
constexpr uint32_t kServeStackRows = 11;
// Default passes per wave per round (about 9 ms at full occupancy). Measured
// on a closed loop of 200 mixed requests: 13.7 requests/s against 10.8 at
// 16,384 passes; every round pays host round trips and the completion flags.
constexpr uint64_t kServeBudget = 65536;
// Default queue memory (each of the two queue buffers): room for N=20 staged
// to row 8 (148 M items, 4.4 GiB), which takes 1.41 s alone through the server
// against 1.99 s at row 7 (1.18 s as a batch).
constexpr uint64_t kServeQueueBytes = 8ull << 30;

__global__ void srv_zero_slots_kernel(const uint32_t* __restrict__ slots, uint32_t n, uint64_t* __restrict__ a,
                                      uint64_t* __restrict__ b) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        a[slots[i]] = 0;
        b[slots[i]] = 0;
    }
}

// queued[s] = 1 if a parked piece of slot s is waiting.
__global__ void srv_parked_kernel(const Cont* __restrict__ conts, uint32_t c, uint32_t* __restrict__ queued) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
        const uint32_t sl = conts[i].solve;
        if (i == 0 || conts[i - 1].solve != sl) queued[sl] = 1u;
    }
}

// queued[s] = 1 if any of items[0..c) belongs to slot s. Only the first item
// of each run of equal slots stores (items of a request sit together), so
// the stores are few and need no atomics.
__global__ void srv_queued_kernel(const Item* __restrict__ items, uint32_t c, uint32_t* __restrict__ queued) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
        const uint32_t sl = items[i].solve;
        if (i == 0 || items[i - 1].solve != sl) queued[sl] = 1u;
    }
}

}  // namespace nqm

struct nqm_server {
    nqm_ctx* ctx = nullptr;
    nqm_server_cfg cfg{};
    uint32_t slots_n = 0;
    uint64_t cap_items = 0;   // admission limit of queued items
    uint64_t buf_items = 0;   // physical capacity of each queue buffer
    nqm::DevBuf q[2], stage[2], part, counts, head, spill_count, err, target, sols, nodes, exec, outst, slot_list,
        lvl_nodes, cont[2], cont_counts;
    int cur = 0;
    int ci = 0;          // cont[ci] holds the parked pieces
    uint32_t conts = 0;  // parked pieces waiting for the next round
    bool fresh = false;  // requests were put in front of the queue since the last round
    uint64_t qstart = 0;  // the queue is q[cur][qstart, qstart + qcount)
    uint32_t qcount = 0;
    // Which requests still have queued items, without reading the queue: the
    // queue holds only requests' staged blocks (contiguous) and the items
    // rounds handed back (appended at the tail). Positions here are logical,
    // physical index in q[cur] + shift (moving the live range keeps them).
    // A block is used up once the head has passed its end; handed-back
    // ranges are kept until the head passes them, and only they and the
    // parked pieces are scanned.
    uint64_t shift = 0;
    std::vector<std::pair<uint64_t, uint64_t>> handed_back;
    struct Slot {
        bool active = false;
        uint64_t tag = 0;
        uint32_t n = 0, row = 0;
        uint64_t host_sols = 0, host_nodes = 0;
        uint32_t rounds = 0;
        double admit_ms = 0;
        uint64_t blk_end = 0;   // logical end of the staged block
        bool blk_done = true;   // the head has passed blk_end
    };
    std::vector<Slot> slots;
    std::vector<uint32_t> free_slots;
    std::vector<nqm_result> done;
    std::chrono::steady_clock::time_point t0;
    uint32_t active = 0;
    uint64_t rounds_total = 0;
    // per round: pieces parked, unstarted items handed back, and the bounds
    uint64_t round_lanes = 0, parked_max = 0, handed_back_max = 0, hand_back_room = 0;
    ~nqm_server() {
        for (nqm::DevBuf* d : {&q[0], &q[1], &stage[0], &stage[1], &part, &counts, &head, &spill_count, &err,
                               &target, &sols, &nodes, &exec, &outst, &slot_list, &lvl_nodes, &cont[0], &cont[1],
                               &cont_counts})
            if (d->ptr) (void)hipFree(d->ptr);
    }
    double clock_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace nqm {

namespace {

// Expand one request's initial items (host) into the staging buffer, one row
// per level up to its target row, but never past `room` items: a level that
// would overflow is dropped and the request stays at the shallower row.
// Returns the item count left in stage[*which].
int stage_request(nqm_server* s, const nqm_solve& q, uint32_t slot, uint64_t room, uint32_t* items_out,
                  int* which) {
    nqm_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    int rc;
    const uint32_t n = q.n;
    const uint32_t mask = (1u << n) - 1u;
    nqm_server::Slot& sl = s->slots[slot];
    // initial items (as nqm_solve_batch builds them)
    std::vector<Item> init;
    uint32_t cols = 0, ld = 0, rd = 0;
    for (uint32_t i = 0; i < q.k; ++i) {
        const uint32_t b = 1u << q.prefix[i];
        cols |= b;
        ld = ((ld | b) << 1) & mask;
        rd = (rd | b) >> 1;
    }
    uint32_t row = q.k;
    const bool mirror = !(s->cfg.flags & NQM_TUNE_NO_MIRROR);
    if (q.k == q.n) {
        sl.host_sols = 1;
    } else if (q.k == 0 && mirror && n >= 2) {
        row = 1;
        for (uint32_t c = 0; c < (n + 1) / 2; ++c) {
            const uint32_t ws = c < n / 2 ? 1u : 0u;
            const uint32_t b = 1u << c;
            Item it{};
            it.cols = b;
            it.ld = (b << 1) & mask;
            it.rd = b >> 1;
            it.f = ~(it.cols | it.ld | it.rd) & mask;
            it.solve = slot;
            it.meta = item_meta(n, 1, ws);
            sl.host_nodes += 1ull << ws;
            if (it.f) init.push_back(it);
        }
    } else {
        Item it{};
        it.cols = cols;
        it.ld = ld;
        it.rd = rd;
        it.f = ~(cols | ld | rd) & mask;
        it.solve = slot;
        it.meta = item_meta(n, q.k, 0);
        if (it.f) init.push_back(it);
    }
    // automatic depth: row 6, or deep enough that a piece needs at most
    // kServeStackRows LDS stack rows (large boards), as far as room allows
    const uint32_t depth = s->cfg.depth ? s->cfg.depth : std::max(6u, n > kServeStackRows + 1u ? n - 1u - kServeStackRows : 0u);
    const uint32_t target = std::min(n - 1, std::max(row, depth));
    uint32_t count = (uint32_t)init.size();
    if ((rc = ensure_dev(ctx, s->stage[0], std::max<size_t>(count, 1) * sizeof(Item))) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, s->part, kExpandPartBytes)) != NQM_OK) return rc;
    if (count)
        NQM_HIP(ctx, hipMemcpyAsync(s->stage[0].ptr, init.data(), count * sizeof(Item), hipMemcpyHostToDevice, st));
    uint32_t* cnt = static_cast<uint32_t*>(s->counts.ptr);
    NQM_HIP(ctx, k2_fill_u32(st, cnt, 1, count));
    int w = 0;
    uint32_t* in_slot = cnt;
    for (uint32_t r = row; r < target && count; ++r) {
        // expand exactly one row: target of this slot = r + 1. The level is
        // counted first and written only if it fits the room left, into a
        // staging buffer of exactly its size.
        uint8_t tgt[1] = {(uint8_t)(r + 1)};
        NQM_HIP(ctx, hipMemcpyAsync(static_cast<uint8_t*>(s->target.ptr) + slot, tgt, 1, hipMemcpyHostToDevice, st));
        uint32_t* out_slot = cnt + 1 + (w ^ 1);
        const uint32_t grid = std::min<uint32_t>(kExpandMaxGrid, (count + 255u) / 256u);
        const Item* const lin = static_cast<Item*>(s->stage[w].ptr);
        const uint8_t* const tgt_dev = static_cast<uint8_t*>(s->target.ptr);
        NQM_HIP(ctx, k2_expand_count(st, grid, lin, in_slot, s->part.ptr, out_slot, tgt_dev));
        uint32_t next = 0;
        NQM_HIP(ctx, hipMemcpyAsync(&next, out_slot, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        if (next > room) break;  // keep the shallower level
        if ((rc = ensure_dev(ctx, s->stage[w ^ 1], std::max<size_t>(next, 1) * sizeof(Item))) != NQM_OK) return rc;
        // the level's placements go to scratch counters first:
        // lvl_nodes[slot] (weighted, per solve) and lvl_nodes[slots_n] (executed)
        uint64_t* const lvl = static_cast<uint64_t*>(s->lvl_nodes.ptr);
        NQM_HIP(ctx, k2_fill_u64(st, lvl + slot, 1));
        NQM_HIP(ctx, k2_fill_u64(st, lvl + s->slots_n, 1));
        NQM_HIP(ctx, k2_expand_write(st, grid, lin, in_slot, s->part.ptr, true,
                                     static_cast<Item*>(s->stage[w ^ 1].ptr), next, tgt_dev, lvl,
                                     lvl + s->slots_n, static_cast<uint32_t*>(s->err.ptr)));
        uint64_t placed = 0;
        NQM_HIP(ctx, hipMemcpyAsync(&placed, lvl + slot, 8, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        sl.host_nodes += placed;
        count = next;
        in_slot = out_slot;
        row = r + 1;
        w ^= 1;
    }
    sl.row = row;
    *items_out = count;
    *which = w;
    return NQM_OK;
}

// Make room for `extra` items after the queue: move the live range to the
// front of the other buffer when the end of this one is near.
int make_room(nqm_server* s, uint64_t extra) {
    if (s->qstart + s->qcount + extra <= s->buf_items) return NQM_OK;
    nqm_ctx* ctx = s->ctx;
    if (s->qcount)
        NQM_HIP(ctx, hipMemcpyAsync(s->q[s->cur ^ 1].ptr, static_cast<Item*>(s->q[s->cur].ptr) + s->qstart,
                                    (size_t)s->qcount * sizeof(Item), hipMemcpyDeviceToDevice, ctx->stream));
    s->shift += s->qstart;  // logical positions stay
    s->cur ^= 1;
    s->qstart = 0;
    return NQM_OK;
}

uint32_t stack_rows_needed(const nqm_server* s) {
    uint32_t rows = 1;
    for (const auto& sl : s->slots)
        if (sl.active && sl.n > sl.row + 1) rows = std::max(rows, sl.n - 1 - sl.row);
    return rows;
}

}  // namespace

}  // namespace nqm

using namespace nqm;

extern "C" {

int nqm_server_create(nqm_ctx* ctx, const nqm_server_cfg* cfg, nqm_server** out) {
    if (!ctx || !out) return NQM_EINVAL;
    *out = nullptr;
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    nqm_server* s = new (std::nothrow) nqm_server();
    if (!s) return NQM_ENOMEM;
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.flags & ~(NQM_TUNE_NO_MIRROR | NQM_TUNE_NO_DONATE)) {
        delete s;
        return fail(ctx, NQM_EINVAL, "unknown server flags");
    }
    s->slots_n = s->cfg.max_active ? std::min(s->cfg.max_active, 1u << 20) : 256u;
    const uint64_t qbytes = s->cfg.queue_bytes ? s->cfg.queue_bytes : kServeQueueBytes;
    s->cap_items = std::max<uint64_t>(qbytes / sizeof(Item), 1024);
    // headroom for one round of spills from every lane (31 rows at most)
    const uint64_t spill = (uint64_t)ctx->prop.multiProcessorCount * 64u * 32u * 32u;
    s->buf_items = s->cap_items + spill;
    if (s->buf_items > 0x7fffffffull) {
        delete s;
        return fail(ctx, NQM_ERANGE, "queue too large");
    }
    int rc = NQM_OK;
    for (DevBuf* d : {&s->q[0], &s->q[1]})
        if (rc == NQM_OK) rc = ensure_dev(ctx, *d, s->buf_items * sizeof(Item));
    for (DevBuf* d : {&s->counts, &s->head, &s->spill_count, &s->err, &s->exec, &s->cont_counts})
        if (rc == NQM_OK) rc = ensure_dev(ctx, *d, 64);
    if (rc == NQM_OK) rc = ensure_dev(ctx, s->lvl_nodes, ((size_t)s->slots_n + 1) * 8);
    if (rc == NQM_OK) rc = ensure_dev(ctx, s->target, s->slots_n);
    for (DevBuf* d : {&s->sols, &s->nodes})
        if (rc == NQM_OK) rc = ensure_dev(ctx, *d, (size_t)s->slots_n * 8);
    if (rc == NQM_OK) rc = ensure_dev(ctx, s->outst, (size_t)s->slots_n * 4);
    if (rc == NQM_OK) rc = ensure_dev(ctx, s->slot_list, (size_t)s->slots_n * 4);
    if (rc != NQM_OK) {
        delete s;
        return rc;
    }
    hipStream_t st = ctx->stream;
    if (k2_fill_u32(st, static_cast<uint32_t*>(s->err.ptr), 1, 0) != hipSuccess ||
        k2_fill_u64(st, static_cast<uint64_t*>(s->exec.ptr), 1) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        delete s;
        return fail(ctx, NQM_EDEVICE, "server init failed");
    }
    s->slots.resize(s->slots_n);
    for (uint32_t i = s->slots_n; i > 0; --i) s->free_slots.push_back(i - 1);
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return NQM_OK;
}

void nqm_server_destroy(nqm_server* s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
}

double nqm_server_clock_ms(const nqm_server* s) { return s ? s->clock_ms() : 0.0; }

int nqm_server_submit(nqm_server* s, const nqm_solve* reqs, const uint64_t* tags, uint32_t nreq,
                      uint32_t* admitted) {
    if (!s || (nreq && (!reqs || !tags)) || !admitted) return NQM_EINVAL;
    nqm_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    *admitted = 0;
    for (uint32_t i = 0; i < nreq; ++i) {
        const nqm_solve& q = reqs[i];
        if (q.n < 1 || q.n > kMaxN || q.k > q.n) return fail(ctx, NQM_EINVAL, "request %u: n=%u k=%u", i, q.n, q.k);
        {  // validate the prefix
            uint32_t c = 0, l = 0, r = 0;
            const uint32_t mask = (1u << q.n) - 1u;
            for (uint32_t j = 0; j < q.k; ++j) {
                if (q.prefix[j] >= q.n) return fail(ctx, NQM_EPREFIX, "request %u: column out of range", i);
                const uint32_t b = 1u << q.prefix[j];
                if (b & (c | l | r)) return fail(ctx, NQM_EPREFIX, "request %u: invalid prefix", i);
                c |= b;
                l = ((l | b) << 1) & mask;
                r = (r | b) >> 1;
            }
        }
        if (s->free_slots.empty()) break;  // every counter slot is in flight
        const uint32_t slot = s->free_slots.back();
        nqm_server::Slot& sl = s->slots[slot];
        sl = nqm_server::Slot();
        sl.n = q.n;
        sl.tag = tags[i];
        // clear the slot's device counters before any expansion adds to them
        uint32_t one[1] = {slot};
        NQM_HIP(ctx, hipMemcpyAsync(s->slot_list.ptr, one, 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(srv_zero_slots_kernel, dim3(1), dim3(64), 0, st,
                           static_cast<const uint32_t*>(s->slot_list.ptr), 1u,
                           static_cast<uint64_t*>(s->sols.ptr), static_cast<uint64_t*>(s->nodes.ptr));
        NQM_HIP(ctx, hipGetLastError());
        uint32_t items = 0;
        int which = 0;
        // room left in the queue; an empty queue always takes a request (its
        // unexpanded items are at most ceil(n/2) <= 16)
        const uint64_t room = s->cap_items > s->qcount ? s->cap_items - s->qcount : 0;
        int rc = stage_request(s, q, slot, room, &items, &which);
        if (rc != NQM_OK) return rc;
        if ((uint64_t)s->qcount + items > s->cap_items && s->qcount > 0) break;  // no room: try later
        // a new request goes in front of the queue when there is room there
        // (the next round serves it first), else behind it
        const bool front = items <= s->qstart;
        if (!front && (rc = make_room(s, items)) != NQM_OK) return rc;
        Item* const q0 = static_cast<Item*>(s->q[s->cur].ptr);
        const uint64_t blk_lo = front ? s->qstart - items : s->qstart + s->qcount;  // physical
        if (items)
            NQM_HIP(ctx, hipMemcpyAsync(q0 + blk_lo, s->stage[which].ptr, (size_t)items * sizeof(Item),
                                        hipMemcpyDeviceToDevice, st));
        sl.blk_end = blk_lo + items + s->shift;
        sl.blk_done = items == 0;
        if (front) s->qstart -= items;
        s->fresh |= front && items != 0;
        s->qcount += items;
        s->free_slots.pop_back();
        sl.active = true;
        sl.admit_ms = s->clock_ms();
        ++s->active;
        ++*admitted;
        if (items == 0) {  // answered on admission (complete or dead prefix)
            NQM_HIP(ctx, hipStreamSynchronize(st));
            uint64_t nd = 0;
            NQM_HIP(ctx, hipMemcpy(&nd, static_cast<uint64_t*>(s->nodes.ptr) + slot, 8, hipMemcpyDeviceToHost));
            nqm_result r{};
            r.tag = sl.tag;
            r.count = sl.host_sols;
            r.nodes = nd + sl.host_nodes;
            r.rounds = 0;
            r.admit_ms = sl.admit_ms;
            r.done_ms = s->clock_ms();
            s->done.push_back(r);
            sl.active = false;
            --s->active;
            s->free_slots.push_back(slot);
        }
    }
    NQM_HIP(ctx, hipStreamSynchronize(st));
    return NQM_OK;
}

int nqm_server_step(nqm_server* s, uint32_t* active_out) {
    if (!s) return NQM_EINVAL;
    nqm_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    int rc;
    if (s->qcount || s->conts) {
        const uint32_t rows = stack_rows_needed(s);
        uint32_t blocks = 0, threads = 0;
        {
            // as many workgroups per CU as LDS allows, of the automatic size
            uint32_t th = 64;
            if ((rc = k2_auto_threads(rows, ctx->lds_per_block, &th)) != NQM_OK)
                return fail(ctx, rc, "occupancy query failed");
            int per_cu = 0;
            if ((rc = k2_max_blocks_per_cu(th, k2_lds_bytes(rows, th), &per_cu)) != NQM_OK)
                return fail(ctx, rc, "occupancy query failed");
            blocks = (uint32_t)std::max(per_cu, 1) * (uint32_t)ctx->prop.multiProcessorCount;
            threads = th;
        }
        // a round hands back the unstarted items its waves hold (unfinished
        // pieces are parked as continuations); the queue itself stays in place
        const uint64_t spill_bound = (uint64_t)blocks * (threads / 64u) * kGrabItems;
        const uint64_t lanes = (uint64_t)blocks * threads;
        if ((rc = ensure_dev(ctx, s->cont[s->ci], std::max<size_t>(s->conts, 1) * sizeof(Cont))) != NQM_OK)
            return rc;
        if ((rc = ensure_dev(ctx, s->cont[s->ci ^ 1], (2 * (size_t)lanes + s->conts) * sizeof(Cont))) != NQM_OK)
            return rc;
        uint32_t* const d_cc = static_cast<uint32_t*>(s->cont_counts.ptr);
        const uint64_t budget = s->cfg.budget ? s->cfg.budget : kServeBudget;
        if ((rc = make_room(s, spill_bound)) != NQM_OK) return rc;
        // with no room for a full round of spills, drain this round to the end
        const bool spill_ok = s->qstart + (uint64_t)s->qcount + spill_bound <= s->buf_items;
        uint32_t* cnt = static_cast<uint32_t*>(s->counts.ptr);
        uint32_t* qslot = cnt + 8;
        NQM_HIP(ctx, k2_fill_u32(st, qslot, 1, s->qcount));
        NQM_HIP(ctx, k2_fill_u32(st, static_cast<uint32_t*>(s->head.ptr), 1, 0));
        NQM_HIP(ctx, k2_fill_u32(st, static_cast<uint32_t*>(s->spill_count.ptr), 1, s->qcount));
        Item* const queue = static_cast<Item*>(s->q[s->cur].ptr) + s->qstart;
        K2DrainParams p;
        p.items = queue;
        p.item_count = qslot;
        p.head = static_cast<uint32_t*>(s->head.ptr);
        p.spill = queue;  // spills go right after the queue
        p.spill_count = static_cast<uint32_t*>(s->spill_count.ptr);
        p.spill_cap = (uint32_t)std::min<uint64_t>(s->buf_items - s->qstart, 0x7fffffffull);
        p.solve_sols = static_cast<uint64_t*>(s->sols.ptr);
        p.solve_nodes = static_cast<uint64_t*>(s->nodes.ptr);
        p.exec_nodes = static_cast<uint64_t*>(s->exec.ptr);
        p.err = static_cast<uint32_t*>(s->err.ptr);
        p.stack_rows = rows;
        p.refill_idle = 8;
        p.donate = (s->cfg.flags & NQM_TUNE_NO_DONATE) ? 0u : 1u;
        p.budget = spill_ok ? budget : 0;
        // Queue items first (new requests are at its front), unless the
        // parked pieces outnumber the lanes: then they go first, so that they
        // cannot pile up. A round right after new requests went in front
        // still takes the queue first while parked pieces stay under 4x the
        // lanes; else a small request could wait rounds behind parked work.
        p.items_first = (s->conts <= lanes || (s->fresh && s->conts <= 4 * lanes)) ? 1u : 0u;
        s->fresh = false;
        NQM_HIP(ctx, k2_fill_u32(st, d_cc, 3, 0));
        NQM_HIP(ctx, k2_fill_u32(st, d_cc, 1, s->conts));
        p.cont_in = static_cast<const Cont*>(s->cont[s->ci].ptr);
        p.cont_in_count = d_cc;
        p.cont_head = d_cc + 1;
        p.cont_out = static_cast<Cont*>(s->cont[s->ci ^ 1].ptr);
        p.cont_out_count = d_cc + 2;
        p.cont_cap = (uint32_t)std::min<uint64_t>(s->cont[s->ci ^ 1].cap / sizeof(Cont), 0x7fffffffull);
        NQM_HIP(ctx, k2_drain(st, blocks, threads, p));
        uint32_t hd = 0, end = 0, err = 0, cc[3] = {0, 0, 0};
        NQM_HIP(ctx, hipMemcpyAsync(&hd, p.head, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(&end, p.spill_count, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(&err, p.err, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(cc, d_cc, 12, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        if (err) return fail(ctx, NQM_ERANGE, "server queue capacity exceeded (err=0x%x)", err);
        const uint32_t h0 = std::min(hd, s->qcount);  // items [h0, qcount) were never taken
        s->round_lanes = std::max(s->round_lanes, lanes);
        s->hand_back_room = std::max(s->hand_back_room, spill_bound);
        s->parked_max = std::max<uint64_t>(s->parked_max, cc[2]);
        s->handed_back_max = std::max<uint64_t>(s->handed_back_max, end > s->qcount ? end - s->qcount : 0u);
        if (end > s->qcount) {  // handed back after the old tail
            const uint64_t lo = s->qstart + s->qcount + s->shift, hi = s->qstart + end + s->shift;
            if (!s->handed_back.empty() && s->handed_back.back().second == lo)
                s->handed_back.back().second = hi;
            else
                s->handed_back.emplace_back(lo, hi);
        }
        s->qstart += h0;
        s->qcount = end - h0;
        // what the head has passed is used up for good
        const uint64_t head = s->qstart + s->shift;
        size_t keep = 0;
        for (auto& r : s->handed_back)
            if (r.second > head) s->handed_back[keep++] = {std::max(r.first, head), r.second};
        s->handed_back.resize(keep);
        for (auto& sl : s->slots)
            if (sl.active && !sl.blk_done && sl.blk_end <= head) sl.blk_done = true;
        // parked pieces no wave resumed this round follow the new ones
        const uint32_t c0 = std::min(cc[1], s->conts);
        if (c0 < s->conts)
            NQM_HIP(ctx, hipMemcpyAsync(static_cast<Cont*>(s->cont[s->ci ^ 1].ptr) + cc[2],
                                        static_cast<const Cont*>(s->cont[s->ci].ptr) + c0,
                                        (size_t)(s->conts - c0) * sizeof(Cont), hipMemcpyDeviceToDevice, st));
        s->conts = cc[2] + (s->conts - c0);
        s->ci ^= 1;
        // which slots still have queued work
        NQM_HIP(ctx, k2_fill_u32(st, static_cast<uint32_t*>(s->outst.ptr), s->slots_n, 0));
        if (s->conts) {
            const uint32_t grid = std::min<uint32_t>(4096u, (s->conts + 255u) / 256u);
            hipLaunchKernelGGL(srv_parked_kernel, dim3(grid), dim3(256), 0, st,
                               static_cast<const Cont*>(s->cont[s->ci].ptr), s->conts,
                               static_cast<uint32_t*>(s->outst.ptr));
            NQM_HIP(ctx, hipGetLastError());
        }
        for (const auto& r : s->handed_back) {  // the only queued items not in a staged block
            const uint32_t c = (uint32_t)(r.second - r.first);
            const uint32_t grid = std::min<uint32_t>(4096u, (c + 255u) / 256u);
            hipLaunchKernelGGL(srv_queued_kernel, dim3(grid), dim3(256), 0, st,
                               static_cast<const Item*>(s->q[s->cur].ptr) + (r.first - s->shift), c,
                               static_cast<uint32_t*>(s->outst.ptr));
            NQM_HIP(ctx, hipGetLastError());
        }
        ++s->rounds_total;
    }
    // completed requests: active slots with nothing left in the queue
    if (s->active) {
        std::vector<uint32_t> outst(s->slots_n);
        std::vector<uint64_t> hs(s->slots_n), hn(s->slots_n);
        NQM_HIP(ctx, hipMemcpyAsync(outst.data(), s->outst.ptr, (size_t)s->slots_n * 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(hs.data(), s->sols.ptr, (size_t)s->slots_n * 8, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(hn.data(), s->nodes.ptr, (size_t)s->slots_n * 8, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        const double now = s->clock_ms();
        for (uint32_t i = 0; i < s->slots_n; ++i) {
            nqm_server::Slot& sl = s->slots[i];
            if (!sl.active) continue;
            ++sl.rounds;
            if ((s->qcount != 0 || s->conts != 0) && (outst[i] != 0 || !sl.blk_done)) continue;
            nqm_result r{};
            r.tag = sl.tag;
            r.count = hs[i] + sl.host_sols;
            r.nodes = hn[i] + sl.host_nodes;
            r.rounds = sl.rounds;
            r.admit_ms = sl.admit_ms;
            r.done_ms = now;
            s->done.push_back(r);
            sl.active = false;
            --s->active;
            s->free_slots.push_back(i);
        }
    }
    if (active_out) *active_out = s->active;
    return NQM_OK;
}

int nqm_server_poll(nqm_server* s, nqm_result* out, uint32_t max, uint32_t* nout) {
    if (!s || !nout || (max && !out)) return NQM_EINVAL;
    const uint32_t k = (uint32_t)std::min<size_t>(max, s->done.size());
    for (uint32_t i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *nout = k;
    return NQM_OK;
}

int nqm_server_round_stats(const nqm_server* s, nqm_stats* out) {
    if (!s || !out) return NQM_EINVAL;
    *out = nqm_stats{};
    out->rounds = (uint32_t)std::min<uint64_t>(s->rounds_total, 0xffffffffu);
    out->lanes = s->round_lanes;
    out->parked_max = s->parked_max;
    out->handed_back_max = s->handed_back_max;
    out->hand_back_room = s->hand_back_room;
    return NQM_OK;
}

int nqm_server_info(const nqm_server* s, uint32_t* active, uint64_t* queued_items, uint64_t* rounds) {
    if (!s) return NQM_EINVAL;
    if (active) *active = s->active;
    if (queued_items) *queued_items = s->qcount;
    if (rounds) *rounds = s->rounds_total;
    return NQM_OK;
}

}  // extern "C"
