// k4_sym.hip -- K4: the symmetry-reduced count on a persistent drain.
//
// The tree is the one of sym_rules.h (corner / non-corner method), searched
// the way K2 searches the plain tree: a queue of open nodes, lanes that draw
// from it, an LDS stack per lane and intra-wave donation. Three steps:
//
//  * items: the host places the first rows, then k4_expand grows them level
//    by level on the device to row R = ceil(n/2) - 1 (so that row 0's side
//    rule, rows < b1 avoid the side columns, is done before the drain).
//
//  * k4_drain: K2's pass plus the row rules left below R, as selects:
//      row b2 (non-corner): with one side column open, only it; with both
//        open, nothing (the exact rule of sym_free, so the node numbers are
//        the oracle's);
//      last row (non-corner): no column within b1 of a corner;
//      row b1 (corner, b1 > R): column 1, kept out of avail above it, is
//        given back.
//    A lane does not know its row; a marker bit in the high, unused bits of
//    its diagonals tells it. The ld marker sits 31 - (b2 - row) bits up, so
//    it reaches bit 31 exactly when the child is at row b2 (and is shifted
//    out after it); the rd marker sits n + (last - row) bits up and reaches
//    bit n at the last row (rd only moves down, and never below n); a corner
//    item's ld marker reaches bit 31 at the child at row b1. They ride along
//    in the stack slots and in donated rows like the real bits. The
//    item's ld is cut to n bits, so real diagonal bits stay below the markers
//    for n <= 21; the stack of last - R rows fits LDS at 16 waves per CU for
//    n <= 19 and at 12 for n = 20 (k4_supports).
//    A full board whose class may have another representative (a tie of the
//    top queen's corner distance: the last queen on column b2, the column-0
//    queen on row b2 -- which puts bit b1 into the last row's ld -- or row
//    b1's queen on column n-1) is not counted in the drain: the lane writes
//    {item, ld and rd of the last row} into its wave's record chunk. Every other
//    full board is its class's only representative and counts 8.
//
//  * ties: k4_rec_compact makes the records one dense list; in
//    k4_research one lane per record searches the item again bottom-up,
//    keeping only the columns whose diagonal bits are in the recorded ld and
//    rd (TieSearch, k4_ties.h), and k4_check runs the rotation check
//    (sym_check_packed) on the boards it finds. Several boards of one item
//    can share an ld and rd, and each of them wrote the same record: each copy
//    counts its m boards with weight 1/m (scaled by lcm(1..16)).
//
// Unique classes = drain leaves + tie classes; total = 8 x drain leaves +
// tie weights; both equal K3's and the oracle's, and so do the nodes.
#include <algorithm>
#include <new>
#include <vector>

#include "nqm_internal.h"
#include "sym_rules.h"
#include "k4_ties.h"

namespace nqm {
namespace {

// An open node of the symmetric search at row `row`.
struct alignas(16) SymQ {
    uint32_t f;       // candidates of `row` under the case's rules (never 0 in a queue)
    uint32_t cols;    // columns of rows 0..row-1
    uint32_t ld, rd;  // relative diagonals at `row`; ld cut to n bits
    uint32_t meta;    // n | row << 8 | b1 << 16 | corner << 24 | top at b1 << 25
    uint32_t pad;
    uint32_t p0, p1;  // columns of rows 0..11, 5 bits each (row r at bit 5 r of p1:p0)
};
static_assert(sizeof(SymQ) == 32, "two 16-byte loads");

__host__ __device__ inline uint32_t symq_meta(uint32_t n, uint32_t row, uint32_t b1, uint32_t corner,
                                              uint32_t top_b1) {
    return (n & 31u) | ((row & 31u) << 8) | ((b1 & 31u) << 16) | ((corner & 1u) << 24) | ((top_b1 & 1u) << 25);
}
__host__ __device__ inline uint32_t symq_n(uint32_t m) { return m & 31u; }
__host__ __device__ inline uint32_t symq_row(uint32_t m) { return (m >> 8) & 31u; }
__host__ __device__ inline uint32_t symq_b1(uint32_t m) { return (m >> 16) & 31u; }
__host__ __device__ inline uint32_t symq_corner(uint32_t m) { return (m >> 24) & 1u; }
__host__ __device__ inline uint32_t symq_top_b1(uint32_t m) { return (m >> 25) & 1u; }

constexpr uint32_t kPrefixRows = 12;  // rows whose columns an item keeps

// The row every item is grown to (the corner rule "rows < b1 avoid column
// 1" that reaches below it is the drain's).
__host__ __device__ inline uint32_t symq_target(uint32_t meta, uint32_t R) {
    (void)meta;
    return R;
}

// The child of item `it` (at row y) for column bit b; f = 0 if it has no
// candidate.
__host__ __device__ inline SymQ symq_child(const SymQ& it, const SymCase& c, uint32_t b) {
    const uint32_t y = symq_row(it.meta);
    SymQ ch;
    ch.cols = it.cols | b;
    ch.ld = ((it.ld | b) << 1) & c.mask;
    ch.rd = (it.rd | b) >> 1;
    ch.f = sym_free(c, y + 1u, ch.ld, ch.cols, ch.rd);
    const uint32_t col = (uint32_t)__builtin_ctz(b);
    uint64_t p = ((uint64_t)it.p1 << 32) | it.p0;
    if (y < kPrefixRows) p |= (uint64_t)col << (5u * y);
    ch.p0 = (uint32_t)p;
    ch.p1 = (uint32_t)(p >> 32);
    const uint32_t top_b1 = symq_top_b1(it.meta) | (!c.corner && y == c.b1 && b == c.top ? 1u : 0u);
    ch.meta = symq_meta(c.n, y + 1u, c.b1, c.corner, top_b1);
    ch.pad = 0;
    return ch;
}

// ---------------------------------------------------------------------------
// expansion: one level of children for every pending item
// ---------------------------------------------------------------------------
// A child at its target row goes to the final queue, one still below it to
// the next level. Three launches per level, none with a same-address
// atomic: k4_expand_count (block b, items [256 b, 256 b + 256): its
// children of each kind and the placements), k4_expand_scan (one block: the
// blocks' offsets, the level's sizes in cnt[0..1], the placements added to
// nodes), k4_expand_write (block b writes its children at its offsets, in
// item order).
__device__ __forceinline__ uint32_t k4_incl_scan(uint32_t x, uint32_t lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// The children of item i of each kind (final, next).
__device__ __forceinline__ void k4_child_counts(const SymQ& it, uint32_t R, uint32_t& nf, uint32_t& nn) {
    const SymCase c = sym_case(symq_n(it.meta), symq_b1(it.meta), symq_corner(it.meta));
    const uint32_t tgt = symq_target(it.meta, R);
    nf = 0;
    nn = 0;
    for (uint32_t f = it.f; f; f &= f - 1u) {
        const SymQ ch = symq_child(it, c, f & (0u - f));
        if (ch.f) {
            if (symq_row(ch.meta) >= tgt)
                ++nf;
            else
                ++nn;
        }
    }
}

// Block sums of 256 per-thread values (a, b) over the block; every thread
// gets its exclusive prefix and the block totals.
__device__ __forceinline__ void k4_block_scan2(uint32_t a, uint32_t b, uint32_t& pa, uint32_t& pb, uint32_t& ta,
                                               uint32_t& tb, uint2* wsum) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t ia = k4_incl_scan(a, lane), ib = k4_incl_scan(b, lane);
    if (lane == 63u) wsum[wave] = make_uint2(ia, ib);
    __syncthreads();
    uint32_t ba = 0, bb = 0;
    ta = 0;
    tb = 0;
    for (uint32_t w = 0; w < 4u; ++w) {
        const uint2 v = wsum[w];
        if (w < wave) {
            ba += v.x;
            bb += v.y;
        }
        ta += v.x;
        tb += v.y;
    }
    pa = ba + ia - a;
    pb = bb + ib - b;
}

__global__ void __launch_bounds__(256)
k4_expand_count_kernel(const SymQ* __restrict__ in, uint32_t count, uint32_t R, uint4* __restrict__ part) {
    __shared__ uint2 wsum[4];
    __shared__ uint32_t wplaced[4];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t nf = 0, nn = 0, placed = 0;
    if (i < count) {
        const SymQ it = in[i];
        k4_child_counts(it, R, nf, nn);
        placed = (uint32_t)__popc(it.f);
    }
    uint32_t pa, pb, ta, tb;
    k4_block_scan2(nf, nn, pa, pb, ta, tb, wsum);
    for (int o = 32; o > 0; o >>= 1) placed += __shfl_xor(placed, o, 64);
    if (lane == 0) wplaced[wave] = placed;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = make_uint4(ta, tb, wplaced[0] + wplaced[1] + wplaced[2] + wplaced[3], 0u);
}

// One block: part[b].x / .y become block b's offsets in the final / next
// queue (exclusive scans); cnt[0], cnt[1] = the level's final and next
// sizes; the level's placements go to nodes.
__global__ void __launch_bounds__(1024)
k4_expand_scan_kernel(uint4* __restrict__ part, uint32_t nblocks, uint32_t* __restrict__ cnt,
                      unsigned long long* __restrict__ nodes) {
    __shared__ uint2 wsum[16];
    __shared__ unsigned long long wpl[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t ca = 0, cb = 0;  // block-uniform carries
    uint64_t placed = 0;
    for (uint32_t base = 0; base < nblocks; base += 1024u) {
        const uint32_t i = base + tid;
        const uint4 v = i < nblocks ? part[i] : make_uint4(0u, 0u, 0u, 0u);
        placed += v.z;
        const uint32_t ia = k4_incl_scan(v.x, lane), ib = k4_incl_scan(v.y, lane);
        if (lane == 63u) wsum[wave] = make_uint2(ia, ib);
        __syncthreads();
        uint32_t ba = 0, bb = 0, ta = 0, tb = 0;
        for (uint32_t w = 0; w < 16u; ++w) {
            const uint2 s = wsum[w];
            if (w < wave) {
                ba += s.x;
                bb += s.y;
            }
            ta += s.x;
            tb += s.y;
        }
        if (i < nblocks) part[i] = make_uint4(ca + ba + ia - v.x, cb + bb + ib - v.y, v.z, 0u);
        ca += ta;
        cb += tb;
        __syncthreads();
    }
    for (int o = 32; o > 0; o >>= 1) placed += __shfl_xor(placed, o, 64);
    if (lane == 0) wpl[wave] = placed;
    __syncthreads();
    if (tid == 0) {
        unsigned long long sp = 0;
        for (uint32_t w = 0; w < 16u; ++w) sp += wpl[w];
        cnt[0] = ca;
        cnt[1] = cb;
        if (sp) atomicAdd(nodes, sp);
    }
}

// Block b writes its children: final ones at fin_q[fin_at + offset] (the
// final queue already holds fin_at items), next ones at next_q[offset].
// Offsets past a capacity set err bit 1 (the level was not the counted one).
// All children of an item go to one queue (they share a row). The block
// stages its items and, per output, the item and column in LDS; then thread k
// writes outputs k, k + 256, ...: consecutive threads, consecutive items.
constexpr uint32_t kK4MaxChildren = 20;  // columns of a row below row 0 (n - 1 <= 19)
__global__ void __launch_bounds__(256)
k4_expand_write_kernel(const SymQ* __restrict__ in, uint32_t count, uint32_t R, const uint4* __restrict__ part,
                       SymQ* __restrict__ fin_q, uint32_t fin_at, uint32_t fin_cap, SymQ* __restrict__ next_q,
                       uint32_t next_cap, uint32_t* __restrict__ err) {
    __shared__ uint2 wsum[4];
    __shared__ SymQ s_item[256];
    __shared__ uint8_t s_map[256u * kK4MaxChildren];  // output -> item of the block
    __shared__ uint8_t s_col[256u * kK4MaxChildren];  // output -> column
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    SymQ it{};
    uint32_t live = 0, nf = 0, nn = 0;
    bool to_fin = false;
    if (i < count) {
        it = in[i];
        const SymCase c = sym_case(symq_n(it.meta), symq_b1(it.meta), symq_corner(it.meta));
        to_fin = symq_row(it.meta) + 1u >= symq_target(it.meta, R);
        for (uint32_t f = it.f; f; f &= f - 1u) {
            const uint32_t b = f & (0u - f);
            if (symq_child(it, c, b).f) live |= b;
        }
        const uint32_t k = (uint32_t)__popc(live);
        nf = to_fin ? k : 0u;
        nn = to_fin ? 0u : k;
    }
    uint32_t pa, pb, ta, tb;
    k4_block_scan2(nf, nn, pa, pb, ta, tb, wsum);
    // outputs of the block: its final children [0, ta), then its next ones
    s_item[threadIdx.x] = it;
    uint32_t o = to_fin ? pa : ta + pb;
    for (uint32_t f = live; f; f &= f - 1u, ++o) {
        s_map[o] = (uint8_t)threadIdx.x;
        s_col[o] = (uint8_t)__builtin_ctz(f);
    }
    __syncthreads();
    const uint4 off = part[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < ta + tb; k += 256u) {
        const SymQ& p = s_item[s_map[k]];
        const SymCase c = sym_case(symq_n(p.meta), symq_b1(p.meta), symq_corner(p.meta));
        const SymQ ch = symq_child(p, c, 1u << s_col[k]);
        if (k < ta) {
            const uint32_t d = fin_at + off.x + k;
            if (d < fin_cap)
                fin_q[d] = ch;
            else
                atomicOr(err, 1u);
        } else {
            const uint32_t d = off.y + (k - ta);
            if (d < next_cap)
                next_q[d] = ch;
            else
                atomicOr(err, 1u);
        }
    }
}

// ---------------------------------------------------------------------------
// drain
// ---------------------------------------------------------------------------
constexpr uint32_t kK4Threads = 256;  // 4 waves per workgroup
constexpr uint32_t kRecChunk = 4096;  // records per chunk of a wave
// A phase ends at a vote with at most this many records in the chunk: the 8
// passes to the next vote add at most 8 x 64.
constexpr uint32_t kRecLimit = kRecChunk - 8u * 64u;
constexpr uint32_t kK4Grab = 128;  // queue items one wave takes per atomic at most
constexpr uint32_t kK4MaxRows = 10;  // rows below an item (last - R, k4_supports)

struct K4Args {
    const SymQ* items;
    uint32_t count;
    uint32_t nseg;
    uint32_t* head;       // nseg heads, kHeadStride u32 apart
    uint32_t n, last, mask, side;
    uint32_t refill_idle;
    uint32_t next_refill;  // a phase ends once this many lanes started their next item
    uint4* recs;          // record chunks (kRecChunk records {item, ld, rd, 0} each); chunk nchunks is the overflow sink
    uint32_t* rec_counts; // records in each chunk, written when the wave closes it
    uint32_t* chunk_off;  // each chunk's place in the dense record list (its share of rec_total)
    uint32_t* chunk_head; // chunks handed out
    uint32_t* rec_total;  // records written
    uint32_t nchunks;
    uint64_t* wave_part;  // per wave {leaves counted 8, executed placements}
    uint32_t* err;
};

// LDS rows as in K2's drain: row 0 = donation scratch and the target of an
// empty stack's pop; slot j of lane t at 16 t + 16 T (j + 1).
constexpr uint32_t kK4Row = 16u * kK4Threads;

typedef uint32_t k4_v4u_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) k4_v4u_t k4_lds_v4u;
typedef __attribute__((address_space(3))) uint32_t k4_lds_u32;
__device__ __forceinline__ uint4 k4_slot_load(uint32_t addr) {
    const k4_v4u_t v = *reinterpret_cast<const k4_lds_v4u*>((size_t)addr);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void k4_slot_store(uint32_t addr, uint4 s) {
    *reinterpret_cast<k4_lds_v4u*>((size_t)addr) = k4_v4u_t{s.x, s.y, s.z, s.w};
}

__device__ __forceinline__ uint32_t k4_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint64_t k4_wave_sum64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ uint32_t k4_wave_max(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// One search phase (asm: it manages the execution mask itself), the pass of
// K2's drain with the row rules and the tie records added. Registers:
// v[88:91] / v[92:95] = {f, avail, ld, rd} of the current row, alternating
// per pass; v[96:99] = the record {item, ld, rd, 0} (v96 holds the lane's
// item).
// A pass:
//   pop    as in K2 (v_cmpx on an empty row, borrow = piece done);
//   step   b = f & -f; a full board when avail == b. A tie board (the last
//          queen on endb, or bit b1 of ld: endb = ~0 makes every board of
//          the item one) leaves the solution mask and is written to the
//          wave's chunk at index cnt + rank; the others add to sols. Push
//          as in K2. The child: ld' = (ld + b) << 1, rd' = (rd | b) >> 1,
//          t1 = ld' >> 31 (arithmetic: the row-b2 marker, or the row-b1
//          marker of a corner item), t2 = bit n of rd' (the last-row
//          marker), avail' = ((avail ^ b) & ~(lm & t2)) | (amask & t1) (a
//          corner item gets column 1 back at row b1), and at row b2 the side
//          rule from the open side columns: m0 / mt = bits o0 / on of avail'
//          as 0 or ~0 (o0 = 0 and on = n-1; 31 for a corner item, whose
//          avail' never has bit 31), blocked = side columns where both are
//          open, all others where either is; the
//          free set is avail' & ~ld' & ~(rd' | (blocked & t1)). The
//          solution and tie masks come last, and only in a pass where some
//          lane completed a board (most passes complete none).
// bitop3 truth tables index (S0 S1 S2) as a 3-bit number, S0 the high bit:
// 0x14 = (a ^ b) & ~c, 0xd4 = c ? a & b : a | b, 0xf8 = a | (b & c),
// 0x10 = a & ~b & ~c.
#define NQM_K4_POP(U, CX, CA, CL, CR, CT)                  \
    "v_cmpx_eq_u32 vcc, 0, " CX "\n"                       \
    "v_subrev_co_u32 %[q], vcc, %[row], %[q]\n"            \
    "ds_read_b128 " CT ", %[q] offset:%[row]\n"            \
    "s_cbranch_vccz .Lk4_nd" #U "_%=\n"                    \
    "s_and_b64 exec, vcc, %[hasnext]\n"                    \
    "s_cbranch_execz .Lk4_nn" #U "_%=\n"                   \
    "s_waitcnt lgkmcnt(0)\n"                               \
    "v_mov_b32 " CX ", v100\n"                             \
    "v_mov_b32 " CA ", v101\n"                             \
    "v_mov_b32 " CL ", v102\n"                             \
    "v_mov_b32 " CR ", v103\n"                             \
    "v_mov_b32 %[lm], v104\n"                              \
    "v_mov_b32 %[bitb1], v105\n"                           \
    "v_mov_b32 %[endb], v106\n"                            \
    "v_mov_b32 %[amask], v107\n"                           \
    "v_mov_b32 %[o0], v108\n"                              \
    "v_mov_b32 %[on], v109\n"                              \
    "v_mov_b32 v96, v110\n"                                \
    "v_add_u32 %[q], %[row], %[q]\n"                       \
    "s_andn2_b64 %[hasnext], %[hasnext], exec\n"           \
    "s_andn2_b64 vcc, vcc, exec\n"                         \
    ".Lk4_nn" #U "_%=:\n"                                  \
    "s_andn2_b64 %[alive], %[alive], vcc\n"                \
    "s_bcnt1_i32_b64 %[t], vcc\n"                          \
    "s_add_u32 %[t2], %[ran], " #U "\n"                    \
    "s_mul_i32 %[t], %[t], %[t2]\n"                        \
    "s_add_u32 %[nodes], %[nodes], %[t]\n"                 \
    ".Lk4_nd" #U "_%=:\n"                                  \
    "s_mov_b64 exec, %[alive]\n"                           \
    "s_waitcnt lgkmcnt(0)\n"
// AM(NA, NL): the corner rule's "column 1 back at row b1" (NQM_K4_AM_ON), or
// nothing in a phase none of whose lanes has that rule (NQM_K4_AM_OFF);
// NQM_K4_CAM_ON is NQM_K4_AM_ON for NQM_K4_CSTEP, which has no t1 of its own
#define NQM_K4_AM_ON(NA, NL) "v_and_or_b32 " NA ", %[amask], %[t1], " NA "\n"
#define NQM_K4_AM_OFF(NA, NL) ""
#define NQM_K4_CAM_ON(NA, NL)                              \
    "v_ashrrev_i32 %[t1], 31, " NL "\n"                    \
    "v_and_or_b32 " NA ", %[amask], %[t1], " NA "\n"
#define NQM_K4_STEP(U, CX, CA, CL, CR, CT, NX, NA, NL, NR, AM) \
    "v_sub_u32 %[ng], 0, " CX "\n"                         \
    "v_and_b32 %[b], " CX ", %[ng]\n"                      \
    "v_or_b32 " NR ", " CR ", %[b]\n"                      \
    "v_add_lshl_u32 " NL ", %[b], " CL ", 1\n"             \
    "v_cmp_eq_u32 %[sm], " CA ", %[b]\n"                   \
    "v_cmpx_ne_u32 vcc, " CX ", %[b]\n"                    \
    "v_xor_b32 " CX ", " CX ", %[b]\n"                     \
    "ds_write_b128 %[q], " CT " offset:%[row]\n"           \
    "v_add_u32 %[q], %[row], %[q]\n"                       \
    "s_mov_b64 exec, %[alive]\n"                           \
    "v_lshrrev_b32 " NR ", 1, " NR "\n"                    \
    "v_ashrrev_i32 %[t1], 31, " NL "\n"                    \
    "v_bfe_i32 %[t2v], " NR ", %[sn], 1\n"                 \
    "v_and_b32 %[t2v], %[lm], %[t2v]\n"                    \
    "v_bitop3_b32 " NA ", " CA ", %[b], %[t2v] bitop3:0x14\n" \
    AM(NA, NL)                                             \
    "v_bfe_i32 %[mt], " NA ", %[on], 1\n"                  \
    "v_bfe_i32 %[m0], " NA ", %[o0], 1\n"                  \
    "v_bitop3_b32 %[mt], %[mt], %[m0], %[sside] bitop3:0xd4\n" \
    "v_bitop3_b32 %[t1], " NR ", %[mt], %[t1] bitop3:0xf8\n" \
    "v_bitop3_b32 " NX ", " NA ", " NL ", %[t1] bitop3:0x10\n" \
    "s_cmp_eq_u64 %[sm], 0\n"                              \
    "s_cbranch_scc1 .Lk4_ne" #U "_%=\n"                    \
    "v_and_b32 %[e], %[b], %[endb]\n"                      \
    "v_and_or_b32 %[e], " CL ", %[bitb1], %[e]\n"          \
    "v_cmp_ne_u32 %[se], 0, %[e]\n"                        \
    "s_and_b64 %[se], %[se], %[sm]\n"                      \
    "s_andn2_b64 %[sm], %[sm], %[se]\n"                    \
    "s_bcnt1_i32_b64 %[t], %[sm]\n"                        \
    "s_add_u32 %[sols], %[sols], %[t]\n"                   \
    "s_cmp_eq_u64 %[se], 0\n"                              \
    "s_cbranch_scc1 .Lk4_ne" #U "_%=\n"                    \
    "s_mov_b64 exec, %[se]\n"                              \
    "v_mbcnt_lo_u32_b32 %[e], exec_lo, 0\n"                \
    "v_mbcnt_hi_u32_b32 %[e], exec_hi, %[e]\n"             \
    "v_add_lshl_u32 %[e], %[e], %[cnt], 4\n"               \
    "v_mov_b32 v97, " CL "\n"                              \
    "v_mov_b32 v98, " CR "\n"                              \
    "global_store_dwordx4 %[e], v[96:99], %[recbase]\n"    \
    "s_bcnt1_i32_b64 %[t], %[se]\n"                        \
    "s_add_u32 %[cnt], %[cnt], %[t]\n"                     \
    "s_mov_b64 exec, %[alive]\n"                           \
    ".Lk4_ne" #U "_%=:\n"
// A step of a phase none of whose lanes holds a piece with the row-b2, last-row
// or tie rules (a corner item: lm = bitb1 = endb = 0, o0 = on = 31): avail' =
// (avail ^ b) | (amask & t1), the free set avail' & ~ld' & ~rd', and every full
// board a non-tie one.
#define NQM_K4_CSTEP(U, CX, CA, CL, CR, CT, NX, NA, NL, NR, AM) \
    "v_sub_u32 %[ng], 0, " CX "\n"                         \
    "v_and_b32 %[b], " CX ", %[ng]\n"                      \
    "v_or_b32 " NR ", " CR ", %[b]\n"                      \
    "v_add_lshl_u32 " NL ", %[b], " CL ", 1\n"             \
    "v_cmp_eq_u32 %[sm], " CA ", %[b]\n"                   \
    "v_cmpx_ne_u32 vcc, " CX ", %[b]\n"                    \
    "v_xor_b32 " CX ", " CX ", %[b]\n"                     \
    "ds_write_b128 %[q], " CT " offset:%[row]\n"           \
    "v_add_u32 %[q], %[row], %[q]\n"                       \
    "s_mov_b64 exec, %[alive]\n"                           \
    "v_lshrrev_b32 " NR ", 1, " NR "\n"                    \
    "v_xor_b32 " NA ", " CA ", %[b]\n"                     \
    AM(NA, NL)                                             \
    "v_bitop3_b32 " NX ", " NA ", " NL ", " NR " bitop3:0x10\n" \
    "s_bcnt1_i32_b64 %[t], %[sm]\n"                        \
    "s_add_u32 %[sols], %[sols], %[t]\n"
#define NQM_K4_PASS_A(U, ST, AM) NQM_K4_POP(U, "v88", "v89", "v90", "v91", "v[88:91]") ST(U, "v88", "v89", "v90", "v91", "v[88:91]", "v92", "v93", "v94", "v95", AM)
#define NQM_K4_PASS_B(U, ST, AM) NQM_K4_POP(U, "v92", "v93", "v94", "v95", "v[92:95]") ST(U, "v92", "v93", "v94", "v95", "v[92:95]", "v88", "v89", "v90", "v91", AM)
#define NQM_K4_PASS8(ST, AM) NQM_K4_PASS_A(0, ST, AM) NQM_K4_PASS_B(1, ST, AM) NQM_K4_PASS_A(2, ST, AM) NQM_K4_PASS_B(3, ST, AM) \
                             NQM_K4_PASS_A(4, ST, AM) NQM_K4_PASS_B(5, ST, AM) NQM_K4_PASS_A(6, ST, AM) NQM_K4_PASS_B(7, ST, AM)

typedef uint32_t k4_v4 __attribute__((ext_vector_type(4)));

typedef uint32_t k4_v8 __attribute__((ext_vector_type(8)));

// The rule constants of a lane's piece (see load_piece in k4_drain_kernel):
// o0 / on = the bits of avail that are the side columns (31: none).
struct K4Rules {
    uint32_t lm, bitb1, endb, amask, o0, on, item;
};

// Runs passes until at most floor_live lanes hold a piece, at most nfloor
// lanes (signed: -1 never) still hold an unstarted next item, lim32 passes
// ran or the chunk holds kRecLimit records (checked every 8 passes). A lane
// in hasnext whose piece ends starts its next item (nst, nr) in the same
// pass, and its bit leaves hasnext. Returns the non-tie full boards (psols)
// and the placements (pnodes) of the wave.
#define NQM_K4_PHASE_ASM(ST, AM) \
    asm volatile( \
        "s_mov_b32 %[sols], 0\n" \
        "s_mov_b32 %[nodes], 0\n" \
        "s_mov_b64 %[live], exec\n" \
        "s_mov_b64 %[alive], exec\n" \
        "s_mov_b32 %[ran], 0\n" \
        ".Lk4_pass_%=:\n" NQM_K4_PASS8(ST, AM) \
        "s_add_u32 %[ran], %[ran], 8\n" \
        "s_bcnt1_i32_b64 %[t], exec\n" \
        "s_cmp_le_u32 %[t], %[floor]\n" \
        "s_cbranch_scc1 .Lk4_end_%=\n" \
        "s_bcnt1_i32_b64 %[t], %[hasnext]\n" \
        "s_cmp_le_i32 %[t], %[nfloor]\n" \
        "s_cbranch_scc1 .Lk4_end_%=\n" \
        "s_cmp_ge_u32 %[cnt], %[clim]\n" \
        "s_cbranch_scc1 .Lk4_end_%=\n" \
        "s_cmp_lt_u32 %[ran], %[lim]\n" \
        "s_cbranch_scc1 .Lk4_pass_%=\n" \
        ".Lk4_end_%=:\n" \
        "s_bcnt1_i32_b64 %[t], exec\n" \
        "s_mul_i32 %[t], %[t], %[ran]\n" \
        "s_add_u32 %[nodes], %[nodes], %[t]\n" \
        "s_mov_b64 exec, %[live]\n" \
        "s_waitcnt lgkmcnt(0)\n" \
        : "+{v[88:91]}"(s4), "={v[92:95]}"(s5), "+{v[96:99]}"(rec), [q] "+v"(q), [b] "=&v"(b), [ng] "=&v"(ng), \
          [e] "=&v"(e), [t1] "=&v"(t1), [t2v] "=&v"(t2v), [m0] "=&v"(m0), [mt] "=&v"(mt), [live] "=&s"(live), \
          [alive] "=&s"(alive), [sm] "=&s"(sm), [se] "=&s"(se), [t] "=&s"(t), [t2] "=&s"(t2), [ran] "=&s"(ran), \
          [sols] "=&s"(sols), [nodes] "=&s"(nodes), [cnt] "+s"(c), [lm] "+v"(lm), [bitb1] "+v"(bitb1), \
          [endb] "+v"(endb), [amask] "+v"(amask), [o0] "+v"(o0), [on] "+v"(on), [hasnext] "+s"(hasnext) \
        : [row] "i"(kK4Row), [floor] "s"(floor_live), [lim] "s"(lim32), [clim] "i"(kRecLimit), [sn] "s"(sn), \
          [sside] "s"(sside), [recbase] "s"(recbase), [nfloor] "s"(nfloor), \
          "{v[100:103]}"(n4), "{v[104:111]}"(n8) \
        : "vcc", "scc", "memory"); \

template <bool kAmask, bool kCorner>
__device__ __forceinline__ void k4_phase(uint4& st, uint32_t& q, K4Rules& rl, const uint4& nst, const K4Rules& nr,
                                         uint64_t& hasnext, int32_t nfloor, uint32_t& cnt,
                                         uint64_t recbase, uint32_t sn, uint32_t sside, uint32_t floor_live,
                                         uint32_t lim32,
                                         uint32_t& psols, uint32_t& pnodes, uint32_t& ran) {
    k4_v4 s4 = {st.x, st.y, st.z, st.w};
    k4_v4 s5;
    k4_v4 rec = {rl.item, 0u, 0u, 0u};
    const k4_v4 n4 = {nst.x, nst.y, nst.z, nst.w};
    const k4_v8 n8 = {nr.lm, nr.bitb1, nr.endb, nr.amask, nr.o0, nr.on, nr.item, 0u};
    uint32_t lm = rl.lm, bitb1 = rl.bitb1, endb = rl.endb, amask = rl.amask, o0 = rl.o0, on = rl.on;
    uint64_t live, alive, sm, se;
    uint32_t b, ng, e, t1, t2v, m0, mt, t, t2;
    uint32_t sols = 0, nodes = 0;
    uint32_t c = cnt;
    if constexpr (kCorner && kAmask) {
        NQM_K4_PHASE_ASM(NQM_K4_CSTEP, NQM_K4_CAM_ON)
    } else if constexpr (kCorner) {
        NQM_K4_PHASE_ASM(NQM_K4_CSTEP, NQM_K4_AM_OFF)
    } else if constexpr (kAmask) {
        NQM_K4_PHASE_ASM(NQM_K4_STEP, NQM_K4_AM_ON)
    } else {
        NQM_K4_PHASE_ASM(NQM_K4_STEP, NQM_K4_AM_OFF)
    }
    (void)s5;
    rl = K4Rules{lm, bitb1, endb, amask, o0, on, rec.x};
    st = make_uint4(s4.x, s4.y, s4.z, s4.w);
    cnt = c;
    psols = sols;
    pnodes = nodes;
}

__global__ void __launch_bounds__(kK4Threads) k4_drain_kernel(K4Args a) {
    extern __shared__ uint4 lds[];
    constexpr uint32_t T = kK4Threads, row = kK4Row;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t base = 16u * tid;
    uint32_t* const wscratch = reinterpret_cast<uint32_t*>(lds) + wave * 64u;  // in row 0
    if ((uint32_t)(size_t)(k4_lds_u32*)lds != 0u) {  // the slots are addressed from 0
        if (tid == 0) atomicOr(a.err, 16u);
        return;
    }
    const uint32_t count = a.count;
    const uint32_t waves = gridDim.x * (T / 64u);

    uint4 st = make_uint4(0u, 0u, 0u, 0u);
    uint32_t q = base;
    K4Rules rl{};  // the rule constants of the lane's piece
    auto holds = [&]() { return st.x != 0u || q != base; };
    // Next item: a lane that holds a piece also holds the queue item it
    // starts when the piece ends (inside the search phase: hasnext).
    uint4 nst = make_uint4(0u, 0u, 0u, 0u);
    K4Rules nrl{};
    bool holdnext = false;
    uint64_t hasnext = 0;  // wave-uniform

    // the queue: nseg segments with a head each (as in K2's drain)
    const uint32_t nseg = a.nseg;
    uint32_t seg = blockIdx.x % nseg, seg_done = 0;
    auto seg_lo = [&](uint32_t sg) { return (uint32_t)((uint64_t)count * sg / nseg); };
    const uint32_t seg_waves = max(1u, waves / nseg);
    bool queue_open = true;
    uint32_t lq = 0, lq_end = 0, rem = seg_lo(seg + 1) - seg_lo(seg);

    // the wave's record chunk (taken before its first phase)
    constexpr uint32_t kNoChunk = 0xffffffffu;
    uint32_t chunk = kNoChunk, cnt = 0;
    auto grab_chunk = [&]() {
        uint32_t ch = 0;
        if (lane == 0) ch = atomicAdd(a.chunk_head, 1u);
        ch = __builtin_amdgcn_readfirstlane(__shfl(ch, 0, 64));
        if (ch >= a.nchunks) {  // out of chunks: the sink chunk (the host retries with more)
            if (lane == 0) atomicOr(a.err, 64u);
            ch = a.nchunks;
        }
        chunk = ch;
        cnt = 0;
    };
    auto close_chunk = [&]() {
        if (lane == 0) {
            a.rec_counts[chunk] = cnt;
            a.chunk_off[chunk] = cnt ? atomicAdd(a.rec_total, cnt) : 0u;
        }
    };

    // the state and rule constants of queue item idx
    auto load_piece = [&](uint32_t idx, uint4& ps, K4Rules& pr) {
        const SymQ it = a.items[idx];
        const uint32_t r = symq_row(it.meta), b1 = symq_b1(it.meta);
        uint32_t ld = it.ld, rd = it.rd, avail = ~it.cols & a.mask;
        pr = K4Rules{0u, 0u, 0u, 0u, 31u, 31u, idx};
        if (symq_corner(it.meta)) {
            if (b1 > r) {  // rows below b1 avoid column 1: out of avail until the child at row b1
                avail &= ~2u;
                ld |= 1u << (31u - (b1 - r));
                pr.amask = 2u;
            }
        } else {
            pr.o0 = 0u;
            pr.on = a.n - 1u;
            const SymCase c = sym_case(a.n, b1, 0u);
            if (c.b2 > r) ld |= 1u << (31u - (c.b2 - r));  // reaches bit 31 at child row b2
            rd |= 1u << (a.n + (a.last - r));              // reaches bit n at the last row
            pr.lm = c.lastmask;
            pr.bitb1 = 1u << b1;
            pr.endb = (b1 < r && !symq_top_b1(it.meta)) ? c.endbit : 0xffffffffu;
        }
        ps = make_uint4(it.f, avail, ld, rd);
    };
    // the queue items for the lanes of `want` (wave-collective): the held
    // range first, then a new grab; returns the lane's item or kNoItem
    constexpr uint32_t kNoItem = 0xffffffffu;
    auto take_queue = [&](uint64_t want) -> uint32_t {
        const uint32_t nw = __popcll(want);
        const uint32_t take_local = min(nw, lq_end - lq);
        uint32_t g = 0, got = 0;
        if (take_local < nw && queue_open) {
            const uint32_t need = nw - take_local;
            uint32_t wn = rem / (2u * seg_waves);
            wn = wn < need ? need : (wn > kK4Grab ? kK4Grab : wn);
            const uint32_t lo = seg_lo(seg), len = seg_lo(seg + 1) - lo;
            if (lane == 0) g = atomicAdd(a.head + seg * kHeadStride, wn);
            g = __shfl(g, 0, 64);
            got = g >= len ? 0u : min(wn, len - g);
            rem = g >= len ? 0u : len - g - got;
            g += lo;
            if (got < wn) {
                if (++seg_done == nseg) queue_open = false;
                seg = seg + 1u == nseg ? 0u : seg + 1u;
                rem = seg_lo(seg + 1) - seg_lo(seg);
            }
        }
        const uint32_t r = k4_rank(want);
        uint32_t idx = kNoItem;
        if ((want >> lane) & 1u) {
            if (r < take_local)
                idx = lq + r;
            else if (r - take_local < got)
                idx = g + (r - take_local);
        }
        lq += take_local;
        if (got) {
            lq = g + min(got, nw - take_local);
            lq_end = g + got;
        }
        return idx;
    };

    uint64_t leaves = 0, exec_nodes = 0;
    for (;;) {
        // a lane whose piece ended outside the phase (its last child was
        // dead and its stack empty when the phase stopped) starts its next
        // item here
        if (!holds() && holdnext) {
            st = nst;
            rl = nrl;
            q = base;
            holdnext = false;
        }
        uint64_t idle = __ballot(!holds());
        if (idle) {
            if (queue_open || lq < lq_end) {
                const uint32_t idx = take_queue(idle);
                if (idx != kNoItem) {
                    load_piece(idx, st, rl);
                    q = base;
                }
                idle = __ballot(!holds());
            }
            if (idle) {
                // donation (as in K2): a lane with an open row gives its
                // bottom one, with its item and rule constants
                const bool is_donor = q != base;
                const uint64_t donors = __ballot(is_donor);
                const uint32_t k = min((uint32_t)__popcll(idle), (uint32_t)__popcll(donors));
                if (k) {
                    const uint32_t dr = k4_rank(donors);
                    const bool gives = is_donor && dr < k;
                    if (gives) wscratch[dr] = lane;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t ir = k4_rank(idle);
                    const bool recv = !holds() && ir < k;
                    const uint32_t dl = recv ? wscratch[ir] : lane;
                    const K4Rules d{(uint32_t)__shfl(rl.lm, dl, 64),    (uint32_t)__shfl(rl.bitb1, dl, 64),
                                    (uint32_t)__shfl(rl.endb, dl, 64),  (uint32_t)__shfl(rl.amask, dl, 64),
                                    (uint32_t)__shfl(rl.o0, dl, 64),    (uint32_t)__shfl(rl.on, dl, 64),
                                    (uint32_t)__shfl(rl.item, dl, 64)};
                    if (recv) {
                        st = k4_slot_load(base - 16u * lane + 16u * dl + row);  // the donor's slot 0
                        rl = d;
                        q = base;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (gives) {
                        for (uint32_t s = base + 2u * row; s <= q; s += row) k4_slot_store(s - row, k4_slot_load(s));
                        q -= row;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        // lanes with a piece and no next item take one
        if (queue_open || lq < lq_end) {
            const uint64_t want = __ballot(holds() && !holdnext);
            if (want) {
                const uint32_t idx = take_queue(want);
                if (idx != kNoItem) {
                    load_piece(idx, nst, nrl);
                    holdnext = true;
                }
            }
        }
        // the phase leaves out the corner rule's instruction when no lane's
        // piece has that rule, and the other rules when every piece is a
        // corner item's; a next item that needs what the phase left out
        // starts after the phase
        const bool amask_wave = __ballot(holds() && rl.amask != 0u) != 0;
        const bool corner_wave = __ballot(holds() && rl.o0 == 0u) == 0;
        hasnext = __ballot(holdnext && (amask_wave || nrl.amask == 0u) && (!corner_wave || nrl.o0 != 0u));

        const uint64_t live_mask = __ballot(holds());
        const bool have_items = queue_open || lq < lq_end;
        if (live_mask == 0 && !have_items) break;

        const uint32_t nlive = __popcll(live_mask);
        const bool can_refill = have_items || __ballot(q != base) != 0;
        uint32_t floor_live = 0, limit = 1u << 25;  // 64 lanes x 2^25 passes fit the u32 counts
        if (can_refill) {
            if (nlive > 64u - a.refill_idle)
                floor_live = 64u - a.refill_idle;
            else
                limit = 256u;
        }
        floor_live = __builtin_amdgcn_readfirstlane(floor_live);
        limit = __builtin_amdgcn_readfirstlane(limit);
        // the phase also ends once next_refill lanes have started their next
        // item (while the queue can give them another)
        const int32_t nfloor = __builtin_amdgcn_readfirstlane(
            have_items ? (int32_t)__popcll(hasnext) - (int32_t)a.next_refill : -1);
        if (chunk == kNoChunk) grab_chunk();
        const uint64_t rb = (uint64_t)(size_t)(a.recs + (size_t)chunk * kRecChunk);
        // (readfirstlane returns int: widen through uint32_t, never by sign)
        const uint64_t recbase = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(rb >> 32)) << 32) |
                                 (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)rb);
        uint32_t psols = 0, pnodes = 0, ran = 0, c = __builtin_amdgcn_readfirstlane(cnt);
        if (holds()) {
            const uint64_t had = hasnext;
            if (corner_wave && amask_wave)
                k4_phase<true, true>(st, q, rl, nst, nrl, hasnext, nfloor, c, recbase, a.n, a.side, floor_live,
                                     limit, psols, pnodes, ran);
            else if (corner_wave)
                k4_phase<false, true>(st, q, rl, nst, nrl, hasnext, nfloor, c, recbase, a.n, a.side, floor_live,
                                      limit, psols, pnodes, ran);
            else if (amask_wave)
                k4_phase<true, false>(st, q, rl, nst, nrl, hasnext, nfloor, c, recbase, a.n, a.side, floor_live,
                                      limit, psols, pnodes, ran);
            else
                k4_phase<false, false>(st, q, rl, nst, nrl, hasnext, nfloor, c, recbase, a.n, a.side, floor_live,
                                       limit, psols, pnodes, ran);
            // the phase cleared the bits of the lanes that started their next item
            if (((had & ~hasnext) >> lane) & 1u) holdnext = false;
            if ((int32_t)q < 0) {  // the piece ended: its last load is void
                q = base;
                st.x = 0;
            }
        }
        c = k4_wave_max(c);
        // the wave's counts come from scalar registers: one lane keeps them
        const uint32_t lead = (uint32_t)__ffsll((long long)live_mask) - 1u;
        psols = __shfl(psols, lead, 64);
        pnodes = __shfl(pnodes, lead, 64);
        leaves += psols;
        exec_nodes += pnodes;
        cnt = c;
        if (cnt >= kRecLimit) {
            close_chunk();
            grab_chunk();
        }
    }
    if (chunk != kNoChunk) close_chunk();
    if (lane == 0) {
        uint64_t* const r = a.wave_part + 2u * (blockIdx.x * (T / 64u) + wave);
        r[0] = leaves;
        r[1] = exec_nodes;
    }
}

// ---------------------------------------------------------------------------
// ties: search the recorded boards again, then run the rotation check
// ---------------------------------------------------------------------------
// A record {item, ld, rd} stands for the boards of the item's subtree whose
// last row has that ld and rd and which the drain records; each such board
// wrote the same record, so a record whose search finds m boards is one of m equal
// records. Instead of removing copies, every copy counts its m boards and
// the sums of records with m boards are divided by m at the end (exact: each
// group adds m equal shares).
constexpr uint32_t kMaxMult = 16;  // boards behind one record (more is an error)
constexpr uint64_t kMultLcm = 720720;  // lcm(1..16): every record adds lcm / m per board, summed once

constexpr uint32_t kBoardChunk = 1024;  // boards per chunk of a wave

// Dense records: chunk ch's records go to [chunk_off[ch], + rec_counts[ch]),
// so the search below takes no empty slot (a wave's last chunk is part full).
// One workgroup per chunk; those past the chunks handed out return. The list
// holds cap records (more means the count changed: the host fails it).
__global__ void __launch_bounds__(256) k4_rec_compact_kernel(const uint4* __restrict__ recs,
                                                            const uint32_t* __restrict__ counts,
                                                            const uint32_t* __restrict__ offs,
                                                            const uint32_t* __restrict__ chunk_head,
                                                            uint32_t nchunks, uint4* __restrict__ dense,
                                                            uint32_t cap) {
    const uint32_t ch = blockIdx.x;
    if (ch >= min(*chunk_head, nchunks)) return;
    const uint32_t cnt = counts[ch], off = offs[ch];
    for (uint32_t i = threadIdx.x; i < cnt; i += 256u)
        if (off + i < cap) dense[off + i] = recs[(size_t)ch * kRecChunk + i];
}

struct K4TieArgs {
    const uint4* recs;         // the dense records
    uint32_t nslots;           // records
    const SymQ* items;
    uint32_t n;
    uint4* boards;             // chunks of {item, slot, path lo, path hi}; chunk nbchunks is the overflow sink
    uint32_t* board_counts;    // boards in each chunk
    uint32_t* bchunk_head;     // chunks handed out
    uint32_t nbchunks;
    uint32_t* mult;            // per record: boards found
    uint32_t* err;
};

constexpr uint32_t kTieThreads = 256;
#ifndef NQM_TIE_REFILL
#define NQM_TIE_REFILL 16
#endif
constexpr uint32_t kTieRefill = NQM_TIE_REFILL;  // idle lanes that make a wave refill
#ifndef NQM_TIE_LIM
#define NQM_TIE_LIM 64
#endif
constexpr uint32_t kTieLim = NQM_TIE_LIM;  // passes of a phase at most
// LDS rows of the tie search: row 0 = the target of an empty stack's pop;
// entry i of lane t at 16 t + 16 T (i + 1)
constexpr uint32_t kTieRow = 16u * kTieThreads;

// The record constants of a lane's tie search (TieSearch's, in the form the
// pass below reads them): cols0m = the prefix's columns and every bit >= n.
struct TieConst {
    uint32_t L, R, cols0m, ld0, rd0, k0, jb2, side, bad, endb, tiel, rdk0;
};

// TieSearch::pass in asm (it manages the execution mask itself), two passes
// per loop, until at most `floor` lanes are still searching or `lim` passes
// ran. Lane registers: v[64:67] = {f, U, UL, UR} of the current row (f: its
// candidates left), j = row - r0, q = the stack top's address, ub2 / cb2,
// v[72:73] = the columns of the rows placed (row r0 + i at bit 5 i: rows are
// placed from the last but one up, each shifted in at the bottom, and a pop
// shifts out the rows it gives back), v70 = the column placed.
// A pass:
//   pop    a lane whose row is spent takes its stack top (v_cmpx on f == 0,
//          borrow of q - row = an empty stack: the lane's search is over,
//          it leaves alive for done);
//   place  b = f & -f, f ^= b, its column shifted into the path; at row b2
//          ub2 = U (the columns of the rows below) and cb2 = b;
//   rows above r0's: push the row if candidates are left ({f | j << 24, U,
//          UL, UR}), then the next row: U |= b, UL = (UL | b) >> 1, UR =
//          (UR + b) << 1, j - 1, and its candidates (k = last - row)
//          (L >> k) & ((R << k) | (2^k - 1)) & ~(cols0m | ld0 << j | rd0 >> j)
//          & ~(U | UL | UR);
//   row r0: the board's test (the last row's column is the one left, ~t
//          with t = cols0m | U | b; its rd (UR + b) >> k0 with the prefix's
//          equals R; the last queen is outside L, R and the corner columns;
//          the side columns above row b2 are some, and both unless row b2
//          took the other; a tie); a board passing it leaves alive for
//          found with c0 = its row-r0 column and lastc = its last-row bit.
//          Row r0's column is shifted back out of the path: the row takes
//          its next candidate in place.
// bitop3 truth tables index (S0 S1 S2) as a 3-bit number, S0 the high bit:
// 0x40 = a & b & ~c, 0x10 = a & ~b & ~c, 0x0c = ~a & b, 0xba = (a & ~b) | c,
// 0xae = (~a & b) | c.
#define NQM_TIE_PASS(U)                                         \
    "v_cmpx_eq_u32 vcc, 0, v64\n"                               \
    "s_cbranch_execz .Ltp_np" #U "_%=\n"                        \
    "v_subrev_co_u32 %[q], vcc, %[row], %[q]\n"                 \
    "ds_read_b128 v[64:67], %[q] offset:%[row]\n"               \
    "s_andn2_b64 %[alive], %[alive], vcc\n"                     \
    "s_or_b64 %[done], %[done], vcc\n"                          \
    "s_waitcnt lgkmcnt(0)\n"                                    \
    "v_lshrrev_b32 %[t], 24, v64\n"                             \
    "v_and_b32 v64, 0xffffff, v64\n"                            \
    "v_sub_u32 %[t2], %[t], %[j]\n"                             \
    "v_mul_u32_u24 %[t2], 5, %[t2]\n"                           \
    "v_lshrrev_b64 v[72:73], %[t2], v[72:73]\n"                 \
    "v_mov_b32 %[j], %[t]\n"                                    \
    ".Ltp_np" #U "_%=:\n"                                       \
    "s_mov_b64 exec, %[alive]\n"                                \
    "s_cbranch_execz .Ltp_end_%=\n"                             \
    "v_sub_u32 %[t], 0, v64\n"                                  \
    "v_and_b32 %[bit], v64, %[t]\n"                             \
    "v_xor_b32 v64, v64, %[bit]\n"                              \
    "v_ffbl_b32 v70, %[bit]\n"                                  \
    "v_lshlrev_b64 v[72:73], 5, v[72:73]\n"                     \
    "v_or_b32 v72, v72, v70\n"                                  \
    "v_cmp_eq_u32 vcc, %[j], %[jb2]\n"                          \
    "v_cndmask_b32 %[cb2], %[cb2], %[bit], vcc\n"               \
    "v_cndmask_b32 %[ub2], %[ub2], v65, vcc\n"                  \
    "v_cmp_ne_u32 %[desc], 0, %[j]\n"                           \
    "s_and_b64 exec, %[alive], %[desc]\n"                       \
    "s_cbranch_execz .Ltp_nd" #U "_%=\n"                        \
    "v_cmpx_ne_u32 vcc, 0, v64\n"                               \
    "s_cbranch_execz .Ltp_npu" #U "_%=\n"                       \
    "v_lshl_or_b32 v64, %[j], 24, v64\n"                        \
    "ds_write_b128 %[q], v[64:67] offset:%[row]\n"              \
    "v_add_u32 %[q], %[row], %[q]\n"                            \
    ".Ltp_npu" #U "_%=:\n"                                      \
    "s_and_b64 exec, %[alive], %[desc]\n"                       \
    "v_or_b32 v65, v65, %[bit]\n"                               \
    "v_or_b32 %[t], v66, %[bit]\n"                              \
    "v_lshrrev_b32 v66, 1, %[t]\n"                              \
    "v_add_lshl_u32 v67, v67, %[bit], 1\n"                      \
    "v_add_u32 %[j], -1, %[j]\n"                                \
    "v_sub_u32 %[k], %[k0], %[j]\n"                             \
    "v_lshrrev_b32 %[a1], %[k], %[L]\n"                         \
    "v_bfm_b32 %[t], %[k], 0\n"                                 \
    "v_lshl_or_b32 %[t], %[R], %[k], %[t]\n"                    \
    "v_lshlrev_b32 %[t2], %[j], %[ld0]\n"                       \
    "v_lshrrev_b32 %[k], %[j], %[rd0]\n"                        \
    "v_or3_b32 %[t2], %[cols0m], %[t2], %[k]\n"                 \
    "v_bitop3_b32 %[a1], %[a1], %[t], %[t2] bitop3:0x40\n"      \
    "v_or3_b32 %[t], v65, v66, v67\n"                           \
    "v_bitop3_b32 v64, %[a1], %[t], 0 bitop3:0x10\n"            \
    ".Ltp_nd" #U "_%=:\n"                                       \
    "s_andn2_b64 exec, %[alive], %[desc]\n"                     \
    "s_cbranch_execz .Ltp_nc" #U "_%=\n"                        \
    "v_or3_b32 %[t], %[cols0m], v65, %[bit]\n"                  \
    "v_add_u32 %[t2], v67, %[bit]\n"                            \
    "v_lshrrev_b32 %[t2], %[k0], %[t2]\n"                       \
    "v_or_b32 %[t2], %[rdk0], %[t2]\n"                          \
    "v_cmp_eq_u32 %[ok], %[t2], %[R]\n"                         \
    "v_bitop3_b32 %[t2], %[t], %[bad], 0 bitop3:0x0c\n"         \
    "v_cmp_eq_u32 %[c1], 0, %[t2]\n"                            \
    "s_and_b64 %[ok], %[ok], %[c1]\n"                           \
    "v_or_b32 %[t2], %[ub2], %[cb2]\n"                          \
    "v_or_b32 %[a1], v65, %[bit]\n"                             \
    "v_bitop3_b32 %[t2], %[a1], %[t2], %[cols0m] bitop3:0xba\n" \
    "v_and_b32 %[t2], %[t2], %[side]\n"                         \
    "v_cmp_ne_u32 %[c1], 0, %[t2]\n"                            \
    "s_and_b64 %[ok], %[ok], %[c1]\n"                           \
    "v_and_b32 %[a1], %[cb2], %[side]\n"                        \
    "v_or_b32 %[t2], %[t2], %[a1]\n"                            \
    "v_cmp_eq_u32 %[c1], %[t2], %[side]\n"                      \
    "s_and_b64 %[ok], %[ok], %[c1]\n"                           \
    "v_bitop3_b32 %[t2], %[t], %[endb], %[tiel] bitop3:0xae\n"  \
    "v_cmp_ne_u32 %[c1], 0, %[t2]\n"                            \
    "s_and_b64 %[ok], %[ok], %[c1]\n"                           \
    "s_and_b64 %[ok], %[ok], exec\n"                            \
    "v_lshrrev_b64 v[72:73], 5, v[72:73]\n"                     \
    "s_cmp_eq_u64 %[ok], 0\n"                                   \
    "s_cbranch_scc1 .Ltp_nc" #U "_%=\n"                         \
    "s_mov_b64 exec, %[ok]\n"                                   \
    "v_not_b32 %[lastc], %[t]\n"                                \
    "v_mov_b32 %[c0], v70\n"                                    \
    "s_andn2_b64 %[alive], %[alive], %[ok]\n"                   \
    "s_or_b64 %[found], %[found], %[ok]\n"                      \
    ".Ltp_nc" #U "_%=:\n"                                       \
    "s_mov_b64 exec, %[alive]\n"

typedef uint32_t k4_v2 __attribute__((ext_vector_type(2)));

// Runs the passes of the searching lanes (the lanes that call it). Returns
// the lanes whose search ended (done) and the lanes that found a board
// (found, with c0 and lastc).
__device__ __forceinline__ void tie_phase(k4_v4& st, uint32_t& j, uint32_t& q, uint32_t& ub2, uint32_t& cb2,
                                          uint64_t& path, const TieConst& c, uint32_t floor, uint32_t lim,
                                          uint64_t& done, uint64_t& found, uint32_t& c0, uint32_t& lastc) {
    uint32_t col = 0;
    k4_v2 pth = {(uint32_t)path, (uint32_t)(path >> 32)};
    uint64_t live, alive, desc, ok, c1, dn, fd;
    uint32_t bit, t, t2, k, a1, ran, cnt;
    asm volatile(
        "s_mov_b64 %[live], exec\n"
        "s_mov_b64 %[alive], exec\n"
        "s_mov_b64 %[done], 0\n"
        "s_mov_b64 %[found], 0\n"
        "s_mov_b32 %[ran], 0\n"
        ".Ltp_pass_%=:\n" NQM_TIE_PASS(0) NQM_TIE_PASS(1)
        "s_add_u32 %[ran], %[ran], 2\n"
        "s_bcnt1_i32_b64 %[cnt], %[alive]\n"
        "s_cmp_le_u32 %[cnt], %[floor]\n"
        "s_cbranch_scc1 .Ltp_end_%=\n"
        "s_cmp_lt_u32 %[ran], %[lim]\n"
        "s_cbranch_scc1 .Ltp_pass_%=\n"
        ".Ltp_end_%=:\n"
        "s_mov_b64 exec, %[live]\n"
        "s_waitcnt lgkmcnt(0)\n"
        : "+{v[64:67]}"(st), "+{v70}"(col), "+{v[72:73]}"(pth), [q] "+v"(q), [j] "+v"(j), [ub2] "+v"(ub2),
          [cb2] "+v"(cb2), [c0] "+v"(c0), [lastc] "+v"(lastc), [bit] "=&v"(bit), [t] "=&v"(t), [t2] "=&v"(t2),
          [k] "=&v"(k), [a1] "=&v"(a1), [live] "=&s"(live), [alive] "=&s"(alive), [desc] "=&s"(desc),
          [ok] "=&s"(ok), [c1] "=&s"(c1), [done] "=&s"(dn), [found] "=&s"(fd), [ran] "=&s"(ran),
          [cnt] "=&s"(cnt)
        : [row] "i"(kTieRow), [floor] "s"(floor), [lim] "s"(lim), [L] "v"(c.L), [R] "v"(c.R),
          [cols0m] "v"(c.cols0m), [ld0] "v"(c.ld0), [rd0] "v"(c.rd0), [k0] "v"(c.k0), [jb2] "v"(c.jb2),
          [side] "v"(c.side), [bad] "v"(c.bad), [endb] "v"(c.endb), [tiel] "v"(c.tiel), [rdk0] "v"(c.rdk0)
        : "vcc", "scc", "memory");
    (void)col;
    path = ((uint64_t)pth.y << 32) | pth.x;
    done = dn;
    found = fd;
}

// Persistent lanes: a lane takes its next record when it is done with its
// record (the records of lane g of the grid are g, g + G, g + 2 G, ...: no
// counter that every wave of the chip would update; a wave refills once
// kTieRefill of its lanes idle, so the loads of a refill are shared), and lanes of a
// wave are not held by the longest search of the wave. The search is
// TieSearch (k4_ties.h) pass for pass, in asm (tie_phase): bottom-up from the
// last row but one, each row's columns limited by the record's ld and rd,
// one queen per pass; its rows with candidates left are kept in LDS
// ([entry][thread], 36 KB per workgroup). A board found is written out
// {item, slot, path} for k4_check into the wave's board chunk (one atomic per
// chunk); the record's count of boards goes to mult[slot] when its search
// ends.
__global__ void __launch_bounds__(kTieThreads) k4_research_kernel(K4TieArgs a) {
    __shared__ uint4 tie_lds[(kTieRows + 1u) * kTieThreads];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if ((uint32_t)(size_t)(k4_lds_u32*)tie_lds != 0u) {  // the stack is addressed from 0
        if (tid == 0) atomicOr(a.err, 16u);
        return;
    }
    const uint32_t n = a.n, mask = (1u << n) - 1u, last = n - 1u;
    const uint32_t gsize = gridDim.x * kTieThreads;
    const uint32_t base = 16u * tid;
    uint32_t next = blockIdx.x * kTieThreads + tid;  // this lane's next record
    bool act = false;
    uint32_t slot = 0, item = 0, m = 0;
    k4_v4 st = {0u, 0u, 0u, 0u};
    uint32_t j = 0, q = base, ub2 = 0, cb2 = 0, c0 = 0, lastc = 0;
    uint64_t path = 0;
    TieConst c{};
    constexpr uint32_t kNoChunk = 0xffffffffu;
    uint32_t bchunk = kNoChunk, bcnt = 0;  // wave-uniform
    for (;;) {
        const uint64_t idle = __ballot(!act);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        const bool exhausted = __ballot(next < a.nslots) == 0;  // no lane has records left
        if (!exhausted && (nidle >= kTieRefill || nidle == 64u)) {
            if (!act && next < a.nslots) {
                const uint4 rc = a.recs[next];
                const SymQ it = a.items[rc.x];
                slot = next;
                item = rc.x;
                next += gsize;
                // the drain's rd carries a marker above bit n-1
                TieSearch ts;
                ts.start(n, rc.y, rc.z & mask, symq_row(it.meta), symq_b1(it.meta), symq_top_b1(it.meta) != 0u,
                         it.cols, it.ld, it.rd);
                c = TieConst{ts.L, ts.R, ts.cols0 | ~mask, ts.ld0, ts.rd0, last - ts.r0, ts.b2 - ts.r0, ts.side,
                             ts.bad, ts.endb, ts.tie_l, ts.rdk0};
                st = k4_v4{ts.f, 0u, 0u, 0u};
                j = ts.y - ts.r0;
                q = base;
                ub2 = cb2 = 0u;
                path = 0;
                m = 0;
                act = true;
            }
        }
        const uint64_t actm = __ballot(act);
        if (exhausted && actm == 0) break;
        // the phase ends once kTieRefill of the searching lanes stopped (done
        // or found), or after kTieLim passes
        const uint32_t nact = (uint32_t)__popcll(actm);
        const uint32_t floor = __builtin_amdgcn_readfirstlane(!exhausted && nact > kTieRefill ? nact - kTieRefill : 0u);
        bool dne = false, fnd = false;
        if (act) {
            uint64_t dm = 0, fm = 0;
            tie_phase(st, j, q, ub2, cb2, path, c, floor, kTieLim, dm, fm, c0, lastc);
            dne = (dm >> lane) & 1u;
            fnd = (fm >> lane) & 1u;
        }
        if (dne) {
            a.mult[slot] = m;
            act = false;
        }
        const uint64_t found = __ballot(fnd);
        if (found) {
            const uint32_t nb = (uint32_t)__popcll(found);
            if (bchunk == kNoChunk || bcnt + nb > kBoardChunk) {
                if (bchunk != kNoChunk && lane == 0) a.board_counts[bchunk] = bcnt;
                uint32_t ch = 0;
                if (lane == 0) ch = atomicAdd(a.bchunk_head, 1u);
                ch = __builtin_amdgcn_readfirstlane(__shfl(ch, 0, 64));
                if (ch >= a.nbchunks) {  // out of chunks: the sink (the count fails)
                    if (lane == 0) atomicOr(a.err, 128u);
                    ch = a.nbchunks;
                }
                bchunk = ch;
                bcnt = 0;
            }
            if (fnd) {
                // rows r0+1.. from the path, row r0's column, the last row's
                const uint64_t board = (path << 5) | c0 | ((uint64_t)__builtin_ctz(lastc) << (5u * c.k0));
                ++m;
                a.boards[(size_t)bchunk * kBoardChunk + bcnt + k4_rank(found)] =
                    make_uint4(item, slot, (uint32_t)board, (uint32_t)(board >> 32));
            }
            bcnt += nb;
        }
    }
    if (bchunk != kNoChunk && lane == 0) a.board_counts[bchunk] = bcnt;
}

#ifndef NQM_CHECK_BLOCKS
#define NQM_CHECK_BLOCKS 16
#endif
constexpr uint32_t kCheckBlocksPerCu = NQM_CHECK_BLOCKS;  // workgroups of 256 per CU

// One thread per board found: the rotation check (sym_check_packed on the
// item's prefix rows and the board's path). A board behind a record with m
// boards is counted m times (once per copy of the record), so it adds
// lcm(1..16) / m; each workgroup writes its sums (classes x lcm, weight x
// lcm) to part[block], k4_sum_kernel adds them up (no same-address atomics).
__global__ void __launch_bounds__(256)
k4_check_kernel(const uint4* __restrict__ boards, const uint32_t* __restrict__ board_counts,
                const uint32_t* __restrict__ bchunk_head, uint32_t nbchunks, const uint32_t* __restrict__ mult,
                const SymQ* __restrict__ items, uint32_t n, ulonglong2* __restrict__ part,
                uint32_t* __restrict__ err) {
    __shared__ ulonglong2 wsum[4];
    const uint32_t nch = min(*bchunk_head, nbchunks);  // past the cap: err bit 128, the count is void
    uint64_t uq = 0, tot = 0;
    for (uint32_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const uint32_t cnt = board_counts[ch];
        for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
        const uint4 e = boards[(size_t)ch * kBoardChunk + j];
        const SymQ it = items[e.x];
        const SymCase c = sym_case(n, symq_b1(it.meta), 0u);
        // rows 0..r0-1 from the prefix, rows r0.. from the path: 5 bits a row,
        // rows below 12 in lo, the others in hi
        const uint32_t sh = 5u * symq_row(it.meta);
        const uint64_t pre = ((uint64_t)it.p1 << 32) | it.p0, path = ((uint64_t)e.w << 32) | e.z;
        const uint64_t flo = pre | (path << sh), fhi = sh ? path >> (64u - sh) : 0ull;
        const uint64_t lo = flo & ((1ull << 60) - 1ull), hi = (flo >> 60) | (fhi << 4);
        const uint32_t w = sym_check_packed(c, lo, hi);
        const uint32_t m = mult[e.y];
        if (m >= 1u && m <= kMaxMult) {
            const uint64_t share = kMultLcm / m;  // exact: m divides the lcm
            uq += w != 0u ? share : 0u;
            tot += share * w;
        } else {
            atomicOr(err, 256u);
        }
        }
    }
    uq = k4_wave_sum64(uq);
    tot = k4_wave_sum64(tot);
    if ((threadIdx.x & 63u) == 0u) wsum[threadIdx.x >> 6] = make_ulonglong2(uq, tot);
    __syncthreads();
    if (threadIdx.x == 0) {
        ulonglong2 b = wsum[0];
        for (uint32_t i = 1; i < 4u; ++i) {
            b.x += wsum[i].x;
            b.y += wsum[i].y;
        }
        part[blockIdx.x] = b;
    }
}

// acc = the sum of the check's per-workgroup sums (one workgroup).
__global__ void __launch_bounds__(256) k4_sum_kernel(const ulonglong2* __restrict__ part, uint32_t count,
                                                    unsigned long long* __restrict__ acc) {
    __shared__ ulonglong2 wsum[4];
    uint64_t a = 0, b = 0;
    for (uint32_t i = threadIdx.x; i < count; i += 256u) {
        a += part[i].x;
        b += part[i].y;
    }
    a = k4_wave_sum64(a);
    b = k4_wave_sum64(b);
    if ((threadIdx.x & 63u) == 0u) wsum[threadIdx.x >> 6] = make_ulonglong2(a, b);
    __syncthreads();
    if (threadIdx.x == 0) {
        acc[0] = wsum[0].x + wsum[1].x + wsum[2].x + wsum[3].x;
        acc[1] = wsum[0].y + wsum[1].y + wsum[2].y + wsum[3].y;
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
struct K4Host {
    std::vector<SymQ> items;
    uint64_t nodes = 0;
};

void k4_host_rec(K4Host& h, const SymCase& c, uint32_t y, uint32_t l, uint32_t d, uint32_t r, uint64_t p,
                 uint32_t top_b1, uint32_t depth) {
    const uint32_t f = sym_free(c, y, l, d, r);
    if (!f) return;
    if (y == depth) {
        SymQ it;
        it.f = f;
        it.cols = d;
        it.ld = l & c.mask;
        it.rd = r;
        it.meta = symq_meta(c.n, y, c.b1, c.corner, top_b1);
        it.pad = 0;
        it.p0 = (uint32_t)p;
        it.p1 = (uint32_t)(p >> 32);
        h.items.push_back(it);
        return;
    }
    for (uint32_t g = f; g; g &= g - 1u) {
        const uint32_t bit = g & (0u - g);
        ++h.nodes;
        const uint64_t p2 = y < kPrefixRows ? p | ((uint64_t)__builtin_ctz(bit) << (5u * y)) : p;
        const uint32_t tb = top_b1 | (!c.corner && y == c.b1 && bit == c.top ? 1u : 0u);
        k4_host_rec(h, c, y + 1u, ((l | bit) << 1) & c.mask, d | bit, (r | bit) >> 1, p2, tb, depth);
    }
}

}  // namespace

// The expansion row: rows < R hold every "rows < b1" side rule of a
// non-corner case (b1 <= (n-3)/2 < R).
uint32_t k4_rows_done(uint32_t n) { return (n & 1u) ? (n - 1u) / 2u : n / 2u - 1u; }

bool k4_supports(uint32_t n) {
    // n = 21 would pass the checks below, but its tie records (about 7.5 times
    // n = 20's 1.3e9) outgrow the 32-bit record indices
    if (n < 10 || n > 20) return false;
    const uint32_t R = k4_rows_done(n), last = n - 1u;
    // the rd marker n + (last - R) fits in 32 bits; real ld bits (at most
    // n - 1 + (last - R)) stay below bit 31, and below the row-b2 marker,
    // which starts 31 - (b2 - R) bits up with b2 <= n - 2; the stack of
    // last - R rows fits LDS (16 waves per CU up to 9 rows, 12 at 10), and a
    // tie search's rows R+1 .. last-1 its kTieRows entries
    return n + (last - R) <= 31u && n - 1u + (last - R) < 31u && n - 2u < 32u - n + R && last - R <= kK4MaxRows &&
           last - R - 1u <= kTieRows;
}

// Grows b to `bytes`, keeping its first `used` bytes (stream-ordered copy).
static int grow_keep(nqm_ctx* ctx, DevBuf& b, size_t used, size_t bytes) {
    if (b.cap >= bytes) return NQM_OK;
    DevBuf nb;
    int rc = ensure_dev(ctx, nb, bytes + bytes / 2);
    if (rc != NQM_OK) return rc;
    if (used) NQM_HIP(ctx, hipMemcpyAsync(nb.ptr, b.ptr, used, hipMemcpyDeviceToDevice, ctx->stream));
    if (b.ptr) {
        NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NQM_HIP(ctx, hipFree(b.ptr));
    }
    b = nb;
    return NQM_OK;
}

int k4_count(nqm_ctx* ctx, uint32_t n, uint64_t* unique, uint64_t* total, nqm_stats* stats) {
    const double t0 = now_ms();
    if (!k4_supports(n)) return fail(ctx, NQM_EUNSUPPORTED, "the symmetric drain takes 10 <= n <= 20, not %u", n);
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    SymWork& w = ctx->sym;
    const uint32_t R = k4_rows_done(n), last = n - 1u;
    const uint32_t depth0 = std::min(R, 4u);

    // host rows: corner (rows 0 and 1 fixed), non-corner (row 0 fixed)
    K4Host h;
    try {
        for (uint32_t b1 = 2; b1 < last; ++b1) {
            const SymCase c = sym_case(n, b1, 1u);
            const uint32_t bit = 1u << b1;
            h.nodes += 2;
            k4_host_rec(h, c, 2, ((2u | bit) << 1) & c.mask, 1u | bit, bit >> 1, (uint64_t)b1 << 5, 0u, depth0);
        }
        for (uint32_t b1 = 1; b1 < last - b1; ++b1) {
            const SymCase c = sym_case(n, b1, 0u);
            const uint32_t bit = 1u << b1;
            h.nodes += 1;
            k4_host_rec(h, c, 1, (bit << 1) & c.mask, bit, bit >> 1, (uint64_t)b1, 0u, depth0);
        }
    } catch (const std::bad_alloc&) {
        return fail(ctx, NQM_ENOMEM, "host items of the symmetric drain");
    }
    // items already at their target row start the final queue, the others
    // are the first pending level
    std::vector<SymQ> hq;
    try {
        hq.reserve(h.items.size());
        for (const SymQ& it : h.items)
            if (symq_row(it.meta) >= symq_target(it.meta, R)) hq.push_back(it);
    } catch (const std::bad_alloc&) {
        return fail(ctx, NQM_ENOMEM, "host items of the symmetric drain");
    }
    const uint32_t host_fin = (uint32_t)hq.size();
    for (const SymQ& it : h.items)
        if (symq_row(it.meta) < symq_target(it.meta, R)) hq.push_back(it);
    const uint32_t host_pend = (uint32_t)hq.size() - host_fin;

    int rc;
    // counters (u32 index): [0..1] level count, [2] final cursor, [3] next
    // cursor, [8..9] u64 expansion nodes, [16] err, [64] chunk head, [65]
    // records written, [67] board chunks,
    // [128..639] queue heads, [640..] u64 tie sums by boards per record
    if ((rc = ensure_dev(ctx, w.ctr, 4096)) != NQM_OK) return rc;
    uint32_t* const d_cnt = static_cast<uint32_t*>(w.ctr.ptr);
    unsigned long long* const d_xnodes = reinterpret_cast<unsigned long long*>(d_cnt + 8);
    uint32_t* const d_err = d_cnt + 16;
    uint32_t* const d_chunk_head = d_cnt + 64;
    uint32_t* const d_rec_total = d_cnt + 65;
    uint32_t* const d_bchunk_head = d_cnt + 67;
    uint32_t* const d_head = d_cnt + 128;
    unsigned long long* const d_acc = reinterpret_cast<unsigned long long*>(d_cnt + 640);
    static_assert(128 + kHeadSegs * kHeadStride <= 640, "heads fit before the tie sums");
    static_assert(640 * 4 + 16 <= 4096, "the tie sums fit the counter page");
    if ((rc = ensure_dev(ctx, w.items, std::max<size_t>(hq.size(), 1) * sizeof(SymQ))) != NQM_OK) return rc;

    const bool known = w.sizes_n == n;
    uint64_t nfin64 = host_fin;
    if (known) {
        for (uint32_t v : w.level_fin) nfin64 += v;
        if ((rc = ensure_dev(ctx, w.fin, (size_t)std::max<uint64_t>(nfin64, 1) * sizeof(SymQ))) != NQM_OK) return rc;
        for (size_t L = 0; L < w.level_in.size(); ++L)
            if ((rc = ensure_dev(ctx, w.pend[(L + 1) & 1], (size_t)std::max<uint32_t>(
                                     L + 1 < w.level_in.size() ? w.level_in[L + 1] : 1u, 1u) * sizeof(SymQ))) != NQM_OK)
                return rc;
    }

    // the host's items go up from pinned memory (an asynchronous copy)
    if ((rc = ensure_host(ctx, w.h_items, std::max<size_t>(hq.size(), 1) * sizeof(SymQ))) != NQM_OK) return rc;
    std::copy(hq.begin(), hq.end(), static_cast<SymQ*>(w.h_items.ptr));
    NQM_HIP(ctx, hipEventRecord(ctx->ev_x0, st));
    NQM_HIP(ctx, hipMemsetAsync(d_cnt, 0, 4096, st));
    NQM_HIP(ctx, hipMemcpyAsync(w.items.ptr, w.h_items.ptr, hq.size() * sizeof(SymQ), hipMemcpyHostToDevice, st));
    // the final queue starts with the host's final items
    if (!known && (rc = grow_keep(ctx, w.fin, 0, (size_t)std::max<uint32_t>(host_fin, 1) * sizeof(SymQ))) != NQM_OK)
        return rc;
    if (host_fin)
        NQM_HIP(ctx, hipMemcpyAsync(w.fin.ptr, w.items.ptr, (size_t)host_fin * sizeof(SymQ),
                                    hipMemcpyDeviceToDevice, st));
    // levels: every pending item grows by one row; children at their target
    // row are appended to the final queue (which holds fin_at items), the
    // others form the next level
    std::vector<uint32_t> level_in, level_fin;
    uint64_t fin_at = host_fin;
    const SymQ* pend = static_cast<const SymQ*>(w.items.ptr) + host_fin;
    uint32_t npend = host_pend;
    uint32_t L = 0;
    while (npend) {
        const uint32_t blocks = (npend + 255u) / 256u;
        if ((rc = ensure_dev(ctx, w.xpart, (size_t)blocks * sizeof(uint4))) != NQM_OK) return rc;
        uint4* const d_xpart = static_cast<uint4*>(w.xpart.ptr);
        hipLaunchKernelGGL(k4_expand_count_kernel, dim3(blocks), dim3(256), 0, st, pend, npend, R, d_xpart);
        NQM_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(k4_expand_scan_kernel, dim3(1), dim3(1024), 0, st, d_xpart, blocks, d_cnt, d_xnodes);
        NQM_HIP(ctx, hipGetLastError());
        uint32_t nfin_l = 0, nnext = 0;
        if (known) {
            nfin_l = w.level_fin[L];
            nnext = L + 1 < w.level_in.size() ? w.level_in[L + 1] : 0u;
        } else {
            uint32_t hc[2] = {0, 0};
            NQM_HIP(ctx, hipMemcpyAsync(hc, d_cnt, 8, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
            nfin_l = hc[0];
            nnext = hc[1];
            if (nfin64 + nfin_l > 0x7fffffffu || nnext > 0x7fffffffu)
                return fail(ctx, NQM_ERANGE, "symmetric expansion exceeds 2^31 items");
            if ((rc = grow_keep(ctx, w.fin, nfin64 * sizeof(SymQ), (size_t)(nfin64 + nfin_l) * sizeof(SymQ))) != NQM_OK)
                return rc;
            if ((rc = ensure_dev(ctx, w.pend[(L + 1) & 1], (size_t)std::max(nnext, 1u) * sizeof(SymQ))) != NQM_OK)
                return rc;
            level_in.push_back(npend);
            level_fin.push_back(nfin_l);
        }
        SymQ* const nxt = static_cast<SymQ*>(w.pend[(L + 1) & 1].ptr);
        hipLaunchKernelGGL(k4_expand_write_kernel, dim3(blocks), dim3(256), 0, st, pend, npend, R, d_xpart,
                           static_cast<SymQ*>(w.fin.ptr), (uint32_t)fin_at, (uint32_t)(fin_at + nfin_l), nxt, nnext,
                           d_err);
        NQM_HIP(ctx, hipGetLastError());
        fin_at += nfin_l;
        if (!known) nfin64 += nfin_l;
        pend = nxt;
        npend = nnext;
        ++L;
    }
    if (!known) {
        w.sizes_n = 0;
        w.host_fin = host_fin;
        w.level_in = level_in;
        w.level_fin = level_fin;
        w.chunks = 0;
        w.nrec_known = false;
    }
    const uint32_t nfin = (uint32_t)nfin64;
    const SymQ* const items = static_cast<const SymQ*>(w.fin.ptr);

    // drain shape: 16 waves per CU when the stack fits (4 workgroups of 4 waves)
    const uint32_t stack_rows = last - R;  // rows R..last-1 push; deeper corner items need fewer
    const size_t lds = (size_t)(stack_rows + 1u) * kK4Row;
    if (lds > ctx->lds_per_block) return fail(ctx, NQM_EUNSUPPORTED, "symmetric stack does not fit LDS");
    const uint32_t per_cu = std::max<uint32_t>(1u, std::min<uint32_t>(4u, (uint32_t)(ctx->lds_per_cu / lds)));
    const uint32_t blocks = (uint32_t)ctx->prop.multiProcessorCount * per_cu;
    const uint32_t nwaves = blocks * (kK4Threads / 64u);
    if ((rc = ensure_dev(ctx, w.wave, (size_t)nwaves * 16u)) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, w.cpart, (size_t)ctx->prop.multiProcessorCount * kCheckBlocksPerCu * 16u)) != NQM_OK)
        return rc;

    // record chunks: as many as the last count of this n used, else an
    // estimate from the queue (a run that runs out sets err bit 64 and is
    // repeated with as many as its chunk grabs asked for, at least twice as
    // many). Records per queue item, measured: 0.4-1.1 for n = 16..19, 2.8
    // at n = 20; a wave's chunks are 67-87% full when it closes them.
    const uint64_t per_item = n >= 20u ? 4u : 1u;
    uint64_t nchunks =
        w.chunks ? w.chunks + w.chunks / 4u + 64u : (uint64_t)nwaves + nfin64 * per_item / (kRecChunk * 3u / 4u) + 64u;
    uint64_t leaves = 0, dnodes = 0, tie_u = 0, tie_t = 0, xnodes = 0;
    float search_ms = 0.f, all_ms = 0.f;
    for (int attempt = 0;; ++attempt) {
        if (nchunks * kRecChunk > 0xffffffffull) return fail(ctx, NQM_ERANGE, "too many tie records");
        if ((rc = ensure_dev(ctx, w.recs, (size_t)(nchunks + 1u) * kRecChunk * sizeof(uint4))) != NQM_OK) return rc;
        // counts, then offsets
        if ((rc = ensure_dev(ctx, w.rec_counts, (size_t)(nchunks + 1u) * 8u)) != NQM_OK) return rc;
        NQM_HIP(ctx, hipMemsetAsync(d_chunk_head, 0, 16, st));  // chunk head, records, slots, boards
        NQM_HIP(ctx, hipMemsetAsync(d_head, 0, kHeadSegs * kHeadStride * 4, st));
        NQM_HIP(ctx, hipMemsetAsync(d_acc, 0, 16, st));
        K4Args a;
        a.items = items;
        a.count = nfin;
        a.nseg = kHeadSegs;
        a.head = d_head;
        a.n = n;
        a.last = last;
        a.mask = (1u << n) - 1u;
        a.side = (1u << last) | 1u;
        a.refill_idle = 12u;
        a.next_refill = 36u;
        a.recs = static_cast<uint4*>(w.recs.ptr);
        a.rec_counts = static_cast<uint32_t*>(w.rec_counts.ptr);
        a.chunk_off = a.rec_counts + nchunks + 1u;
        a.chunk_head = d_chunk_head;
        a.rec_total = d_rec_total;
        a.nchunks = (uint32_t)nchunks;
        a.wave_part = static_cast<uint64_t*>(w.wave.ptr);
        a.err = d_err;
        NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
        if (nfin) {
            hipLaunchKernelGGL(k4_drain_kernel, dim3(blocks), dim3(kK4Threads), lds, st, a);
            NQM_HIP(ctx, hipGetLastError());
        }
        NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        // the tie stage's sizes: from the last count of this n (its records
        // are the same), else read back here
        const bool rec_known = w.nrec_known;
        uint32_t hc[2] = {(uint32_t)nchunks, w.nrec}, err = 0;  // chunks used (a bound when known), records
        if (!rec_known) {
            NQM_HIP(ctx, hipMemcpyAsync(hc, d_chunk_head, 8, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
            if (err & 1u) return fail(ctx, NQM_EDEVICE, "symmetric expansion overflowed its counted size");
            if (err & 16u) return fail(ctx, NQM_EDEVICE, "symmetric drain: dynamic LDS not at address 0");
            if (err & 64u) {  // the chunk head counts every grab, those past the end too
                if (attempt >= 3) return fail(ctx, NQM_ERANGE, "tie records exceed %llu chunks", (unsigned long long)nchunks);
                nchunks = std::max<uint64_t>(nchunks * 2u, (uint64_t)hc[0] + hc[0] / 8u + 64u);
                NQM_HIP(ctx, hipMemsetAsync(d_err, 0, 4, st));
                continue;
            }
        }
        const uint32_t used = hc[0], nrec = hc[1];
        // ties: the records made dense, every record searched again, boards checked
        const uint32_t tie_blocks = (uint32_t)ctx->prop.multiProcessorCount * 4u;  // LDS: 4 workgroups per CU
        // boards: about one per record; every research wave holds a chunk
        const uint32_t nbchunks = 2u * (nrec / kBoardChunk) + tie_blocks * (kTieThreads / 64u) + 64u;
        if ((rc = ensure_dev(ctx, w.dense, (size_t)std::max<uint32_t>(nrec, 1u) * sizeof(uint4))) != NQM_OK) return rc;
        if ((rc = ensure_dev(ctx, w.mult, (size_t)std::max<uint32_t>(nrec, 1u) * 4u)) != NQM_OK) return rc;
        if ((rc = ensure_dev(ctx, w.boards, ((size_t)nbchunks + 1u) * kBoardChunk * sizeof(uint4) +
                                                ((size_t)nbchunks + 1u) * 4u)) != NQM_OK)
            return rc;
        uint4* const d_boards = static_cast<uint4*>(w.boards.ptr);
        uint32_t* const d_board_counts = reinterpret_cast<uint32_t*>(d_boards + ((size_t)nbchunks + 1u) * kBoardChunk);
        if (nrec) {
            const uint32_t* const d_counts = static_cast<const uint32_t*>(w.rec_counts.ptr);
            hipLaunchKernelGGL(k4_rec_compact_kernel, dim3(used), dim3(256), 0, st, static_cast<const uint4*>(w.recs.ptr),
                               d_counts, d_counts + nchunks + 1u, d_chunk_head, (uint32_t)nchunks,
                               static_cast<uint4*>(w.dense.ptr), nrec);
            NQM_HIP(ctx, hipGetLastError());
            K4TieArgs ta;
            ta.recs = static_cast<const uint4*>(w.dense.ptr);
            ta.nslots = nrec;
            ta.items = items;
            ta.n = n;
            ta.boards = d_boards;
            ta.board_counts = d_board_counts;
            ta.bchunk_head = d_bchunk_head;
            ta.nbchunks = nbchunks;
            ta.mult = static_cast<uint32_t*>(w.mult.ptr);
            ta.err = d_err;
            hipLaunchKernelGGL(k4_research_kernel, dim3(tie_blocks), dim3(kTieThreads), 0, st, ta);
            NQM_HIP(ctx, hipGetLastError());
            const uint32_t check_blocks = (uint32_t)ctx->prop.multiProcessorCount * kCheckBlocksPerCu;
            ulonglong2* const d_cpart = static_cast<ulonglong2*>(w.cpart.ptr);
            hipLaunchKernelGGL(k4_check_kernel, dim3(check_blocks), dim3(256), 0, st, d_boards, d_board_counts,
                               d_bchunk_head, nbchunks, static_cast<const uint32_t*>(w.mult.ptr), items, n, d_cpart,
                               d_err);
            NQM_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k4_sum_kernel, dim3(1), dim3(256), 0, st, d_cpart, check_blocks, d_acc);
            NQM_HIP(ctx, hipGetLastError());
        }
        NQM_HIP(ctx, hipEventRecord(ctx->ev_x1, st));
        std::vector<uint64_t> wp((size_t)nwaves * 2u);
        uint64_t hacc[2] = {};
        unsigned long long hx = 0;
        uint32_t err2 = 0, hc2[2] = {0, 0};
        NQM_HIP(ctx, hipMemcpyAsync(wp.data(), w.wave.ptr, wp.size() * 8u, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(hacc, d_acc, sizeof(hacc), hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(&hx, d_xnodes, 8, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(&err2, d_err, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(hc2, d_chunk_head, 8, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        NQM_HIP(ctx, hipEventElapsedTime(&all_ms, ctx->ev_x0, ctx->ev_x1));
        NQM_HIP(ctx, hipEventElapsedTime(&search_ms, ctx->ev_start, ctx->ev_stop));
        if (rec_known) {  // the drain's checks, made here
            if (err2 & 1u) return fail(ctx, NQM_EDEVICE, "symmetric expansion overflowed its counted size");
            if (err2 & 16u) return fail(ctx, NQM_EDEVICE, "symmetric drain: dynamic LDS not at address 0");
            if (err2 & 64u) {  // out of record chunks: the tie stage ran on partial records; again with more
                if (attempt >= 3) return fail(ctx, NQM_ERANGE, "tie records exceed %llu chunks", (unsigned long long)nchunks);
                nchunks = std::max<uint64_t>(nchunks * 2u, (uint64_t)hc2[0] + hc2[0] / 8u + 64u);
                w.nrec_known = false;
                NQM_HIP(ctx, hipMemsetAsync(d_err, 0, 4, st));
                continue;
            }
            if (hc2[1] != nrec) return fail(ctx, NQM_EDEVICE, "tie records %u, %u at the last count", hc2[1], nrec);
        }
        if (err2 & 128u) return fail(ctx, NQM_ERANGE, "tie boards exceed %u chunks", nbchunks);
        if (err2 & 256u) return fail(ctx, NQM_ERANGE, "more than %u tie boards behind one record", kMaxMult);
        // every board was counted lcm / m times m (its m records)
        if (hacc[0] % kMultLcm || hacc[1] % kMultLcm) return fail(ctx, NQM_EDEVICE, "tie sums not whole");
        const uint64_t hs[2] = {hacc[0] / kMultLcm, hacc[1] / kMultLcm};
        for (uint32_t wv = 0; wv < nwaves; ++wv) {
            leaves += wp[2u * wv];
            dnodes += wp[2u * wv + 1u];
        }
        tie_u = hs[0];
        tie_t = hs[1];
        xnodes = hx;
        w.chunks = hc2[0];
        w.nrec = hc2[1];
        w.nrec_known = true;
        break;
    }
    w.sizes_n = n;  // the sizes (and the chunk count) of this n are now known
    const uint64_t u = leaves + tie_u, t = 8u * leaves + tie_t;
    if (unique) *unique = u;
    if (total) *total = t;
    if (stats) {
        stats->total = t;
        stats->nodes = h.nodes + xnodes + dnodes;
        stats->kernel_ms = all_ms;
        stats->search_ms = search_ms;
        stats->wall_ms = now_ms() - t0;
        stats->launches = 3u * L + 5u;
        stats->blocks = blocks;
        stats->rounds = 1;
        stats->lanes = (uint64_t)blocks * kK4Threads;
    }
    return NQM_OK;
}

}  // namespace nqm
