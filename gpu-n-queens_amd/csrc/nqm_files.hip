// nqm_files.hip -- .nqsf pool files and the reference's nq_state_t layout.
//
// Host-only code with no HIP dependency (it also builds with g++ for the
// host sanitizer run of the parser). The file layout and the state layout
// are the reference's (nq_utils.cu:127-259, n_queens.cuh:100-113), so pools
// move between the two programs unchanged; the code is written for this
// library: bounded reads, status codes instead of asserts, no N fixed at
// compile time.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/nqm.h"
#include "nqm_limits.h"

namespace {

constexpr uint32_t kNqsfVersion = 1;  // nq_utils.cu:13

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

bool put(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n; }
bool get(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

// Rebuild the reference's state for one prefix (nq_utils.cu:244-251 does it
// with place_queen_at): absolute diagonals bit c+r and c+n-r.
void make_state(uint32_t n, uint32_t lock, const uint8_t* pre, uint8_t* st) {
    uint64_t diag = 0, anti = 0;
    uint32_t cols = 0;
    for (uint32_t r = 0; r < lock; ++r) {
        const uint32_t c = pre[r];
        cols |= 1u << c;
        diag |= (uint64_t)(1u << c) << r;
        anti |= (uint64_t)(1u << c) << (n - r);
    }
    std::memset(st, 0, nqm_state_size(n));
    std::memcpy(st, &diag, 8);
    std::memcpy(st + 8, &anti, 8);
    std::memcpy(st + 16, &cols, 4);
    st[20] = (uint8_t)lock;
    std::memset(st + 21, 0xFF, n);
    std::memcpy(st + 21, pre, lock);
}

}  // namespace

extern "C" {

size_t nqm_state_size(uint32_t n) {
    // n_queens.cuh:100-113: diagonals_t (16) + u32 (4) + char (1) + u8[n],
    // aligned to the 8-byte alignment of its u64 members.
    if (n < 1 || n > nqm::kMaxN) return 0;
    return (size_t)((21 + n + 7) & ~7u);
}

void nqm_free(void* p) { std::free(p); }

int nqm_nqsf_write(const char* path, uint32_t n, uint32_t lock, const uint8_t* pool, uint64_t count,
                   uint64_t flags) {
    if (!path || n < 1 || n > nqm::kMaxN || lock > n || (count && lock && !pool)) return NQM_EINVAL;
    File file;
    file.f = std::fopen(path, "wb");
    if (!file.f) return NQM_EIO;
    const uint8_t en = (uint8_t)n, lk = (uint8_t)lock;
    bool ok = put(file.f, &kNqsfVersion, 4) && put(file.f, &flags, 8) && put(file.f, &en, 1) &&
              put(file.f, &lk, 1) && put(file.f, &count, 8);
    if (ok && count) {  // one segment: every prefix has `lock` rows
        ok = put(file.f, &lk, 1) && put(file.f, &count, 8) &&
             (lock == 0 || put(file.f, pool, (size_t)count * lock));
    }
    if (!ok || std::fflush(file.f) != 0) return NQM_EIO;
    return NQM_OK;
}

int nqm_nqsf_read(const char* path, uint32_t expect_n, uint64_t expect_flags, uint8_t** pool,
                  uint64_t* count, uint32_t* n_out, uint32_t* lock_out, uint64_t* flags_out) {
    if (!path || !pool || !count) return NQM_EINVAL;
    *pool = nullptr;
    *count = 0;
    File file;
    file.f = std::fopen(path, "rb");
    if (!file.f) return NQM_EIO;
    uint32_t version = 0;
    uint64_t flags = 0, total = 0;
    uint8_t en = 0, lk = 0;
    if (!get(file.f, &version, 4) || !get(file.f, &flags, 8)) return NQM_EFORMAT;
    if (version != kNqsfVersion) return NQM_EFORMAT;               // nq_utils.cu:205-207
    if (expect_flags != UINT64_MAX && flags != expect_flags) return NQM_EFLAGS;  // :208-210
    if (!get(file.f, &en, 1) || !get(file.f, &lk, 1) || !get(file.f, &total, 8)) return NQM_EFORMAT;
    if (en < 1 || en > nqm::kMaxN || lk > en) return NQM_EFORMAT;
    if (expect_n && en != expect_n) return NQM_EINVAL;  // the reference only asserts (:217)
    if (lk == 0 && total) return NQM_EFORMAT;
    // refuse sizes the file cannot hold before allocating
    long here = std::ftell(file.f);
    if (std::fseek(file.f, 0, SEEK_END) != 0) return NQM_EIO;
    const long end = std::ftell(file.f);
    if (here < 0 || end < 0 || std::fseek(file.f, here, SEEK_SET) != 0) return NQM_EIO;
    if (lk && total > (uint64_t)(end - here) / lk) return NQM_EFORMAT;
    uint8_t* out = static_cast<uint8_t*>(std::malloc(total ? (size_t)total * lk : 1));
    if (!out) return NQM_ENOMEM;
    uint64_t filled = 0;
    for (;;) {
        uint8_t row = 0;
        if (!get(file.f, &row, 1)) break;  // end of file ends the segments
        uint64_t cnt = 0;
        if (!get(file.f, &cnt, 8) || row != lk || cnt > total - filled ||
            (cnt && !get(file.f, out + filled * lk, (size_t)cnt * lk))) {
            std::free(out);
            return NQM_EFORMAT;  // a pool mixing rows is not a single-lock pool
        }
        filled += cnt;
    }
    if (filled != total) {
        std::free(out);
        return NQM_EFORMAT;
    }
    for (uint64_t i = 0; i < total * lk; ++i) {
        if (out[i] >= en) {
            std::free(out);
            return NQM_EFORMAT;
        }
    }
    *pool = out;
    *count = total;
    if (n_out) *n_out = en;
    if (lock_out) *lock_out = lk;
    if (flags_out) *flags_out = flags;
    return NQM_OK;
}

int nqm_states_to_pool(uint32_t n, const void* states, uint64_t count, uint32_t lock, uint8_t* pool_out) {
    if (n < 1 || n > nqm::kMaxN || lock > n || (count && (!states || (lock && !pool_out)))) return NQM_EINVAL;
    const size_t sz = nqm_state_size(n);
    const uint8_t* s = static_cast<const uint8_t*>(states);
    for (uint64_t i = 0; i < count; ++i) {
        const uint8_t* st = s + i * sz;
        if ((uint32_t)st[20] != lock) return NQM_EPREFIX;
        std::memcpy(pool_out + i * lock, st + 21, lock);
    }
    return NQM_OK;
}

int nqm_pool_to_states(uint32_t n, uint32_t lock, const uint8_t* pool, uint64_t count, void* states_out) {
    if (n < 1 || n > nqm::kMaxN || lock > n || (count && (!states_out || (lock && !pool)))) return NQM_EINVAL;
    const size_t sz = nqm_state_size(n);
    uint8_t* o = static_cast<uint8_t*>(states_out);
    for (uint64_t i = 0; i < count; ++i) {
        const uint8_t* pre = pool + i * lock;
        for (uint32_t r = 0; r < lock; ++r)
            if (pre[r] >= n) return NQM_EPREFIX;
        make_state(n, lock, pre, o + i * sz);
    }
    return NQM_OK;
}

}  // extern "C"
