// emu_batch.cpp -- runs the device batch loop (ua::group_block from search.hpp) for one
// one-block team of 4, 8 or 16 waves on host threads, for each kernel set. Built by
// tests/test_emu_search.py.
#include "hip_emu.hpp"

#include <vector>

#include "../../uastar_amd/csrc/search.hpp"

template <int SET>
static int run_batch(const ua::QueryDesc *qs, ua::QueryResult *rs, int nq, int waves) {
    if (waves != 4 && waves != 8 && waves != 16) return -1;
    const int segcap = waves * 64 * 8;
    std::vector<uint64_t> segs((size_t)2 * segcap, 0);
    ua::TeamVars tv{};
    uint32_t qhead[ua::kClasses] = {};
    ua::LaunchShape shape{};
    shape.groups[ua::CLS_GRID] = 1;  // one one-block group takes every query
    for (int k = 0; k < ua::kClasses; ++k) shape.bpg[k] = 1;
    shape.nq[ua::CLS_GRID] = nq;
    shape.segcap[ua::CLS_GRID] = segcap;
    std::barrier<> bar(waves * 64);
    std::vector<emu::Wave> ws(waves);
    emu::Block blk{&bar, ws.data(), (unsigned)(waves * 64)};
    emu::block = &blk;
    std::vector<std::thread> th;
    for (int t = 0; t < waves * 64; ++t)
        th.emplace_back([&, t] {
            emu::tid = (unsigned)t;
            ua::group_block<SET>(qs, rs, shape, qhead, &tv, segs.data());
        });
    for (auto &x : th) x.join();
    emu::block = nullptr;
    return 0;
}

extern "C" int emu_batch(const ua::QueryDesc *qs, ua::QueryResult *rs, int nq, int waves) {
    return run_batch<ua::SET_MAIN>(qs, rs, nq, waves);
}
// grids on per-wave heaps (k_teams<., SET_GHEAP>)
extern "C" int emu_batch_gheap(const ua::QueryDesc *qs, ua::QueryResult *rs, int nq, int waves) {
    return run_batch<ua::SET_GHEAP>(qs, rs, nq, waves);
}
// the 5x5 problem set (k_teams<., SET_FIVE>)
extern "C" int emu_batch5(const ua::QueryDesc *qs, ua::QueryResult *rs, int nq, int waves) {
    return run_batch<ua::SET_FIVE>(qs, rs, nq, waves);
}

// 5x5 board helpers of the device search, for direct checks against the oracle
extern "C" uint32_t emu_rank5(uint64_t w0, uint64_t w1, int p) { return ua::rank5(w0, w1, p); }
extern "C" int emu_pos5(uint64_t w0, uint64_t w1, int t) { return ua::pos5(w0, w1, t); }
extern "C" int emu_cell5(uint64_t w0, uint64_t w1, int c) { return (int)ua::cell5(w0, w1, c); }
extern "C" void emu_move5(uint64_t *w0, uint64_t *w1, int bl, int nb) {
    uint64_t t = ua::cell5(*w0, *w1, nb);
    ua::xor5(*w0, *w1, bl, t);
    ua::xor5(*w0, *w1, nb, t);
}

// The device heap self-test (k_heap_selftest) on one emulated wave: ops[i] >= 0 inserts
// keys[k .. k + ops[i]), ops[i] < 0 pops 64 keys into out; *status = final size or -1.
extern "C" int emu_heap_selftest(const int *ops, int nops, const uint64_t *keys, uint64_t *storage, int cap,
                                 uint64_t *out, int *status) {
    std::barrier<> bar(64);
    emu::Wave w;
    emu::Block blk{&bar, &w, 64u};
    emu::block = &blk;
    std::vector<std::thread> th;
    for (int t = 0; t < 64; ++t)
        th.emplace_back([&, t] {
            emu::tid = (unsigned)t;
            ua::WaveHeap h;
            h.init(storage, cap);
            const int l = ua::lane_id();
            int k = 0, o = 0;
            bool ok = true;
            for (int i = 0; i < nops; ++i) {
                const int c = ops[i];
                if (c >= 0) {
                    uint64_t v = l < c ? keys[k + l] : ua::SENT;
                    k += c;
                    v = ua::wave_sort(v);
                    ok &= h.insert(v, c);
                } else {
                    out[(size_t)o * 64 + l] = h.pop();
                    ++o;
                }
            }
            if (l == 0) *status = ok ? h.size() : -1;
        });
    for (auto &x : th) x.join();
    emu::block = nullptr;
    return 0;
}
