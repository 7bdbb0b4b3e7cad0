// emu_batch.cpp -- runs the device batch loop (ua::group_block from search.hpp) for one
// one-block team of 4, 8 or 16 waves on host threads. Built by tests/test_emu_search.py.
#include "hip_emu.hpp"

#include <vector>

#include "../../uastar_amd/csrc/search.hpp"

extern "C" int emu_batch(const ua::QueryDesc *qs, ua::QueryResult *rs, int nq, int waves, const uint8_t *pdb3,
                         const uint8_t *pdb4) {
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
            ua::group_block(qs, rs, shape, qhead, &tv, segs.data(), pdb3, pdb4);
        });
    for (auto &x : th) x.join();
    emu::block = nullptr;
    return 0;
}
