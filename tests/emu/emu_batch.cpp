// emu_batch.cpp -- runs the device batch loop (ua::batch_block from search.hpp) for one
// workgroup on host threads. Built by tests/test_emu_search.py into tests/emu/libemu.so.
#include "hip_emu.hpp"

#include <vector>

#include "../../uastar_amd/csrc/search.hpp"

extern "C" int emu_batch(const ua::QueryDesc *qs, ua::QueryResult *rs, int nq, int waves, const uint8_t *pdb3,
                         const uint8_t *pdb4) {
    if (waves < 1 || waves > 8) return -1;
    const int segcap = waves * 64 * 8;
    std::vector<uint64_t> lds((size_t)ua::wg_lds_bytes(segcap) / 8 + 1, 0);
    uint32_t qhead = 0;
    std::barrier<> bar(waves * 64);
    std::vector<emu::Wave> ws(waves);
    emu::Block blk{&bar, ws.data(), (unsigned)(waves * 64)};
    emu::block = &blk;
    std::vector<std::thread> th;
    for (int t = 0; t < waves * 64; ++t)
        th.emplace_back([&, t] {
            emu::tid = (unsigned)t;
            ua::batch_block(qs, rs, nq, &qhead, segcap, pdb3, pdb4, lds.data());
        });
    for (auto &x : th) x.join();
    emu::block = nullptr;
    return 0;
}
