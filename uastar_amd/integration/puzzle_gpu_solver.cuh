// puzzle_gpu_solver.cuh -- uAstar's gpusolver::GPUPuzzleSolver<N> (src/puzzle/GPU-solver.cuh:
// 55-294) over libuastar's C ABI, for N = 3, 4 and 5. In a uAstar build this file's
// content replaces the body of src/puzzle/GPU-solver.cuh (which puzzle.cu includes);
// link libuastar.so.
//
//   initialize()  -> ua_create on vm_options["ordinal"] (GPU-solver.cuh:142); the pattern
//                    databases come from uAstar's cache files database_<n>_<tiles>.bin in
//                    the working directory when all are present and intact
//                    (PatternDatabase::fetchDatabase, database.cpp:86-122), uploaded with
//                    ua_puzzle_load_pdb; otherwise they are built on the device
//                    (ua_puzzle_build_pdb) and written to those files for later runs
//   solve()       -> ua_puzzle_solve (GPU-solver.cuh:197-274); an unsolvable board
//                    returns false at once (the reference loops forever on one)
//   getSolution() -> ua_puzzle_moves (GPU-solver.cuh:276-287): blank moves 0..3
//
// Every status is checked; errors throw std::runtime_error with ua_last_error's message.
#pragma once
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "puzzle/puzzle.cuh"
#include "uastar.h"

namespace gpusolver {

// The additive disjoint partitions of puzzle/CPU-solver.hpp:44-104 (their order fixes the
// tables' layout and the cache file names).
inline std::vector<std::vector<int>> uastar_partitions(int n) {
    if (n == 3) return {{1, 2, 3, 4, 5, 6, 7, 8}};
    if (n == 4) return {{1, 2, 3, 4, 5, 6, 9, 13}, {7, 8, 10, 11, 12, 14, 15}};
    return {{3, 4, 5, 10, 15, 20}, {2, 1, 6, 11, 16, 21}, {7, 8, 9, 12, 17, 22}, {13, 14, 18, 19, 23, 24}};
}

template <int N>
class GPUPuzzleSolver {
public:
    GPUPuzzleSolver(Puzzle *puzzle) : p(puzzle) {}
    ~GPUPuzzleSolver() {
        if (ctx) ua_destroy(ctx);
    }

    void initialize() {
        check(ua_create(vm_options["ordinal"].as<int>(), nullptr, &ctx), "ua_create");
        std::vector<std::string> files;
        std::vector<size_t> sizes;
        size_t total = 0;
        for (const auto &t : uastar_partitions(N)) {
            std::string f = "database_" + std::to_string(N);
            for (int v : t) f += "_" + std::to_string(v);
            size_t s = 1;
            for (size_t i = 0; i < t.size(); ++i) s *= (size_t)(N * N) - i;
            files.push_back(f + ".bin");
            sizes.push_back(s);
            total += s;
        }
        std::vector<uint8_t> pdb(total);
        bool cached = true;
        for (size_t i = 0, off = 0; i < files.size(); off += sizes[i++]) {
            FILE *f = std::fopen(files[i].c_str(), "rb");
            cached = cached && f && std::fread(pdb.data() + off, 1, sizes[i], f) == sizes[i];
            if (f) std::fclose(f);
        }
        if (cached) {
            check(ua_puzzle_load_pdb(ctx, N, pdb.data(), total), "ua_puzzle_load_pdb");
            return;
        }
        std::printf("\tGenerating pattern databse\n");  // the message of database.cpp:106
        check(ua_puzzle_build_pdb(ctx, N), "ua_puzzle_build_pdb");
        check(ua_puzzle_download_pdb(ctx, N, pdb.data(), total), "ua_puzzle_download_pdb");
        for (size_t i = 0, off = 0; i < files.size(); off += sizes[i++]) {
            FILE *f = std::fopen(files[i].c_str(), "wb");
            const bool ok = f && std::fwrite(pdb.data() + off, 1, sizes[i], f) == sizes[i];
            if (f) std::fclose(f);
            if (!ok) std::printf("Cannot write to file %s\n", files[i].c_str());
        }
    }

    bool solve() {
        std::vector<uint8_t> s;
        p->initialState(s);  // row-major, 0 = blank
        const int st = ua_puzzle_solve(ctx, N, s.data(), &res);
        check(st, "ua_puzzle_solve");
        if (st == UA_OK)  // the line of GPU-solver.cuh:246
            std::printf("\t\t\t Number of nodes expanded: %d\n", (int)res.expanded);
        m_optimalStep = (uint32_t)res.steps;
        return st == UA_OK;
    }

    void getSolution(int *optimal, std::vector<int> *pathList) {
        int64_t len = 0;
        check(ua_puzzle_moves(ctx, nullptr, 0, &len), "ua_puzzle_moves");
        std::vector<int32_t> dirs((size_t)len);
        check(ua_puzzle_moves(ctx, dirs.data(), len, &len), "ua_puzzle_moves");
        *optimal = (int)m_optimalStep;
        pathList->assign(dirs.begin(), dirs.begin() + len);  // blank moves in the DX / DY order
    }

private:
    void check(int st, const char *what) const {
        if (st < 0)
            throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(st) +
                                     "): " + (ctx ? ua_last_error(ctx) : "no context"));
    }
    Puzzle *p;
    ua_ctx *ctx = nullptr;
    ua_result res{};
    uint32_t m_optimalStep = 0;
};

}  // namespace gpusolver
