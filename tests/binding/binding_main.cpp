// binding_main.cpp -- drives uAstar's solver classes as implemented over libuastar by
// uastar_amd/integration/, the way Pathway / Puzzle call them (gpuInitialize, gpuSolve,
// output -> getSolution; src/pathway/pathway.cpp:72-110, src/puzzle/puzzle.cu:115-220).
//
//   binding_main pathway H W sx sy ex ey <graph file: H*W bytes> [ordinal]
//   binding_main puzzle N t0 t1 ... t(N*N-1) [ordinal]
//
// Prints "found 0|1", then "cost %.6f" and "path id id ..." (pathway) or "steps k" and
// "moves d d ..." (puzzle). Exit status 0 on a completed search, 2 when a call threw.
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "pathway/GPU-solver.hpp"
#include "puzzle_gpu_solver.cuh"

OptionMap vm_options;

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    const std::string kind = argv[1];
    try {
        if (kind == "pathway" && argc >= 9) {
            Pathway pw;
            pw.m_height = std::atoi(argv[2]);
            pw.m_width = std::atoi(argv[3]);
            pw.m_sx = std::atoi(argv[4]), pw.m_sy = std::atoi(argv[5]);
            pw.m_ex = std::atoi(argv[6]), pw.m_ey = std::atoi(argv[7]);
            vm_options["ordinal"].v = argc >= 10 ? std::atoi(argv[9]) : 0;
            pw.m_graph.resize((size_t)pw.m_height * pw.m_width);
            FILE *f = std::fopen(argv[8], "rb");
            if (!f || std::fread(pw.m_graph.data(), 1, pw.m_graph.size(), f) != pw.m_graph.size()) return 1;
            std::fclose(f);
            GPUPathwaySolver solver(&pw);
            solver.initialize();
            const bool ok = solver.solve();
            std::printf("found %d\n", ok ? 1 : 0);
            if (ok) {
                float cost = 0;
                vector<int> path;
                solver.getSolution(&cost, &path);
                std::printf("cost %.6f\npath", cost);
                for (int v : path) std::printf(" %d", v);
                std::printf("\n");
            }
            return 0;
        }
        if (kind == "puzzle" && argc >= 3) {
            Puzzle pz;
            pz.n = std::atoi(argv[2]);
            if (argc < 3 + pz.n * pz.n) return 1;
            for (int i = 0; i < pz.n * pz.n; ++i) pz.m_initialState.push_back((uint8_t)std::atoi(argv[3 + i]));
            vm_options["ordinal"].v = argc >= 4 + pz.n * pz.n ? std::atoi(argv[3 + pz.n * pz.n]) : 0;
            bool ok = false;
            int steps = 0;
            vector<int> moves;
            auto run = [&](auto &solver) {
                solver.initialize();
                ok = solver.solve();
                if (ok) solver.getSolution(&steps, &moves);
            };
            if (pz.n == 3) {
                gpusolver::GPUPuzzleSolver<3> s(&pz);
                run(s);
            } else if (pz.n == 4) {
                gpusolver::GPUPuzzleSolver<4> s(&pz);
                run(s);
            } else if (pz.n == 5) {
                gpusolver::GPUPuzzleSolver<5> s(&pz);
                run(s);
            } else {
                return 1;
            }
            std::printf("found %d\n", ok ? 1 : 0);
            if (ok) {
                std::printf("steps %d\nmoves", steps);
                for (int d : moves) std::printf(" %d", d);
                std::printf("\n");
            }
            return 0;
        }
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
    return 1;
}
