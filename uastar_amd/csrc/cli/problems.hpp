// problems.hpp -- the command line's problems: the same life cycle as uAstar's Problem
// (problem.hpp:6-24) -- prepare, CPU / GPU initialise, solve, output -- for grid
// pathfinding and the sliding-tile puzzle, with the GPU side on libuastar's C ABI.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include <uastar.h>

namespace cli {

struct Options {
    bool pathway = false, puzzle = false;
    int height = 0, width = 0;  // 0 = not given
    std::string input_module = "custom";
    std::string plot;           // BMP file of the grid and the GPU path (pathway only)
    int block_rate = 30;        // random module: percentage of blocked cells
    bool no_cpu = false, no_gpu = false;
    int ordinal = 0;
    uint64_t seed = 1;
    std::string pdb_dir = ".";  // where the database_*.bin caches are read / written
};

class Problem {
public:
    virtual ~Problem() = default;
    virtual std::string problemName() const = 0;
    virtual bool prepare() = 0;  // false: the input was unusable (message printed)
    virtual void cpuInitialize() = 0;
    virtual bool gpuInitialize() = 0;
    virtual void cpuSolve() = 0;
    virtual bool gpuSolve() = 0;
    virtual bool output() = 0;  // false: CPU and GPU disagree
};

std::unique_ptr<Problem> make_pathway(const Options &o);
std::unique_ptr<Problem> make_puzzle(const Options &o);

// A GPU context for the problems; nullptr (with the message printed) if none.
ua_ctx *open_gpu(const Options &o);

}  // namespace cli
