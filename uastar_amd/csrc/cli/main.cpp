// main.cpp -- the uastar command line: the options and run sequence of uAstar's main
// (main.cpp:60-177) on libuastar.
//
//   uastar --pathway -H 5 -W 5 --input-module custom   < grid.txt
//   uastar --pathway -H 1000 -W 1000 --input-module random --block-rate 30 --seed 1
//   uastar --pathway -H 301 -W 101 --input-module zigzag  < gap.txt
//   uastar --puzzle -W 4 < board.txt
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "problems.hpp"

using namespace cli;

static const char *kUsage =
    "uastar -- A* search on an MI355X (GA*: many per-wavefront heaps searched in rounds).\n"
    "Two problems:\n"
    "  --pathway   shortest 8-direction path on a grid (orthogonal step 1, diagonal sqrt 2)\n"
    "  --puzzle    sliding-tile puzzle (3x3, 4x4 or 5x5) with an additive pattern database\n"
    "\n"
    "Options:\n"
    "  -h, --help                 this message\n"
    "  --pathway | --puzzle       the problem to solve\n"
    "  -H, --height N             grid height (puzzle: side)\n"
    "  -W, --width N              grid width (puzzle: side)\n"
    "  --input-module M           pathway: custom (stdin: sx sy ex ey, then H rows of W\n"
    "                             cells, 0 = obstacle), zigzag (stdin: the row gap) or\n"
    "                             random (--block-rate, --seed); puzzle: custom (stdin:\n"
    "                             n*n tiles row by row, 0 = blank)\n"
    "  -b, --block-rate R         random module: percentage of blocked cells (1-99)\n"
    "  -s, --seed S               random module: generator seed\n"
    "  -p, --plot FILE            pathway: write a BMP of the grid and the GPU path\n"
    "  -G, --no-cpu               skip the sequential CPU A*\n"
    "  -C, --no-gpu               skip the GPU search\n"
    "  -o, --ordinal N            GPU ordinal\n"
    "  --pdb-dir DIR              puzzle: directory of the database_*.bin caches (default .)\n";

[[noreturn]] static void usage(const char *why) {
    if (why) std::printf("%s\n%s\n\n", why, std::string(std::strlen(why), '=').c_str());
    std::printf("%s", kUsage);
    std::exit(1);
}

static std::string stamp(std::chrono::steady_clock::time_point t0) {
    long ms = (long)std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    char b[32];
    std::snprintf(b, sizeof b, "[%07ld]", ms);
    return b;
}

static int run(Problem &p, const Options &o) {
    auto t0 = std::chrono::steady_clock::now();
    const bool cpu = !o.no_cpu, gpu = !o.no_gpu;
    std::printf("%s Generating input data ......\n", stamp(t0).c_str());
    if (!p.prepare()) return 1;
    if (cpu) {
        std::printf("%s Initializing CPU data structure ......\n", stamp(t0).c_str());
        p.cpuInitialize();
    }
    if (gpu) {
        std::printf("%s Initializing GPU data structure ......\n", stamp(t0).c_str());
        if (!p.gpuInitialize()) return 1;
    }
    if (cpu) {
        std::printf("%s Solving the problem on a pure CPU platform ......\n", stamp(t0).c_str());
        p.cpuSolve();
    }
    if (gpu) {
        std::printf("%s Solving the problem with GPU acceleration ......\n", stamp(t0).c_str());
        if (!p.gpuSolve()) return 1;
    }
    if (cpu || gpu) {
        std::printf("%s Checking the result ......\n", stamp(t0).c_str());
        if (!p.output()) {
            std::printf("%s ERROR: Output of the CPU and GPU is not consistent!\n", stamp(t0).c_str());
            return 1;
        }
    }
    return 0;
}

int main(int argc, char **argv) {
    Options o;
    bool any = false;
    auto value = [&](int &i) -> const char * {
        if (i + 1 >= argc) usage((std::string("missing value for ") + argv[i]).c_str());
        return argv[++i];
    };
    auto number = [&](int &i) -> long long {
        const char *opt = argv[i], *v = value(i);
        char *end = nullptr;
        long long x = std::strtoll(v, &end, 10);
        if (!*v || *end) usage((std::string("not a number for ") + opt + ": " + v).c_str());
        return x;
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        any = true;
        if (a == "-h" || a == "--help") usage(nullptr);
        else if (a == "--pathway") o.pathway = true;
        else if (a == "--puzzle") o.puzzle = true;
        else if (a == "-H" || a == "--height") o.height = (int)number(i);
        else if (a == "-W" || a == "--width") o.width = (int)number(i);
        else if (a == "--input-module") o.input_module = value(i);
        else if (a == "-p" || a == "--plot") o.plot = value(i);
        else if (a == "-b" || a == "--block-rate") o.block_rate = (int)number(i);
        else if (a == "-s" || a == "--seed") o.seed = (uint64_t)number(i);
        else if (a == "-G" || a == "--no-cpu") o.no_cpu = true;
        else if (a == "-C" || a == "--no-gpu") o.no_gpu = true;
        else if (a == "-o" || a == "--ordinal") o.ordinal = (int)number(i);
        else if (a == "--pdb-dir") o.pdb_dir = value(i);
        else usage(("unknown option " + a).c_str());
    }
    if (!any) usage(nullptr);
    if (o.pathway == o.puzzle) usage("Please select the type of problem you want to solve!");
    if (o.block_rate < 1 || o.block_rate > 99) usage("The block rate must be within 1-99.");
    std::unique_ptr<Problem> p = o.pathway ? make_pathway(o) : make_puzzle(o);
    if (!p) return 1;
    return run(*p, o);
}
