// puzzle.cpp -- the command line's sliding-tile puzzle: uAstar's Puzzle (puzzle.cuh,
// puzzle.cu) for n = 3, 4 and 5 on libuastar, with the pattern database shared between
// the CPU and the GPU side through uAstar's database_*.bin cache files.
#include <cstdio>
#include <iostream>

#include "cpu_solvers.hpp"
#include "problems.hpp"

namespace cli {

namespace {

const int DX[4] = {1, -1, 0, 0};
const int DY[4] = {0, 0, 1, -1};

class Puzzle : public Problem {
public:
    Puzzle(const Options &o, int n) : o_(o), n_(n) {}
    ~Puzzle() override {
        if (ctx_) ua_destroy(ctx_);
    }
    std::string problemName() const override { return "Tile Puzzle"; }

    bool prepare() override {
        board_.resize((size_t)n_ * n_);
        std::vector<int> seen(n_ * n_, 0);
        for (auto &t : board_) {
            int v;
            if (!(std::cin >> v) || v < 0 || v >= n_ * n_ || seen[v]++) {
                std::printf("The board must hold each of 0..%d once (0 = blank).\n", n_ * n_ - 1);
                return false;
            }
            t = (uint8_t)v;
        }
        return true;
    }

    // The CPU side uses the cached databases when they are there; otherwise it takes the
    // GPU-built ones (gpuInitialize) or, without a GPU, builds them itself.
    void cpuInitialize() override {
        if (!o_.no_gpu) {
            bool cached = false;
            pdb_ = try_cache(&cached);
            if (!cached) pdb_.clear();
            return;
        }
        bool cached = false;
        pdb_ = pdb_load_or_build(n_, o_.pdb_dir, true, &cached);
    }

    bool gpuInitialize() override {
        if (!(ctx_ = open_gpu(o_))) return false;
        int rc;
        if (!pdb_.empty()) {
            rc = ua_puzzle_load_pdb(ctx_, n_, pdb_.data(), pdb_.size());  // the cached tables
        } else {
            rc = ua_puzzle_build_pdb(ctx_, n_);  // built on the device
            if (!rc && !o_.no_cpu) {
                pdb_.resize(pdb_size());
                rc = ua_puzzle_download_pdb(ctx_, n_, pdb_.data(), pdb_.size());
                if (!rc && !pdb_write_cache(n_, o_.pdb_dir, pdb_))
                    std::printf("Warning: cannot write the database files to %s\n", o_.pdb_dir.c_str());
            }
        }
        if (rc) std::printf("GPU pattern database failed: %s\n", ua_last_error(ctx_));
        return rc == UA_OK;
    }

    void cpuSolve() override {
        int64_t expanded = 0, dedup = 0;
        cpu_ok_ = cpu_puzzle_astar(n_, board_, pdb_, &cpu_steps_, &cpu_moves_, &expanded, &dedup);
        cpu_done_ = true;
        if (cpu_ok_) {  // the lines CPUPuzzleSolver::solve prints at the goal (puzzle/CPU-solver.hpp:169-170)
            std::printf("\t\tNumber of nodes deduplicated: %d\n", (int)dedup);
            std::printf("\t\tNumber of nodes expanded: %d\n", (int)expanded);
        }
    }

    bool gpuSolve() override {
        int rc = ua_puzzle_solve(ctx_, n_, board_.data(), &res_);
        if (rc < 0) {
            std::printf("GPU search failed: %s\n", ua_last_error(ctx_));
            return false;
        }
        gpu_ok_ = rc == UA_OK;
        gpu_done_ = true;
        if (gpu_ok_) {
            // the line GPUPuzzleSolver<N>::solve prints at termination (puzzle/GPU-solver.cuh:246)
            std::printf("\t\t\t Number of nodes expanded: %d\n", (int)res_.expanded);
            int64_t len = 0;
            if (ua_puzzle_moves(ctx_, nullptr, 0, &len)) return false;
            std::vector<int32_t> m(len > 0 ? len : 1);
            if (ua_puzzle_moves(ctx_, m.data(), len, &len)) return false;
            gpu_moves_.assign(m.begin(), m.begin() + len);
            gpu_steps_ = (int)res_.steps;
        }
        return true;
    }

    bool output() override {
        if (cpu_done_ && gpu_done_ && cpu_ok_ != gpu_ok_) return false;
        if (cpu_done_) report(cpu_ok_, cpu_moves_, "CPU", "solutionCPU.txt");
        if (gpu_done_) report(gpu_ok_, gpu_moves_, "GPU", "solutionGPU.txt");
        if (cpu_done_ && cpu_ok_) std::printf(" > Optimal steps from CPU: %d\n", cpu_steps_);
        if (gpu_done_ && gpu_ok_) {
            std::printf(" > Optimal steps from GPU: %d\n", gpu_steps_);
            std::printf(" > GPU search: %lld rounds, %lld nodes expanded, %.2f ms on the device\n",
                        (long long)res_.rounds, (long long)res_.expanded, res_.device_ms);
        }
        return !(cpu_done_ && gpu_done_ && cpu_ok_ && cpu_steps_ != gpu_steps_);
    }

private:
    size_t pdb_size() const {
        size_t s = 0;
        for (const auto &t : pdb_partitions(n_)) s += pdb_table_size(n_, t.size());
        return s;
    }

    std::vector<uint8_t> try_cache(bool *cached) const {
        std::vector<uint8_t> pdb(pdb_size());
        size_t off = 0;
        *cached = true;
        for (const auto &t : pdb_partitions(n_)) {
            size_t size = pdb_table_size(n_, t.size());
            std::string path = o_.pdb_dir + "/" + pdb_file_name(n_, t);
            FILE *f = std::fopen(path.c_str(), "rb");
            size_t got = f ? std::fread(pdb.data() + off, 1, size, f) : 0;
            if (f) std::fclose(f);
            if (got != size) *cached = false;
            off += size;
        }
        return pdb;
    }

    // Every state from the start along the moves: n rows of "%4d" cells and a blank line.
    void report(bool ok, const std::vector<int> &moves, const char *who, const char *file) const {
        if (!ok) {
            std::printf("No solution from %s.\n", who);
            return;
        }
        FILE *f = std::fopen(file, "w");
        if (!f) {
            std::printf("ERROR: %s cannot be open for writing.\n", file);
            return;
        }
        std::vector<uint8_t> s = board_;
        auto dump = [&] {
            for (int i = 0; i < n_; ++i) {
                for (int j = 0; j < n_; ++j) std::fprintf(f, "%4d", s[i * n_ + j]);
                std::fprintf(f, "\n");
            }
            std::fprintf(f, "\n");
        };
        dump();
        for (int d : moves) {
            int z = 0;
            while (s[z]) ++z;
            int nz = (z / n_ + DX[d]) * n_ + (z % n_ + DY[d]);
            std::swap(s[z], s[nz]);
            dump();
        }
        std::fclose(f);
    }

    Options o_;
    int n_;
    std::vector<uint8_t> board_, pdb_;
    ua_ctx *ctx_ = nullptr;
    ua_result res_{};
    bool cpu_done_ = false, gpu_done_ = false, cpu_ok_ = false, gpu_ok_ = false;
    int cpu_steps_ = 0, gpu_steps_ = 0;
    std::vector<int> cpu_moves_, gpu_moves_;
};

}  // namespace

std::unique_ptr<Problem> make_puzzle(const Options &o) {
    if (o.height <= 0 && o.width <= 0) {
        std::printf("Please set the width or height for your puzzle.\n");
        return nullptr;
    }
    if (o.height > 0 && o.width > 0 && o.height != o.width) {
        std::printf("The puzzle must be square: width must equal height.\n");
        return nullptr;
    }
    int n = o.width > 0 ? o.width : o.height;
    if (n < 3 || n > 5) {
        std::printf("Currently we can only solve N=3~5 puzzle problem\n");
        return nullptr;
    }
    return std::make_unique<Puzzle>(o, n);
}

}  // namespace cli
