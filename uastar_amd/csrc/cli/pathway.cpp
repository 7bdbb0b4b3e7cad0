// pathway.cpp -- the command line's grid problem: uAstar's Pathway (pathway.hpp,
// pathway.cpp) with its input modules (custom, zigzag; input/custom.cpp:14-26,
// input/zigzag.cpp:14-40), the random module its main advertises (main.cpp:28-29),
// solution files and an optional BMP plot, on libuastar for the GPU side.
#include <cmath>
#include <cstdio>
#include <iostream>

#include "cpu_solvers.hpp"
#include "problems.hpp"

namespace cli {

namespace {

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The random module's cell rule, identical to ua_grid_generate's.
bool random_open(uint64_t seed, uint64_t thr, uint64_t id) {
    return (mix64(mix64(seed) + id * 0x9E3779B97F4A7C15ull) >> 32) >= thr;
}

class Pathway : public Problem {
public:
    explicit Pathway(const Options &o) : o_(o), H_(o.height), W_(o.width) {}
    ~Pathway() override {
        if (grid_) ua_grid_free(ctx_, grid_);
        if (ctx_) ua_destroy(ctx_);
    }
    std::string problemName() const override { return "Pathway Finding"; }

    bool prepare() override {
        graph_.assign((size_t)H_ * W_, 0xFF);
        if (o_.input_module == "custom") {
            if (!(std::cin >> sx_ >> sy_ >> ex_ >> ey_)) return bad("cannot read the start and end points");
            for (size_t i = 0; i < graph_.size(); ++i) {
                int t;
                if (!(std::cin >> t)) return bad("cannot read the grid cells");
                graph_[i] = t ? 0xFF : 0;
            }
        } else if (o_.input_module == "zigzag") {
            int gap = 0;
            if (!(std::cin >> gap) || gap < 1) return bad("cannot read the row gap");
            sx_ = 0, sy_ = W_ / 2, ex_ = H_ - 1, ey_ = W_ / 2;
            bool right = true;  // the first wall is open at its right end, then alternately
            for (int i = 1; i < H_; ++i) {
                if (i % gap) continue;
                for (int j = 0; j < W_; ++j) graph_[(size_t)i * W_ + j] = 0;
                graph_[(size_t)i * W_ + (right ? W_ - 1 : 0)] = 0xFF;
                right = !right;
            }
        } else if (o_.input_module == "random") {
            sx_ = 0, sy_ = 0, ex_ = H_ - 1, ey_ = W_ - 1;
            const uint64_t thr = (uint64_t)std::floor(o_.block_rate / 100.0 * 4294967296.0);
            for (size_t i = 0; i < graph_.size(); ++i) graph_[i] = random_open(o_.seed, thr, i) ? 0xFF : 0;
            graph_[(size_t)sx_ * W_ + sy_] = graph_[(size_t)ex_ * W_ + ey_] = 0xFF;
        } else {
            return bad("Please set your input-module parameter correctly (custom, zigzag or random).");
        }
        if (sx_ < 0 || sx_ >= H_ || sy_ < 0 || sy_ >= W_ || ex_ < 0 || ex_ >= H_ || ey_ < 0 || ey_ >= W_)
            return bad("start or end point outside the grid");
        return true;
    }

    void cpuInitialize() override {}

    bool gpuInitialize() override {
        if (!(ctx_ = open_gpu(o_))) return false;
        int rc;
        if (o_.input_module == "random") {  // generated on the device by the same cell rule
            rc = ua_grid_generate(ctx_, H_, W_, o_.seed, o_.block_rate / 100.0, sx_, sy_, ex_, ey_, &grid_);
        } else {
            rc = ua_grid_from_bytes(ctx_, H_, W_, graph_.data(), &grid_);
        }
        if (rc) std::printf("GPU grid upload failed: %s\n", ua_last_error(ctx_));
        return rc == UA_OK;
    }

    void cpuSolve() override {
        cpu_ok_ = cpu_grid_astar(graph_, H_, W_, sx_, sy_, ex_, ey_, &cpu_cost_, &cpu_path_, &cpu_expanded_);
        cpu_done_ = true;
    }

    bool gpuSolve() override {
        int rc = ua_grid_solve(ctx_, grid_, sx_, sy_, ex_, ey_, &res_);
        if (rc < 0) {
            std::printf("GPU search failed: %s\n", ua_last_error(ctx_));
            return false;
        }
        gpu_ok_ = rc == UA_OK;
        gpu_done_ = true;
        if (gpu_ok_) {
            // the line GPUPathwaySolver::solve prints at termination (pathway/GPU-solver.cu:200)
            std::printf("\t\t\t Number of nodes expanded: %d\n", (int)res_.expanded);
            int64_t len = 0;
            if (ua_grid_path(ctx_, nullptr, 0, &len) || len < 1) return false;
            std::vector<int32_t> ids(len);
            if (ua_grid_path(ctx_, ids.data(), len, &len)) return false;
            gpu_path_.assign(ids.begin(), ids.end());
            gpu_cost_.orth = res_.orth;
            gpu_cost_.diag = res_.diag;
        }
        return true;
    }

    // The order and text of Pathway::output (pathway/pathway.cpp:89-130): the solution files
    // with their "-" / "+" counts (the CPU solver's getSolution first prints its node
    // count, pathway/CPU-solver.cpp:94), then the optimal distances. The exact step counts
    // follow on lines of their own.
    bool output() override {
        if (cpu_done_ && gpu_done_ && cpu_ok_ != gpu_ok_) return false;
        if (cpu_done_) {
            if (cpu_ok_) std::printf("\t\t\tNumber of nodes expanded: %d\n", (int)cpu_expanded_);
            report(cpu_ok_, cpu_path_, cpu_cost_, "CPU", "solutionCPU.txt");
        }
        if (gpu_done_) report(gpu_ok_, gpu_path_, gpu_cost_, "GPU", "solutionGPU.txt");
        if (cpu_done_ && cpu_ok_) std::printf(" > Optimal distance from CPU: %.3f\n", cpu_cost_.value());
        if (gpu_done_ && gpu_ok_) std::printf(" > Optimal distance from GPU: %.3f\n", gpu_cost_.value());
        if (cpu_done_ && cpu_ok_)
            std::printf(" > Exact distance from CPU: %lld + %lld * sqrt(2)\n", (long long)cpu_cost_.orth,
                        (long long)cpu_cost_.diag);
        if (gpu_done_ && gpu_ok_) {
            std::printf(" > Exact distance from GPU: %lld + %lld * sqrt(2)\n", (long long)gpu_cost_.orth,
                        (long long)gpu_cost_.diag);
            std::printf(" > GPU search: %lld rounds, %lld nodes expanded, %.2f ms on the device\n",
                        (long long)res_.rounds, (long long)res_.expanded, res_.device_ms);
        }
        if (!o_.plot.empty()) plot(gpu_done_ && gpu_ok_ ? gpu_path_ : cpu_path_);
        // exact comparison: the (orth, diag) step counts are the cost
        if (cpu_done_ && gpu_done_ && cpu_ok_ &&
            (cpu_cost_.orth != gpu_cost_.orth || cpu_cost_.diag != gpu_cost_.diag))
            return false;
        return true;
    }

private:
    bool bad(const char *msg) {
        std::printf("%s\n", msg);
        return false;
    }

    void report(bool ok, const std::vector<int> &path, const GridCost &c, const char *who, const char *file) const {
        if (!ok) {
            std::printf("No solution from %s.\n", who);
            return;
        }
        if (FILE *f = std::fopen(file, "w")) {  // "(x y) -> (x y) -> ...", six cells a line
            for (size_t i = 0; i < path.size(); ++i) {
                if (i) std::fprintf(f, " -> ");
                if ((i + 1) % 6 == 0) std::fprintf(f, "\n\t");
                std::fprintf(f, "(%d %d)", path[i] / W_, path[i] % W_);
            }
            std::fprintf(f, "\n");
            std::fclose(f);
        } else {
            std::printf("ERROR: %s cannot be open for writing.\n", file);
        }
        std::printf(" > Number of -: %lld, +: %lld\n", (long long)c.orth, (long long)c.diag);
    }

    // A 24-bit BMP with one px-by-px square per cell: open white, blocked grey, path green.
    void plot(const std::vector<int> &path) const {
        const int px = std::max(1, std::min(8, 1024 / std::max(H_, W_)));
        const int w = W_ * px, h = H_ * px, row = (w * 3 + 3) & ~3;
        if ((int64_t)row * h > (256ll << 20)) {
            std::printf("Warning: the grid is too large to plot\n");
            return;
        }
        std::vector<uint8_t> img((size_t)row * h, 0);
        std::vector<uint8_t> on_path((size_t)H_ * W_, 0);
        for (int v : path) on_path[v] = 1;
        for (int x = 0; x < H_; ++x)
            for (int y = 0; y < W_; ++y) {
                size_t id = (size_t)x * W_ + y;
                uint8_t b = 255, g = 255, r = 255;
                if (on_path[id]) b = 0, g = 200, r = 0;
                else if (!graph_[id]) b = g = r = 96;
                for (int i = 0; i < px; ++i)
                    for (int j = 0; j < px; ++j) {
                        uint8_t *p = &img[(size_t)(h - 1 - (x * px + i)) * row + (size_t)(y * px + j) * 3];
                        p[0] = b, p[1] = g, p[2] = r;
                    }
            }
        FILE *f = std::fopen(o_.plot.c_str(), "wb");
        if (!f) {
            std::printf("ERROR: %s cannot be open for writing.\n", o_.plot.c_str());
            return;
        }
        auto u32 = [&](uint32_t v) { std::fwrite(&v, 4, 1, f); };
        auto u16 = [&](uint16_t v) { std::fwrite(&v, 2, 1, f); };
        std::fwrite("BM", 1, 2, f);
        u32(54 + (uint32_t)img.size()), u32(0), u32(54);
        u32(40), u32((uint32_t)w), u32((uint32_t)h), u16(1), u16(24), u32(0), u32((uint32_t)img.size());
        u32(2835), u32(2835), u32(0), u32(0);
        std::fwrite(img.data(), 1, img.size(), f);
        std::fclose(f);
    }

    Options o_;
    int H_, W_;
    int sx_ = 0, sy_ = 0, ex_ = 0, ey_ = 0;
    std::vector<uint8_t> graph_;
    ua_ctx *ctx_ = nullptr;
    ua_grid *grid_ = nullptr;
    ua_result res_{};
    bool cpu_done_ = false, gpu_done_ = false, cpu_ok_ = false, gpu_ok_ = false;
    GridCost cpu_cost_, gpu_cost_;
    int64_t cpu_expanded_ = 0;
    std::vector<int> cpu_path_, gpu_path_;
};

}  // namespace

ua_ctx *open_gpu(const Options &o) {
    ua_ctx *c = nullptr;
    int rc = ua_create(o.ordinal, nullptr, &c);
    if (rc) {
        std::printf("No usable GPU at ordinal %d (status %d)\n", o.ordinal, rc);
        return nullptr;
    }
    return c;
}

std::unique_ptr<Problem> make_pathway(const Options &o) {
    if (o.height <= 0 || o.width <= 0) {
        std::printf("Please set the width and height for your graph.\n");
        return nullptr;
    }
    return std::make_unique<Pathway>(o);
}

}  // namespace cli
