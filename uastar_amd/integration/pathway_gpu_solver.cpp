// pathway_gpu_solver.cpp -- uAstar's GPUPathwaySolver (src/pathway/GPU-solver.hpp:20-37)
// implemented over libuastar's C ABI. It replaces src/pathway/GPU-solver.cu in a uAstar
// build: compile it with uAstar's include path and link libuastar.so; no CUDA, moderngpu
// or nvcc is needed. The class declaration stays uAstar's own.
//
//   initialize()  -> ua_create on vm_options["ordinal"] (GPU-solver.cu:96), then the
//                    byte graph of Pathway::graph() to the device bitmap (GPU-solver.cu:103)
//   solve()       -> ua_grid_solve (GPU-solver.cu:142-290); false = no path
//   getSolution() -> ua_grid_path (GPU-solver.cu:292-316): cost and start..goal cell ids
//
// Every status is checked: UA_NO_PATH is `false`, any error throws std::runtime_error
// with ua_last_error's message (the reference ignores CUDA errors).
#include "pathway/GPU-solver.hpp"

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "uastar.h"

class DeviceData {  // forward-declared by GPU-solver.hpp
public:
    ua_ctx *ctx = nullptr;
    ua_grid *grid = nullptr;
    ua_result res{};
};

namespace {
void check(ua_ctx *c, int st, const char *what) {
    if (st < 0) throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(st) + "): " +
                                         (c ? ua_last_error(c) : "no context"));
}
}  // namespace

GPUPathwaySolver::GPUPathwaySolver(Pathway *pathway) : p(pathway), d(new DeviceData) {}

GPUPathwaySolver::~GPUPathwaySolver() {
    if (d->grid) ua_grid_free(d->ctx, d->grid);
    if (d->ctx) ua_destroy(d->ctx);
    delete d;
}

void GPUPathwaySolver::initialize() {
    check(nullptr, ua_create(vm_options["ordinal"].as<int>(), nullptr, &d->ctx), "ua_create");
    // one byte per cell, 0 blocked / 0xFF open (pathway.hpp:116-119); a partial out-move
    // mask is rejected with UA_ERR_INVALID
    check(d->ctx, ua_grid_from_bytes(d->ctx, p->height(), p->width(), p->graph(), &d->grid), "ua_grid_from_bytes");
}

bool GPUPathwaySolver::solve() {
    const int st = ua_grid_solve(d->ctx, d->grid, p->sx(), p->sy(), p->ex(), p->ey(), &d->res);
    check(d->ctx, st, "ua_grid_solve");
    if (st == UA_OK)  // the line of GPU-solver.cu:200
        std::printf("\t\t\t Number of nodes expanded: %d\n", (int)d->res.expanded);
    m_optimalNodeAddr = 0;
    m_optimalDistance = (float)d->res.cost;  // exact orth + diag * sqrt(2), rounded once
    return st == UA_OK;
}

void GPUPathwaySolver::getSolution(float *optimal, vector<int> *pathList) {
    int64_t len = 0;
    check(d->ctx, ua_grid_path(d->ctx, nullptr, 0, &len), "ua_grid_path");
    std::vector<int32_t> ids((size_t)len);
    check(d->ctx, ua_grid_path(d->ctx, ids.data(), len, &len), "ua_grid_path");
    *optimal = m_optimalDistance;
    pathList->assign(ids.begin(), ids.begin() + len);  // start..goal cell ids x * W + y
}
