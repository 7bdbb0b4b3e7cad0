// puzzle.cuh -- the accessors of uAstar's Puzzle (src/puzzle/puzzle.cuh:6-38) that
// GPUPuzzleSolver<N> uses, for the binding test.
#pragma once
#include "problem.hpp"

class Puzzle : public Problem {
public:
    string problemName() const override { return "Tile Puzzle"; }
    void prepare() override {}
    void cpuInitialize() override {}
    void gpuInitialize() override {}
    void cpuSolve() override {}
    void gpuSolve() override {}
    bool output() override { return true; }

    int length() { return n; }
    int size() { return n * n; }
    void initialState(vector<uint8_t> &data) { data = m_initialState; }

    int n = 0;
    vector<uint8_t> m_initialState;
};
