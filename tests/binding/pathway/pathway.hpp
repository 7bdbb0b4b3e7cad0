// pathway.hpp -- the accessors of uAstar's Pathway (src/pathway/pathway.hpp:10-33) that
// GPUPathwaySolver uses, for the binding test. The test sets the fields directly instead
// of running an input module.
#pragma once
#include "problem.hpp"

class GPUPathwaySolver;

class Pathway : public Problem {
public:
    string problemName() const override { return "Pathway Finding"; }
    void prepare() override {}
    void cpuInitialize() override {}
    void gpuInitialize() override {}
    void cpuSolve() override {}
    void gpuSolve() override {}
    bool output() override { return true; }

    int sx() const { return m_sx; }
    int sy() const { return m_sy; }
    int ex() const { return m_ex; }
    int ey() const { return m_ey; }
    int size() const { return m_width * m_height; }
    int width() const { return m_width; }
    int height() const { return m_height; }
    int toID(int x, int y) const { return x * m_width + y; }
    const uint8_t *graph() const { return m_graph.data(); }

    int m_sx = 0, m_sy = 0, m_ex = 0, m_ey = 0, m_width = 0, m_height = 0;
    vector<uint8_t> m_graph;
};
