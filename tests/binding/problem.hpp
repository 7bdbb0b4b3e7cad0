// problem.hpp -- uAstar's Problem interface (src/problem.hpp:6-24), declared for the
// binding test.
#pragma once
#include "utils.hpp"

class Problem {
public:
    virtual ~Problem() {}
    virtual string problemName() const = 0;
    virtual void prepare() = 0;
    virtual void cpuInitialize() = 0;
    virtual void gpuInitialize() = 0;
    virtual void cpuSolve() = 0;
    virtual void gpuSolve() = 0;
    virtual bool output() = 0;
};
