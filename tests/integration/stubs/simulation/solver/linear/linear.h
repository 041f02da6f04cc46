// The shape of LinearSolver<T> (src/simulation/solver/linear/linear.h:52-70)
// that the COO adapter derives from; syntax-check input only.
#pragma once
#include <cstdio>

template <typename T> class LinearSolver {
public:
    LinearSolver() = default;
    virtual ~LinearSolver() = default;
    virtual void Solve(int N, T* d_b, T* d_x, T* A = nullptr, int nz = 0, int* rowIdx = nullptr,
                       int* colIdx = nullptr, T* d_guess = nullptr) = 0;
};
