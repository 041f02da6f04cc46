// The shape of Solver<Scalar> (src/simulation/solver/solver.h:12-28) that the
// adapters derive from; syntax-check input only.
#pragma once
#include "../../def.h"

template <typename Scalar> class Solver {
public:
    explicit Solver(int threadsPerBlock) : threadsPerBlock(threadsPerBlock) {}
    virtual ~Solver() = default;
    virtual void Update(SolverData<Scalar>& solverData, const SolverParams<Scalar>& solverParams) = 0;
    virtual void Reset() {}
    void SetPerf(bool val) { perf = val; }

protected:
    virtual void SolverPrepare(SolverData<Scalar>& solverData, const SolverParams<Scalar>& solverParams) = 0;
    virtual bool SolverStep(SolverData<Scalar>& solverData, const SolverParams<Scalar>& solverParams) = 0;
    int threadsPerBlock;
    bool perf = false;
    std::vector<std::pair<std::string, float>> performanceData;
};
