// Declarations the adapters in INTEGRATION.md compile against: the shapes of
// the reference's glm vectors, SolverData<Scalar> and SolverParams<Scalar>
// (src/def.h:22-100) reduced to the members the adapters use.  Test input for
// a syntax check only; nothing here is linked or run.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

using indexType = uint32_t;

namespace glm {
template <typename T> struct tvec3 { T x, y, z; };
using vec3 = tvec3<float>;
using dvec3 = tvec3<double>;
struct mat4 { float m[16]; };
inline const float* value_ptr(const mat4& a) { return a.m; }
}  // namespace glm

class FixedBodyData;

template <typename Scalar> struct SolverData {
    int numVerts = 0, numTets = 0, numTris = 0, numDBC = 0;
    glm::tvec3<Scalar> *X = nullptr, *V = nullptr, *X0 = nullptr, *XTilde = nullptr, *DBCX = nullptr;
    indexType *Tet = nullptr, *Tri = nullptr, *dev_TriFathers = nullptr;
    Scalar *mass = nullptr, *mu = nullptr, *lambda = nullptr, *V0 = nullptr, *DBC = nullptr, *moreDBC = nullptr;
    FixedBodyData* pFixedBodies = nullptr;
};

template <typename Scalar> struct SolverParams {
    int numIterations = 1, maxIterations = 100;
    Scalar dt = 0.01, damp = 0.999, muN = 0.5, muT = 0.5, gravity = 9.8, dhat = 1e-3, tol = 1e-2, rho = 0.9992;
    bool handleCollision = false;
};
