// The shapes of FixedBody, Sphere, Cylinder (src/collision/rigid/*.h) and of
// FixedBodyData (src/simulation/fixedBodyData.h:10-30) as the adapter in
// INTEGRATION.md changes it (an SbsFixed* in place of its device arrays);
// syntax-check input only.
#pragma once
#include <sbs/sbs.h>
#include "../def.h"

enum class BodyType { Sphere, Plane, Cylinder };
class FixedBody {
public:
    virtual ~FixedBody() = default;
    virtual BodyType getType() const = 0;
    glm::mat4 m_model{};
};
class Sphere : public FixedBody {
public:
    BodyType getType() const override { return BodyType::Sphere; }
    float m_radius = 1.f;
};
class Cylinder : public FixedBody {
public:
    BodyType getType() const override { return BodyType::Cylinder; }
    float m_radius = 1.f;
};

class FixedBodyData {
public:
    FixedBodyData(int threadsPerBlock, const std::vector<FixedBody*>& bodies);
    ~FixedBodyData();
    template <typename Scalar>
    void HandleCollisions(glm::tvec3<Scalar>* X, glm::tvec3<Scalar>* V, int numVerts, Scalar muT, Scalar muN);
    SbsFixed* sbs() const;

private:
    int threadsPerBlock;
    SbsFixed* h = nullptr;
};
