// Continuous collision detection between surface triangles on gfx950:
// LBVH broad phase, vertex-face / edge-edge CCD narrow phase, hit resolution
// and the velocity response of the PD step.
//
// Reference behaviour reproduced:
//   CollisionDetection::DetectCollision      src/collision/bvh.cu:171-181
//   CollisionDetection::ComputeMinStepSize   src/collision/bvh.cu:183-190
//   BroadPhaseCCD (leaf boxes, pair filters, 12 queries per pair, canonical
//   form, unique)                            src/collision/broadphase.cu:540-554
//   NarrowPhase + storeTi                    src/collision/narrowphase.cu:115-132
//   ccdCollisionTest                         src/collision/intersections.cu:286-337
//   CCDKernel                                src/simulation/collisionUtil.cu:50-70
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace sbs {

// Canonical query: type 1 = vertex v0 against triangle (v1, v2, v3), type 2 =
// edge (v0, v1) against edge (v2, v3); the reference's Query (aabb.h:67-77).
struct QueryKey {
    uint32_t type, v0, v1, v2, v3;
};

template <class T> class CollisionDetector {
public:
    // tri: 3*num_tris u32 and fathers: num_tris u32 (body of each triangle),
    // device arrays owned by the caller and read at every detect.
    CollisionDetector(int num_verts, int num_tris, const uint32_t* tri, const uint32_t* fathers, hipStream_t stream);
    ~CollisionDetector();

    // DetectCollision: tI[v] = 1 for every vertex, 0.5 for the vertices of
    // every resolved hit; normals (3 floats per vertex) written for hit
    // vertices only.  Synchronises the stream (query counts come to the host).
    void detect(const T* X, const T* XT, bool ignore_self, T* tI, float* normals);
    // ComputeMinStepSize: the smallest time of impact over all queries (1 when
    // there is none).
    double min_step(const T* X, const T* XT, bool ignore_self);

    // UpdateQueries' candidates (broadphase.cu:557-577): the canonical queries
    // of the triangle pairs whose boxes at X, grown by `bound`, overlap; left
    // on the device (queries()), count returned.  Synchronises the stream.
    long long proximity(const T* X, T bound, bool ignore_self);
    const QueryKey* queries() const { return qunique_.get(); }
    long long num_queries() const { return nq_; }
    int num_hits() const { return hits_; }
    long long num_pairs() const { return npairs_; }
    // Host copy of the last query list (sorted, unique) with toi and normals.
    void get_queries(uint32_t* q5, double* toi, double* n3) const;
    void set_stream(hipStream_t s) { stream_ = s; }
    // SetBuildType / GetBuildType (bvh.h:67-68): BVH<Scalar>::BuildType
    // Serial (0), Atomic (1, the default), Cooperative (2) choose how the
    // boxes of the internal nodes are fitted (bvh.cu:118-135).  Serial runs
    // one launch per tree level until the root is fitted; Atomic and
    // Cooperative run the one-launch bottom-up climb (a grid-wide barrier
    // costs more than that climb on this GPU).  Internal boxes are exact
    // unions either way, so every type gives the same queries.
    void set_build_type(int t);
    int build_type() const { return build_type_; }

private:
    void broad_phase(const T* X, const T* XT, bool ignore_self, T bound = 0);
    void narrow_phase(const T* X, const T* XT);

    int nv_ = 0, nt_ = 0;
    const uint32_t* tri_ = nullptr;
    const uint32_t* fathers_ = nullptr;
    hipStream_t stream_ = nullptr;
    // tree
    DevBuf<T> leaf_box_;                 // 6 per triangle (caller order)
    DevBuf<unsigned long long> keys_, keys_sorted_;
    DevBuf<uint32_t> order_in_, order_;  // sorted position -> triangle
    DevBuf<int> left_, right_, parent_, flags_;
    DevBuf<T> node_box_;                 // 6 per node: internal 0..n-2, leaves n-1..2n-2
    DevBuf<T> scene_;                    // per-block partial bounds, then the scene box
    DevBuf<unsigned char> sort_tmp_;
    size_t sort_tmp_bytes_ = 0;
    // pairs and queries
    DevBuf<uint2> pairs_;
    size_t pair_cap_ = 0;
    DevBuf<unsigned long long> counters_;  // [0] pairs, [1] stack overflow flag, [2] hits
    DevBuf<QueryKey> qraw_, qsorted_, qunique_;
    DevBuf<int> qflag_, qscan_;
    DevBuf<unsigned char> qtmp_;
    size_t qtmp_bytes_ = 0;
    size_t qcap_ = 0;
    DevBuf<T> toi_;
    DevBuf<T> qnormal_;                  // 3 per unique query
    DevBuf<int> owner_;                  // per vertex: winning query of its normal
    DevBuf<T> red_;                      // min-toi partials
    long long nq_ = 0, npairs_ = 0;
    int hits_ = 0;
    int build_type_ = 1;
    unsigned long long* host_ = nullptr;  // pinned scratch for counts
};

// Fixed (kinematic) bodies: FixedBody and its subclasses Sphere, Plane and
// Cylinder (src/collision/rigid/rigid.h:8-28, sphere.h, plane.h, cylinder.h)
// as FixedBodyData holds them (src/simulation/fixedBodyData.cu:8-61), with
// HandleCollisions (fixedBodyData.cu:68-148).
enum FixedType : int { kFixedSphere = 0, kFixedPlane = 1, kFixedCylinder = 2 };  // BodyType order
struct FixedBodySpec {
    int type = kFixedPlane;
    float model[16] = {};  // FixedBody::m_model, glm column-major: model[4 * col + row]
    float radius = 0.f;    // Sphere::m_radius; Cylinder::m_radius (= scale[0]); unused for planes
};
constexpr int kFixedFloats = 8;  // per body: p[3], u[3], r, 0

class FixedBodies {
public:
    // Derives each body's parameters on the host, in float as the reference
    // holds them: p = the model's translation (the sphere/cylinder centre, a
    // point of the plane); plane: u = m_floorUp = normalize(row 1 of the
    // inverse of the model's 3x3 part) (plane.cpp:9, the y column of the
    // inverse transpose); cylinder: u = normalize(model * (0,1,0,0))
    // (fixedBodyData.cu:121); r = the radius.  Bodies are kept grouped by
    // type in the order HandleCollisions visits them: spheres, planes,
    // cylinders, each group in the caller's order.
    explicit FixedBodies(const std::vector<FixedBodySpec>& bodies);
    // HandleCollisions on positions X and velocities V (3*nv each, device).
    template <class T> void handle(T* X, T* V, int nv, T mu_t, T mu_n, hipStream_t stream) const;
    int spheres() const { return ns_; }
    int planes() const { return np_; }
    int cylinders() const { return nc_; }
    // kFixedFloats per body, in the caller's order (tests).
    const std::vector<float>& params() const { return params_; }
    // the device records, grouped spheres, planes, cylinders
    const float* device() const { return dev_.get(); }

private:
    int ns_ = 0, np_ = 0, nc_ = 0;
    std::vector<float> params_;
    DevBuf<float> dev_;  // kFixedFloats per body, grouped by type
};

// CCDKernel (collisionUtil.cu:50-70): hit vertices (tI < 1) get
// V = -(d.n) n with d = XTilde - X and keep X; the others take X = XTilde.
template <class T>
void ccd_response(int nv, T* X, const T* XT, T* V, const T* tI, const float* normals, hipStream_t stream);

}  // namespace sbs
