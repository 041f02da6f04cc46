// Host loading of soft bodies from TetGen files, as the reference's DataLoader
// does for one body (src/simulation/dataLoader.cu):
//   loadNodeFile        dataLoader.cu:130-170   vertices, optional centring with y/z swap
//   loadEleFaceFile     dataLoader.cu:39-127    tets, the .face triangles or the
//                                               boundary faces extracted from the tets
//   CollectEdges        dataLoader.cu:172-199   unique surface edges
//   CollectData         dataLoader.cu:204-237   model transform T(pos) S(scale) Rx Ry Rz
//                                               applied to the vertices (TransformVertices,
//                                               src/simulation/utilities.cu:57-64)
// Plain C++17 (no HIP).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sbs {

struct BodyMesh {
    std::vector<double> X;          // 3*nv positions (values of the precision the transform ran in)
    std::vector<uint32_t> tet;      // 4*nt, 0-based
    std::vector<uint32_t> tri;      // 3*ntri, 0-based
    std::vector<uint32_t> edges;    // 2*nedge, (min, max) pairs in ascending order
    bool tri_from_file = false;     // false: extracted from the tets
};

struct TetgenOptions {
    int start_index = 0;            // index of the first vertex in the files (0 or 1)
    bool centralize = false;        // subtract the centroid, then swap y and z
    float pos[3] = {0.f, 0.f, 0.f};
    float scale[3] = {1.f, 1.f, 1.f};
    float rot[3] = {0.f, 0.f, 0.f}; // degrees about x, then y, then z
    int precision = 32;             // 32: float arithmetic (PD scenes); 64: double (IPC scenes)
};

// Throws std::invalid_argument on an unreadable node or ele file, malformed
// lines, or vertex indices outside the node list (the reference prints and
// continues).  A face file that is not given or does not open means "extract
// the boundary", as in the reference.
BodyMesh load_tetgen(const std::string& node_file, const std::string& ele_file, const std::string& face_file,
                     const TetgenOptions& opt);

// Boundary faces of a tet list: every tet contributes (v0 v1 v2), (v0 v2 v3),
// (v0 v3 v1), (v1 v3 v2); a face (as a vertex set) seen an odd number of times
// is kept, with the winding of its last occurrence, in ascending order of its
// sorted vertex triple (the std::set/std::map walk of dataLoader.cu:92-127).
std::vector<uint32_t> boundary_faces(const std::vector<uint32_t>& tet);

// Unique undirected edges of a triangle list, ascending (dataLoader.cu:172-199).
std::vector<uint32_t> surface_edges(const std::vector<uint32_t>& tri);

// The 4x4 model matrix of CollectData (dataLoader.cu:215-221), column-major
// (m[4*col + row]), in float: T(pos) S(scale) Rx(rot.x) Ry(rot.y) Rz(rot.z).
void body_model_matrix(const float pos[3], const float scale[3], const float rot[3], float m[16]);

}  // namespace sbs
