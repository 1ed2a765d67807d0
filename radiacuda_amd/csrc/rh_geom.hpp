// rh_geom.hpp -- host-side packing of magnetizable elements into the flat
// records the kernels read (see rh_math.hpp for the record layouts).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/radhip.h"

namespace rh {

// Per transform copy: Minv (9), V (3), F = s*M (9), pad (3).
constexpr int kCopySize = 24;

// Conforming-mesh form of the polyhedra for the field sweep (rh_mesh.hpp):
// unique edges and faces with their element incidences, and per-element
// plane sets for the inside test.  Every transform copy of an element is a
// mesh element in the laboratory frame.  Built when every polyhedron has at
// most kMeshPlanes faces.
struct MeshForm {
    bool active = false;
    std::vector<double> edge_geo;        // kMeshEdgeGeo per unique edge
    std::vector<int64_t> edge_inc_off;   // [E+1]
    std::vector<int32_t> edge_inc_elem;  // element of each incidence
    std::vector<double> edge_inc_w;      // F_c W_e of that element copy, 9 per incidence (row-major)
    std::vector<double> face_geo;        // kMeshFaceGeo per unique face
    std::vector<int64_t> face_inc_off;   // [F+1]
    std::vector<int32_t> face_inc_elem;
    std::vector<double> face_inc_v;      // 3 per incidence: the charge of face f gains v . M_j
    std::vector<int32_t> in_elem;        // element of each inside record
    std::vector<double> in_geo;          // 1 + 4 kMeshPlanes per element copy: count, planes
    std::vector<double> in_F;            // 9 per inside record: F_c of the copy
    std::vector<double> tile_box;        // 6 per tile of kMeshTile inside records
    std::vector<double> edge_cont;       // 1: the edge starts at the previous edge's end
    std::vector<double> face_cont;       // 0: triangle, 1: triangle whose vertices 0, 1 are the
                                         // previous face's 1, 2; 2: quadrilateral
    // Edges owned by faces: edge records [ne_free, ne) are evaluated by the
    // face holding both their vertices (face_edge: per face, the record index
    // of the edge on sides (v0,v1), (v1,v2), (v2,v0) of a triangle or (v0,v1),
    // (v1,v2), (v2,v3) of a quadrilateral, or -1); records [0, ne_free) are
    // swept on their own.
    int64_t ne_free = 0;
    std::vector<int64_t> face_edge;      // 3 per face
    int64_t nf_tri = 0;                  // faces [0, nf_tri) are triangles, the rest quadrilaterals
    int64_t n_edges() const { return int64_t(edge_inc_off.empty() ? 0 : edge_inc_off.size() - 1); }
    int64_t n_faces() const { return int64_t(face_inc_off.empty() ? 0 : face_inc_off.size() - 1); }
    int64_t n_in() const { return int64_t(in_elem.size()); }
};
constexpr int kMeshTile = 32;

struct PackedModel {
    int64_t n = 0;
    std::vector<int32_t> kind;      // [n] EK_*
    std::vector<int64_t> off;       // [n] offset of the element record in data
    std::vector<double> data;       // records
    std::vector<int64_t> copy_off;  // [n+1]
    std::vector<double> copies;     // [kCopySize * C]
    std::vector<double> colloc;     // [3n] collocation points, lab frame
    std::vector<double> rowt;       // [9n] row transforms s*Minv of the main copy
    std::vector<int32_t> kind_idx;  // element indices grouped by kind slot (block, tet, general)
    int64_t kind_off[4] = {0, 0, 0, 0};
    bool sym = false;               // some element has a non-identity or extra copy
    bool rowt_identity = true;      // every row transform is the identity
    int64_t counts[8] = {0};        // see rh_model_counts
    // adjoint geometry (rh_adjoint.hpp): per element nv, ntri, vertices in the
    // caller's order, fan triangles of the outward-oriented faces; blocks get
    // an empty record (nv = ntri = 0)
    std::vector<int64_t> adj_off;   // [n+1]
    std::vector<double> adj;
    std::vector<int64_t> vert_off;  // [n+1] first vertex of each element in the caller's list
    MeshForm mesh;
};

// Builds P.mesh from the packed records (no-op when the model does not
// qualify).
void build_mesh_form(PackedModel& P);

// Returns RH_OK or an error status with a message in *err.
rh_status pack_model(const rh_model_desc* d, PackedModel* out, std::string* err);

// 3x3 helpers (row-major).
void inv3(const double* m, double* o);

}  // namespace rh
