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
};

// Returns RH_OK or an error status with a message in *err.
rh_status pack_model(const rh_model_desc* d, PackedModel* out, std::string* err);

// 3x3 helpers (row-major).
void inv3(const double* m, double* o);

}  // namespace rh
