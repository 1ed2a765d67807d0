// rh_hostcheck.cpp -- host build of the packing and of the exact device math,
// for CPU unit tests of the formulation (no GPU).  This library is a test
// aid: the product entry points in libradhip.so never call it.
#include <string>
#include <vector>

#include "rh_adjoint.hpp"
#include "rh_current.hpp"
#include "rh_eval.hpp"
#include "rh_geom.hpp"
#include "rh_mesh.hpp"
#include "rh_opcount.hpp"

namespace {
thread_local std::string g_err;

rh::ModelView view(const rh::PackedModel& P)
{
    return rh::ModelView{P.kind.data(), P.off.data(), P.data.data(), P.copy_off.data(),
                         P.copies.data()};
}
}  // namespace

extern "C" {

const char* rhc_last_error(void) { return g_err.c_str(); }

// Interaction matrix rows [i0, i1) (3 rows each) via the host build: fp32 0
// FP64 integrals, 1 FP32 integrals, 2 the mixed assembly's choice (FP32 for
// far blocks, FP64 near).
int rhc_assemble(const rh_model_desc* d, int64_t i0, int64_t i1, int32_t fp32, double* out)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    rh::ModelView mv = view(P);
    const int64_t n3 = 3 * P.n;
    for (int64_t i = i0; i < i1; i++) {
        const double* x = P.colloc.data() + 3 * i;
        for (int64_t j = 0; j < P.n; j++) {
            double q[9];
            if (fp32 == 2) rh::block_Q_mixed<true, -1>(mv, j, x, q);
            else if (fp32) rh::block_Q<float, true>(mv, j, x, q);
            else rh::block_Q<double, true>(mv, j, x, q);
            const double* rt = P.rowt.data() + 9 * i;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    out[(3 * (i - i0) + a) * n3 + 3 * j + b] =
                        rt[3 * a] * q[b] + rt[3 * a + 1] * q[3 + b] + rt[3 * a + 2] * q[6 + b];
        }
    }
    return RH_OK;
}

int rhc_field(const rh_model_desc* d, const double* M, const double* pts, int64_t np, int32_t fp32,
              double* B, double* H)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    rh::ModelView mv = view(P);
    for (int64_t p = 0; p < np; p++) {
        double b[3] = {0, 0, 0}, h[3] = {0, 0, 0};
        if (fp32) rh::field_add<float, true>(mv, M, nullptr, 0, P.n, pts + 3 * p, b, h);
        else rh::field_add<double, true>(mv, M, nullptr, 0, P.n, pts + 3 * p, b, h);
        for (int k = 0; k < 3; k++) {
            if (B) B[3 * p + k] = b[k];
            if (H) H[3 * p + k] = h[k];
        }
    }
    return RH_OK;
}

// Field through the per-kind path the kernels use (blocks, tetrahedra with
// charged records, general polyhedra).
int rhc_field_kinds(const rh_model_desc* d, const double* M, const double* pts, int64_t np,
                    int32_t fp32, double* B, double* H)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    rh::ModelView mv = view(P);
    const int64_t t0 = P.kind_off[1], t1 = P.kind_off[2];
    std::vector<double> charged(size_t(rh::kChargedTet * (t1 - t0)) + 1);
    for (int64_t k = t0; k < t1; k++) {
        const int64_t j = P.kind_idx[k];
        rh::charge_tet(P.data.data() + P.off[j], M + 3 * j, charged.data() + rh::kChargedTet * (k - t0));
    }
    mv.charged = charged.data();
    mv.charged_base = t0;
    const int32_t* idx = P.kind_idx.data();
    for (int64_t p = 0; p < np; p++) {
        double b[3] = {0, 0, 0}, h[3] = {0, 0, 0};
        const double* x = pts + 3 * p;
        if (fp32) {
            rh::field_add<float, true, 0>(mv, M, idx, P.kind_off[0], P.kind_off[1], x, b, h);
            rh::field_add<float, true, 1>(mv, M, idx, P.kind_off[1], P.kind_off[2], x, b, h);
            rh::field_add<float, true, 2>(mv, M, idx, P.kind_off[2], P.kind_off[3], x, b, h);
        } else {
            rh::field_add<double, true, 0>(mv, M, idx, P.kind_off[0], P.kind_off[1], x, b, h);
            rh::field_add<double, true, 1>(mv, M, idx, P.kind_off[1], P.kind_off[2], x, b, h);
            rh::field_add<double, true, 2>(mv, M, idx, P.kind_off[2], P.kind_off[3], x, b, h);
        }
        for (int k = 0; k < 3; k++) {
            if (B) B[3 * p + k] = b[k];
            if (H) H[3 * p + k] = h[k];
        }
    }
    return RH_OK;
}

int rhc_colloc(const rh_model_desc* d, double* X)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    for (size_t k = 0; k < P.colloc.size(); k++) X[k] = P.colloc[k];
    return RH_OK;
}

// Current sources: kind[ns], par[12 ns], copies[24 ns] (Minv, V, s*M).
int rhc_current(int64_t ns, const int32_t* kind, const double* par, const double* cp,
                const double* pts, int64_t np, double* B)
{
    for (int64_t p = 0; p < np; p++) {
        double b[3] = {0, 0, 0};
        for (int64_t s = 0; s < ns; s++) rh::current_add(kind[s], par + 12 * s, cp + 24 * s, pts + 3 * p, b);
        for (int k = 0; k < 3; k++) B[3 * p + k] = b[k];
    }
    return 0;
}

// H at np points from the triangulated adjoint geometry (value only, FP64,
// no symmetry copies): the function whose derivatives the adjoint kernels take.
int rhc_adj_field(const rh_model_desc* d, const double* M, const double* pts, int64_t np, double* H)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    for (int64_t p = 0; p < np; p++) {
        double h[3] = {0, 0, 0};
        for (int64_t j = 0; j < P.n; j++) {
            double hj[3];
            rh::adj_field_value(P.adj.data() + P.adj_off[j], pts + 3 * p, M + 3 * j, hj);
            for (int k = 0; k < 3; k++) h[k] += rh::kInv4Pi * hj[k];
        }
        for (int k = 0; k < 3; k++) H[3 * p + k] = h[k];
    }
    return RH_OK;
}

// Host version of the vertex and point gradients for weights g (dL/dH), no
// symmetry copies: gV [3 nvert], gX [3 np].
int rhc_adj_grad(const rh_model_desc* d, const double* M, const double* pts, int64_t np,
                 const double* g, double* gV, double* gX)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    const int64_t nvert = P.vert_off.back();
    for (int64_t k = 0; k < 3 * nvert; k++) gV[k] = 0.0;
    for (int64_t p = 0; p < np; p++) {
        const double gl[3] = {rh::kInv4Pi * g[3 * p], rh::kInv4Pi * g[3 * p + 1], rh::kInv4Pi * g[3 * p + 2]};
        double gx[3] = {0, 0, 0};
        for (int64_t j = 0; j < P.n; j++) {
            const double* rec = P.adj.data() + P.adj_off[j];
            for (int t = 0; t < rh::adj_ntri(rec); t++)
                for (int k = 0; k < 3; k++) {
                    double o[3];
                    int vid = 0;
                    rh::adj_tri_corner_grad(rec, t, k, pts + 3 * p, M + 3 * j, gl, o, &vid);
                    for (int a = 0; a < 3; a++) {
                        gV[3 * (P.vert_off[j] + vid) + a] += o[a];
                        gx[a] -= o[a];
                    }
                }
        }
        for (int a = 0; a < 3; a++) gX[3 * p + a] = gx[a];
    }
    return RH_OK;
}

}  // extern "C"

namespace {
// Charged mesh records (the device's k_mesh_charge_* output) for M.
void charge_mesh(const rh::MeshForm& F, const double* M, std::vector<double>& ce, std::vector<double>& cf,
                 std::vector<double>& ci)
{
    const int64_t ne = F.n_edges(), nf = F.n_faces(), ni = F.n_in();
    ce.assign(size_t(rh::kMeshEdge * ne), 0.0);
    cf.assign(size_t(rh::kMeshFace * nf), 0.0);
    ci.assign(size_t(rh::kMeshIn * ni), 0.0);
    for (int64_t i = 0; i < ne; i++) {
        double w[3] = {0, 0, 0};
        for (int64_t k = F.edge_inc_off[i]; k < F.edge_inc_off[i + 1]; k++) {
            const double* W = F.edge_inc_w.data() + 9 * k;
            const double* m = M + 3 * F.edge_inc_elem[k];
            w[0] += W[0] * m[0] + W[1] * m[1] + W[2] * m[2];
            w[1] += W[3] * m[0] + W[4] * m[1] + W[5] * m[2];
            w[2] += W[6] * m[0] + W[7] * m[1] + W[8] * m[2];
        }
        double* o = ce.data() + rh::kMeshEdge * i;
        for (int k = 0; k < rh::kMeshEdgeGeo; k++) o[k] = F.edge_geo[rh::kMeshEdgeGeo * i + k];
        o[10] = w[0];
        o[11] = w[1];
        o[12] = w[2];
    }
    for (int64_t i = 0; i < nf; i++) {
        const double* g = F.face_geo.data() + rh::kMeshFaceGeo * i;
        double s = 0;
        for (int64_t k = F.face_inc_off[i]; k < F.face_inc_off[i + 1]; k++) {
            const double* m = M + 3 * F.face_inc_elem[k];
            const double* v = F.face_inc_v.data() + 3 * k;
            s += v[0] * m[0] + v[1] * m[1] + v[2] * m[2];
        }
        double* o = cf.data() + rh::kMeshFace * i;
        for (int k = 0; k < rh::kMeshFaceGeo; k++) o[k] = g[k];
        o[18] = s * g[12];
        o[19] = s * g[13];
        o[20] = s * g[14];
    }
    for (int64_t i = 0; i < ni; i++) {
        const double* g = F.in_geo.data() + (1 + 4 * rh::kMeshPlanes) * i;
        double* o = ci.data() + rh::kMeshIn * i;
        o[0] = g[0];
        const double* Fc = F.in_F.data() + 9 * i;
        const double* m = M + 3 * F.in_elem[i];
        for (int k = 0; k < 3; k++) o[1 + k] = Fc[3 * k] * m[0] + Fc[3 * k + 1] * m[1] + Fc[3 * k + 2] * m[2];
        for (int k = 0; k < 4 * rh::kMeshPlanes; k++) o[4 + k] = g[1 + k];
    }
}

void put_counts(const opc::Counts& c, double evals, double* o)
{
    const double e = evals > 0 ? evals : 1.0;
    o[0] = c.add / e;
    o[1] = c.mul / e;
    o[2] = c.div / e;
    o[3] = c.sqrt / e;
    o[4] = c.fma / e;
    o[5] = c.log / e;
    o[6] = c.atan2 / e;
}
}  // namespace

extern "C" {

// Operation counts of the FP64 integrals as written, from the kernel math
// instantiated with an op-counting scalar (rh_opcount.hpp), averaged over
// the np points and the model's entities.  out (28 doubles), four groups of
// [add/sub, mul, div, sqrt, fma, log, atan2] per evaluation:
//   0: (point, unique mesh edge)   mesh_edge_L
//   1: (point, unique mesh face)   mesh_face_omega (triangles and quadrilaterals)
//   2: (point, tetrahedron) block  eval_tet (assembly)
//   3: (point, tetrahedron) field  field_tet (per-element field sweep)
// Operations the kernels perform in FP64 outside the templated integrals
// (vertex offsets, accumulation of the weighted terms) are not included.
// Returns 2 when the model has no mesh form.
int rhc_opcount(const rh_model_desc* d, const double* M, const double* pts, int64_t np, double* out)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    const rh::MeshForm& F = P.mesh;
    if (!F.active) return 2;
    std::vector<double> ce, cf, ci;
    charge_mesh(F, M, ce, cf, ci);
    const int64_t ne = F.n_edges(), nf = F.n_faces();
    opc::Counts c[4];
    double evals[4] = {0, 0, 0, 0};
    for (int64_t p = 0; p < np; p++) {
        const double* x = pts + 3 * p;
        opc::counter() = &c[0];
        for (int64_t i = 0; i < ne; i++) (void)rh::mesh_edge_L<opc::OpC>(ce.data() + rh::kMeshEdge * i, x);
        evals[0] += double(ne);
        opc::counter() = &c[1];
        for (int64_t i = 0; i < nf; i++) (void)rh::mesh_face_omega<opc::OpC>(cf.data() + rh::kMeshFace * i, x);
        evals[1] += double(nf);
        for (int64_t j = 0; j < P.n; j++) {
            if (P.kind[j] != rh::EK_TET) continue;
            const double* rec = P.data.data() + P.off[j];
            double q6[6], H[3], ch[rh::kChargedTet];
            opc::counter() = &c[2];
            (void)rh::eval_tet<opc::OpC>(rec, x, q6);
            evals[2] += 1.0;
            opc::counter() = nullptr;
            rh::charge_tet(rec, M + 3 * j, ch);
            opc::counter() = &c[3];
            (void)rh::field_tet<opc::OpC>(ch, x, H);
            evals[3] += 1.0;
        }
    }
    opc::counter() = nullptr;
    for (int g = 0; g < 4; g++) put_counts(c[g], evals[g], out + 7 * g);
    return RH_OK;
}

// Field from the conforming-mesh form (host build of the device sweep):
// returns 2 when the model has no mesh form.  counts3 (may be null) gets the
// numbers of unique edges, unique faces and inside records.
int rhc_field_mesh(const rh_model_desc* d, const double* M, const double* pts, int64_t np,
                   int32_t fp32, double* B, double* H, int64_t* counts3)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    const rh::MeshForm& F = P.mesh;
    if (!F.active) return 2;
    const int64_t ne = F.n_edges(), nf = F.n_faces(), ni = F.n_in();
    if (counts3) {
        counts3[0] = ne;
        counts3[1] = nf;
        counts3[2] = ni;
    }
    std::vector<double> ce(size_t(rh::kMeshEdge * ne)), cf(size_t(rh::kMeshFace * nf)),
        ci(size_t(rh::kMeshIn * ni));
    for (int64_t i = 0; i < ne; i++) {
        double w[3] = {0, 0, 0};
        for (int64_t k = F.edge_inc_off[i]; k < F.edge_inc_off[i + 1]; k++) {
            const double* W = F.edge_inc_w.data() + 9 * k;
            const double* m = M + 3 * F.edge_inc_elem[k];
            w[0] += W[0] * m[0] + W[1] * m[1] + W[2] * m[2];
            w[1] += W[3] * m[0] + W[4] * m[1] + W[5] * m[2];
            w[2] += W[6] * m[0] + W[7] * m[1] + W[8] * m[2];
        }
        double* o = ce.data() + rh::kMeshEdge * i;
        for (int k = 0; k < rh::kMeshEdgeGeo; k++) o[k] = F.edge_geo[rh::kMeshEdgeGeo * i + k];
        o[10] = w[0];
        o[11] = w[1];
        o[12] = w[2];
    }
    for (int64_t i = 0; i < nf; i++) {
        const double* g = F.face_geo.data() + rh::kMeshFaceGeo * i;
        double s = 0;
        for (int64_t k = F.face_inc_off[i]; k < F.face_inc_off[i + 1]; k++) {
            const double* m = M + 3 * F.face_inc_elem[k];
            const double* v = F.face_inc_v.data() + 3 * k;
            s += v[0] * m[0] + v[1] * m[1] + v[2] * m[2];
        }
        double* o = cf.data() + rh::kMeshFace * i;
        for (int k = 0; k < rh::kMeshFaceGeo; k++) o[k] = g[k];
        o[18] = s * g[12];
        o[19] = s * g[13];
        o[20] = s * g[14];
    }
    for (int64_t i = 0; i < ni; i++) {
        const double* g = F.in_geo.data() + (1 + 4 * rh::kMeshPlanes) * i;
        double* o = ci.data() + rh::kMeshIn * i;
        o[0] = g[0];
        const double* Fc = F.in_F.data() + 9 * i;
        const double* m = M + 3 * F.in_elem[i];
        for (int k = 0; k < 3; k++) o[1 + k] = Fc[3 * k] * m[0] + Fc[3 * k + 1] * m[1] + Fc[3 * k + 2] * m[2];
        for (int k = 0; k < 4 * rh::kMeshPlanes; k++) o[4 + k] = g[1 + k];
    }
    // blocks through the per-element path
    rh::ModelView mv = view(P);
    for (int64_t p = 0; p < np; p++) {
        const double* x = pts + 3 * p;
        double h[3] = {0, 0, 0}, b[3] = {0, 0, 0}, mi[3] = {0, 0, 0};
        for (int64_t i = 0; i < ne; i++) {
            const double* r = ce.data() + rh::kMeshEdge * i;
            const double L = fp32 ? rh::mesh_edge_L<float>(r, x) : rh::mesh_edge_L<double>(r, x);
            for (int k = 0; k < 3; k++) h[k] += L * r[10 + k];
        }
        for (int64_t i = 0; i < nf; i++) {
            const double* r = cf.data() + rh::kMeshFace * i;
            const double om = fp32 ? rh::mesh_face_omega<float>(r, x) : rh::mesh_face_omega<double>(r, x);
            for (int k = 0; k < 3; k++) h[k] -= om * r[18 + k];
        }
        for (int64_t i = 0; i < ni; i++) {
            const double* r = ci.data() + rh::kMeshIn * i;
            if (rh::mesh_inside(r, x))
                for (int k = 0; k < 3; k++) mi[k] += r[1 + k];
        }
        for (int k = 0; k < 3; k++) {
            h[k] *= rh::kInv4Pi;
            b[k] = h[k] + mi[k];
        }
        double hb[3] = {0, 0, 0}, bb[3] = {0, 0, 0};
        rh::field_add<double, true, 0>(mv, M, P.kind_idx.data(), P.kind_off[0], P.kind_off[1], x, bb, hb);
        for (int k = 0; k < 3; k++) {
            B[3 * p + k] = b[k] + bb[k];
            H[3 * p + k] = h[k] + hb[k];
        }
    }
    return RH_OK;
}

// Sweep-order statistics of the mesh form: out[0..6] = unique edges, edges
// swept on their own (owned by no face), unique faces, triangles,
// continued triangles, quadrilaterals, continued quadrilaterals.  Returns 2
// when the model has no mesh form.
int rhc_mesh_stats(const rh_model_desc* d, int64_t* out)
{
    rh::PackedModel P;
    rh_status st = rh::pack_model(d, &P, &g_err);
    if (st != RH_OK) return st;
    const rh::MeshForm& F = P.mesh;
    if (!F.active) return 2;
    int64_t cont[4] = {0, 0, 0, 0};
    for (double c : F.face_cont) cont[int(c) & 3]++;
    out[0] = F.n_edges();
    out[1] = F.ne_free;
    out[2] = F.n_faces();
    out[3] = F.nf_tri;
    out[4] = cont[1];
    out[5] = F.n_faces() - F.nf_tri;
    out[6] = cont[3];
    return RH_OK;
}

}  // extern "C"
