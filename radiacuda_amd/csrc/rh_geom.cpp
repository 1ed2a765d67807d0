// rh_geom.cpp -- packing of blocks and convex polyhedra into kernel records.
//
// Polyhedron preparation restates what the reference does when it builds a
// polyhedron (cpp/src/core/radvlpgn.cpp:98-345): face normals, outward
// orientation, the convexity check (Radia::Error106) and the centroid as the
// mean of the given vertices (radvlpgn.h:321-327).  The per-edge tables are
// this implementation's own (rh_math.hpp).
#include "rh_geom.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>

#include "rh_math.hpp"
#include "rh_mesh.hpp"

namespace rh {

void inv3(const double* m, double* o)
{
    double c[9] = {m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                   m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                   m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
    double det = m[0] * c[0] + m[1] * c[3] + m[2] * c[6];
    double inv = 1.0 / det;
    for (int i = 0; i < 9; i++) o[i] = c[i] * inv;
}

namespace {

struct Vec {
    double x, y, z;
};
Vec sub(const Vec& a, const Vec& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec crs(const Vec& a, const Vec& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double dt(const Vec& a, const Vec& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double nrm(const Vec& a) { return std::sqrt(dt(a, a)); }
Vec scl(const Vec& a, double s) { return {a.x * s, a.y * s, a.z * s}; }

// Accumulates m n^T (symmetrized) into w6 = xx yy zz xy xz yz.
void add_mn(double* w6, const Vec& m, const Vec& n)
{
    w6[0] += m.x * n.x;
    w6[1] += m.y * n.y;
    w6[2] += m.z * n.z;
    w6[3] += 0.5 * (m.x * n.y + m.y * n.x);
    w6[4] += 0.5 * (m.x * n.z + m.z * n.x);
    w6[5] += 0.5 * (m.y * n.z + m.z * n.y);
}

struct Face {
    std::vector<int> v;  // counter-clockwise about n
    Vec n;
};

// Normals (Newell), outward orientation and convexity of one polyhedron.
rh_status orient_faces(const std::vector<Vec>& V, std::vector<Face>& F, std::string* err)
{
    Vec c{0, 0, 0};
    for (const Vec& p : V) c = {c.x + p.x, c.y + p.y, c.z + p.z};
    c = scl(c, 1.0 / double(V.size()));
    double size = 0;
    for (const Vec& p : V) size = std::max(size, nrm(sub(p, c)));
    const double tol = 1e-9 * size;
    for (Face& f : F) {
        Vec n{0, 0, 0};
        const size_t k = f.v.size();
        for (size_t i = 0; i < k; i++) {
            const Vec& a = V[f.v[i]];
            const Vec& b = V[f.v[(i + 1) % k]];
            n.x += (a.y - b.y) * (a.z + b.z);
            n.y += (a.z - b.z) * (a.x + b.x);
            n.z += (a.x - b.x) * (a.y + b.y);
        }
        double ln = nrm(n);
        if (!(ln > 1e-300)) {
            *err = "degenerate polyhedron face (zero area)";
            return RH_ERR_GEOMETRY;
        }
        n = scl(n, 1.0 / ln);
        // planarity (Radia::Error047)
        for (size_t i = 1; i < k; i++) {
            if (std::fabs(dt(sub(V[f.v[i]], V[f.v[0]]), n)) > 1e-6 * size) {
                *err = "polyhedron face is not planar";
                return RH_ERR_GEOMETRY;
            }
        }
        if (dt(sub(V[f.v[0]], c), n) < 0) {
            n = scl(n, -1.0);
            std::reverse(f.v.begin(), f.v.end());
        }
        f.n = n;
        // convexity (Radia::Error106): every vertex on the inner side, to the
        // planarity tolerance
        for (const Vec& p : V) {
            if (dt(sub(p, V[f.v[0]]), n) > tol + 1e-6 * size) {
                *err = "polyhedron is not convex";
                return RH_ERR_GEOMETRY;
            }
        }
    }
    return RH_OK;
}

void append(std::vector<double>& d, const Vec& v)
{
    d.push_back(v.x);
    d.push_back(v.y);
    d.push_back(v.z);
}

// Canonical tetrahedron record (rh_math.hpp: kTetSize doubles).
void pack_tet(std::vector<Vec> V, std::vector<double>& d)
{
    if (dt(crs(sub(V[1], V[0]), sub(V[2], V[0])), sub(V[3], V[0])) < 0) std::swap(V[2], V[3]);
    for (int v = 0; v < 4; v++) append(d, V[v]);
    Vec n[4];
    for (int f = 0; f < 4; f++) {
        const int* fv = kTetFace[f];
        Vec c = crs(sub(V[fv[1]], V[fv[0]]), sub(V[fv[2]], V[fv[0]]));
        n[f] = scl(c, 1.0 / nrm(c));
    }
    for (int e = 0; e < 6; e++) {
        const int a = kTetEdgeA[e], b = kTetEdgeB[e];
        Vec t = sub(V[b], V[a]);
        double len = nrm(t);
        t = scl(t, 1.0 / len);
        double w6[6] = {0, 0, 0, 0, 0, 0};
        for (int f = 0; f < 4; f++) {
            const int* fv = kTetFace[f];
            for (int i = 0; i < 3; i++) {
                int p = fv[i], q = fv[(i + 1) % 3];
                if ((p == a && q == b) || (p == b && q == a)) {
                    Vec tf = (p == a) ? t : scl(t, -1.0);
                    add_mn(w6, crs(tf, n[f]), n[f]);
                }
            }
        }
        append(d, t);
        d.push_back(len);
        for (int k = 0; k < 6; k++) d.push_back(w6[k]);
    }
    for (int f = 0; f < 4; f++) {
        const int* fv = kTetFace[f];
        Vec c = crs(sub(V[fv[1]], V[fv[0]]), sub(V[fv[2]], V[fv[0]]));
        append(d, n[f]);
        d.push_back(nrm(c));  // twice the face area
    }
}

void pack_gen(const std::vector<Vec>& V, const std::vector<Face>& F, std::vector<double>& d)
{
    // unique undirected edges with their W
    std::map<std::pair<int, int>, int> eid;
    std::vector<std::pair<int, int>> edges;
    std::vector<std::array<double, 6>> W;
    for (const Face& f : F) {
        const size_t k = f.v.size();
        for (size_t i = 0; i < k; i++) {
            int p = f.v[i], q = f.v[(i + 1) % k];
            auto key = std::make_pair(std::min(p, q), std::max(p, q));
            auto it = eid.find(key);
            int id;
            if (it == eid.end()) {
                id = int(edges.size());
                eid[key] = id;
                edges.push_back(key);
                W.push_back({0, 0, 0, 0, 0, 0});
            } else {
                id = it->second;
            }
            Vec tf = sub(V[q], V[p]);
            tf = scl(tf, 1.0 / nrm(tf));
            add_mn(W[id].data(), crs(tf, f.n), f.n);
        }
    }
    int nfv = 0;
    for (const Face& f : F) nfv += int(f.v.size());
    d.push_back(double(V.size()));
    d.push_back(double(edges.size()));
    d.push_back(double(F.size()));
    d.push_back(double(nfv));
    for (const Vec& v : V) append(d, v);
    for (size_t e = 0; e < edges.size(); e++) {
        int a = edges[e].first, b = edges[e].second;
        Vec t = sub(V[b], V[a]);
        double len = nrm(t);
        t = scl(t, 1.0 / len);
        d.push_back(double(a));
        d.push_back(double(b));
        append(d, t);
        d.push_back(len);
        for (int k = 0; k < 6; k++) d.push_back(W[e][k]);
    }
    int b0 = 0;
    for (const Face& f : F) {
        append(d, f.n);
        d.push_back(double(b0));
        d.push_back(double(f.v.size()));
        b0 += int(f.v.size());
    }
    for (const Face& f : F)
        for (int v : f.v) d.push_back(double(v));
}

}  // namespace

rh_status pack_model(const rh_model_desc* d, PackedModel* out, std::string* err)
{
    if (!d || d->n_elem < 0) {
        *err = "null model description";
        return RH_ERR_ARG;
    }
    const int64_t n = d->n_elem;
    PackedModel& P = *out;
    P = PackedModel();
    P.n = n;
    P.kind.resize(n);
    P.off.resize(n);
    P.copy_off.assign(d->copy_off, d->copy_off + n + 1);
    P.colloc.resize(3 * n);
    P.rowt.resize(9 * n);
    std::vector<double> cent(3 * n);
    const int64_t nf_all = n > 0 ? d->face_off[n] : 0;
    std::vector<int64_t> fstart(nf_all + 1, 0);
    for (int64_t f = 0; f < nf_all; f++) fstart[f + 1] = fstart[f] + d->face_len[f];
    P.adj_off.assign(size_t(n + 1), 0);
    P.vert_off.assign(size_t(n + 1), 0);
    for (int64_t e = 0; e < n; e++) {
        P.off[e] = int64_t(P.data.size());
        P.adj_off[e] = int64_t(P.adj.size());
        P.vert_off[e + 1] = P.vert_off[e];
        if (d->elem_type[e] == 0) {
            P.adj.push_back(0.0);
            P.adj.push_back(0.0);
            P.kind[e] = EK_REC;
            const double* c = d->rec_center + 3 * e;
            const double* w = d->rec_dims + 3 * e;
            if (!(w[0] > 0 && w[1] > 0 && w[2] > 0)) {
                *err = "block with non-positive dimension";
                return RH_ERR_GEOMETRY;
            }
            for (int k = 0; k < 3; k++) P.data.push_back(c[k]);
            for (int k = 0; k < 3; k++) P.data.push_back(0.5 * w[k]);
            for (int k = 0; k < 3; k++) cent[3 * e + k] = c[k];
            P.counts[5]++;
            continue;
        }
        if (d->elem_type[e] != 1) {
            *err = "unknown element type";
            return RH_ERR_UNSUPPORTED;
        }
        const int64_t v0 = d->vert_off[e], v1 = d->vert_off[e + 1];
        std::vector<Vec> V;
        for (int64_t v = v0; v < v1; v++)
            V.push_back({d->verts[3 * v], d->verts[3 * v + 1], d->verts[3 * v + 2]});
        // centroid: mean of the vertices in the order given
        Vec s{0, 0, 0};
        for (const Vec& p : V) s = {s.x + p.x, s.y + p.y, s.z + p.z};
        double b = 1.0 / double(V.size());
        cent[3 * e] = b * s.x;
        cent[3 * e + 1] = b * s.y;
        cent[3 * e + 2] = b * s.z;
        // faces of this element (face_idx is flat over all faces of all elements)
        std::vector<Face> F;
        for (int64_t f = d->face_off[e]; f < d->face_off[e + 1]; f++) {
            Face fc;
            if (d->face_len[f] < 3) {
                *err = "face with fewer than 3 vertices";
                return RH_ERR_GEOMETRY;
            }
            for (int k = 0; k < d->face_len[f]; k++) {
                int vi = d->face_idx[fstart[f] + k];
                if (vi < 0 || vi >= int(V.size())) {
                    *err = "face vertex index out of range";
                    return RH_ERR_ARG;
                }
                fc.v.push_back(vi);
            }
            F.push_back(fc);
        }
        rh_status st = orient_faces(V, F, err);
        if (st != RH_OK) return st;
        {
            std::vector<double> tri;
            for (const Face& f : F)
                for (size_t k = 1; k + 1 < f.v.size(); k++) {
                    tri.push_back(double(f.v[0]));
                    tri.push_back(double(f.v[k]));
                    tri.push_back(double(f.v[k + 1]));
                }
            P.adj.push_back(double(V.size()));
            P.adj.push_back(double(tri.size() / 3));
            for (const Vec& v : V) append(P.adj, v);
            P.adj.insert(P.adj.end(), tri.begin(), tri.end());
            P.vert_off[e + 1] = P.vert_off[e] + int64_t(V.size());
        }
        bool tet = (V.size() == 4 && F.size() == 4);
        for (const Face& f : F) tet = tet && (f.v.size() == 3);
        if (tet) {
            P.kind[e] = EK_TET;
            pack_tet(V, P.data);
            P.counts[2]++;
            P.counts[6] += 6;
            P.counts[7] += 4;
        } else {
            P.kind[e] = EK_GEN;
            // A face that is planar only to the input tolerance is replaced by its
            // fan of triangles (each with its own normal): the field is then that
            // of the closed triangulated surface, which is smooth in the vertex
            // positions and equals the face form when the face is planar.
            {
                double size = 0;
                for (const Vec& p : V) size = std::max(size, nrm(sub(p, V[0])));
                std::vector<Face> F2;
                for (const Face& f : F) {
                    bool flat = true;
                    for (size_t i = 1; i < f.v.size(); i++)
                        flat = flat && std::fabs(dt(sub(V[f.v[i]], V[f.v[0]]), f.n)) <= 1e-12 * size;
                    if (flat || f.v.size() == 3) {
                        F2.push_back(f);
                        continue;
                    }
                    for (size_t k = 1; k + 1 < f.v.size(); k++) {
                        Face t;
                        t.v = {f.v[0], f.v[k], f.v[k + 1]};
                        Vec c = crs(sub(V[t.v[1]], V[t.v[0]]), sub(V[t.v[2]], V[t.v[0]]));
                        t.n = scl(c, 1.0 / nrm(c));
                        F2.push_back(t);
                    }
                }
                F.swap(F2);
            }
            size_t before = P.data.size();
            pack_gen(V, F, P.data);
            P.counts[4]++;
            P.counts[6] += int64_t(P.data[before + 1]);
            P.counts[7] += int64_t(F.size());
        }
    }
    P.adj_off[n] = int64_t(P.adj.size());
    // element lists per kind slot: 0 blocks, 1 tetrahedra, 2 general polyhedra
    for (int slot = 0; slot < 3; slot++) {
        const int32_t want = (slot == 0) ? EK_REC : (slot == 1 ? EK_TET : EK_GEN);
        P.kind_off[slot] = int64_t(P.kind_idx.size());
        for (int64_t e = 0; e < n; e++)
            if (P.kind[e] == want) P.kind_idx.push_back(int32_t(e));
    }
    P.kind_off[3] = int64_t(P.kind_idx.size());
    // copies
    const int64_t C = d->copy_off[n];
    P.counts[0] = n;
    P.counts[1] = C;
    P.copies.assign(size_t(kCopySize * C), 0.0);
    for (int64_t c = 0; c < C; c++) {
        const double* m = d->copy_m + 9 * c;
        double* o = P.copies.data() + kCopySize * c;
        inv3(m, o);
        for (int k = 0; k < 3; k++) o[9 + k] = d->copy_v[3 * c + k];
        for (int k = 0; k < 9; k++) o[12 + k] = d->copy_s[c] * m[k];
        bool ident = d->copy_s[c] == 1.0 && d->copy_v[3 * c] == 0 && d->copy_v[3 * c + 1] == 0 &&
                     d->copy_v[3 * c + 2] == 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) ident = ident && (m[3 * i + j] == (i == j ? 1.0 : 0.0));
        if (!ident) P.sym = true;
    }
    for (int64_t e = 0; e < n; e++) {
        if (d->copy_off[e + 1] - d->copy_off[e] != 1) P.sym = true;
        if (d->copy_off[e + 1] <= d->copy_off[e]) {
            *err = "element without a transform copy";
            return RH_ERR_ARG;
        }
        const int64_t c = d->copy_off[e];
        const double* m = d->copy_m + 9 * c;
        const double* v = d->copy_v + 3 * c;
        const double* ce = cent.data() + 3 * e;
        for (int i = 0; i < 3; i++)
            P.colloc[3 * e + i] = m[3 * i] * ce[0] + m[3 * i + 1] * ce[1] + m[3 * i + 2] * ce[2] + v[i];
        double mi[9];
        inv3(m, mi);
        for (int k = 0; k < 9; k++) {
            P.rowt[9 * e + k] = d->copy_s[c] * mi[k];
            if (P.rowt[9 * e + k] != ((k % 4 == 0) ? 1.0 : 0.0)) P.rowt_identity = false;
        }
    }
    build_mesh_form(P);
    return RH_OK;
}

namespace {
using VKey = std::array<uint64_t, 3>;

VKey vkey(const double* v)
{
    VKey k;
    std::memcpy(k.data(), v, 3 * sizeof(double));
    return k;
}

// One polyhedron seen through its packed record: vertices, edges (vertex
// pair and W6), faces (vertex loop CCW about the outward normal, n).
struct ElemView {
    int nv = 0;
    const double* verts = nullptr;
    std::vector<std::array<int, 2>> edges;
    std::vector<const double*> edge_w;
    std::vector<std::vector<int>> faces;
    std::vector<const double*> face_n;
};

ElemView view_elem(const PackedModel& P, int64_t j)
{
    ElemView v;
    const double* rec = P.data.data() + P.off[j];
    if (P.kind[j] == EK_TET) {
        v.nv = 4;
        v.verts = rec;
        for (int e = 0; e < 6; e++) {
            v.edges.push_back({kTetEdgeA[e], kTetEdgeB[e]});
            v.edge_w.push_back(rec + 12 + kTetEdge * e + 4);
        }
        for (int f = 0; f < 4; f++) {
            v.faces.push_back({kTetFace[f][0], kTetFace[f][1], kTetFace[f][2]});
            v.face_n.push_back(rec + 12 + 6 * kTetEdge + kTetFaceRec * f);
        }
        return v;
    }
    const int nv = int(rec[0]), ne = int(rec[1]), nf = int(rec[2]);
    v.nv = nv;
    v.verts = rec + kGenHdr;
    const double* E = v.verts + 3 * nv;
    const double* F = E + kGenEdge * ne;
    const double* fv = F + kGenFace * nf;
    for (int e = 0; e < ne; e++) {
        v.edges.push_back({int(E[kGenEdge * e]), int(E[kGenEdge * e + 1])});
        v.edge_w.push_back(E + kGenEdge * e + 6);
    }
    for (int f = 0; f < nf; f++) {
        const double* fr = F + kGenFace * f;
        std::vector<int> loop;
        for (int k = 0; k < int(fr[4]); k++) loop.push_back(int(fv[int(fr[3]) + k]));
        v.faces.push_back(loop);
        v.face_n.push_back(fr);
    }
    return v;
}
}  // namespace

void build_mesh_form(PackedModel& P)
{
    MeshForm& M = P.mesh;
    M = MeshForm();
    if (P.kind_off[3] - P.kind_off[1] < 2) return;
    std::vector<int64_t> elems(P.kind_idx.begin() + P.kind_off[1], P.kind_idx.begin() + P.kind_off[3]);
    std::vector<ElemView> views;
    for (int64_t j : elems) {
        views.push_back(view_elem(P, j));
        if (int(views.back().faces.size()) > kMeshPlanes) return;
    }
    // Every element copy is a mesh element in the laboratory frame: copy c
    // maps points x -> M_c x + V_c and fields by F_c = s_c M_c (M_c
    // orthogonal), so the copy's H is F_c times the element's H at the mapped
    // point.  L_e and a face's solid angle (signed by its transformed normal)
    // are unchanged by the isometry, hence an edge incidence weighs M_j by
    // F_c W_e and a face incidence adds (F_c^T n) . M_j to the charge of a
    // face with normal n.  Without copies F_c = I and the weights are W_e and
    // n, exactly.
    struct Inst {
        int32_t j;
        const double* F;               // F_c, row-major
        std::vector<double> verts;     // laboratory frame
        std::vector<double> normals;   // laboratory frame, 3 per face
        const ElemView* v;
    };
    std::vector<Inst> inst;
    for (size_t q = 0; q < elems.size(); q++) {
        const int64_t j = elems[q];
        const ElemView& v = views[q];
        for (int64_t c = P.copy_off[size_t(j)]; c < P.copy_off[size_t(j) + 1]; c++) {
            const double* cp = P.copies.data() + kCopySize * c;  // Minv, V, F
            double Mc[9];
            inv3(cp, Mc);
            Inst I;
            I.j = int32_t(j);
            I.F = cp + 12;
            I.v = &v;
            for (int k = 0; k < v.nv; k++) {
                const double* x = v.verts + 3 * k;
                for (int r = 0; r < 3; r++)
                    // + 0.0: a vertex mirrored onto a symmetry plane gets +0, not -0,
                    // so that it keys equal to the original
                    I.verts.push_back(Mc[3 * r] * x[0] + Mc[3 * r + 1] * x[1] + Mc[3 * r + 2] * x[2] + cp[9 + r] + 0.0);
            }
            for (size_t f = 0; f < v.faces.size(); f++) {
                const double* n = v.face_n[f];
                double nl[3];
                for (int r = 0; r < 3; r++) nl[r] = cp[r] * n[0] + cp[3 + r] * n[1] + cp[6 + r] * n[2];
                const double ln = std::sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
                for (int r = 0; r < 3; r++) I.normals.push_back(nl[r] / ln);
            }
            inst.push_back(std::move(I));
        }
    }
    std::map<VKey, int32_t> vid;
    std::vector<const double*> vpos;
    std::map<std::pair<int32_t, int32_t>, int64_t> eid;
    std::vector<std::vector<std::pair<int32_t, std::array<double, 9>>>> einc;
    std::map<std::vector<int32_t>, int64_t> fid;
    std::vector<std::vector<std::pair<int32_t, std::array<double, 3>>>> finc;
    std::vector<std::vector<int32_t>> floop;  // unique-vertex loop of the first incidence
    std::vector<const double*> fnorm;
    for (const Inst& I : inst) {
        const ElemView& v = *I.v;
        const double* F = I.F;
        std::vector<int32_t> u(size_t(v.nv));
        for (int k = 0; k < v.nv; k++) {
            const double* x = I.verts.data() + 3 * k;
            auto it = vid.find(vkey(x));
            if (it == vid.end()) {
                it = vid.emplace(vkey(x), int32_t(vpos.size())).first;
                vpos.push_back(x);
            }
            u[size_t(k)] = it->second;
        }
        for (size_t e = 0; e < v.edges.size(); e++) {
            int32_t a = u[size_t(v.edges[e][0])], b = u[size_t(v.edges[e][1])];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = eid.find(key);
            if (it == eid.end()) {
                it = eid.emplace(key, int64_t(einc.size())).first;
                einc.emplace_back();
            }
            // F W with W symmetric (xx yy zz xy xz yz)
            const double* w = v.edge_w[e];
            const double W[9] = {w[0], w[3], w[4], w[3], w[1], w[5], w[4], w[5], w[2]};
            std::array<double, 9> FW;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    FW[size_t(3 * r + c)] = F[3 * r] * W[c] + F[3 * r + 1] * W[3 + c] + F[3 * r + 2] * W[6 + c];
            einc[size_t(it->second)].push_back({I.j, FW});
        }
        for (size_t f = 0; f < v.faces.size(); f++) {
            const std::vector<int>& loop = v.faces[f];
            const double* n1 = I.normals.data() + 3 * f;
            // polygons of more than 4 vertices become fan pieces of at most 4
            for (size_t s = 0; s + 2 < loop.size(); s += 2) {
                std::vector<int32_t> piece{u[size_t(loop[0])], u[size_t(loop[s + 1])],
                                           u[size_t(loop[s + 2])]};
                if (s + 3 < loop.size()) piece.push_back(u[size_t(loop[s + 3])]);
                std::vector<int32_t> key = piece;
                std::sort(key.begin(), key.end());
                auto it = fid.find(key);
                double sign = 1.0;
                if (it == fid.end()) {
                    it = fid.emplace(key, int64_t(finc.size())).first;
                    finc.emplace_back();
                    floop.push_back(piece);
                    fnorm.push_back(n1);
                } else {
                    const double* n0 = fnorm[size_t(it->second)];
                    sign = (n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2] > 0) ? 1.0 : -1.0;
                }
                // sign F^T n of the unique face's normal n
                const double* n = fnorm[size_t(it->second)];
                std::array<double, 3> gv;
                for (int c = 0; c < 3; c++) gv[size_t(c)] = sign * (F[c] * n[0] + F[3 + c] * n[1] + F[6 + c] * n[2]);
                finc[size_t(it->second)].push_back({I.j, gv});
            }
        }
    }
    // the regrouped sweep pays off only when faces are shared: a face
    // evaluated on its own recomputes the vertex distances the per-element
    // sweep shares between the faces of an element
    int64_t face_incidences = 0;
    for (const auto& f : finc) face_incidences += int64_t(f.size());
    if (double(finc.size()) > 0.8 * double(face_incidences)) return;
    // Edge paths: each unique edge oriented and ordered so that, where a path
    // continues, its first vertex is the previous edge's second vertex (L_e
    // and W_e do not depend on the orientation).  Triangle strips: where a
    // strip continues, a triangle's first two vertices are the previous
    // triangle's last two (the solid angle of a triangle does not depend on
    // the order of its vertices once the normal is fixed).  The sweeps then
    // compute one vertex distance per continued edge or triangle.
    std::vector<int32_t> forder;
    std::vector<std::vector<int32_t>> fverts;
    std::vector<char> fcont;
    {
        std::map<std::pair<int32_t, int32_t>, std::vector<int32_t>> tri_at;
        for (size_t f = 0; f < floop.size(); f++) {
            if (floop[f].size() != 3) continue;
            for (int k = 0; k < 3; k++) {
                const int32_t a = floop[f][size_t(k)], b = floop[f][size_t((k + 1) % 3)];
                tri_at[{std::min(a, b), std::max(a, b)}].push_back(int32_t(f));
            }
        }
        std::vector<char> used(floop.size(), 0);
        auto other = [&](int32_t a, int32_t b) -> int32_t {
            auto it = tri_at.find({std::min(a, b), std::max(a, b)});
            if (it == tri_at.end()) return -1;
            for (int32_t g : it->second)
                if (!used[size_t(g)]) return g;
            return -1;
        };
        for (size_t f0 = 0; f0 < floop.size(); f0++) {
            if (used[f0] || floop[f0].size() != 3) continue;
            std::vector<int32_t> v = floop[f0];
            used[f0] = 1;
            // start with the rotation whose last edge continues the strip
            for (int r = 0; r < 3; r++) {
                if (other(v[1], v[2]) >= 0) break;
                std::rotate(v.begin(), v.begin() + 1, v.end());
            }
            forder.push_back(int32_t(f0));
            fverts.push_back(v);
            fcont.push_back(0);
            // among the triangles on (p, q), take one whose far edge (q, r)
            // still continues the strip, preferring the one with the fewest
            // continuations (strip starts on the C-yoke: 13 % -> 9 %)
            auto third = [&](int32_t g, int32_t a, int32_t b) {
                int32_t r = -1;
                for (int32_t x : floop[size_t(g)])
                    if (x != a && x != b) r = x;
                return r;
            };
            auto next = [&](int32_t a, int32_t b) -> int32_t {
                auto it = tri_at.find({std::min(a, b), std::max(a, b)});
                if (it == tri_at.end()) return -1;
                int32_t best = -1;
                int best_n = 0;
                for (int32_t g : it->second) {
                    if (used[size_t(g)]) continue;
                    const int32_t r = third(g, a, b);
                    int n = 0;
                    auto jt = tri_at.find({std::min(b, r), std::max(b, r)});
                    if (jt != tri_at.end())
                        for (int32_t h : jt->second)
                            if (h != g && !used[size_t(h)]) n++;
                    // score: continuing candidates first, then fewer continuations
                    const bool better = best < 0 || (n > 0 && (best_n == 0 || n < best_n));
                    if (better) {
                        best = g;
                        best_n = n;
                    }
                }
                return best;
            };
            int32_t p = v[1], q = v[2];
            for (int32_t g = next(p, q); g >= 0; g = next(p, q)) {
                used[size_t(g)] = 1;
                const int32_t r = third(g, p, q);
                forder.push_back(g);
                fverts.push_back({p, q, r});
                fcont.push_back(1);
                p = q;
                q = r;
            }
        }
        M.nf_tri = int64_t(forder.size());
        // Quadrilateral strips: where a strip continues, a quadrilateral's
        // first two vertices are the previous one's last two, its loop read
        // from there in either direction.
        std::map<std::pair<int32_t, int32_t>, std::vector<int32_t>> quad_at;
        for (size_t f = 0; f < floop.size(); f++) {
            if (floop[f].size() != 4) continue;
            for (int k = 0; k < 4; k++) {
                const int32_t a = floop[f][size_t(k)], b = floop[f][size_t((k + 1) % 4)];
                quad_at[{std::min(a, b), std::max(a, b)}].push_back(int32_t(f));
            }
        }
        auto next_quad = [&](int32_t a, int32_t b) -> int32_t {
            auto it = quad_at.find({std::min(a, b), std::max(a, b)});
            if (it == quad_at.end()) return -1;
            for (int32_t g : it->second)
                if (!used[size_t(g)]) return g;
            return -1;
        };
        for (size_t f0 = 0; f0 < floop.size(); f0++) {
            if (used[f0] || floop[f0].size() != 4) continue;
            std::vector<int32_t> v = floop[f0];
            used[f0] = 1;
            // start with the rotation whose side (v2, v3) continues the strip
            for (int r = 0; r < 4; r++) {
                if (next_quad(v[2], v[3]) >= 0) break;
                std::rotate(v.begin(), v.begin() + 1, v.end());
            }
            forder.push_back(int32_t(f0));
            fverts.push_back(v);
            fcont.push_back(2);  // quadrilateral
            int32_t p = v[2], q = v[3];
            for (int32_t g = next_quad(p, q); g >= 0; g = next_quad(p, q)) {
                used[size_t(g)] = 1;
                const std::vector<int32_t>& l = floop[size_t(g)];
                int i = 0;
                while (l[size_t(i)] != p) i++;
                // around the loop from p through q
                const int dir = (l[size_t((i + 1) % 4)] == q) ? 1 : 3;
                std::vector<int32_t> w(4);
                for (int k = 0; k < 4; k++) w[size_t(k)] = l[size_t((i + dir * k) % 4)];
                forder.push_back(g);
                fverts.push_back(w);
                fcont.push_back(3);  // quadrilateral continuing the strip
                p = w[2];
                q = w[3];
            }
        }
        for (size_t f = 0; f < floop.size(); f++)
            if (!used[f]) {
                forder.push_back(int32_t(f));
                fverts.push_back(floop[f]);
                fcont.push_back(2);
            }
    }
    // Each unique edge that is a side of a face is owned by the first face in
    // sweep order that has it (a triangle any of its sides, a quadrilateral
    // its sides (v0,v1), (v1,v2), (v2,v3)): that face's sweep step already
    // holds both vertex offsets, distances and their dot product.
    std::vector<std::pair<int32_t, int32_t>> ekeys(einc.size());
    for (const auto& kv : eid) ekeys[size_t(kv.second)] = kv.first;
    std::vector<int64_t> owner(einc.size(), -1);  // 3 * face position + side
    for (size_t o = 0; o < forder.size(); o++) {
        const size_t nv = fverts[o].size();
        for (int k = 0; k < 3; k++) {
            const int32_t a = fverts[o][size_t(k)], b = fverts[o][size_t(k + 1) % nv];
            const auto it = eid.find({std::min(a, b), std::max(a, b)});
            if (it == eid.end()) continue;  // a diagonal of a fanned polygon
            if (owner[size_t(it->second)] < 0) owner[size_t(it->second)] = int64_t(3 * o + k);
        }
    }
    std::vector<int32_t> eorder;
    std::vector<char> eflip, econt;
    {
        std::vector<std::vector<int32_t>> vadj(vpos.size());
        for (size_t e = 0; e < ekeys.size(); e++) {
            vadj[size_t(ekeys[e].first)].push_back(int32_t(e));
            vadj[size_t(ekeys[e].second)].push_back(int32_t(e));
        }
        std::vector<char> used(ekeys.size(), 0);
        std::vector<size_t> next(vpos.size(), 0);
        auto take = [&](int32_t v) -> int32_t {
            std::vector<int32_t>& a = vadj[size_t(v)];
            while (next[size_t(v)] < a.size() && used[size_t(a[next[size_t(v)]])]) next[size_t(v)]++;
            return next[size_t(v)] < a.size() ? a[next[size_t(v)]] : -1;
        };
        for (size_t e = 0; e < ekeys.size(); e++)
            if (owner[e] >= 0) used[e] = 1;  // evaluated by its face
        for (size_t e0 = 0; e0 < ekeys.size(); e0++) {
            if (used[e0]) continue;
            int32_t cur = ekeys[e0].first, e = int32_t(e0);
            bool first = true;
            while (e >= 0) {
                used[size_t(e)] = 1;
                const bool flip = ekeys[size_t(e)].first != cur;
                eorder.push_back(e);
                eflip.push_back(flip ? 1 : 0);
                econt.push_back(first ? 0 : 1);
                first = false;
                cur = flip ? ekeys[size_t(e)].first : ekeys[size_t(e)].second;
                e = take(cur);
            }
        }
        M.ne_free = int64_t(eorder.size());
        for (size_t e = 0; e < ekeys.size(); e++)
            if (owner[e] >= 0) {
                eorder.push_back(int32_t(e));
                eflip.push_back(0);
                econt.push_back(0);
            }
    }
    // edges
    M.edge_inc_off.push_back(0);
    for (size_t o = 0; o < eorder.size(); o++) {
        const size_t e = size_t(eorder[o]);
        const double* a = vpos[size_t(eflip[o] ? ekeys[e].second : ekeys[e].first)];
        const double* b = vpos[size_t(eflip[o] ? ekeys[e].first : ekeys[e].second)];
        double t[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double len = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int k = 0; k < 3; k++) t[k] /= len;
        for (int k = 0; k < 3; k++) M.edge_geo.push_back(a[k]);
        for (int k = 0; k < 3; k++) M.edge_geo.push_back(b[k]);
        for (int k = 0; k < 3; k++) M.edge_geo.push_back(t[k]);
        M.edge_geo.push_back(len);
        M.edge_cont.push_back(double(econt[o]));
        for (const auto& inc : einc[e]) {
            M.edge_inc_elem.push_back(inc.first);
            for (int k = 0; k < 9; k++) M.edge_inc_w.push_back(inc.second[size_t(k)]);
        }
        M.edge_inc_off.push_back(int64_t(M.edge_inc_elem.size()));
    }
    // owned edges: record index per face side
    M.face_edge.assign(3 * forder.size(), -1);
    for (size_t o = 0; o < eorder.size(); o++) {
        const int64_t w = owner[size_t(eorder[o])];
        if (w >= 0) M.face_edge[size_t(w)] = int64_t(o);
    }
    // faces
    M.face_inc_off.push_back(0);
    for (size_t o = 0; o < forder.size(); o++) {
        const size_t f = size_t(forder[o]);
        const std::vector<int32_t>& loop = fverts[o];
        const double* n = fnorm[f];
        double vv[4][3] = {{0}};
        for (size_t k = 0; k < loop.size(); k++)
            for (int c = 0; c < 3; c++) vv[k][c] = vpos[size_t(loop[k])][c];
        auto area2 = [&](int i0, int i1, int i2) {
            const double e1[3] = {vv[i1][0] - vv[i0][0], vv[i1][1] - vv[i0][1], vv[i1][2] - vv[i0][2]};
            const double e2[3] = {vv[i2][0] - vv[i0][0], vv[i2][1] - vv[i0][1], vv[i2][2] - vv[i0][2]};
            const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                 e1[0] * e2[1] - e1[1] * e2[0]};
            return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        };
        for (int k = 0; k < 4; k++)
            for (int c = 0; c < 3; c++) M.face_geo.push_back(vv[k][c]);
        for (int c = 0; c < 3; c++) M.face_geo.push_back(n[c]);
        M.face_geo.push_back(area2(0, 1, 2));
        M.face_geo.push_back(loop.size() > 3 ? area2(0, 2, 3) : 0.0);
        M.face_geo.push_back(double(loop.size()));
        M.face_cont.push_back(double(fcont[o]));
        for (const auto& inc : finc[f]) {
            M.face_inc_elem.push_back(inc.first);
            for (int k = 0; k < 3; k++) M.face_inc_v.push_back(inc.second[size_t(k)]);
        }
        M.face_inc_off.push_back(int64_t(M.face_inc_elem.size()));
    }
    // inside records (one per element copy, with its field transform) and tile boxes
    for (const Inst& I : inst) {
        const ElemView& v = *I.v;
        M.in_elem.push_back(I.j);
        for (int k = 0; k < 9; k++) M.in_F.push_back(I.F[k]);
        M.in_geo.push_back(double(v.faces.size()));
        for (int f = 0; f < kMeshPlanes; f++) {
            if (f < int(v.faces.size())) {
                const double* n = I.normals.data() + 3 * f;
                const double* p0 = I.verts.data() + 3 * v.faces[size_t(f)][0];
                for (int c = 0; c < 3; c++) M.in_geo.push_back(n[c]);
                M.in_geo.push_back(n[0] * p0[0] + n[1] * p0[1] + n[2] * p0[2]);
            } else {
                for (int c = 0; c < 4; c++) M.in_geo.push_back(0.0);
            }
        }
    }
    for (size_t t0 = 0; t0 < inst.size(); t0 += kMeshTile) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t q = t0; q < std::min(inst.size(), t0 + kMeshTile); q++)
            for (int k = 0; k < inst[q].v->nv; k++)
                for (int c = 0; c < 3; c++) {
                    lo[c] = std::min(lo[c], inst[q].verts[size_t(3 * k + c)]);
                    hi[c] = std::max(hi[c], inst[q].verts[size_t(3 * k + c)]);
                }
        for (int c = 0; c < 3; c++) M.tile_box.push_back(lo[c]);
        for (int c = 0; c < 3; c++) M.tile_box.push_back(hi[c]);
    }
    M.active = true;
}

}  // namespace rh
