// rh_mesh.hpp -- field of a conforming polyhedral mesh as line terms on its
// unique edges and charge sheets on its unique faces.
//
// Summed over the elements of a mesh, the per-element form
//     4 pi H = sum_e L_e W_e M - sum_f Omega_f (n_f . M) n_f
// regroups by mesh entity.  L_e depends only on the edge and the point, so
// every edge is evaluated once with the net weight w_e = sum_{elements at e}
// W_{e,j} M_j.  A face shared by two elements appears with opposite outward
// normals and opposite signed solid angles, so it is evaluated once with the
// net charge g_f = n_f sum_j s_j (n_f . M_j), s_j = +-1 the orientation of
// element j's copy of the face.  The sum is the same; on the C-yoke mesh the
// work per tetrahedron drops from 6 logs and 3 atan2 to about 1.4 and 2.2.
// B = H + M of the element containing the point, found by plane tests.
#pragma once

#include "rh_math.hpp"

namespace rh {

// Charged edge record: a(3), b(3), t(3), len, w(3), cont (1: a is the
// previous record's b).  The FP32 sweep's records hold t(3), len and the
// near distance as floats in slots 6-8.
constexpr int kMeshEdge = 14;
// Charged face record: v0..v3 (12), n(3), 2*area(v0,v1,v2), 2*area(v0,v2,v3),
// nv, g(3), code (0 triangle, 1 triangle whose v0, v1 are the previous
// record's v1, v2, 2 quadrilateral, 3 quadrilateral whose v0, v1 are the
// previous record's v2, v3), then the edges the face owns: for the sides
// (v0,v1), (v1,v2), (v2,v0) of a triangle or (v0,v1), (v1,v2), (v2,v3) of a
// quadrilateral in turn len and w(3) (slots 22-33), and
// the mask of owned sides as an integer in the low word of slot 34; slot 35:
// the FP32 sweep's near distance (a float).  A face's vertices may come in
// either direction of its loop: its solid angle below takes the fan
// triangles' triple products from n and their areas, not from the order.
constexpr int kMeshFace = 36;
constexpr int kMeshFaceEdges = 22;
// Inside-test record: nplanes, M(3), up to 8 planes (n(3), d) with outward n:
// inside iff n.p < d for every plane.
constexpr int kMeshPlanes = 8;
constexpr int kMeshIn = 4 + 4 * kMeshPlanes;  // 36
// Geometry-only sizes (host arrays the charged records are built from).
constexpr int kMeshEdgeGeo = 10;  // a, b, t, len
constexpr int kMeshFaceGeo = 18;  // v0..v3, n, a1, a2, nv

// L_e of one charged edge at p.
template <class R>
RH_HD double mesh_edge_L(const double* __restrict__ rec, const double p[3])
{
    const V3<R> ra = rel<R>(rec, p), rb = rel<R>(rec + 3, p);
    const R Ra = rh_sqrt(dot(ra, ra)), Rb = rh_sqrt(dot(rb, rb));
    return double(edge_L(ra, Ra, rb, Rb, R(rec[9])));
}

// Signed solid angle of one charged face at p: the fan triangles (v0,v1,v2)
// and (v0,v2,v3) multiplied as complex numbers, one atan2 per face.
template <class R>
RH_HD double mesh_face_omega(const double* __restrict__ rec, const double p[3])
{
    const V3<R> r0 = rel<R>(rec, p), r1 = rel<R>(rec + 3, p), r2 = rel<R>(rec + 6, p);
    const R R0 = rh_sqrt(dot(r0, r0)), R1 = rh_sqrt(dot(r1, r1)), R2 = rh_sqrt(dot(r2, r2));
    const V3<R> n = v3(R(rec[12]), R(rec[13]), R(rec[14]));
    const R h = dot(r0, n);  // r0 . n, the same for every fan triangle
    R im = R(rec[15]) * h;
    R re = R0 * R1 * R2 + dot(r0, r1) * R2 + dot(r0, r2) * R1 + dot(r1, r2) * R0;
    if (rec[17] > 3.5) {
        const V3<R> r3 = rel<R>(rec + 9, p);
        const R R3 = rh_sqrt(dot(r3, r3));
        const R im2 = R(rec[16]) * h;
        const R re2 = R0 * R2 * R3 + dot(r0, r2) * R3 + dot(r0, r3) * R2 + dot(r2, r3) * R0;
        const R pr = re * re2 - im * im2;
        const R pi = re * im2 + im * re2;
        re = pr;
        im = pi;
    }
    return 2.0 * double(rh_atan2_wave(im, re));
}

// Wave-uniform integer of a 0/1/2/3 code stored as a double: its high word
// (0, 0x3ff00000, 0x40000000, 0x40080000) read into a scalar register.
RH_HD int mesh_code(const double* __restrict__ slot)
{
#ifdef __HIP_DEVICE_COMPILE__
    const int hi = __builtin_amdgcn_readfirstlane(reinterpret_cast<const int*>(slot)[1]);
#else
    const int hi = int(uint64_t(__builtin_bit_cast(uint64_t, *slot)) >> 32);
#endif
    return hi == 0 ? 0 : (hi == 0x3ff00000 ? 1 : (hi == 0x40000000 ? 2 : 3));
}

// The owned-side mask of a face record (an integer in the slot's low word),
// wave-uniform on the device.
RH_HD int mesh_mask(const double* __restrict__ slot)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_readfirstlane(reinterpret_cast<const int*>(slot)[0]);
#else
    return int(uint32_t(__builtin_bit_cast(uint64_t, *slot) & 0xffffffffu));
#endif
}

// A vertex relative to the observation point and its distance.
template <class R>
struct MeshVtx {
    V3<R> r;
    R d;
};
template <class R>
RH_HD MeshVtx<R> mesh_vtx(const double* __restrict__ v, const double p[3])
{
    const V3<R> r = rel<R>(v, p);
    return MeshVtx<R>{r, rh_sqrt(dot(r, r))};
}

// Signed solid angle of a charged triangle from the offsets and distances of
// its vertices and r0 . r1.
template <class R>
RH_HD double mesh_tri_omega_from(const V3<R>& r0, R R0, const V3<R>& r1, R R1, const V3<R>& r2, R R2, R d01,
                                 R d12, R d20, const double* __restrict__ rec)
{
    const V3<R> n = v3(R(rec[12]), R(rec[13]), R(rec[14]));
    const R im = R(rec[15]) * dot(r0, n);
    const R re = R0 * R1 * R2 + d01 * R2 + d20 * R1 + d12 * R0;
    return 2.0 * double(rh_atan2_wave(im, re));
}

// FP32 sweep records: the geometry the lanes use in FP32 is stored as floats
// when the records are charged (edge slots 6-7: t(3), len; face slots 12-14:
// n(3), the two doubled areas), so the sweep converts none of it per record;
// vertices (offsets are formed in FP64) and the weights w and g (the sums are
// accumulated in FP64) stay FP64.
RH_HD const float* mesh32(const double* __restrict__ rec, int slot)
{
    return reinterpret_cast<const float*>(rec + slot);
}

RH_HD float mesh32_tri_omega(const V3<float>& r0, float R0, const V3<float>& r1, float R1, const V3<float>& r2,
                             float R2, float d01, float d12, float d20, const double* __restrict__ rec)
{
    const float* f = mesh32(rec, 12);
    const float im = f[3] * dot(r0, v3(f[0], f[1], f[2]));
    const float re = R0 * R1 * R2 + d01 * R2 + d20 * R1 + d12 * R0;
    return 2.0f * rh_atan2_wave(im, re);
}

// FP64 solid angle of a face of an FP32 sweep record, whose normal and areas
// are stored as floats: the fan triangles' triple products r0 . (r1 x r2)
// are taken as r0 . ((v1 - v0) x (v2 - v0)) from the FP64 vertices, signed by
// the stored normal (the same quantity as area * (r0 . n) of the FP64 sweep).
RH_HD double mesh_face_omega_near(const double* __restrict__ rec, const double p[3])
{
    const float* f = mesh32(rec, 12);
    const V3<double> e1 = v3(rec[3] - rec[0], rec[4] - rec[1], rec[5] - rec[2]);
    const V3<double> e2 = v3(rec[6] - rec[0], rec[7] - rec[1], rec[8] - rec[2]);
    const V3<double> c1 = cross(e1, e2);
    const double sg = (c1.x * f[0] + c1.y * f[1] + c1.z * f[2] >= 0.0) ? 1.0 : -1.0;
    const V3<double> r0 = rel<double>(rec, p), r1 = rel<double>(rec + 3, p), r2 = rel<double>(rec + 6, p);
    const double R0 = rh_sqrt(dot(r0, r0)), R1 = rh_sqrt(dot(r1, r1)), R2 = rh_sqrt(dot(r2, r2));
    double im = sg * dot(r0, c1);
    double re = R0 * R1 * R2 + dot(r0, r1) * R2 + dot(r0, r2) * R1 + dot(r1, r2) * R0;
    if (rec[17] > 3.5) {
        const V3<double> r3 = rel<double>(rec + 9, p);
        const double R3 = rh_sqrt(dot(r3, r3));
        const V3<double> c2 = cross(e2, v3(rec[9] - rec[0], rec[10] - rec[1], rec[11] - rec[2]));
        const double im2 = sg * dot(r0, c2);
        const double re2 = R0 * R2 * R3 + dot(r0, r2) * R3 + dot(r0, r3) * R2 + dot(r2, r3) * R0;
        const double pr = re * re2 - im * im2;
        const double pi = re * im2 + im * re2;
        re = pr;
        im = pi;
    }
    return 2.0 * rh_atan2_wave(im, re);
}

// Near test of the FP32 mesh sweep: FP32 offsets carry eps32 of the distance,
// so a term's relative error grows like eps32 * size / distance.  A wave with
// a point within kNear32 entity sizes of a vertex (faces: the longest side;
// edges: the length) evaluates that entity in FP64.  thr: kNear32 sizes, as
// the record carries it (faces: slot 35, edges: slot 8).
RH_HD bool mesh32_near(float dmin, float thr)
{
    bool near = dmin < thr;
#ifdef __HIP_DEVICE_COMPILE__
    near = __any(near);
#endif
    return near;
}

// L_e of an FP32 sweep edge record in FP64 (vertices in slots 0-5, the FP64
// length in slot 9).
RH_HD double mesh_edge_L_near(const double* __restrict__ rec, const double p[3])
{
    const MeshVtx<double> a = mesh_vtx<double>(rec, p), b = mesh_vtx<double>(rec + 3, p);
    return edge_L(a.r, a.d, b.r, b.d, rec[9]);
}

// A face of an FP32 sweep record in FP64 (triangle or quadrilateral): its
// solid angle and the sum of L_e w_e over the sides it owns.
struct MeshFace64 {
    double om, h0, h1, h2;
};
RH_HD MeshFace64 mesh_face_near(const double* __restrict__ rec, int mask, const double p[3])
{
    const int nv = rec[17] > 3.5 ? 4 : 3;
    MeshFace64 t{mesh_face_omega_near(rec, p), 0.0, 0.0, 0.0};
    for (int k = 0; k < 3; k++) {
        if (!(mask & (1 << k))) continue;
        const double* e = rec + kMeshFaceEdges + 4 * k;
        const int k1 = (k + 1 == nv) ? 0 : k + 1;
        const MeshVtx<double> a = mesh_vtx<double>(rec + 3 * k, p), b = mesh_vtx<double>(rec + 3 * k1, p);
        const double L = edge_L(a.r, a.d, b.r, b.d, dot(a.r, b.r), e[0]);
        t.h0 += L * e[1];
        t.h1 += L * e[2];
        t.h2 += L * e[3];
    }
    return t;
}

// A quadrilateral face record at p from its vertex offsets, distances and
// dot products: the signed solid angle (fans (v0,v1,v2) and (v0,v2,v3)
// multiplied as complex numbers).  FP64 records hold n and the doubled fan
// areas in slots 12-16, FP32 records as floats in slots 12-14.
template <class R>
RH_HD R mesh_quad_omega_from(const MeshVtx<R> v[4], R d01, R d02, R d03, R d12, R d23,
                             const double* __restrict__ rec)
{
    R n0, n1, n2, a1, a2;
    if constexpr (sizeof(R) == 4) {
        const float* f = mesh32(rec, 12);
        n0 = f[0], n1 = f[1], n2 = f[2], a1 = f[3], a2 = f[4];
    } else {
        n0 = R(rec[12]), n1 = R(rec[13]), n2 = R(rec[14]), a1 = R(rec[15]), a2 = R(rec[16]);
    }
    const R h = dot(v[0].r, v3(n0, n1, n2));
    const R im = a1 * h, im2 = a2 * h;
    const R re = v[0].d * v[1].d * v[2].d + d01 * v[2].d + d02 * v[1].d + d12 * v[0].d;
    const R re2 = v[0].d * v[2].d * v[3].d + d02 * v[3].d + d03 * v[2].d + d23 * v[0].d;
    return R(2) * rh_atan2_wave(re * im2 + im * re2, re * re2 - im * im2);
}

// Plane test of one inside record.
RH_HD bool mesh_inside(const double* __restrict__ rec, const double p[3])
{
    const int np = int(rec[0]);
    bool in = true;
    for (int k = 0; k < kMeshPlanes; k++) {
        if (k < np) {
            const double* q = rec + 4 + 4 * k;
            in = in && (q[0] * p[0] + q[1] * p[1] + q[2] * p[2] < q[3]);
        }
    }
    return in;
}

}  // namespace rh
