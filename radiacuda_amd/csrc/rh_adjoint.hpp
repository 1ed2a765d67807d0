// rh_adjoint.hpp -- derivatives of the element fields for the adjoint kernels.
//
// dB/dM is the field block itself (B is linear in M).  For dB/d(vertex) the
// element's surface is taken as its faces fanned into triangles (v0, vk, vk+1),
// each CCW about the outward normal, and the field of the uniformly
// magnetized body is written per triangle t with unit normal n_t and surface
// charge sigma_t = n_t . M:
//
//     4 pi H(p) = sum_t sigma_t [ sum_{e in t} L_e (t_e x n_t) - Omega_t n_t ]
//
// L_e = ln((R_b + s_b)/(R_a + s_a)) over the triangle's edge a -> b and
// Omega_t the signed solid angle (Van Oosterom & Strackee).  For planar faces
// the internal fan edges cancel pairwise and this equals the face form used by
// the forward kernels; for displaced vertices it stays a closed triangulated
// surface, so its derivative is that of a well-defined field.  Derivatives are
// forward-mode (dual numbers) with the three coordinates of one triangle
// corner seeded at a time: a triangle depends only on its own corners.
#pragma once

#include "rh_math.hpp"

namespace rh {

template <int K>
struct Dual {
    double v;
    double d[K];
    RH_HD Dual() : v(0.0)
    {
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
    RH_HD Dual(double x) : v(x)  // NOLINT: implicit constant
    {
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
};

template <int K>
RH_HD Dual<K> operator+(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r;
    r.v = a.v + b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K>
RH_HD Dual<K> operator-(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r;
    r.v = a.v - b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K>
RH_HD Dual<K> operator-(const Dual<K>& a)
{
    Dual<K> r;
    r.v = -a.v;
    for (int k = 0; k < K; k++) r.d[k] = -a.d[k];
    return r;
}
template <int K>
RH_HD Dual<K> operator*(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r;
    r.v = a.v * b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K>
RH_HD Dual<K> operator/(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r;
    const double ib = 1.0 / b.v;
    r.v = a.v * ib;
    for (int k = 0; k < K; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int K>
RH_HD bool operator<(const Dual<K>& a, const Dual<K>& b) { return a.v < b.v; }
template <int K>
RH_HD bool operator>(const Dual<K>& a, const Dual<K>& b) { return a.v > b.v; }
template <int K>
RH_HD bool operator>=(const Dual<K>& a, const Dual<K>& b) { return a.v >= b.v; }
template <int K>
RH_HD bool operator<=(const Dual<K>& a, const Dual<K>& b) { return a.v <= b.v; }

template <int K>
RH_HD Dual<K> rh_sqrt(const Dual<K>& a)
{
    Dual<K> r;
    r.v = sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * h;
    return r;
}
template <int K>
RH_HD Dual<K> rh_log(const Dual<K>& a)
{
    Dual<K> r;
    r.v = log(a.v);
    const double ia = 1.0 / a.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * ia;
    return r;
}
template <int K>
RH_HD Dual<K> rh_atan2(const Dual<K>& y, const Dual<K>& x)
{
    Dual<K> r;
    r.v = atan2(y.v, x.v);
    const double q = 1.0 / (x.v * x.v + y.v * y.v);
    for (int k = 0; k < K; k++) r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) * q;
    return r;
}

template <class T>
RH_HD V3<T> vsub(const V3<T>& a, const V3<T>& b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
RH_HD V3<T> vscale(const V3<T>& a, const T& s) { return V3<T>{a.x * s, a.y * s, a.z * s}; }

// Adds 4 pi H of triangle (a, b, c) -- corners relative to the observation
// point, CCW about the outward normal -- for magnetization M.
template <class T>
RH_HD void tri_field(const V3<T>& a, const V3<T>& b, const V3<T>& c, const double M[3], V3<T>& H)
{
    const V3<T> nn = cross(vsub(b, a), vsub(c, a));
    const T an = rh_sqrt(dot(nn, nn));
    const T ian = T(1.0) / an;
    const V3<T> n = vscale(nn, ian);
    const T sigma = n.x * T(M[0]) + n.y * T(M[1]) + n.z * T(M[2]);
    const T Ra = rh_sqrt(dot(a, a)), Rb = rh_sqrt(dot(b, b)), Rc = rh_sqrt(dot(c, c));
    const V3<T>* P[3] = {&a, &b, &c};
    const T Rv[3] = {Ra, Rb, Rc};
    V3<T> s{T(0.0), T(0.0), T(0.0)};
    for (int e = 0; e < 3; e++) {
        const V3<T>& u = *P[e];
        const V3<T>& w = *P[(e + 1) % 3];
        const V3<T> d = vsub(w, u);
        const T len = rh_sqrt(dot(d, d));
        const V3<T> t = vscale(d, T(1.0) / len);
        T num, den;
        edge_ratio(u, Rv[e], w, Rv[(e + 1) % 3], t, num, den);
        const T L = rh_log(num / den);
        const V3<T> m = cross(t, n);
        s.x = s.x + L * m.x;
        s.y = s.y + L * m.y;
        s.z = s.z + L * m.z;
    }
    const T im = dot(a, cross(b, c));
    const T re = Ra * Rb * Rc + dot(a, b) * Rc + dot(a, c) * Rb + dot(b, c) * Ra;
    const T om = T(2.0) * rh_atan2(im, re);
    H.x = H.x + sigma * (s.x - om * n.x);
    H.y = H.y + sigma * (s.y - om * n.y);
    H.z = H.z + sigma * (s.z - om * n.z);
}

// Adjoint geometry record of one element: nv, ntri, vertices (3 nv, element
// frame, in the caller's vertex order), triangles (3 ntri vertex indices).
RH_HD int adj_nv(const double* rec) { return int(rec[0]); }
RH_HD int adj_ntri(const double* rec) { return int(rec[1]); }

// 4 pi H at p (value only) from the triangulated surface.
RH_HD void adj_field_value(const double* rec, const double p[3], const double M[3], double H[3])
{
    const int nv = adj_nv(rec), nt = adj_ntri(rec);
    const double* V = rec + 2;
    const double* T = V + 3 * nv;
    V3<double> h{0.0, 0.0, 0.0};
    for (int t = 0; t < nt; t++) {
        const int i0 = int(T[3 * t]), i1 = int(T[3 * t + 1]), i2 = int(T[3 * t + 2]);
        const V3<double> a = rel<double>(V + 3 * i0, p), b = rel<double>(V + 3 * i1, p),
                         c = rel<double>(V + 3 * i2, p);
        tri_field(a, b, c, M, h);
    }
    H[0] = h.x;
    H[1] = h.y;
    H[2] = h.z;
}

// Derivative of g . (4 pi H_t(p)) with respect to the three coordinates of
// corner k of triangle t; returns the triangle's indices' vertex in *vid.
RH_HD void adj_tri_corner_grad(const double* rec, int t, int k, const double p[3],
                               const double M[3], const double g[3], double out[3], int* vid)
{
    const int nv = adj_nv(rec);
    const double* V = rec + 2;
    const double* T = V + 3 * nv;
    const int id[3] = {int(T[3 * t]), int(T[3 * t + 1]), int(T[3 * t + 2])};
    V3<Dual<3>> r[3];
    for (int q = 0; q < 3; q++) {
        const double* v = V + 3 * id[q];
        r[q] = V3<Dual<3>>{Dual<3>(v[0] - p[0]), Dual<3>(v[1] - p[1]), Dual<3>(v[2] - p[2])};
    }
    r[k].x.d[0] = 1.0;
    r[k].y.d[1] = 1.0;
    r[k].z.d[2] = 1.0;
    V3<Dual<3>> h{Dual<3>(0.0), Dual<3>(0.0), Dual<3>(0.0)};
    tri_field(r[0], r[1], r[2], M, h);
    for (int c = 0; c < 3; c++) out[c] = g[0] * h.x.d[c] + g[1] * h.y.d[c] + g[2] * h.z.d[c];
    *vid = id[k];
}

}  // namespace rh
