// rh_eval.hpp -- interaction blocks and point fields over transform copies.
//
// block_Q:  Q_ij = rowT_i * sum_c F_c Q_j(Minv_c (x_i - V_c))
//           (cpp/src/core/radintrc.cpp:831-878: TrMatrix of each symmetry copy,
//           TrMatrix_inv of the row element's main transform)
// field:    B(x) = sum_j sum_c F_c [Q_j(x') M_j + inside * M_j],
//           x' = Minv_c (x - V_c)   (cpp/src/core/radg3d.cpp:57-115)
#pragma once

#include <type_traits>

#include "rh_math.hpp"

namespace rh {

struct ModelView {
    const int32_t* kind;
    const int64_t* off;
    const double* data;
    const int64_t* copy_off;
    const double* copies;  // kCopySize (24) doubles each: Minv, V, F, pad
    // optional field-only tetrahedron records for the current magnetization,
    // kChargedTet doubles each, indexed by position in the tetrahedron list
    const double* charged = nullptr;
    int64_t charged_base = 0;
};

RH_HD void apply_inv(const double* c, const double x[3], double o[3])
{
    const double d0 = x[0] - c[9], d1 = x[1] - c[10], d2 = x[2] - c[11];
    o[0] = c[0] * d0 + c[1] * d1 + c[2] * d2;
    o[1] = c[3] * d0 + c[4] * d1 + c[5] * d2;
    o[2] = c[6] * d0 + c[7] * d1 + c[8] * d2;
}

RH_HD void mat3_vec(const double* m, const double v[3], double o[3])
{
    o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

// Element evaluation specialised on the kind slot (0 block, 1 tetrahedron,
// 2 general polyhedron, -1 = read the kind at run time).
template <class R, int KIND>
RH_HD bool eval_k(int kind, const double* __restrict__ rec, const double p[3], double q6[6])
{
    if (KIND == 0) return eval_rec<R>(rec, p, q6);
    if (KIND == 1) return eval_tet<R>(rec, p, q6);
    if (KIND == 2) return eval_gen<R>(rec, p, q6);
    return eval_elem<R>(kind, rec, p, q6);
}

// Sum over copies of F_c Q_j(x'), without the row transform.
template <class R, bool SYM, int KIND = -1>
RH_HD void block_Q(const ModelView& mv, int64_t j, const double x[3], double q9[9])
{
    const int kind = (KIND < 0) ? mv.kind[j] : 0;
    const double* rec = mv.data + mv.off[j];
    if (!SYM) {
        double q6[6];
        eval_k<R, KIND>(kind, rec, x, q6);
        q6_to_q9(q6, q9);
        return;
    }
    for (int k = 0; k < 9; k++) q9[k] = 0.0;
    for (int64_t c = mv.copy_off[j]; c < mv.copy_off[j + 1]; c++) {
        const double* cp = mv.copies + 24 * c;
        double xl[3], q6[6], ql[9];
        apply_inv(cp, x, xl);
        eval_k<R, KIND>(kind, rec, xl, q6);
        q6_to_q9(q6, ql);
        const double* F = cp + 12;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                q9[3 * a + b] += F[3 * a] * ql[b] + F[3 * a + 1] * ql[3 + b] + F[3 * a + 2] * ql[6 + b];
    }
}

// Near test of the mixed-precision assembly: the point (element frame) lies
// within kNearRadii circumradii of the element's centre (blocks: the half
// diagonal about the centre; tetrahedra: the farthest vertex from the vertex
// mean).  General polyhedra always count as near.
// The near tests spell out their fused multiply-adds (sq3), so that two
// kernels that must take the same decision (the FP32 field sweep's far and
// near passes) compute the same bits whatever the compiler contracts.
constexpr double kNearRadii = 2.0;
template <int KIND>
RH_HD bool near_element(int kind, const double* __restrict__ rec, const double x[3])
{
    const int K = (KIND >= 0) ? KIND : (kind == EK_REC ? 0 : (kind == EK_TET ? 1 : 2));
    if (K == 0) {
        const double dx = x[0] - rec[0], dy = x[1] - rec[1], dz = x[2] - rec[2];
        const double r2 = sq3(rec[3], rec[4], rec[5]);
        return sq3(dx, dy, dz) < kNearRadii * kNearRadii * r2;
    }
    if (K == 1) {
        // offsets from the vertex mean 0.25 s as fma(-0.25, s, .)
        double s4[3];
        for (int k = 0; k < 3; k++) s4[k] = (rec[k] + rec[3 + k]) + (rec[6 + k] + rec[9 + k]);
        double r2 = 0.0;
        for (int v = 0; v < 4; v++)
            r2 = fmax(r2, sq3(fma(-0.25, s4[0], rec[3 * v]), fma(-0.25, s4[1], rec[3 * v + 1]),
                              fma(-0.25, s4[2], rec[3 * v + 2])));
        return sq3(fma(-0.25, s4[0], x[0]), fma(-0.25, s4[1], x[1]), fma(-0.25, s4[2], x[2])) <
               kNearRadii * kNearRadii * r2;
    }
    return true;
}

// One element's Q at xl (element frame) in mixed precision: FP64 integrals
// when the point is near (on the device: when any lane of the wave is near,
// so the wave takes one path), FP32 integrals otherwise.
template <int KIND>
RH_HD void eval_mixed(int kind, const double* __restrict__ rec, const double xl[3], double q6[6])
{
    bool near = near_element<KIND>(kind, rec, xl);
#ifdef __HIP_DEVICE_COMPILE__
    near = __any(near);
#endif
    if (near)
        eval_k<double, KIND>(kind, rec, xl, q6);
    else
        eval_k<float, KIND>(kind, rec, xl, q6);
}

// block_Q with FP32 integrals for the far elements and FP64 for the near
// ones (self and neighbours), copy by copy.
template <bool SYM, int KIND>
RH_HD void block_Q_mixed(const ModelView& mv, int64_t j, const double x[3], double q9[9])
{
    const int kind = (KIND < 0) ? mv.kind[j] : 0;
    const double* rec = mv.data + mv.off[j];
    if (!SYM) {
        double q6[6];
        eval_mixed<KIND>(kind, rec, x, q6);
        q6_to_q9(q6, q9);
        return;
    }
    for (int k = 0; k < 9; k++) q9[k] = 0.0;
    for (int64_t c = mv.copy_off[j]; c < mv.copy_off[j + 1]; c++) {
        const double* cp = mv.copies + 24 * c;
        double xl[3], q6[6], ql[9];
        apply_inv(cp, x, xl);
        eval_mixed<KIND>(kind, rec, xl, q6);
        q6_to_q9(q6, ql);
        const double* F = cp + 12;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                q9[3 * a + b] += F[3 * a] * ql[b] + F[3 * a + 1] * ql[3 + b] + F[3 * a + 2] * ql[6 + b];
    }
}

// Near test of the FP32 field sweeps (element frame): blocks and tetrahedra
// as the assembly's; general polyhedra: a vertex within kNear32 longest edges
// of the point.
template <int KIND>
RH_HD bool near_field(int kind, const double* __restrict__ rec, const double x[3])
{
    const int K = (KIND >= 0) ? KIND : (kind == EK_REC ? 0 : (kind == EK_TET ? 1 : 2));
    if (K != 2) return near_element<KIND>(kind, rec, x);
    const int nv = int(rec[0]), ne = int(rec[1]);
    const double* vt = rec + kGenHdr;
    const double* ed = vt + 3 * nv;
    double d2 = 1e300, l = 0.0;
    for (int v = 0; v < nv; v++) d2 = fmin(d2, sq3(vt[3 * v] - x[0], vt[3 * v + 1] - x[1], vt[3 * v + 2] - x[2]));
    for (int e = 0; e < ne; e++) l = fmax(l, ed[kGenEdge * e + 5]);
    const double t = double(kNear32) * l;
    return d2 < t * t;
}

// Passes of a field sweep over a range of elements: every element at R
// (kPassAll; FP32 takes FP64 for the near elements inline), the FP32 sweep's
// far elements only (kPassFar), or its near elements in FP64 (kPassNear), run
// after kPassFar over the same elements and waves and added to its result.
// Split in two kernels, the FP64 code does not weigh on the FP32 sweep's
// registers.
enum : int { kPassAll = 0, kPassFar = 1, kPassNear = 2 };

// The near test of the FP32 sweeps for one element at x (element frame):
// wave-uniform on the device.
template <int KIND>
RH_HD bool near32(int kind, const double* __restrict__ rec, const double* __restrict__ crec,
                  const double x[3])
{
    if (crec) return tet_near32(crec, x);
    bool near = near_field<KIND>(kind, rec, x);
#ifdef __HIP_DEVICE_COMPILE__
    near = __any(near);
#endif
    return near;
}

// One element's H at x (element frame) from its record (or its charged
// record crec) and magnetization m; returns the inside flag.
template <class R, int KIND>
RH_HD bool elem_H_at(int kind, const double* __restrict__ rec, const double* __restrict__ crec,
                     const double m[3], const double x[3], double hl[3])
{
    if (crec) return field_tet<R>(crec, x, hl);
    double q6[6];
    const bool in = eval_k<R, KIND>(kind, rec, x, q6);
    hl[0] = q6[0] * m[0] + q6[3] * m[1] + q6[4] * m[2];
    hl[1] = q6[3] * m[0] + q6[1] * m[1] + q6[5] * m[2];
    hl[2] = q6[4] * m[0] + q6[5] * m[1] + q6[2] * m[2];
    return in;
}

// elem_H_at for pass PASS: false when the pass skips the element.
template <class R, int KIND, int PASS>
RH_HD bool elem_H(int kind, const double* __restrict__ rec, const double* __restrict__ crec,
                  const double m[3], const double x[3], double hl[3], bool& in)
{
    if constexpr (PASS == kPassAll && std::is_same<R, float>::value) {
        in = near32<KIND>(kind, rec, crec, x) ? elem_H_at<double, KIND>(kind, rec, crec, m, x, hl)
                                               : elem_H_at<float, KIND>(kind, rec, crec, m, x, hl);
        return true;
    }
    if constexpr (PASS != kPassAll) {
        if (near32<KIND>(kind, rec, crec, x) != (PASS == kPassNear)) return false;
    }
    in = elem_H_at<R, KIND>(kind, rec, crec, m, x, hl);
    return true;
}

// Adds the B and H of the elements idx[k0..k1) (all of kind slot KIND, or of
// any kind when KIND < 0; idx == nullptr means the identity list) at x, for
// pass PASS.
template <class R, bool SYM, int KIND = -1, int PASS = kPassAll>
RH_HD void field_add(const ModelView& mv, const double* M, const int32_t* idx, int64_t k0,
                     int64_t k1, const double x[3], double B[3], double H[3])
{
    for (int64_t k = k0; k < k1; k++) {
        const int64_t j = idx ? int64_t(idx[k]) : k;
        const int kind = (KIND < 0) ? mv.kind[j] : 0;
        const double* rec = mv.data + mv.off[j];
        const double m[3] = {M[3 * j], M[3 * j + 1], M[3 * j + 2]};
        const bool charged = (KIND == 1) && mv.charged != nullptr;
        const double* crec = charged ? mv.charged + kChargedTet * (k - mv.charged_base) : nullptr;
        if (!SYM) {
            double hl[3];
            bool in;
            if (!elem_H<R, KIND, PASS>(kind, rec, crec, m, x, hl, in)) continue;
            H[0] += hl[0]; H[1] += hl[1]; H[2] += hl[2];
            B[0] += hl[0] + (in ? m[0] : 0.0);
            B[1] += hl[1] + (in ? m[1] : 0.0);
            B[2] += hl[2] + (in ? m[2] : 0.0);
            continue;
        }
        for (int64_t c = mv.copy_off[j]; c < mv.copy_off[j + 1]; c++) {
            const double* cp = mv.copies + 24 * c;
            double xl[3], hl[3];
            apply_inv(cp, x, xl);
            bool in;
            if (!elem_H<R, KIND, PASS>(kind, rec, crec, m, xl, hl, in)) continue;
            double bl[3] = {hl[0] + (in ? m[0] : 0.0), hl[1] + (in ? m[1] : 0.0),
                            hl[2] + (in ? m[2] : 0.0)};
            double hg[3], bg[3];
            mat3_vec(cp + 12, hl, hg);
            mat3_vec(cp + 12, bl, bg);
            for (int q = 0; q < 3; q++) { H[q] += hg[q]; B[q] += bg[q]; }
        }
    }
}

}  // namespace rh
