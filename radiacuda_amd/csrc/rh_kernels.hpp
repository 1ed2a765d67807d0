// rh_kernels.hpp -- launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace rh {

// Device arrays of the conforming-mesh field form (rh_mesh.hpp); charged
// records ch_* are rebuilt from M before each field sweep.
struct MeshDev {
    bool active = false;
    int64_t ne = 0, nf = 0, ni = 0, ntiles = 0;
    const double* edge_geo = nullptr;
    const int64_t* edge_inc_off = nullptr;
    const int32_t* edge_inc_elem = nullptr;
    const double* edge_inc_w = nullptr;
    const double* face_geo = nullptr;
    const int64_t* face_inc_off = nullptr;
    const int32_t* face_inc_elem = nullptr;
    const double* face_inc_v = nullptr;  // 3 per face incidence
    const int32_t* in_elem = nullptr;
    const double* in_geo = nullptr;
    const double* in_F = nullptr;        // 9 per inside record
    const double* tile_box = nullptr;
    const double* edge_cont = nullptr;
    const double* face_cont = nullptr;
    int64_t ne_free = 0;                // edges [0, ne_free) are swept on their own
    int64_t nf_tri = 0;                 // faces [0, nf_tri) are triangles, the rest quadrilaterals
    const int64_t* face_edge = nullptr;  // 3 per face: owned edge record or -1
    double* ch_edge = nullptr;
    double* ch_face = nullptr;
    double* ch_in = nullptr;
};

struct DevModel {
    int64_t n = 0;
    const int32_t* kind = nullptr;
    const int64_t* off = nullptr;
    const double* data = nullptr;
    const int64_t* copy_off = nullptr;
    const double* copies = nullptr;
    const double* colloc = nullptr;  // [3n] lab
    const double* rowt = nullptr;    // [9n]
    const int32_t* kidx = nullptr;   // element indices grouped by kind slot
    int64_t koff[4] = {0, 0, 0, 0};  // slot ranges in kidx (block, tet, general)
    const double* charged = nullptr;  // field-only tetrahedron records (per call)
    bool sym = false;
    bool rowt_identity = true;
    MeshDev mesh;
    bool use_mesh = false;  // field sweeps use the mesh form (when active)
};

// Interaction-matrix rows of elements [r0, r1) into Q (row-major, leading
// dimension ld elements of elem_bytes = 4 or 8; Q points at row 3*r0); the
// 3x3 diagonal blocks go to qdiag [9n] in FP64 and are stored as zeros in Q.
// far32 (FP32 storage only): blocks and tetrahedra far from the collocation
// point (beyond kNearRadii circumradii) are integrated in FP32, near ones and
// general polyhedra in FP64.
//
// planes = 5 or 6 (models without transform copies or row transforms only):
// the symmetric-block layout instead, element row i as planes (xx, xy, xz,
// yy, yz[, zz]) of ld values from Q + planes*(i - r0)*ld, plane k holding
// entry k of the blocks (i, 0 .. n-1).  With 5 planes zz = -(xx + yy), and a
// block whose trace exceeds 1/4 in magnitude (a collocation point inside
// another element) sets *flag to 1: the caller must then use 6 planes.
hipError_t launch_assemble(const DevModel& m, void* Q, int elem_bytes, int64_t ld, double* qdiag,
                           int64_t r0, int64_t r1, hipStream_t s, bool far32 = false, int planes = 0,
                           int32_t* flag = nullptr);

// Number of partial-sum slices the field kernels use for np points (>= 1);
// the work buffer needs 6 * np * slices doubles when slices > 1.  work_cap >= 0
// bounds that buffer (in doubles): the element splits are reduced until the
// partials fit, and the result exceeds the cap only when every kind slot is
// already down to one split (the caller must then refuse the call).  A negative
// work_cap leaves the splits as chosen for occupancy.
int64_t field_slices(const DevModel& m, int64_t np, int64_t work_cap = -1);
// B, H (either may be null) at np points.  charged: kChargedTet doubles per
// tetrahedron, rebuilt from M before the sweep.  work: 6 * np * slices doubles
// (field_slices with the same work_cap) when slices > 1.
hipError_t launch_field(const DevModel& m, const double* M, const double* pts, int64_t np,
                        double* B, double* H, double* work, double* charged, bool fp32,
                        hipStream_t s, int64_t work_cap = -1);

// y = Q x (+ add), Q: nrows x ld (ld multiple of 256, padded columns zero,
// x padded with zeros to ld).  skip (may be null): the pass returns at once
// when *skip is non-zero (a device-driven loop that has finished).
hipError_t launch_matvec(const void* Q, int elem_bytes, int64_t nrows, int64_t ld, const double* x,
                         const double* add, double* y, hipStream_t s, const int32_t* skip = nullptr);
// y = Q x (+ add) over nel element rows of the symmetric-block layout (planes
// 5 or 6, plane stride ldp a multiple of 256, x padded with zeros to 3*ldp).
hipError_t launch_matvec_sym(const void* Q, int elem_bytes, int planes, int64_t nel, int64_t ldp,
                             const double* x, const double* add, double* y, hipStream_t s,
                             const int32_t* skip = nullptr);

// Per-element implicit update: proposal M_prop = M_mat(H) with
// H = y + hext + Qii M_prop solved locally; F = M_prop - M.  partials gets
// per-block [sum |F|^2, max |M_prop|, max |H|].
constexpr int kRelaxLanes = 64;  // lanes per workgroup of the per-element relaxation kernels
int64_t relax_blocks(int64_t n);
// tab: node triples of tabulated laws (material kind 3), may be null otherwise.
hipError_t launch_relax_update(int64_t n, const double* mats, const double* tab, const double* hext,
                               const double* qdiag, const double* y, const double* M, double* F,
                               double* Hout, double* partials, hipStream_t s);

// The relaxation step with the Anderson bookkeeping in the same pass: the
// update above, then (when dFw is set) dFw = F - Fp and dXw = M - xp, then
// Fp = F and xp = M, and the dot products of the newest difference (dF[nh-1],
// which is dFw) with every dF[k] and of every dF[k] with F.  partials holds
// kRelaxPart doubles per block: [sum |F|^2, max |M_prop|, max |H|, nh dots
// with the newest difference, padding to kHistMax, nh dots with F, padding].
// Anderson depth: on the 40,000-tetrahedron yoke (prec 1e-5) 16 differences
// took 55-59 iterations where 8 took 60-85 (over mixing factors 0.3-0.9)
#ifndef RH_HIST
#define RH_HIST 16
#endif
constexpr int kHistMax = RH_HIST;
constexpr int kRelaxPart = 3 + 2 * kHistMax;
// dot products and combined vectors of one host-driven Anderson iteration
// (launch_dots, launch_lincomb)
constexpr int kDotMax = 2 * kHistMax + 1;
constexpr int kCombMax = 2 * kHistMax + 2;
struct HistPack {
    double* Fp = nullptr;
    double* xp = nullptr;
    double* dFw = nullptr;  // difference slot written in this pass (null: none)
    double* dXw = nullptr;
    const double* dF[kHistMax] = {nullptr};  // oldest first; dF[nh-1] == dFw
    int nh = 0;
};
hipError_t launch_relax_step(int64_t n, const double* mats, const double* tab, const double* hext,
                             const double* qdiag, const double* y, const double* M, double* F,
                             double* Hout, double* partials, const HistPack& hp, hipStream_t s);

// ------------------------------------------------------------ device-driven loop
// State of the Anderson relaxation loop kept on the device, so that the host
// enqueues iterations without waiting for each one: a pass of the matrix, the
// relaxation step with the Anderson step in its last block (residual,
// convergence test, history Gram matrix, small least-squares solve, next
// coefficients) and the update.  Once the Anderson step finds the loop
// finished it sets done, and the kernels of iterations enqueued after that
// return at once.
struct RelaxDevState {
    int32_t it;         // iterations completed
    int32_t done;       // 0 running; 1 finished, the update writes M = x + F; 2 finished, M kept
    int32_t fail;       // non-finite residual
    int32_t hist;       // differences in use in the current iteration
    int32_t head;       // next difference slot (the newest is head - 1)
    int32_t ext;        // the current iteration writes a difference into slot head - 1
    int32_t max_iter;
    int32_t keep_if_converged;  // a first iterate that meets prec is kept as it is
    int32_t arrive;             // blocks of the current relaxation step that have finished
    int32_t pad_;
    double prec, beta, best, misfit, maxM, maxH, n;
    double G[kHistMax * kHistMax];  // dF_a . dF_b of the differences in use, oldest first
    double cF;                      // update: x += cF F + sum_b cdX[b] dX[b] + cdF[b] dF[b]
    double cdX[kHistMax], cdF[kHistMax];
};
struct HistBufs {
    double* Fp;
    double* xp;
    double* dF[kHistMax];
    double* dX[kHistMax];
};
// relaxation step of the device-driven loop (slots and history from st); the
// last of its blocks to finish runs the Anderson step (reduction of the nb
// block partials, residual, convergence test, history Gram matrix, small
// least-squares solve, next coefficients) and advances st.
hipError_t launch_relax_step_dev(int64_t n, const double* mats, const double* tab, const double* hext,
                                 const double* qdiag, const double* y, const double* M, double* F,
                                 double* Hout, double* partials, const HistBufs& hb, RelaxDevState* st,
                                 hipStream_t s);
// x += cF F + sum_b (cdX[b] dX[b] + cdF[b] dF[b]); when st->done == 1, M_out = x + F.
hipError_t launch_update_dev(int64_t n3, double* x, const double* F, const HistBufs& hb, double* M_out,
                             const RelaxDevState* st, hipStream_t s);

// out = hext + rowT*H64 - y - Qii*M per element (rowt may be null = identity).
hipError_t launch_refine_defect(int64_t n, const double* H64, const double* rowt, const double* y,
                                const double* qdiag, const double* M, const double* hext,
                                double* out, hipStream_t s);

// ---------------------------------------------------------------- Newton-Krylov
// Coefficients passed by value to the basis kernels (at most kCoefMax vectors).
constexpr int kCoefMax = 64;
struct Coef64 {
    double c[kCoefMax];
};
// F = M_mat(H) - M, D = dM_mat/dH, P = (I - D Qii)^{-1}, Hout = H at
// H = y + Qii M + hext; partials as launch_relax_update.
hipError_t launch_nk_material(int64_t n, const double* mats, const double* tab, const double* hext,
                              const double* qdiag, const double* y, const double* M, double* F, double* D,
                              double* P, double* Hout, double* partials, hipStream_t s);
// z = P v (3x3 blocks).
hipError_t launch_nk_prec(int64_t n, const double* P, const double* v, double* z, hipStream_t s);
// w = z - D (y + Qii z).
hipError_t launch_nk_apply(int64_t n, const double* D, const double* qdiag, const double* z, const double* y,
                           double* w, hipStream_t s);
// partials[j * dot_blocks(len) + b]: w . V_j (j < nv) and w . w (j = nv).
hipError_t launch_mdot(int nv, const double* V, int64_t ldv, const double* w, int64_t len, double* partials,
                       hipStream_t s);
// out = cb * base + sum_j c_j V_j (base may be null; out may alias base).
hipError_t launch_vcomb(int nv, const double* V, int64_t ldv, const double* c, const double* base, double cb,
                        int64_t len, double* out, hipStream_t s);

// Batched dot products: out[k*nb + b] = partial of sum_i a_k[i]*b_k[i] over block b.
int64_t dot_blocks(int64_t len);
hipError_t launch_dots(int nd, const double* const* a, const double* const* b, int64_t len,
                       double* partials, hipStream_t s);

// y = sum_k c_k v_k (k < nv <= 34); y may alias v_0.
hipError_t launch_lincomb(int nv, const double* const* v, const double* c, int64_t len, double* y,
                          hipStream_t s);

// B += field of current sources (kind [ns], par [12 ns], copies [24 ns]) at np points.
hipError_t launch_current(int64_t ns, const int32_t* kind, const double* par, const double* cp,
                          const double* pts, int64_t np, double* B, hipStream_t s);

// ---------------------------------------------------------------- adjoint
// Point groups (workgroups along the point axis, each striding over tiles of
// 256 points) for `slices` element slices.
int64_t adj_point_groups(int64_t np, int64_t slices);
// Element slices of the dL/dM sweep.
int64_t adj_m_slices(int64_t n);
// Element slice bounds of the dL/dvertex sweep (<= 128 vertices per slice);
// empty with *err set if an element has more vertices.
std::vector<int64_t> adj_geom_slices(const std::vector<int64_t>& vert_off, std::string* err);
// gM [3n] = sum_p sum_c (Q_j(x'_pc) + [addin and inside] I) F_c^T g_p.
// part: groups * 3n doubles.
hipError_t launch_adj_m(const DevModel& m, const double* pts, int64_t np, const double* g, bool addin,
                        double* part, int64_t groups, double* gM, hipStream_t s);
// gV [3 nvert] = dL/d(vertex), gX [3 np] = dL/d(point) (either may be null).
// partv: groups * 3 nvert doubles; partx: nslices * 3 np doubles.
hipError_t launch_adj_geom(const DevModel& m, const double* adj, const int64_t* adj_off,
                           const int64_t* vert_off, const int64_t* sl, int64_t nslices,
                           int64_t nvert, const double* M, const double* pts, int64_t np,
                           const double* g, double* partv, double* partx, int64_t groups,
                           double* gV, double* gX, hipStream_t s);
// y = Q^T x over the stored rows (diagonal blocks stored as zeros);
// part: chunks * ld doubles.
int64_t matvec_t_chunks(int64_t nrows, int64_t ld);
hipError_t launch_matvec_t(const void* Q, int elem_bytes, int64_t nrows, int64_t ld, const double* x,
                           double* part, int64_t chunks, double* y, hipStream_t s);
// The same over nel element rows of the symmetric-block layout (part: chunks *
// 3*ldp doubles).
hipError_t launch_matvec_t_sym(const void* Q, int elem_bytes, int planes, int64_t nel, int64_t ldp,
                               const double* x, double* part, int64_t chunks, double* y, hipStream_t s);
// mu_i = rowT_i^T K_i^T lam_i (rowt may be null).
hipError_t launch_adj_scale(int64_t n, const double* mats, const double* rowt, const double* lam,
                            double* mu, hipStream_t s);
// Block-Jacobi update of (I - Q^T K^T) lam = rhs (linear materials).
hipError_t launch_adj_update(int64_t n, const double* mats, const double* rhs, const double* qdiag,
                             const double* y, const double* lam, double* F, double* partials,
                             hipStream_t s);

}  // namespace rh
