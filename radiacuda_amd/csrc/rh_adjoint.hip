// rh_adjoint.hip -- gfx950 adjoint kernels: dL/dM and dL/d(vertex) of the
// field at observation points for given weights g = dL/dB, the transposed
// matrix pass and the per-element update of the linear-material adjoint
// relaxation.
//
// Mapping: as in the forward field sweep, each lane owns one observation
// point and every lane of the workgroup visits the same source elements
// (wave-uniform records).  A lane's contribution to an element (or vertex)
// gradient is summed over the wavefront with cross-lane shuffles and
// accumulated per wave in LDS; each workgroup writes one partial per element
// slot and a second kernel sums the partials in a fixed order, so gradients
// are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "rh_adjoint.hpp"
#include "rh_eval.hpp"
#include "rh_kernels.hpp"

namespace rh {

namespace {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int kAdjSliceE = 64;   // elements per workgroup slice (dL/dM)
constexpr int kAdjSliceV = 128;  // vertices per workgroup slice (dL/dvertex)

// dL/dM partials: part[bx][j][3] for the elements of slice blockIdx.y.
template <bool SYM>
__global__ __launch_bounds__(256) void k_adj_m(DevModel m, const double* __restrict__ pts,
                                               int64_t np, const double* __restrict__ g,
                                               int addin, double* __restrict__ part)
{
    __shared__ double acc[4][kAdjSliceE * 3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 4 * kAdjSliceE * 3; i += 256) (&acc[0][0])[i] = 0.0;
    __syncthreads();
    const int64_t j0 = int64_t(blockIdx.y) * kAdjSliceE;
    const int64_t j1 = (j0 + kAdjSliceE < m.n) ? (j0 + kAdjSliceE) : m.n;
    const int64_t tiles = (np + 255) / 256;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t p = tile * 256 + threadIdx.x;
        const bool live = p < np;
        const double x[3] = {live ? pts[3 * p] : 0.0, live ? pts[3 * p + 1] : 0.0,
                             live ? pts[3 * p + 2] : 0.0};
        const double gp[3] = {live ? g[3 * p] : 0.0, live ? g[3 * p + 1] : 0.0,
                              live ? g[3 * p + 2] : 0.0};
        for (int64_t j = j0; j < j1; j++) {
            const int kind = m.kind[j];
            const double* rec = m.data + m.off[j];
            double v[3] = {0.0, 0.0, 0.0};
            if (live) {
                for (int64_t c = m.copy_off[j]; c < m.copy_off[j + 1]; c++) {
                    const double* cp = m.copies + 24 * c;
                    double xl[3], q6[6], gl[3];
                    if (SYM) {
                        apply_inv(cp, x, xl);
                        // F^T g
                        for (int k = 0; k < 3; k++)
                            gl[k] = cp[12 + k] * gp[0] + cp[15 + k] * gp[1] + cp[18 + k] * gp[2];
                    } else {
                        for (int k = 0; k < 3; k++) {
                            xl[k] = x[k];
                            gl[k] = gp[k];
                        }
                    }
                    const bool in = eval_elem<double>(kind, rec, xl, q6);
                    const double e = (in && addin) ? 1.0 : 0.0;
                    v[0] += (q6[0] + e) * gl[0] + q6[3] * gl[1] + q6[4] * gl[2];
                    v[1] += q6[3] * gl[0] + (q6[1] + e) * gl[1] + q6[5] * gl[2];
                    v[2] += q6[4] * gl[0] + q6[5] * gl[1] + (q6[2] + e) * gl[2];
                }
            }
            for (int k = 0; k < 3; k++) {
                const double s = wsum(v[k]);
                if (lane == 0) acc[wave][3 * (j - j0) + k] += s;
            }
        }
    }
    __syncthreads();
    const int64_t cnt = 3 * (j1 - j0);
    for (int64_t i = threadIdx.x; i < cnt; i += 256)
        part[int64_t(blockIdx.x) * 3 * m.n + 3 * j0 + i] = acc[0][i] + acc[1][i] + acc[2][i] + acc[3][i];
}

// dL/dvertex partials: partv[bx][vertex slot][3] for the vertices of element
// slice [sl[y], sl[y+1]); optional per-point dL/dx partials partx[y][p][3].
template <bool SYM>
__global__ __launch_bounds__(256) void k_adj_geom(DevModel m, const double* __restrict__ adj,
                                                  const int64_t* __restrict__ adj_off,
                                                  const int64_t* __restrict__ vert_off,
                                                  const int64_t* __restrict__ sl, int64_t nvert,
                                                  const double* __restrict__ M,
                                                  const double* __restrict__ pts, int64_t np,
                                                  const double* __restrict__ g,
                                                  double* __restrict__ partv,
                                                  double* __restrict__ partx)
{
    __shared__ double acc[4][kAdjSliceV * 3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 4 * kAdjSliceV * 3; i += 256) (&acc[0][0])[i] = 0.0;
    __syncthreads();
    const int64_t j0 = sl[blockIdx.y], j1 = sl[blockIdx.y + 1];
    const int64_t v0 = vert_off[j0];
    const int64_t tiles = (np + 255) / 256;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t p = tile * 256 + threadIdx.x;
        const bool live = p < np;
        const double x[3] = {live ? pts[3 * p] : 0.0, live ? pts[3 * p + 1] : 0.0,
                             live ? pts[3 * p + 2] : 0.0};
        const double gp[3] = {live ? g[3 * p] : 0.0, live ? g[3 * p + 1] : 0.0,
                              live ? g[3 * p + 2] : 0.0};
        double gx[3] = {0.0, 0.0, 0.0};
        for (int64_t j = j0; j < j1; j++) {
            const double* rec = adj + adj_off[j];
            const int nt = adj_ntri(rec);
            const double Mj[3] = {M[3 * j], M[3 * j + 1], M[3 * j + 2]};
            const int64_t slot0 = vert_off[j] - v0;
            for (int64_t c = m.copy_off[j]; c < m.copy_off[j + 1]; c++) {
                const double* cp = m.copies + 24 * c;
                double xl[3], gl[3];
                if (SYM) {
                    apply_inv(cp, x, xl);
                    for (int k = 0; k < 3; k++)
                        gl[k] = kInv4Pi * (cp[12 + k] * gp[0] + cp[15 + k] * gp[1] + cp[18 + k] * gp[2]);
                } else {
                    for (int k = 0; k < 3; k++) {
                        xl[k] = x[k];
                        gl[k] = kInv4Pi * gp[k];
                    }
                }
                const double* tri = rec + 2 + 3 * adj_nv(rec);
                for (int t = 0; t < nt; t++) {
                    for (int k = 0; k < 3; k++) {
                        double d[3] = {0.0, 0.0, 0.0};
                        const int vid = int(tri[3 * t + k]);  // wave-uniform
                        int vid_lane = vid;
                        if (live) adj_tri_corner_grad(rec, t, k, xl, Mj, gl, d, &vid_lane);
                        // translation invariance: d/dxl = -sum over corners of d/dV
                        if (SYM) {
                            for (int a = 0; a < 3; a++)
                                gx[a] -= cp[a] * d[0] + cp[3 + a] * d[1] + cp[6 + a] * d[2];
                        } else {
                            for (int a = 0; a < 3; a++) gx[a] -= d[a];
                        }
                        for (int a = 0; a < 3; a++) {
                            const double s = wsum(d[a]);
                            if (lane == 0) acc[wave][3 * (slot0 + vid) + a] += s;
                        }
                    }
                }
            }
        }
        if (partx && live)
            for (int a = 0; a < 3; a++) partx[(int64_t(blockIdx.y) * np + p) * 3 + a] = gx[a];
    }
    __syncthreads();
    const int64_t cnt = 3 * (vert_off[j1] - v0);
    for (int64_t i = threadIdx.x; i < cnt; i += 256)
        partv[int64_t(blockIdx.x) * 3 * nvert + 3 * v0 + i] =
            acc[0][i] + acc[1][i] + acc[2][i] + acc[3][i];
}

// out[i] = sum_k part[k * stride + i], i < len
__global__ __launch_bounds__(256) void k_sum_parts(int64_t len, int64_t ns, int64_t stride,
                                                   const double* __restrict__ part,
                                                   double* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int64_t k = 0; k < ns; k++) s += part[k * stride + i];
    out[i] = s;
}

// Transposed matrix pass: part[rc][col] = sum over rows of chunk rc of
// Q[r][col] x[r]; each lane owns 4 consecutive columns (16-byte FP32 or 2x16
// byte FP64 loads along the row), rows of a chunk are streamed in order.
template <class T>
__global__ __launch_bounds__(256) void k_matvec_t(const T* __restrict__ Q, int64_t nrows, int64_t ld,
                                                  int64_t rows_per_chunk, const double* __restrict__ x,
                                                  double* __restrict__ part)
{
    const int64_t c0 = (int64_t(blockIdx.x) * 256 + threadIdx.x) * 4;
    const int64_t r0 = int64_t(blockIdx.y) * rows_per_chunk;
    const int64_t r1 = (r0 + rows_per_chunk < nrows) ? (r0 + rows_per_chunk) : nrows;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (c0 < ld) {
        for (int64_t r = r0; r < r1; r++) {
            const T* row = Q + r * ld + c0;
            const double xr = x[r];
#pragma unroll
            for (int k = 0; k < 4; k++) acc[k] += double(row[k]) * xr;
        }
        double* o = part + int64_t(blockIdx.y) * ld + c0;
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = acc[k];
    }
}

// Transposed pass over the symmetric-block layout: every stored block is
// symmetric, so (Q^T x)_j = sum_i Q_ij x_i; part[rc][3j + a] sums the element
// rows of chunk rc.  Each lane owns two consecutive column elements.
template <class T, int P>
__global__ __launch_bounds__(256) void k_matvec_t_sym(const T* __restrict__ Q, int64_t nel, int64_t ldp,
                                                      int64_t el_per_chunk, const double* __restrict__ x,
                                                      double* __restrict__ part, int64_t ld)
{
    const int64_t j0 = (int64_t(blockIdx.x) * 256 + threadIdx.x) * 2;
    const int64_t i0 = int64_t(blockIdx.y) * el_per_chunk;
    const int64_t i1 = (i0 + el_per_chunk < nel) ? (i0 + el_per_chunk) : nel;
    if (j0 >= ldp) return;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = i0; i < i1; i++) {
        const T* b = Q + i * P * ldp + j0;
        const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const double xx = double(b[e]), xy = double(b[ldp + e]), xz = double(b[2 * ldp + e]);
            const double yy = double(b[3 * ldp + e]), yz = double(b[4 * ldp + e]);
            const double zz = (P == 6) ? double(b[5 * ldp + e]) : -(xx + yy);
            acc[3 * e] += xx * x0 + xy * x1 + xz * x2;
            acc[3 * e + 1] += xy * x0 + yy * x1 + yz * x2;
            acc[3 * e + 2] += xz * x0 + yz * x1 + zz * x2;
        }
    }
    double* o = part + int64_t(blockIdx.y) * ld + 3 * j0;
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = acc[k];
}

// mu_i = K_i^T lam_i (linear susceptibility tensor of element i, element frame)
// followed by rowT_i^T when the rows carry a transform.
__device__ __forceinline__ void lin_tensor(const double* mat, double K[9])
{
    if (int(mat[0]) == 0) {
        for (int k = 0; k < 9; k++) K[k] = 0.0;
        K[0] = K[4] = K[8] = mat[1];
    } else {
        for (int k = 0; k < 9; k++) K[k] = mat[4 + k];
    }
}

__global__ __launch_bounds__(256) void k_adj_scale(int64_t n, const double* __restrict__ mats,
                                                   const double* __restrict__ rowt,
                                                   const double* __restrict__ lam,
                                                   double* __restrict__ mu)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    double K[9];
    lin_tensor(mats + 16 * i, K);
    const double l[3] = {lam[3 * i], lam[3 * i + 1], lam[3 * i + 2]};
    double u[3];
    for (int a = 0; a < 3; a++) u[a] = K[a] * l[0] + K[3 + a] * l[1] + K[6 + a] * l[2];
    if (rowt) {
        const double* t = rowt + 9 * i;
        double w[3];
        for (int a = 0; a < 3; a++) w[a] = t[a] * u[0] + t[3 + a] * u[1] + t[6 + a] * u[2];
        for (int a = 0; a < 3; a++) u[a] = w[a];
    }
    for (int a = 0; a < 3; a++) mu[3 * i + a] = u[a];
}

// Block-Jacobi update of the transposed system (I - Q^T K^T) lam = rhs:
// lam_new_i = (I - Qii^T K_i^T)^{-1} (rhs_i + y_i), y = offdiag(Q)^T K^T lam;
// F = lam_new - lam, per-block [sum |F|^2, max |lam_new|, 0].
__global__ __launch_bounds__(kRelaxLanes) void k_adj_update(int64_t n, const double* __restrict__ mats,
                                                    const double* __restrict__ rhs,
                                                    const double* __restrict__ qdiag,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ lam,
                                                    double* __restrict__ F,
                                                    double* __restrict__ partials)
{
    __shared__ double red[2][kRelaxLanes];
    const int64_t i = int64_t(blockIdx.x) * kRelaxLanes + threadIdx.x;
    double f2 = 0.0, mm = 0.0;
    if (i < n) {
        double K[9];
        lin_tensor(mats + 16 * i, K);
        const double* D = qdiag + 9 * i;
        // A = I - D^T K^T = I - (K D)^T
        double KD[9], A[9], Ai[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                KD[3 * a + b] = K[3 * a] * D[b] + K[3 * a + 1] * D[3 + b] + K[3 * a + 2] * D[6 + b];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) A[3 * a + b] = (a == b ? 1.0 : 0.0) - KD[3 * b + a];
        const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                           A[2] * (A[3] * A[7] - A[4] * A[6]);
        const double id = 1.0 / det;
        Ai[0] = (A[4] * A[8] - A[5] * A[7]) * id;
        Ai[1] = (A[2] * A[7] - A[1] * A[8]) * id;
        Ai[2] = (A[1] * A[5] - A[2] * A[4]) * id;
        Ai[3] = (A[5] * A[6] - A[3] * A[8]) * id;
        Ai[4] = (A[0] * A[8] - A[2] * A[6]) * id;
        Ai[5] = (A[2] * A[3] - A[0] * A[5]) * id;
        Ai[6] = (A[3] * A[7] - A[4] * A[6]) * id;
        Ai[7] = (A[1] * A[6] - A[0] * A[7]) * id;
        Ai[8] = (A[0] * A[4] - A[1] * A[3]) * id;
        const double r[3] = {rhs[3 * i] + y[3 * i], rhs[3 * i + 1] + y[3 * i + 1],
                             rhs[3 * i + 2] + y[3 * i + 2]};
        double nl[3];
        for (int a = 0; a < 3; a++) nl[a] = Ai[3 * a] * r[0] + Ai[3 * a + 1] * r[1] + Ai[3 * a + 2] * r[2];
        for (int a = 0; a < 3; a++) {
            const double d = nl[a] - lam[3 * i + a];
            F[3 * i + a] = d;
            f2 += d * d;
        }
        mm = sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
    }
    red[0][threadIdx.x] = f2;
    red[1][threadIdx.x] = mm;
    __syncthreads();
    for (int s = kRelaxLanes / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = red[0][0];
        partials[3 * blockIdx.x + 1] = red[1][0];
        partials[3 * blockIdx.x + 2] = 0.0;
    }
}

}  // namespace

int64_t adj_point_groups(int64_t np, int64_t slices)
{
    const int64_t tiles = (np + 255) / 256;
    int64_t pg = std::max<int64_t>(1, (2048 + slices - 1) / std::max<int64_t>(1, slices));
    return std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(pg, tiles), 64));
}

int64_t adj_m_slices(int64_t n) { return (n + kAdjSliceE - 1) / kAdjSliceE; }

std::vector<int64_t> adj_geom_slices(const std::vector<int64_t>& vert_off, std::string* err)
{
    std::vector<int64_t> sl{0};
    const int64_t n = int64_t(vert_off.size()) - 1;
    for (int64_t j = 0; j < n; j++) {
        const int64_t nv = vert_off[j + 1] - vert_off[j];
        if (nv > kAdjSliceV) {
            if (err) *err = "element with more than 128 vertices";
            return {};
        }
        if (vert_off[j + 1] - vert_off[sl.back()] > kAdjSliceV) sl.push_back(j);
    }
    sl.push_back(n);
    return sl;
}

hipError_t launch_adj_m(const DevModel& m, const double* pts, int64_t np, const double* g, bool addin,
                        double* part, int64_t groups, double* gM, hipStream_t s)
{
    const int64_t ns = adj_m_slices(m.n);
    const dim3 grid{unsigned(groups), unsigned(ns), 1u};
    if (m.sym)
        hipLaunchKernelGGL(k_adj_m<true>, grid, dim3(256), 0, s, m, pts, np, g, addin ? 1 : 0, part);
    else
        hipLaunchKernelGGL(k_adj_m<false>, grid, dim3(256), 0, s, m, pts, np, g, addin ? 1 : 0, part);
    const int64_t len = 3 * m.n;
    hipLaunchKernelGGL(k_sum_parts, dim3(unsigned((len + 255) / 256)), dim3(256), 0, s, len, groups,
                       len, part, gM);
    return hipGetLastError();
}

hipError_t launch_adj_geom(const DevModel& m, const double* adj, const int64_t* adj_off,
                           const int64_t* vert_off, const int64_t* sl, int64_t nslices,
                           int64_t nvert, const double* M, const double* pts, int64_t np,
                           const double* g, double* partv, double* partx, int64_t groups,
                           double* gV, double* gX, hipStream_t s)
{
    const dim3 grid{unsigned(groups), unsigned(nslices), 1u};
    if (m.sym)
        hipLaunchKernelGGL(k_adj_geom<true>, grid, dim3(256), 0, s, m, adj, adj_off, vert_off, sl,
                           nvert, M, pts, np, g, partv, gX ? partx : nullptr);
    else
        hipLaunchKernelGGL(k_adj_geom<false>, grid, dim3(256), 0, s, m, adj, adj_off, vert_off, sl,
                           nvert, M, pts, np, g, partv, gX ? partx : nullptr);
    if (gV) {
        const int64_t len = 3 * nvert;
        hipLaunchKernelGGL(k_sum_parts, dim3(unsigned((len + 255) / 256)), dim3(256), 0, s, len,
                           groups, len, partv, gV);
    }
    if (gX) {
        const int64_t len = 3 * np;
        hipLaunchKernelGGL(k_sum_parts, dim3(unsigned((len + 255) / 256)), dim3(256), 0, s, len,
                           nslices, len, partx, gX);
    }
    return hipGetLastError();
}

int64_t matvec_t_chunks(int64_t nrows, int64_t ld)
{
    const int64_t ctiles = (ld + 1023) / 1024;
    int64_t ch = std::max<int64_t>(1, (4096 + ctiles - 1) / ctiles);
    ch = std::min<int64_t>(ch, std::max<int64_t>(1, nrows / 16));
    return std::min<int64_t>(ch, 256);
}

hipError_t launch_matvec_t(const void* Q, int elem_bytes, int64_t nrows, int64_t ld, const double* x,
                           double* part, int64_t chunks, double* y, hipStream_t s)
{
    const int64_t rpc = (nrows + chunks - 1) / chunks;
    const dim3 grid{unsigned((ld + 1023) / 1024), unsigned(chunks), 1u};
    if (elem_bytes == 4)
        hipLaunchKernelGGL(k_matvec_t<float>, grid, dim3(256), 0, s, static_cast<const float*>(Q),
                           nrows, ld, rpc, x, part);
    else
        hipLaunchKernelGGL(k_matvec_t<double>, grid, dim3(256), 0, s, static_cast<const double*>(Q),
                           nrows, ld, rpc, x, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(unsigned((nrows + 255) / 256)), dim3(256), 0, s, nrows,
                       chunks, ld, part, y);
    return hipGetLastError();
}

hipError_t launch_matvec_t_sym(const void* Q, int elem_bytes, int planes, int64_t nel, int64_t ldp,
                               const double* x, double* part, int64_t chunks, double* y, hipStream_t s)
{
    if ((planes != 5 && planes != 6) || ldp % 256 != 0 || chunks < 1) return hipErrorInvalidValue;
    const int64_t epc = (nel + chunks - 1) / chunks;
    const int64_t ld = 3 * ldp;
    const dim3 grid{unsigned((ldp / 2 + 255) / 256), unsigned(chunks), 1u};
    if (elem_bytes == 4) {
        const float* q = static_cast<const float*>(Q);
        if (planes == 5)
            hipLaunchKernelGGL((k_matvec_t_sym<float, 5>), grid, dim3(256), 0, s, q, nel, ldp, epc, x, part, ld);
        else
            hipLaunchKernelGGL((k_matvec_t_sym<float, 6>), grid, dim3(256), 0, s, q, nel, ldp, epc, x, part, ld);
    } else {
        const double* q = static_cast<const double*>(Q);
        if (planes == 5)
            hipLaunchKernelGGL((k_matvec_t_sym<double, 5>), grid, dim3(256), 0, s, q, nel, ldp, epc, x, part, ld);
        else
            hipLaunchKernelGGL((k_matvec_t_sym<double, 6>), grid, dim3(256), 0, s, q, nel, ldp, epc, x, part, ld);
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(unsigned((3 * nel + 255) / 256)), dim3(256), 0, s, 3 * nel,
                       chunks, ld, part, y);
    return hipGetLastError();
}

hipError_t launch_adj_scale(int64_t n, const double* mats, const double* rowt, const double* lam,
                            double* mu, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_scale, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, mats, rowt,
                       lam, mu);
    return hipGetLastError();
}

hipError_t launch_adj_update(int64_t n, const double* mats, const double* rhs, const double* qdiag,
                             const double* y, const double* lam, double* F, double* partials,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_update, dim3(unsigned(relax_blocks(n))), dim3(kRelaxLanes), 0, s, n, mats, rhs,
                       qdiag, y, lam, F, partials);
    return hipGetLastError();
}

}  // namespace rh
