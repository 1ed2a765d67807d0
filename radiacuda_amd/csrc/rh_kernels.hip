// rh_kernels.hip -- gfx950 kernels: interaction-matrix assembly, field
// evaluation, the relaxation's matrix pass and per-element update, and
// current-source fields.
//
// Mapping (wave64): in the assembly and field kernels each lane owns one
// observation point (a matrix row's collocation point, or a field point) and
// every lane of the workgroup sweeps the same source elements, so element
// records are read with wave-uniform addresses (scalar-cache loads, no
// per-lane gathers) and the per-element branches are uniform.  These kernels
// are bound by FP64 transcendental/VALU throughput.  The matrix pass streams
// the matrix once per iteration with 16-byte non-temporal loads and FP64
// accumulation; it is bound by HBM bandwidth.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rh_current.hpp"
#include "rh_eval.hpp"
#include "rh_geom.hpp"
#include "rh_kernels.hpp"
#include "rh_mesh.hpp"

namespace rh {

namespace {

__host__ __device__ __forceinline__ ModelView view(const DevModel& m)
{
    return ModelView{m.kind, m.off, m.data, m.copy_off, m.copies, m.charged, m.koff[1]};
}

// Field-only tetrahedron records for magnetization M (one lane per tetrahedron).
__global__ __launch_bounds__(256) void k_charge_tets(DevModel m, const double* __restrict__ M,
                                                     double* __restrict__ out)
{
    const int64_t k = m.koff[1] + int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (k >= m.koff[2]) return;
    const int64_t j = m.kidx[k];
    const double mm[3] = {M[3 * j], M[3 * j + 1], M[3 * j + 2]};
    charge_tet(m.data + m.off[j], mm, out + kChargedTet * (k - m.koff[1]));
}

template <class T>
__device__ __forceinline__ void store_out(T* p, double v) { *p = T(v); }

// ---------------------------------------------------------------- assembly
// Rows of elements [r0, r1) into Q (Q holds row 3*r0 first); block columns
// are the elements kidx[k0 .. k1) of one kind slot.  Each lane owns one row
// element and the wave sweeps the sources with uniform records.  A lane's
// 3x3 blocks for S consecutive sources are staged in LDS and the wave then
// writes the 3*64 matrix rows of the tile as contiguous row segments of 96
// bytes (4 sources in FP64, 8 in FP32), so that a cache line is written by
// at most two consecutive tiles of one wave instead of by 64 rows' worth of
// interleaved 24-byte stores; the tile (20 KB) still allows two waves per
// SIMD.  Measured on the 8,000-tetrahedron yoke: 2.95 ms against 4.15 ms
// unstaged (FP64), 2.41 against 3.03 ms (FP32); wider tiles cost occupancy.
// Sources whose columns are not consecutive (mixed kinds) are stored
// directly.
//
// Symmetric-block layout (PK = 5 or 6): without transform copies and row
// transforms every block is the symmetric tensor Q(x_i) of source j, so an
// element row is stored as PK planes (xx, xy, xz, yy, yz[, zz]) of ld values,
// plane k holding entry k of the blocks (i, 0 .. n-1); with PK = 5 the zz
// entry is -(xx + yy), exact for a collocation point outside the source.  A
// block with a trace beyond 1/4 (a collocation point inside another element)
// sets *flag, and the caller stores the matrix with 6 planes instead.
#ifndef RH_ASM_S64
#define RH_ASM_S64 4
#endif
#ifndef RH_ASM_S32
#define RH_ASM_S32 8
#endif
// symmetric-block tiles: 8 (FP64) or 16 (FP32) sources, 64-byte plane
// segments; on the 40,000-tetrahedron yoke 49.2 / 38.5 ms against 55.1 / 40.4
// with 4 / 8 sources and 80.7 / 48.1 with 2 / 4 (padded tiles); unpadded
// swizzled tiles 46.2 / 37.2 against 47.4 / 39.0 ms padded.  Three waves
// per SIMD (RH_ASM_WAVES=3: 154 VGPRs) with 4 / 8 sources measured 58-63 /
// 40 ms: the wider plane segments matter more than the occupancy
#ifndef RH_ASM_P64
#define RH_ASM_P64 8
#endif
#ifndef RH_ASM_P32
#define RH_ASM_P32 16
#endif
template <class OUT, int PK>
struct AsmTile {
    static constexpr int RPE = PK ? PK : 3;  // stored rows per element row
    static constexpr int VPS = PK ? 1 : 3;   // values per source in a stored row
    static constexpr int S = PK ? ((sizeof(OUT) == 8) ? RH_ASM_P64 : RH_ASM_P32)
                                : ((sizeof(OUT) == 8) ? RH_ASM_S64 : RH_ASM_S32);  // sources per tile
    static constexpr int U = VPS * S * int(sizeof(OUT)) / 8;  // 8-byte units per row segment
    static constexpr int EPU = 8 / int(sizeof(OUT));          // elements per 8-byte unit
    // LDS row stride (elements): row-major tiles pad a unit against bank
    // conflicts; plane tiles are unpadded (20 KB per wave, so that two waves
    // per SIMD fit the CU's LDS) with the units of the rows of lane l stored
    // at unit index u ^ (l & SWZ) instead
    static constexpr int W = (PK ? U : U + 1) * EPU;
    static constexpr int SWZ = PK ? U - 1 : 0;
    static_assert(!PK || (U & (U - 1)) == 0, "plane tiles swizzle a power-of-two number of units");
    // element index of element c of tile row r (rows RPE l .. RPE l + RPE - 1 belong to lane l)
    __device__ static int at(int r, int c) { return r * W + (((c / EPU) ^ ((r / RPE) & SWZ)) * EPU + c % EPU); }
};

#ifdef RH_ASM_WAVES
#define RH_ASM_ATTR __attribute__((amdgpu_waves_per_eu(RH_ASM_WAVES)))
#else
#define RH_ASM_ATTR
#endif
// FAR32: far blocks in FP32 arithmetic, near ones in FP64 (block_Q_mixed).
// PK: 0 row-major 3x3 blocks, 5 or 6 the symmetric-block planes (then ld is
// the plane stride and SYM, ROWT are false).
template <class R, bool SYM, bool ROWT, class OUT, int KIND, bool FAR32 = false, int PK = 0>
__global__ __launch_bounds__(64) RH_ASM_ATTR void k_assemble(DevModel m, OUT* __restrict__ Q, int64_t ld,
                                                 double* __restrict__ qdiag, int64_t jt, int64_t r0,
                                                 int64_t r1, int32_t* __restrict__ flag)
{
    static_assert(PK == 0 || (!SYM && !ROWT), "the symmetric-block layout needs lab-frame blocks");
    using TL = AsmTile<OUT, PK>;
    constexpr int S = TL::S, W = TL::W, RPE = TL::RPE, VPS = TL::VPS;
    __shared__ OUT tile[RPE * 64 * W];
    const int lane = threadIdx.x;
    const int64_t i0 = r0 + int64_t(blockIdx.x) * 64;
    const int64_t i = i0 + lane;
    const bool live = i < r1;
    const int nrow = int((r1 - i0 < 64) ? (r1 - i0) : 64);
    const int64_t kend = m.koff[KIND + 1];
    const int64_t k0 = m.koff[KIND] + int64_t(blockIdx.y) * jt;
    const int64_t k1 = (k0 + jt < kend) ? (k0 + jt) : kend;
    const ModelView mv = view(m);
    const int64_t ic = live ? i : i0;  // idle lanes evaluate row i0 and store nothing
    const double x[3] = {m.colloc[3 * ic], m.colloc[3 * ic + 1], m.colloc[3 * ic + 2]};
    double rt[9];
    if (ROWT)
        for (int k = 0; k < 9; k++) rt[k] = m.rowt[9 * ic + k];
    OUT* row = Q + (RPE * (ic - r0)) * ld;
    bool traced = false;
    for (int64_t kb = k0; kb < k1; kb += S) {
        const int cnt = int((k1 - kb < S) ? (k1 - kb) : S);
        const int64_t jb = m.kidx[kb];
        const bool run = (m.kidx[kb + cnt - 1] - jb == cnt - 1);  // consecutive columns
        for (int s = 0; s < cnt; s++) {
            const int64_t j = m.kidx[kb + s];
            double q[9];
            if constexpr (FAR32)
                block_Q_mixed<SYM, KIND>(mv, j, x, q);
            else
                block_Q<R, SYM, KIND>(mv, j, x, q);
            if (ROWT) {
                double t[9];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        t[3 * a + b] = rt[3 * a] * q[b] + rt[3 * a + 1] * q[3 + b] + rt[3 * a + 2] * q[6 + b];
                for (int k = 0; k < 9; k++) q[k] = t[k];
            }
            if (live && j == i) {
                for (int k = 0; k < 9; k++) {
                    qdiag[9 * i + k] = q[k];
                    q[k] = 0.0;
                }
            }
            // stored values: the 3x3 block row by row, or the planes
            double v[RPE * VPS];
            if constexpr (PK == 0) {
#pragma unroll
                for (int k = 0; k < 9; k++) v[k] = q[k];
            } else {
                if (PK == 5) traced |= fabs(q[0] + q[4] + q[8]) > 0.25;
                v[0] = q[0];
                v[1] = 0.5 * (q[1] + q[3]);
                v[2] = 0.5 * (q[2] + q[6]);
                v[3] = q[4];
                v[4] = 0.5 * (q[5] + q[7]);
                if constexpr (PK == 6) v[5] = q[8];
            }
            if (run) {
#pragma unroll
                for (int a = 0; a < RPE; a++)
#pragma unroll
                    for (int b = 0; b < VPS; b++) tile[TL::at(RPE * lane + a, VPS * s + b)] = OUT(v[VPS * a + b]);
            } else if (live) {
                OUT* o = row + VPS * j;
#pragma unroll
                for (int a = 0; a < RPE; a++)
#pragma unroll
                    for (int b = 0; b < VPS; b++) store_out(o + a * ld + b, v[VPS * a + b]);
            }
        }
        if (run) {
            __syncthreads();
            OUT* base = Q + (RPE * (i0 - r0)) * ld + VPS * jb;
            constexpr int U = TL::U, RP = 64 / U;  // rows per pass
            constexpr int TR = RPE * 64;            // stored rows of a full tile
            if (cnt == S && nrow == 64 && ((VPS * jb * int64_t(sizeof(OUT))) & 7) == 0) {
                // full tile: 8-byte units, RP rows of U units per pass, offsets
                // in 32 bits from a uniform base (the tile spans < 2^32 bytes)
                const int lr = lane / U, lc = lane - U * lr;
                if (lr < RP) {
                    constexpr int WU = W / TL::EPU;
                    const uint64_t* src = reinterpret_cast<const uint64_t*>(tile);
                    // unit lc of tile row r
                    auto unit = [&](int r) { return src[r * WU + (lc ^ ((r / RPE) & TL::SWZ))]; };
                    char* dst = reinterpret_cast<char*>(base);  // uniform, advanced per pass
                    const int64_t step = int64_t(RP) * ld * int64_t(sizeof(OUT));
                    const uint32_t off = uint32_t(lr) * uint32_t(ld * sizeof(OUT)) + uint32_t(8 * lc);
                    constexpr int full = TR / RP;  // passes with every lane's row in the tile
#pragma unroll 4
                    for (int it = 0; it < full; it++) {
                        *reinterpret_cast<uint64_t*>(dst + off) = unit(lr + RP * it);
                        dst += step;
                    }
                    if (RP * full + lr < TR) *reinterpret_cast<uint64_t*>(dst + off) = unit(lr + RP * full);
                }
            } else {
                // partial tile: the RPE*nrow stored rows, VPS*cnt values each
                const int cols = VPS * cnt, total = RPE * nrow * cols;
                const int step_r = 64 / cols, step_c = 64 - step_r * cols;
                int r = lane / cols, c = lane - r * cols;
                for (int e = lane; e < total; e += 64) {
                    base[int64_t(r) * ld + c] = tile[TL::at(r, c)];
                    c += step_c;
                    r += step_r;
                    if (c >= cols) {
                        c -= cols;
                        r++;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (PK == 5 && live && traced) *flag = 1;
}

template <bool SYM, bool ROWT, class OUT, int PK = 0>
hipError_t asm_dispatch(const DevModel& m, OUT* Q, int64_t ld, double* qdiag, int64_t r0, int64_t r1,
                        hipStream_t s, bool far32, int32_t* flag)
{
    if (r1 <= r0) return hipSuccess;
    const int64_t gx = (r1 - r0 + 63) / 64;
    for (int slot = 0; slot < 3; slot++) {
        const int64_t cnt = m.koff[slot + 1] - m.koff[slot];
        if (cnt == 0) continue;
        // enough workgroups to fill 256 CUs several times over; column tiles
        // are whole LDS tiles
        constexpr int S = AsmTile<OUT, PK>::S;
        const int64_t gy_target = std::max<int64_t>(1, (8192 + gx - 1) / gx);
        int64_t jt = std::max<int64_t>(4, (cnt + gy_target - 1) / gy_target);
        jt = std::min<int64_t>((jt + S - 1) / S * S, 256);
        const dim3 grid{unsigned(gx), unsigned((cnt + jt - 1) / jt), 1u};
        // FP32 storage: far blocks in FP32 arithmetic (general polyhedra stay FP64)
        constexpr bool mixed = sizeof(OUT) == 4;
        if (slot == 0 && mixed && far32)
            hipLaunchKernelGGL((k_assemble<double, SYM, ROWT, OUT, 0, mixed, PK>), grid, dim3(64), 0, s, m, Q, ld, qdiag, jt, r0, r1, flag);
        else if (slot == 0)
            hipLaunchKernelGGL((k_assemble<double, SYM, ROWT, OUT, 0, false, PK>), grid, dim3(64), 0, s, m, Q, ld, qdiag, jt, r0, r1, flag);
        else if (slot == 1 && mixed && far32)
            hipLaunchKernelGGL((k_assemble<double, SYM, ROWT, OUT, 1, mixed, PK>), grid, dim3(64), 0, s, m, Q, ld, qdiag, jt, r0, r1, flag);
        else if (slot == 1)
            hipLaunchKernelGGL((k_assemble<double, SYM, ROWT, OUT, 1, false, PK>), grid, dim3(64), 0, s, m, Q, ld, qdiag, jt, r0, r1, flag);
        else
            hipLaunchKernelGGL((k_assemble<double, SYM, ROWT, OUT, 2, false, PK>), grid, dim3(64), 0, s, m, Q, ld, qdiag, jt, r0, r1, flag);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------- field
// Sums the elements kidx[koff[KIND] + y*js ...] of one kind slot; slice
// `slice0 + blockIdx.y` of the partial buffer (or B/H directly).
// The FP32 sweep runs as two launches over the same grid: kPassFar (FP32,
// the elements no lane of the wave is near) stores the sums, kPassNear (FP64,
// the others) adds to them.
template <int PASS>
__device__ __forceinline__ void field_store(double* __restrict__ o, const double b[3], const double h[3])
{
    for (int k = 0; k < 3; k++) {
        if (PASS == kPassNear) {
            o[k] += b[k];
            o[3 + k] += h[k];
        } else {
            o[k] = b[k];
            o[3 + k] = h[k];
        }
    }
}
template <int PASS>
__device__ __forceinline__ void field_store(double* __restrict__ B, double* __restrict__ H, int64_t p,
                                            const double b[3], const double h[3])
{
    for (int k = 0; k < 3; k++) {
        if (B) B[3 * p + k] = (PASS == kPassNear) ? B[3 * p + k] + b[k] : b[k];
        if (H) H[3 * p + k] = (PASS == kPassNear) ? H[3 * p + k] + h[k] : h[k];
    }
}

template <class R, bool SYM, int KIND, int PASS>
__global__ __launch_bounds__(256) void k_field(DevModel m, const double* __restrict__ M,
                                               const double* __restrict__ pts, int64_t np,
                                               int64_t js, int64_t slice0, double* __restrict__ B,
                                               double* __restrict__ H, double* __restrict__ part)
{
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= np) return;
    const int64_t kend = m.koff[KIND + 1];
    const int64_t k0 = m.koff[KIND] + int64_t(blockIdx.y) * js;
    const int64_t k1 = (k0 + js < kend) ? (k0 + js) : kend;
    const double x[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    double b[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    field_add<R, SYM, KIND, PASS>(view(m), M, m.kidx, k0, k1, x, b, h);
    if (part)
        field_store<PASS>(part + ((slice0 + int64_t(blockIdx.y)) * np + p) * 6, b, h);
    else
        field_store<PASS>(B, H, p, b, h);
}

// Tetrahedron field sweep over charged records staged in LDS: the workgroup
// copies kTetTile records (86 doubles each) with coalesced 16-byte loads, and
// every lane then reads them at wave-uniform LDS addresses (broadcast, no bank
// conflicts) while it evaluates its own point.  Staging keeps the record out of
// scalar registers, whose pressure otherwise spills through VGPR lanes.
constexpr int kTetTile = 32;
// occupancy floor for the tetrahedron sweep (waves per SIMD): bounds
// the VGPRs the scheduler may use for hoisted LDS reads
#ifndef RH_TET_WAVES
#define RH_TET_WAVES 3
#endif
// mesh sweep: occupancy floor (waves per SIMD) and edge and face records
// staged per LDS tile.  FP64: 6 waves (80 VGPRs, 8 bytes of spill outside
// the record loops) measured 50.9 / 566 ms on the C-yoke / Halbach ring
// against 51.3 / 569 ms at 5 waves (85 VGPRs); FP32: 7 waves (72 VGPRs),
// the records near a wave evaluated in FP64 after each tile's loop.
#ifndef RH_MESH_WAVES32
#define RH_MESH_WAVES32 7
#endif
#ifndef RH_MESH_WAVES
#define RH_MESH_WAVES 6
#endif
#ifndef RH_MESH_STAGE32
#define RH_MESH_STAGE32 64
#endif
#ifndef RH_MESH_STAGE
#define RH_MESH_STAGE 64
#endif

template <class R, bool SYM, int PASS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH_TET_WAVES))) void k_field_tet(DevModel m, const double* __restrict__ pts,
                                                   int64_t np, int64_t js, int64_t slice0,
                                                   double* __restrict__ B, double* __restrict__ H,
                                                   double* __restrict__ part)
{
    __shared__ __align__(16) double tile[kTetTile * kChargedTet];
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const bool live = p < np;
    const int64_t ntet = m.koff[2] - m.koff[1];
    const int64_t t0 = int64_t(blockIdx.y) * js;
    const int64_t t1 = (t0 + js < ntet) ? (t0 + js) : ntet;
    double x[3] = {0.0, 0.0, 0.0};
    if (live) {
        x[0] = pts[3 * p];
        x[1] = pts[3 * p + 1];
        x[2] = pts[3 * p + 2];
    }
    double b[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    for (int64_t base = t0; base < t1; base += kTetTile) {
        const int cnt = int((t1 - base < kTetTile) ? (t1 - base) : kTetTile);
        const double2* src = reinterpret_cast<const double2*>(m.charged + kChargedTet * base);
        double2* dst = reinterpret_cast<double2*>(tile);
        __syncthreads();  // the previous tile has been consumed
        for (int i = threadIdx.x; i < cnt * (kChargedTet / 2); i += 256) dst[i] = src[i];
        __syncthreads();
        if (!live) continue;
#pragma unroll 1
        for (int q = 0; q < cnt; q++) {
            const double* rec = tile + kChargedTet * q;
            const double mm[3] = {rec[kChargedM], rec[kChargedM + 1], rec[kChargedM + 2]};
            if (!SYM) {
                double hl[3];
                if (PASS != kPassAll && tet_near32(rec, x) != (PASS == kPassNear)) continue;
                const bool in = field_tet<R>(rec, x, hl);
                for (int k = 0; k < 3; k++) {
                    h[k] += hl[k];
                    b[k] += hl[k] + (in ? mm[k] : 0.0);
                }
                continue;
            }
            const int64_t j = m.kidx[m.koff[1] + base + q];
            for (int64_t c = m.copy_off[j]; c < m.copy_off[j + 1]; c++) {
                const double* cp = m.copies + 24 * c;
                double xl[3], hl[3], hg[3], bg[3];
                apply_inv(cp, x, xl);
                if (PASS != kPassAll && tet_near32(rec, xl) != (PASS == kPassNear)) continue;
                const bool in = field_tet<R>(rec, xl, hl);
                const double bl[3] = {hl[0] + (in ? mm[0] : 0.0), hl[1] + (in ? mm[1] : 0.0),
                                      hl[2] + (in ? mm[2] : 0.0)};
                mat3_vec(cp + 12, hl, hg);
                mat3_vec(cp + 12, bl, bg);
                for (int k = 0; k < 3; k++) {
                    h[k] += hg[k];
                    b[k] += bg[k];
                }
            }
        }
    }
    if (!live) return;
    if (part)
        field_store<PASS>(part + ((slice0 + int64_t(blockIdx.y)) * np + p) * 6, b, h);
    else
        field_store<PASS>(B, H, p, b, h);
}

// ---------------------------------------------------------------- mesh form
// Charged records of the mesh form for magnetization M (one lane per entity).
__global__ __launch_bounds__(256) void k_mesh_charge_edges(MeshDev md, const double* __restrict__ M, bool pack32)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= md.ne) return;
    double w[3] = {0.0, 0.0, 0.0};
    for (int64_t k = md.edge_inc_off[i]; k < md.edge_inc_off[i + 1]; k++) {
        const int64_t j = md.edge_inc_elem[k];
        const double* W = md.edge_inc_w + 9 * k;  // F_c W_e, row-major
        const double m0 = M[3 * j], m1 = M[3 * j + 1], m2 = M[3 * j + 2];
        w[0] += W[0] * m0 + W[1] * m1 + W[2] * m2;
        w[1] += W[3] * m0 + W[4] * m1 + W[5] * m2;
        w[2] += W[6] * m0 + W[7] * m1 + W[8] * m2;
    }
    const double* g = md.edge_geo + kMeshEdgeGeo * i;
    double* o = md.ch_edge + kMeshEdge * i;
    for (int k = 0; k < kMeshEdgeGeo; k++) o[k] = g[k];
    o[10] = w[0];
    o[11] = w[1];
    o[12] = w[2];
    o[13] = md.edge_cont[i];
    if (pack32) {
        float* f = reinterpret_cast<float*>(o + 6);
        const float t0 = float(g[6]), t1 = float(g[7]), t2 = float(g[8]), len = float(g[9]);
        f[0] = t0;
        f[1] = t1;
        f[2] = t2;
        f[3] = len;
        // slot 8: the near distance of the FP32 sweep (kNear32 lengths)
        f[4] = kNear32 * len;
        f[5] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_mesh_charge_faces(MeshDev md, const double* __restrict__ M, bool pack32)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= md.nf) return;
    const double* g = md.face_geo + kMeshFaceGeo * i;
    const double n0 = g[12], n1 = g[13], n2 = g[14];
    double s = 0.0;
    for (int64_t k = md.face_inc_off[i]; k < md.face_inc_off[i + 1]; k++) {
        const int64_t j = md.face_inc_elem[k];
        const double* v = md.face_inc_v + 3 * k;
        s += v[0] * M[3 * j] + v[1] * M[3 * j + 1] + v[2] * M[3 * j + 2];
    }
    double* o = md.ch_face + kMeshFace * i;
    for (int k = 0; k < kMeshFaceGeo; k++) o[k] = g[k];
    o[18] = s * n0;
    o[19] = s * n1;
    o[20] = s * n2;
    o[21] = md.face_cont[i];
    // owned edges: length and the net weight of the charged edge record
    // (charged before the faces, on the same stream)
    int32_t mask = 0;
    for (int k = 0; k < 3; k++) {
        const int64_t e = md.face_edge[3 * i + k];
        double* d = o + kMeshFaceEdges + 4 * k;
        if (e >= 0) {
            const double* ce = md.ch_edge + kMeshEdge * e;
            d[0] = md.edge_geo[kMeshEdgeGeo * e + 9];
            d[1] = ce[10];
            d[2] = ce[11];
            d[3] = ce[12];
            mask |= 1 << k;
        } else {
            d[0] = d[1] = d[2] = d[3] = 0.0;
        }
    }
    int32_t* mk = reinterpret_cast<int32_t*>(o + kMeshFaceEdges + 12);
    mk[0] = mask;
    mk[1] = 0;
    o[kMeshFaceEdges + 13] = 0.0;
    if (pack32) {
        // slot 35: the near distance of the FP32 sweep (kNear32 longest sides)
        const int nv = g[17] > 3.5 ? 4 : 3;
        double lmax = 0.0;
        for (int k = 0; k < nv; k++) {
            const double* a = g + 3 * k;
            const double* b = g + 3 * ((k + 1) % nv);
            const double d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
            lmax = fmax(lmax, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
        }
        reinterpret_cast<float*>(o + kMeshFaceEdges + 13)[0] = float(double(kNear32) * lmax);
    }
    if (pack32) {
        float* f = reinterpret_cast<float*>(o + 12);
        const float a1 = float(g[15]), a2 = float(g[16]);
        f[0] = float(n0);
        f[1] = float(n1);
        f[2] = float(n2);
        f[3] = a1;
        f[4] = a2;
        f[5] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_mesh_charge_in(MeshDev md, const double* __restrict__ M)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= md.ni) return;
    const int64_t j = md.in_elem[i];
    const double* g = md.in_geo + (1 + 4 * kMeshPlanes) * i;
    double* o = md.ch_in + kMeshIn * i;
    const double* F = md.in_F + 9 * i;  // the copy's field transform
    const double m0 = M[3 * j], m1 = M[3 * j + 1], m2 = M[3 * j + 2];
    o[0] = g[0];
    o[1] = F[0] * m0 + F[1] * m1 + F[2] * m2;
    o[2] = F[3] * m0 + F[4] * m1 + F[5] * m2;
    o[3] = F[6] * m0 + F[7] * m1 + F[8] * m2;
    for (int k = 0; k < 4 * kMeshPlanes; k++) o[4 + k] = g[1 + k];
}

// Field sweep of the mesh form: slice blockIdx.y of the unique edges, unique
// faces and inside tiles, each staged through LDS in tiles of kMeshTile
// records read at wave-uniform addresses.  Inside tiles whose bounding box
// holds no point of the wave are skipped (wave-uniform vote).
template <class R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? RH_MESH_WAVES32 : RH_MESH_WAVES))) void
k_field_mesh(DevModel m, const double* __restrict__ pts, int64_t np, int64_t S, int64_t slice0,
             double* __restrict__ B, double* __restrict__ H, double* __restrict__ part)
{
    constexpr int kStage = (sizeof(R) == 4) ? RH_MESH_STAGE32 : RH_MESH_STAGE;
    __shared__ __align__(16) double tile[kStage * kMeshFace > kMeshTile * kMeshIn ? kStage * kMeshFace
                                                                                     : kMeshTile * kMeshIn];
    const MeshDev& md = m.mesh;
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const bool live = p < np;
    const int64_t y = blockIdx.y;
    double x[3] = {0.0, 0.0, 0.0};
    if (live) {
        x[0] = pts[3 * p];
        x[1] = pts[3 * p + 1];
        x[2] = pts[3 * p + 2];
    }
    double h[3] = {0.0, 0.0, 0.0}, min[3] = {0.0, 0.0, 0.0};
    double2* dst = reinterpret_cast<double2*>(tile);
    // unique edges that no triangle owns (edges of quadrilaterals only):
    // + L_e w_e, along paths: a continued edge takes its first vertex's
    // offset and distance from the previous edge.  The loop is unrolled by
    // two with the vertex slots P, Q swapping roles, so the carried values
    // stay in place (no register copies per edge).
    // FP32: the records of a tile that the wave is near (mesh32_near) are
    // marked in `defer` and evaluated in FP64 after the tile's loop, where
    // the loop's carried values are dead, so the FP64 code does not add to
    // the loop's registers
    static_assert(kStage <= 64, "one bit per staged record");
    uint64_t defer = 0;
    {
        MeshVtx<R> P{}, Q{};
        const int64_t e0 = md.ne_free * y / S, e1 = md.ne_free * (y + 1) / S;
        auto edge = [&](int64_t g, int q, MeshVtx<R>& a, MeshVtx<R>& b) {
            const double* rec = tile + kMeshEdge * q;
            if (!(mesh_code(rec + 13) == 1 && g != e0)) a = mesh_vtx<R>(rec, x);
            b = mesh_vtx<R>(rec + 3, x);
            // L_e from the end offsets, distances and the length (the
            // FP32 records carry the length as a float in slot 7, and keep
            // it in FP64 in slot 9 for the waves near the edge)
            if constexpr (sizeof(R) == 4) {
                if (mesh32_near(fminf(a.d, b.d), mesh32(rec, 8)[0])) {
                    defer |= uint64_t(1) << q;
                } else {
                    const double L = double(edge_L(a.r, a.d, b.r, b.d, mesh32(rec, 6)[3]));
                    h[0] += L * rec[10];
                    h[1] += L * rec[11];
                    h[2] += L * rec[12];
                }
            } else {
                const double L = double(edge_L(a.r, a.d, b.r, b.d, R(rec[9])));
                h[0] += L * rec[10];
                h[1] += L * rec[11];
                h[2] += L * rec[12];
            }
        };
        for (int64_t base = e0; base < e1; base += kStage) {
            const int cnt = int((e1 - base < kStage) ? (e1 - base) : kStage);
            const double2* src = reinterpret_cast<const double2*>(md.ch_edge + kMeshEdge * base);
            __syncthreads();
            for (int i = threadIdx.x; i < cnt * (kMeshEdge / 2); i += 256) dst[i] = src[i];
            __syncthreads();
            if (!live) continue;
            int q = 0;
#pragma unroll 1
            for (; q + 2 <= cnt; q += 2) {
                edge(base + q, q, P, Q);
                edge(base + q + 1, q + 1, Q, P);
            }
            if (q < cnt) {
                edge(base + q, q, P, Q);
                P = Q;
            }
            if constexpr (sizeof(R) == 4) {
                for (; defer; defer &= defer - 1) {
                    const double* rec = tile + kMeshEdge * __builtin_ctzll(defer);
                    const double L = mesh_edge_L_near(rec, x);
                    h[0] += L * rec[10];
                    h[1] += L * rec[11];
                    h[2] += L * rec[12];
                }
            }
        }
    }
    // unique faces: - Omega_f g_f, triangles along strips: a continued
    // triangle takes its first two vertices (and their dot product) from the
    // previous one.  Unrolled by three with the slots A, B, C rotating.  A
    // triangle also evaluates the edges it owns (+ L_e w_e) from the
    // offsets, distances and dot products of its vertices.
    {
        MeshVtx<R> A{}, Bv{}, C{}, QA{}, QB{}, QC{}, QD{};
        R dAB = R(0), dBC = R(0), dCA = R(0), dQAB = R(0), dQCD = R(0);
        // this slice's faces: triangles [f0, ft), then quadrilaterals [fq, f1)
        const int64_t f0 = md.nf * y / S, f1 = md.nf * (y + 1) / S;
        const int64_t ft = f1 < md.nf_tri ? f1 : md.nf_tri, fq = f0 > md.nf_tri ? f0 : md.nf_tri;
        // an edge owned by the current face: the offsets and distances of its
        // ends and their dot product are already at hand
        auto owned = [&](const MeshVtx<R>& u, const MeshVtx<R>& v, R duv, const double* e) {
            const double L = double(edge_L(u.r, u.d, v.r, v.d, duv, R(e[0])));
            h[0] += L * e[1];
            h[1] += L * e[2];
            h[2] += L * e[3];
        };
        // FP64 for the records of the tile the wave is near
        auto deferred = [&]() {
            if constexpr (sizeof(R) == 4) {
                for (; defer; defer &= defer - 1) {
                    const double* rec = tile + kMeshFace * __builtin_ctzll(defer);
                    const MeshFace64 t = mesh_face_near(rec, mesh_mask(rec + kMeshFaceEdges + 12), x);
                    h[0] += t.h0 - t.om * rec[18];
                    h[1] += t.h1 - t.om * rec[19];
                    h[2] += t.h2 - t.om * rec[20];
                }
            }
        };
        auto tri = [&](int64_t g, int q, MeshVtx<R>& a, MeshVtx<R>& b, MeshVtx<R>& c, R& dab, R& dbc) {
            const double* rec = tile + kMeshFace * q;
            if (!(mesh_code(rec + 21) == 1 && g != f0)) {
                a = mesh_vtx<R>(rec, x);
                b = mesh_vtx<R>(rec + 3, x);
                dab = dot(a.r, b.r);
            }
            c = mesh_vtx<R>(rec + 6, x);
            dbc = dot(b.r, c.r);
            double om;
            if constexpr (sizeof(R) == 4) {
                // a near record only sets its bit; the carried values
                // (b, c, dbc) are in place either way
                if (mesh32_near(fminf(fminf(a.d, b.d), c.d), mesh32(rec, kMeshFaceEdges + 13)[0])) {
                    defer |= uint64_t(1) << q;
                    return;
                }
                const R dca = dot(c.r, a.r);
                const int mask = mesh_mask(rec + kMeshFaceEdges + 12);
                om = double(mesh32_tri_omega(a.r, a.d, b.r, b.d, c.r, c.d, dab, dbc, dca, rec));
                if (mask & 1) owned(a, b, dab, rec + kMeshFaceEdges);
                if (mask & 2) owned(b, c, dbc, rec + kMeshFaceEdges + 4);
                if (mask & 4) owned(c, a, dca, rec + kMeshFaceEdges + 8);
            } else {
                const R dca = dot(c.r, a.r);
                const int mask = mesh_mask(rec + kMeshFaceEdges + 12);
                om = mesh_tri_omega_from<R>(a.r, a.d, b.r, b.d, c.r, c.d, dab, dbc, dca, rec);
                if (mask & 1) owned(a, b, dab, rec + kMeshFaceEdges);
                if (mask & 2) owned(b, c, dbc, rec + kMeshFaceEdges + 4);
                if (mask & 4) owned(c, a, dca, rec + kMeshFaceEdges + 8);
            }
            h[0] -= om * rec[18];
            h[1] -= om * rec[19];
            h[2] -= om * rec[20];
        };
        // a quadrilateral with the sides (v0,v1), (v1,v2), (v2,v3) it owns,
        // along strips: a continued quadrilateral takes v0, v1 and their dot
        // product from the previous one's v2, v3 (slots a, b, dab)
        auto quad = [&](int64_t g, int q, MeshVtx<R>& a, MeshVtx<R>& b, MeshVtx<R>& c, MeshVtx<R>& d, R& dab,
                        R& dcd) {
            const double* rec = tile + kMeshFace * q;
            if (!(mesh_code(rec + 21) == 3 && g != fq)) {
                a = mesh_vtx<R>(rec, x);
                b = mesh_vtx<R>(rec + 3, x);
                dab = dot(a.r, b.r);
            }
            c = mesh_vtx<R>(rec + 6, x);
            d = mesh_vtx<R>(rec + 9, x);
            dcd = dot(c.r, d.r);
            if constexpr (sizeof(R) == 4) {
                if (mesh32_near(fminf(fminf(a.d, b.d), fminf(c.d, d.d)), mesh32(rec, kMeshFaceEdges + 13)[0])) {
                    defer |= uint64_t(1) << q;
                    return;
                }
            }
            const MeshVtx<R> v[4] = {a, b, c, d};
            const R dbc = dot(b.r, c.r);
            const double om =
                double(mesh_quad_omega_from<R>(v, dab, dot(a.r, c.r), dot(a.r, d.r), dbc, dcd, rec));
            const int mask = mesh_mask(rec + kMeshFaceEdges + 12);
            if (mask & 1) owned(a, b, dab, rec + kMeshFaceEdges);
            if (mask & 2) owned(b, c, dbc, rec + kMeshFaceEdges + 4);
            if (mask & 4) owned(c, d, dcd, rec + kMeshFaceEdges + 8);
            h[0] -= om * rec[18];
            h[1] -= om * rec[19];
            h[2] -= om * rec[20];
        };
        for (int64_t base = f0; base < ft; base += kStage) {
            const int cnt = int((ft - base < kStage) ? (ft - base) : kStage);
            const double2* src = reinterpret_cast<const double2*>(md.ch_face + kMeshFace * base);
            __syncthreads();
            for (int i = threadIdx.x; i < cnt * (kMeshFace / 2); i += 256) dst[i] = src[i];
            __syncthreads();
            if (!live) continue;
            int q = 0;
#pragma unroll 1
            for (; q + 3 <= cnt; q += 3) {
                tri(base + q, q, A, Bv, C, dAB, dBC);
                tri(base + q + 1, q + 1, Bv, C, A, dBC, dCA);
                tri(base + q + 2, q + 2, C, A, Bv, dCA, dAB);
            }
            // remainder, then the carried pair back into (A, B)
            if (q < cnt) {
                tri(base + q, q, A, Bv, C, dAB, dBC);
                if (q + 1 < cnt) {
                    tri(base + q + 1, q + 1, Bv, C, A, dBC, dCA);
                    Bv = A;
                    A = C;
                    dAB = dCA;
                } else {
                    A = Bv;
                    Bv = C;
                    dAB = dBC;
                }
            }
            deferred();
        }
        for (int64_t base = fq; base < f1; base += kStage) {
            const int cnt = int((f1 - base < kStage) ? (f1 - base) : kStage);
            const double2* src = reinterpret_cast<const double2*>(md.ch_face + kMeshFace * base);
            __syncthreads();
            for (int i = threadIdx.x; i < cnt * (kMeshFace / 2); i += 256) dst[i] = src[i];
            __syncthreads();
            if (!live) continue;
            int q = 0;
#pragma unroll 1
            for (; q + 2 <= cnt; q += 2) {
                quad(base + q, q, QA, QB, QC, QD, dQAB, dQCD);
                quad(base + q + 1, q + 1, QC, QD, QA, QB, dQCD, dQAB);
            }
            // remainder, then the carried pair back into (QA, QB)
            if (q < cnt) {
                quad(base + q, q, QA, QB, QC, QD, dQAB, dQCD);
                QA = QC;
                QB = QD;
                dQAB = dQCD;
            }
            deferred();
        }
    }
    // inside tests: B gets the magnetization of the element holding the point
    {
        const int64_t t0 = md.ntiles * y / S, t1 = md.ntiles * (y + 1) / S;
        for (int64_t t = t0; t < t1; t++) {
            const double* bx = md.tile_box + 6 * t;
            const bool near = live && x[0] >= bx[0] && x[1] >= bx[1] && x[2] >= bx[2] &&
                              x[0] <= bx[3] && x[1] <= bx[4] && x[2] <= bx[5];
            if (!__syncthreads_or(near ? 1 : 0)) continue;
            const int64_t base = t * kMeshTile;
            const int cnt = int((md.ni - base < kMeshTile) ? (md.ni - base) : kMeshTile);
            const double2* src = reinterpret_cast<const double2*>(md.ch_in + kMeshIn * base);
            for (int i = threadIdx.x; i < cnt * (kMeshIn / 2); i += 256) dst[i] = src[i];
            __syncthreads();
            if (near) {
                for (int q = 0; q < cnt; q++) {
                    const double* rec = tile + kMeshIn * q;
                    if (mesh_inside(rec, x)) {
                        min[0] += rec[1];
                        min[1] += rec[2];
                        min[2] += rec[3];
                    }
                }
            }
            __syncthreads();
        }
    }
    if (!live) return;
    for (int k = 0; k < 3; k++) h[k] *= kInv4Pi;
    if (part) {
        double* o = part + ((slice0 + y) * np + p) * 6;
        for (int k = 0; k < 3; k++) {
            o[k] = h[k] + min[k];
            o[3 + k] = h[k];
        }
    } else {
        for (int k = 0; k < 3; k++) {
            if (B) B[3 * p + k] = h[k] + min[k];
            if (H) H[3 * p + k] = h[k];
        }
    }
}

__global__ __launch_bounds__(256) void k_field_reduce(int64_t np, int64_t splits,
                                                      const double* __restrict__ part,
                                                      double* __restrict__ B, double* __restrict__ H)
{
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= np) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t s = 0; s < splits; s++) {
        const double* o = part + (s * np + p) * 6;
        for (int k = 0; k < 6; k++) acc[k] += o[k];
    }
    for (int k = 0; k < 3; k++) {
        if (B) B[3 * p + k] = acc[k];
        if (H) H[3 * p + k] = acc[3 + k];
    }
}

// ---------------------------------------------------------------- matrix pass
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef double nd2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int RW>
__global__ __launch_bounds__(256) void k_matvec_f32(const float* __restrict__ Q, int64_t nrows,
                                                    int64_t ld, const double* __restrict__ x,
                                                    const double* __restrict__ add,
                                                    double* __restrict__ y,
                                                    const int32_t* __restrict__ skip)
{
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const int64_t r0 = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * RW;
    if (r0 >= nrows) return;
    double acc[RW];
#pragma unroll
    for (int r = 0; r < RW; r++) acc[r] = 0.0;
    const nf4* rows[RW];
#pragma unroll
    for (int r = 0; r < RW; r++) {
        const int64_t rr = (r0 + r < nrows) ? (r0 + r) : r0;
        rows[r] = reinterpret_cast<const nf4*>(Q + rr * ld);
    }
    const int64_t nv = ld / 4;  // float4 per row (ld multiple of 256)
    const nd2* x2 = reinterpret_cast<const nd2*>(x);
    for (int64_t c = lane; c < nv; c += 64) {
        const nd2 xa = x2[2 * c], xb = x2[2 * c + 1];
#pragma unroll
        for (int r = 0; r < RW; r++) {
            const nf4 q = __builtin_nontemporal_load(rows[r] + c);
            acc[r] += double(q.x) * xa.x + double(q.y) * xa.y + double(q.z) * xb.x + double(q.w) * xb.y;
        }
    }
#pragma unroll
    for (int r = 0; r < RW; r++) {
        const double v = wave_sum(acc[r]);
        if (lane == 0 && r0 + r < nrows) y[r0 + r] = v + (add ? add[r0 + r] : 0.0);
    }
}

template <int RW>
__global__ __launch_bounds__(256) void k_matvec_f64(const double* __restrict__ Q, int64_t nrows,
                                                    int64_t ld, const double* __restrict__ x,
                                                    const double* __restrict__ add,
                                                    double* __restrict__ y,
                                                    const int32_t* __restrict__ skip)
{
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const int64_t r0 = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * RW;
    if (r0 >= nrows) return;
    double acc[RW];
#pragma unroll
    for (int r = 0; r < RW; r++) acc[r] = 0.0;
    const nd2* rows[RW];
#pragma unroll
    for (int r = 0; r < RW; r++) {
        const int64_t rr = (r0 + r < nrows) ? (r0 + r) : r0;
        rows[r] = reinterpret_cast<const nd2*>(Q + rr * ld);
    }
    const int64_t nv = ld / 2;
    const nd2* x2 = reinterpret_cast<const nd2*>(x);
#ifdef RH_MV_UNROLL
#pragma unroll RH_MV_UNROLL
#endif
    for (int64_t c = lane; c < nv; c += 64) {
        const nd2 xa = x2[c];
#pragma unroll
        for (int r = 0; r < RW; r++) {
            const nd2 q = __builtin_nontemporal_load(rows[r] + c);
            acc[r] += q.x * xa.x + q.y * xa.y;
        }
    }
#pragma unroll
    for (int r = 0; r < RW; r++) {
        const double v = wave_sum(acc[r]);
        if (lane == 0 && r0 + r < nrows) y[r0 + r] = v + (add ? add[r0 + r] : 0.0);
    }
}

// Matrix pass over the symmetric-block layout: ER element rows per wave, their
// P planes (xx, xy, xz, yy, yz[, zz]) streamed with 16-byte non-temporal loads
// (two column elements per lane in FP64, four in FP32), the iterate read once
// per column chunk as the interleaved (x, y, z) triples of those columns and
// shared by the ER rows; FP64 accumulation.
// element rows per wave: on the 40,000-tetrahedron yoke (5 planes) 4 rows
// measured 9.57 ms FP64 / 4.87 ms FP32 per pass, 2 rows 9.75 / 5.04, 1 row
// 10.70 / 6.18 (the iterate's reads per matrix byte fall as 1/rows); a
// twofold unroll of the column loop gained nothing
#ifndef RH_MVS_ER
#define RH_MVS_ER 4
#endif
#ifndef RH_MVS_UNROLL
#define RH_MVS_UNROLL 1
#endif
template <int P, int ER>
__global__ __launch_bounds__(256) void k_matvec_sym_f64(const double* __restrict__ Q, int64_t nel,
                                                        int64_t ldp, const double* __restrict__ x,
                                                        const double* __restrict__ add,
                                                        double* __restrict__ y,
                                                        const int32_t* __restrict__ skip)
{
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * ER;
    if (i0 >= nel) return;
    const nd2* pl[ER][P];
#pragma unroll
    for (int r = 0; r < ER; r++) {
        const int64_t i = (i0 + r < nel) ? (i0 + r) : i0;  // rows past the end repeat row i0, unstored
#pragma unroll
        for (int k = 0; k < P; k++) pl[r][k] = reinterpret_cast<const nd2*>(Q + (i * P + k) * ldp);
    }
    const nd2* x2 = reinterpret_cast<const nd2*>(x);
    double h[ER][3];
#pragma unroll
    for (int r = 0; r < ER; r++) h[r][0] = h[r][1] = h[r][2] = 0.0;
    const int64_t nv = ldp / 2;
#pragma unroll RH_MVS_UNROLL
    for (int64_t c = lane; c < nv; c += 64) {
        nd2 q[ER][P];
#pragma unroll
        for (int r = 0; r < ER; r++)
#pragma unroll
            for (int k = 0; k < P; k++) q[r][k] = __builtin_nontemporal_load(pl[r][k] + c);
        // columns 2c and 2c+1: (a.x, a.y, b.x) and (b.y, d.x, d.y)
        const nd2 a = x2[3 * c], b = x2[3 * c + 1], d = x2[3 * c + 2];
#pragma unroll
        for (int r = 0; r < ER; r++) {
            const nd2* v = q[r];
            const nd2 zz = (P == 6) ? v[P - 1] : -(v[0] + v[3]);
            h[r][0] += v[0].x * a.x + v[1].x * a.y + v[2].x * b.x + v[0].y * b.y + v[1].y * d.x + v[2].y * d.y;
            h[r][1] += v[1].x * a.x + v[3].x * a.y + v[4].x * b.x + v[1].y * b.y + v[3].y * d.x + v[4].y * d.y;
            h[r][2] += v[2].x * a.x + v[4].x * a.y + zz.x * b.x + v[2].y * b.y + v[4].y * d.x + zz.y * d.y;
        }
    }
#pragma unroll
    for (int r = 0; r < ER; r++) {
        const double hx = wave_sum(h[r][0]), hy = wave_sum(h[r][1]), hz = wave_sum(h[r][2]);
        const int64_t i = i0 + r;
        if (lane == 0 && i < nel) {
            y[3 * i] = hx + (add ? add[3 * i] : 0.0);
            y[3 * i + 1] = hy + (add ? add[3 * i + 1] : 0.0);
            y[3 * i + 2] = hz + (add ? add[3 * i + 2] : 0.0);
        }
    }
}

template <int P, int ER>
__global__ __launch_bounds__(256) void k_matvec_sym_f32(const float* __restrict__ Q, int64_t nel,
                                                        int64_t ldp, const double* __restrict__ x,
                                                        const double* __restrict__ add,
                                                        double* __restrict__ y,
                                                        const int32_t* __restrict__ skip)
{
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * ER;
    if (i0 >= nel) return;
    const nf4* pl[ER][P];
#pragma unroll
    for (int r = 0; r < ER; r++) {
        const int64_t i = (i0 + r < nel) ? (i0 + r) : i0;  // rows past the end repeat row i0, unstored
#pragma unroll
        for (int k = 0; k < P; k++) pl[r][k] = reinterpret_cast<const nf4*>(Q + (i * P + k) * ldp);
    }
    const nd2* x2 = reinterpret_cast<const nd2*>(x);
    double h[ER][3];
#pragma unroll
    for (int r = 0; r < ER; r++) h[r][0] = h[r][1] = h[r][2] = 0.0;
    const int64_t nv = ldp / 4;
#pragma unroll RH_MVS_UNROLL
    for (int64_t c = lane; c < nv; c += 64) {
        nf4 q[ER][P];
#pragma unroll
        for (int r = 0; r < ER; r++)
#pragma unroll
            for (int k = 0; k < P; k++) q[r][k] = __builtin_nontemporal_load(pl[r][k] + c);
        // columns 4c .. 4c+3: twelve doubles of x
        nd2 m[6];
#pragma unroll
        for (int k = 0; k < 6; k++) m[k] = x2[6 * c + k];
        const double mv[12] = {m[0].x, m[0].y, m[1].x, m[1].y, m[2].x, m[2].y,
                               m[3].x, m[3].y, m[4].x, m[4].y, m[5].x, m[5].y};
#pragma unroll
        for (int r = 0; r < ER; r++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const double xx = q[r][0][e], xy = q[r][1][e], xz = q[r][2][e], yy = q[r][3][e],
                             yz = q[r][4][e];
                const double zz = (P == 6) ? double(q[r][P - 1][e]) : -(xx + yy);
                const double mx = mv[3 * e], my = mv[3 * e + 1], mz = mv[3 * e + 2];
                h[r][0] += xx * mx + xy * my + xz * mz;
                h[r][1] += xy * mx + yy * my + yz * mz;
                h[r][2] += xz * mx + yz * my + zz * mz;
            }
    }
#pragma unroll
    for (int r = 0; r < ER; r++) {
        const double hx = wave_sum(h[r][0]), hy = wave_sum(h[r][1]), hz = wave_sum(h[r][2]);
        const int64_t i = i0 + r;
        if (lane == 0 && i < nel) {
            y[3 * i] = hx + (add ? add[3 * i] : 0.0);
            y[3 * i + 1] = hy + (add ? add[3 * i + 1] : 0.0);
            y[3 * i + 2] = hz + (add ? add[3 * i + 2] : 0.0);
        }
    }
}

// ---------------------------------------------------------------- relaxation update
__device__ __forceinline__ void inv3d(const double* m, double* o)
{
    const double c0 = m[4] * m[8] - m[5] * m[7], c1 = m[2] * m[7] - m[1] * m[8],
                 c2 = m[1] * m[5] - m[2] * m[4];
    const double c3 = m[5] * m[6] - m[3] * m[8], c4 = m[0] * m[8] - m[2] * m[6],
                 c5 = m[2] * m[3] - m[0] * m[5];
    const double c6 = m[3] * m[7] - m[4] * m[6], c7 = m[1] * m[6] - m[0] * m[7],
                 c8 = m[0] * m[4] - m[1] * m[3];
    const double inv = 1.0 / (m[0] * c0 + m[1] * c3 + m[2] * c6);
    o[0] = c0 * inv; o[1] = c1 * inv; o[2] = c2 * inv;
    o[3] = c3 * inv; o[4] = c4 * inv; o[5] = c5 * inv;
    o[6] = c6 * inv; o[7] = c7 * inv; o[8] = c8 * inv;
}

// H = (I - A)^{-1} b with A = chi_s * Q (iso) or Q * K (aniso).
__device__ __forceinline__ void solve_IminusA(const double* A, const double* b, double* H)
{
    double m[9];
    for (int k = 0; k < 9; k++) m[k] = -A[k];
    m[0] += 1.0; m[4] += 1.0; m[8] += 1.0;
    double inv[9];
    inv3d(m, inv);
    mat3_vec(inv, b, H);
}

// secant susceptibility f(h)/h of the tanh-sum law (radmater.h:644-657), h >= 0
__device__ __forceinline__ double frm_chi(const double* mat, double h)
{
    const int nt = int(mat[1]);
    double chi = 0.0;
    for (int t = 0; t < nt; t++) {
        const double ms = mat[2 + t], ks = mat[5 + t];
        if (ms == 0.0) continue;
        if (h > 0.0) {
            const double u = ks * h / ms;
            chi += (u < 1e-8) ? ks : ms * tanh(u) / h;
        } else {
            chi += ks;
        }
    }
    return chi;
}

// |M|(h) of a tabulated law (node triples x, y, d): cubic Hermite between
// nodes, y[0] below the first, the last slope beyond the last
// (radmater.cpp:619-646)
__device__ __forceinline__ double tab_f(const double* t, int L, double h)
{
    if (h < t[0]) return t[1];
    if (h >= t[3 * (L - 1)]) return t[3 * (L - 1) + 1] + (h - t[3 * (L - 1)]) * t[3 * (L - 1) + 2];
    int lo = 0, hi = L - 1;  // t[3 lo] <= h < t[3 hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t[3 * mid] <= h) lo = mid;
        else hi = mid;
    }
    const double* a = t + 3 * lo;
    const double s = a[3] - a[0], u = h - a[0];
    const double sl = (a[4] - a[1]) / s;
    const double c2 = (3.0 * sl - 2.0 * a[2] - a[5]) / s;
    const double c3 = (-2.0 * sl + a[2] + a[5]) / (s * s);
    return a[1] + u * (a[2] + u * (c2 + u * c3));
}

// secant susceptibility |M|(h)/h of an isotropic nonlinear law (kind 2: tanh
// sum, kind 3: table)
__device__ __forceinline__ double law_chi(const double* mat, const double* tab, double h)
{
    if (int(mat[0]) == 3) {
        const double* t = tab + 3 * int64_t(mat[1]);
        return (h > 0.0) ? tab_f(t, int(mat[2]), h) / h : t[2];
    }
    return frm_chi(mat, h);
}

// |M|(h) and d|M|/dh of an isotropic nonlinear law (kind 2 tanh sum, kind 3
// table), h > 0.
__device__ __forceinline__ void law_f_df(const double* mat, const double* tab, double h, double& f,
                                         double& df)
{
    if (int(mat[0]) == 3) {
        const double* t = tab + 3 * int64_t(mat[1]);
        const int L = int(mat[2]);
        if (h < t[0]) {
            f = t[1];
            df = 0.0;
            return;
        }
        if (h >= t[3 * (L - 1)]) {
            f = t[3 * (L - 1) + 1] + (h - t[3 * (L - 1)]) * t[3 * (L - 1) + 2];
            df = t[3 * (L - 1) + 2];
            return;
        }
        int lo = 0, hi = L - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (t[3 * mid] <= h) lo = mid;
            else hi = mid;
        }
        const double* a = t + 3 * lo;
        const double sg = a[3] - a[0], u = h - a[0];
        const double sl = (a[4] - a[1]) / sg;
        const double c2 = (3.0 * sl - 2.0 * a[2] - a[5]) / sg;
        const double c3 = (-2.0 * sl + a[2] + a[5]) / (sg * sg);
        f = a[1] + u * (a[2] + u * (c2 + u * c3));
        df = a[2] + u * (2.0 * c2 + 3.0 * c3 * u);
        return;
    }
    const int nt = int(mat[1]);
    f = 0.0;
    df = 0.0;
    for (int t = 0; t < nt; t++) {
        const double ms = mat[2 + t], ks = mat[5 + t];
        if (ms == 0.0) continue;
        const double u = ks * h / ms;
        // tanh u and sech^2 u from one expm1: with e = exp(-2|u|) - 1,
        // tanh|u| = -e / (2 + e) and sech^2 u = 4 (1 + e) / (2 + e)^2
        const double e = expm1(-2.0 * fabs(u));
        const double d = 2.0 + e;
        f += ms * copysign(-e / d, u);
        df += ks * (4.0 * (1.0 + e) / (d * d));
    }
}

__device__ void local_solve(const double* mat, const double* q, const double* heff, double* Mp,
                            double* Hl, const double* tab, double warm)
{
    const int kind = int(mat[0]);
    const double* mr = mat + 13;
    double qmr[3];
    mat3_vec(q, mr, qmr);
    const double b[3] = {heff[0] + qmr[0], heff[1] + qmr[1], heff[2] + qmr[2]};
    if (kind == 0) {
        double A[9];
        for (int k = 0; k < 9; k++) A[k] = mat[1] * q[k];
        solve_IminusA(A, b, Hl);
        for (int k = 0; k < 3; k++) Mp[k] = mat[1] * Hl[k] + mr[k];
        return;
    }
    if (kind == 1) {
        const double* K = mat + 4;
        double A[9];
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++)
                A[3 * a + c] = q[3 * a] * K[c] + q[3 * a + 1] * K[3 + c] + q[3 * a + 2] * K[6 + c];
        solve_IminusA(A, b, Hl);
        double kh[3];
        mat3_vec(K, Hl, kh);
        for (int k = 0; k < 3; k++) Mp[k] = kh[k] + mr[k];
        return;
    }
    // nonlinear isotropic: find h = |H(h)|, H(h) = (I - chi(h) Q)^{-1} b with
    // the secant susceptibility chi(h) = |M|(h)/h.  Safeguarded Newton on
    // g(h) = |H(h)| - h inside a bracket g(lo) > 0 > g(hi), started from the
    // previous iteration's |H| (warm) when it lies in the bracket; where a
    // Newton step leaves the bracket, false position, and bisection after two
    // false-position steps in a row.
    const double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (!(bn > 0.0)) {
        Hl[0] = Hl[1] = Hl[2] = 0.0;
        for (int k = 0; k < 3; k++) Mp[k] = mr[k];
        return;
    }
    const double chi0 = law_chi(mat, tab, 0.0);
    // |H(h)| into the return value, H(h) into Hv, dg/dh into *dg (if given)
    auto phi = [&](double h, double* Hv, double* dg) -> double {
        double chi = chi0, dchi = 0.0;
        if (h > 0.0) {
            double f, df;
            law_f_df(mat, tab, h, f, df);
            chi = f / h;
            dchi = (df - chi) / h;
        }
        double A[9];
        for (int k = 0; k < 9; k++) A[k] = -chi * q[k];
        A[0] += 1.0;
        A[4] += 1.0;
        A[8] += 1.0;
        double Ai[9];
        inv3d(A, Ai);
        mat3_vec(Ai, b, Hv);
        const double hv = sqrt(Hv[0] * Hv[0] + Hv[1] * Hv[1] + Hv[2] * Hv[2]);
        if (dg) {
            double qh[3], dH[3];
            mat3_vec(q, Hv, qh);
            mat3_vec(Ai, qh, dH);
            *dg = (hv > 0.0 ? dchi * (Hv[0] * dH[0] + Hv[1] * dH[1] + Hv[2] * dH[2]) / hv : 0.0) - 1.0;
        }
        return hv;
    };
    double Hv[3];
    double lo = 0.0, glo = phi(0.0, Hv, nullptr);
    double hi = bn, ghi = phi(hi, Hv, nullptr) - hi;
    for (int k = 0; k < 60 && ghi > 0.0; k++) {
        lo = hi;
        glo = ghi;
        hi *= 2.0;
        ghi = phi(hi, Hv, nullptr) - hi;
    }
    double h;
    if (!(glo > 0.0) || !(ghi < 0.0)) {
        h = (glo <= 0.0) ? lo : hi;
    } else {
        h = (warm > lo && warm < hi) ? warm : (lo * ghi - hi * glo) / (ghi - glo);
        int fp = 0;
        for (int it = 0; it < 60; it++) {
            double dg;
            const double g = phi(h, Hv, &dg) - h;
            if (g == 0.0) break;
            if (g > 0.0) {
                lo = h;
                glo = g;
            } else {
                hi = h;
                ghi = g;
            }
            double hn = (dg < 0.0) ? h - g / dg : -1.0;
            if (hn > lo && hn < hi) {
                fp = 0;
            } else {
                hn = (fp < 2) ? (lo * ghi - hi * glo) / (ghi - glo) : 0.5 * (lo + hi);
                fp = (fp < 2) ? fp + 1 : 0;
                if (!(hn > lo && hn < hi)) hn = 0.5 * (lo + hi);
            }
            const bool done = fabs(hn - h) <= 1e-15 * bn || hi - lo <= 1e-15 * bn;
            h = hn;
            if (done) break;
        }
    }
    phi(h, Hv, nullptr);
    const double hn = sqrt(Hv[0] * Hv[0] + Hv[1] * Hv[1] + Hv[2] * Hv[2]);
    const double chi = law_chi(mat, tab, hn);
    for (int k = 0; k < 3; k++) {
        Hl[k] = Hv[k];
        Mp[k] = chi * Hv[k] + mr[k];
    }
}

// FP64 defect of the stored-matrix field: out = hext + rowT*H64 - y - Qii M,
// with H64 the FP64 field of all elements at the collocation points (lab
// frame) and y the stored-matrix product (diagonal blocks excluded).
__global__ __launch_bounds__(256) void k_refine_defect(int64_t n, const double* __restrict__ H64,
                                                       const double* __restrict__ rowt,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ qdiag,
                                                       const double* __restrict__ M,
                                                       const double* __restrict__ hext,
                                                       double* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    double h[3] = {H64[3 * i], H64[3 * i + 1], H64[3 * i + 2]};
    if (rowt) {
        const double* t = rowt + 9 * i;
        const double g[3] = {h[0], h[1], h[2]};
        for (int k = 0; k < 3; k++) h[k] = t[3 * k] * g[0] + t[3 * k + 1] * g[1] + t[3 * k + 2] * g[2];
    }
    const double* D = qdiag + 9 * i;
    const double m[3] = {M[3 * i], M[3 * i + 1], M[3 * i + 2]};
    for (int k = 0; k < 3; k++)
        out[3 * i + k] = hext[3 * i + k] + (h[k] - y[3 * i + k]) -
                         (D[3 * k] * m[0] + D[3 * k + 1] * m[1] + D[3 * k + 2] * m[2]);
}

__global__ __launch_bounds__(kRelaxLanes) void k_relax_update(int64_t n, const double* __restrict__ mats,
                                                              const double* __restrict__ tab,
                                                              const double* __restrict__ hext,
                                                              const double* __restrict__ qdiag,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ M,
                                                              double* __restrict__ F,
                                                              double* __restrict__ Hout,
                                                              double* __restrict__ partials)
{
    __shared__ double red[3][kRelaxLanes];
    const int64_t i = int64_t(blockIdx.x) * kRelaxLanes + threadIdx.x;
    double f2 = 0.0, mm = 0.0, hh = 0.0;
    if (i < n) {
        const double heff[3] = {y[3 * i] + hext[3 * i], y[3 * i + 1] + hext[3 * i + 1],
                                y[3 * i + 2] + hext[3 * i + 2]};
        // the previous local field's magnitude starts the nonlinear solve
        const double warm = sqrt(Hout[3 * i] * Hout[3 * i] + Hout[3 * i + 1] * Hout[3 * i + 1] +
                                 Hout[3 * i + 2] * Hout[3 * i + 2]);
        double mp[3], hl[3];
        local_solve(mats + 16 * i, qdiag + 9 * i, heff, mp, hl, tab, warm);
        for (int k = 0; k < 3; k++) {
            const double d = mp[k] - M[3 * i + k];
            F[3 * i + k] = d;
            Hout[3 * i + k] = hl[k];
            f2 += d * d;
        }
        mm = sqrt(mp[0] * mp[0] + mp[1] * mp[1] + mp[2] * mp[2]);
        hh = sqrt(hl[0] * hl[0] + hl[1] * hl[1] + hl[2] * hl[2]);
    }
    red[0][threadIdx.x] = f2;
    red[1][threadIdx.x] = mm;
    red[2][threadIdx.x] = hh;
    __syncthreads();
    for (int st = kRelaxLanes / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            red[0][threadIdx.x] += red[0][threadIdx.x + st];
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + st]);
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = red[0][0];
        partials[3 * blockIdx.x + 1] = red[1][0];
        partials[3 * blockIdx.x + 2] = red[2][0];
    }
}

// ---------------------------------------------------------------- Newton-Krylov
// Material law at the total field of the current magnetization: H = y + Qii M
// + hext, F = M_mat(H) - M, tangent D = dM_mat/dH (D = mp hh^T + sc (I - hh^T)
// for isotropic laws, hh = H/|H|, sc = |M|/|H|, mp = d|M|/d|H|; chi0 I at
// H = 0; chi I or K for linear laws) and the block preconditioner
// P = (I - D Qii)^{-1}.  partials per block: [sum |F|^2, max |M_mat|, max |H|].
__global__ __launch_bounds__(kRelaxLanes) void k_nk_material(int64_t n, const double* __restrict__ mats,
                                                     const double* __restrict__ tab,
                                                     const double* __restrict__ hext,
                                                     const double* __restrict__ qdiag,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ M,
                                                     double* __restrict__ F, double* __restrict__ D,
                                                     double* __restrict__ P, double* __restrict__ Hout,
                                                     double* __restrict__ partials)
{
    __shared__ double red[3][kRelaxLanes];
    const int64_t i = int64_t(blockIdx.x) * kRelaxLanes + threadIdx.x;
    double f2 = 0.0, mm = 0.0, hh = 0.0;
    if (i < n) {
        const double* mat = mats + 16 * i;
        const double* q = qdiag + 9 * i;
        const double m[3] = {M[3 * i], M[3 * i + 1], M[3 * i + 2]};
        double qm[3];
        mat3_vec(q, m, qm);
        double H[3];
        for (int k = 0; k < 3; k++) H[k] = y[3 * i + k] + qm[k] + hext[3 * i + k];
        const int kind = int(mat[0]);
        double d[9], mn[3];
        if (kind == 1) {
            for (int k = 0; k < 9; k++) d[k] = mat[4 + k];
            mat3_vec(d, H, mn);
        } else {
            double sc, mp;
            const double h = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
            if (kind == 0) {
                sc = mp = mat[1];
            } else if (h > 1e-300) {
                double f, df;
                law_f_df(mat, tab, h, f, df);
                sc = f / h;
                mp = df;
            } else {
                sc = mp = law_chi(mat, tab, 0.0);
            }
            const double e[3] = {h > 1e-300 ? H[0] / h : 0.0, h > 1e-300 ? H[1] / h : 0.0,
                                 h > 1e-300 ? H[2] / h : 0.0};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) d[3 * a + b] = (mp - sc) * e[a] * e[b] + (a == b ? sc : 0.0);
            for (int k = 0; k < 3; k++) mn[k] = sc * H[k];
        }
        for (int k = 0; k < 3; k++) mn[k] += mat[13 + k];
        // P = (I - D Qii)^{-1}
        double a[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                a[3 * r + c] = (r == c ? 1.0 : 0.0) -
                               (d[3 * r] * q[c] + d[3 * r + 1] * q[3 + c] + d[3 * r + 2] * q[6 + c]);
        double pinv[9];
        inv3d(a, pinv);
        for (int k = 0; k < 9; k++) {
            D[9 * i + k] = d[k];
            P[9 * i + k] = pinv[k];
        }
        for (int k = 0; k < 3; k++) {
            const double fk = mn[k] - m[k];
            F[3 * i + k] = fk;
            Hout[3 * i + k] = H[k];
            f2 += fk * fk;
        }
        mm = sqrt(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
        hh = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    }
    red[0][threadIdx.x] = f2;
    red[1][threadIdx.x] = mm;
    red[2][threadIdx.x] = hh;
    __syncthreads();
    for (int st = kRelaxLanes / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            red[0][threadIdx.x] += red[0][threadIdx.x + st];
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + st]);
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = red[0][0];
        partials[3 * blockIdx.x + 1] = red[1][0];
        partials[3 * blockIdx.x + 2] = red[2][0];
    }
}

// z = P v per element (3x3 blocks).
__global__ __launch_bounds__(256) void k_nk_prec(int64_t n, const double* __restrict__ P,
                                                 const double* __restrict__ v, double* __restrict__ z)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double x[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    double o[3];
    mat3_vec(P + 9 * i, x, o);
    for (int k = 0; k < 3; k++) z[3 * i + k] = o[k];
}

// w = z - D (y + Qii z): the Newton operator (I - D Q) applied to z, with
// y = offdiag(Q) z from the matrix pass.
__global__ __launch_bounds__(256) void k_nk_apply(int64_t n, const double* __restrict__ D,
                                                  const double* __restrict__ qdiag,
                                                  const double* __restrict__ z,
                                                  const double* __restrict__ y, double* __restrict__ w)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double zz[3] = {z[3 * i], z[3 * i + 1], z[3 * i + 2]};
    double qz[3];
    mat3_vec(qdiag + 9 * i, zz, qz);
    const double g[3] = {y[3 * i] + qz[0], y[3 * i + 1] + qz[1], y[3 * i + 2] + qz[2]};
    double dg[3];
    mat3_vec(D + 9 * i, g, dg);
    for (int k = 0; k < 3; k++) w[3 * i + k] = zz[k] - dg[k];
}

// Dot products of w with nv basis vectors (V, leading dimension ldv) and with
// itself: grid (chunks, nv + 1), partials[j * chunks + chunk].
__global__ __launch_bounds__(256) void k_mdot(int nv, const double* __restrict__ V, int64_t ldv,
                                              const double* __restrict__ w, int64_t len,
                                              double* __restrict__ partials)
{
    __shared__ double red[4];
    const int j = blockIdx.y;
    const double* a = (j < nv) ? V + int64_t(j) * ldv : w;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += int64_t(gridDim.x) * 256)
        acc += a[i] * w[i];
    const double v = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[int64_t(j) * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// out = cb * base + sum_j c_j V_j (base may be null; out may alias base).
__global__ __launch_bounds__(256) void k_vcomb(int nv, const double* __restrict__ V, int64_t ldv,
                                               Coef64 c, const double* base, double cb, int64_t len,
                                               double* out)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < len; i += int64_t(gridDim.x) * 256) {
        double acc = base ? cb * base[i] : 0.0;
        for (int j = 0; j < nv; j++) acc += c.c[j] * V[int64_t(j) * ldv + i];
        out[i] = acc;
    }
}

// ---------------------------------------------------------------- vector helpers
struct DotPack {
    const double* a[kDotMax];
    const double* b[kDotMax];
};

// All nd dot products in one pass: every lane keeps nd running sums, waves
// reduce by shuffles, the workgroup writes one partial per dot.
__global__ __launch_bounds__(256) void k_dots(int nd, DotPack pk, int64_t len,
                                              double* __restrict__ partials)
{
    __shared__ double red[4][kDotMax];
    double acc[kDotMax];
#pragma unroll
    for (int d = 0; d < kDotMax; d++) acc[d] = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < len;
         i += int64_t(gridDim.x) * 256) {
#pragma unroll
        for (int d = 0; d < kDotMax; d++)
            if (d < nd) acc[d] += pk.a[d][i] * pk.b[d][i];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < kDotMax; d++) {
        if (d < nd) {
            const double v = wave_sum(acc[d]);
            if (lane == 0) red[wave][d] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < nd)
        partials[threadIdx.x * gridDim.x + blockIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

struct CombPack {
    const double* v[kCombMax];
    double c[kCombMax];
};

__global__ __launch_bounds__(256) void k_lincomb(int nv, CombPack pk, int64_t len,
                                                 double* __restrict__ y)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < len;
         i += int64_t(gridDim.x) * 256) {
        double acc = 0.0;
        for (int k = 0; k < nv; k++) acc += pk.c[k] * pk.v[k][i];
        y[i] = acc;
    }
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// One relaxation step with its Anderson bookkeeping (one wave per block):
// lane i updates element i; dF[k] (k < nh, oldest first) are the stored
// differences, the newest of which (dF[nh - 1]) is the one written here when
// dFw is set.
static_assert(kRelaxLanes == 64, "k_relax_step reduces over one wave per block");
__device__ __forceinline__ void relax_step_block(int64_t n, const double* __restrict__ mats,
                                                 const double* __restrict__ tab,
                                                 const double* __restrict__ hext,
                                                 const double* __restrict__ qdiag,
                                                 const double* __restrict__ y, const double* __restrict__ M,
                                                 double* __restrict__ F, double* __restrict__ Hout,
                                                 double* __restrict__ partials, double* Fp, double* xp,
                                                 double* dFw, double* dXw, const double* const* dF, int nh)
{
    const int64_t i = int64_t(blockIdx.x) * kRelaxLanes + threadIdx.x;
    double f2 = 0.0, mm = 0.0, hh = 0.0;
    double dd[2 * kHistMax];
#pragma unroll
    for (int k = 0; k < 2 * kHistMax; k++) dd[k] = 0.0;
    if (i < n) {
        const double heff[3] = {y[3 * i] + hext[3 * i], y[3 * i + 1] + hext[3 * i + 1],
                                y[3 * i + 2] + hext[3 * i + 2]};
        const double warm = sqrt(Hout[3 * i] * Hout[3 * i] + Hout[3 * i + 1] * Hout[3 * i + 1] +
                                 Hout[3 * i + 2] * Hout[3 * i + 2]);
        double mp[3], hl[3];
        local_solve(mats + 16 * i, qdiag + 9 * i, heff, mp, hl, tab, warm);
        double fi[3], dfn[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 3; k++) {
            const int64_t j = 3 * i + k;
            const double xk = M[j];
            const double d = mp[k] - xk;
            fi[k] = d;
            F[j] = d;
            Hout[j] = hl[k];
            f2 += d * d;
            if (dFw) {
                dfn[k] = d - Fp[j];
                dFw[j] = dfn[k];
                dXw[j] = xk - xp[j];
            }
            Fp[j] = d;
            xp[j] = xk;
        }
        mm = sqrt(mp[0] * mp[0] + mp[1] * mp[1] + mp[2] * mp[2]);
        hh = sqrt(hl[0] * hl[0] + hl[1] * hl[1] + hl[2] * hl[2]);
#pragma unroll
        for (int h = 0; h < kHistMax; h++) {
            if (h < nh) {
                double a = 0.0, b = 0.0;
                for (int k = 0; k < 3; k++) {
                    const double v = (h == nh - 1) ? dfn[k] : dF[h][3 * i + k];
                    a += dfn[k] * v;
                    b += v * fi[k];
                }
                dd[h] = a;
                dd[kHistMax + h] = b;
            }
        }
    }
    f2 = wave_sum(f2);
    mm = wave_max(mm);
    hh = wave_max(hh);
#pragma unroll
    for (int k = 0; k < 2 * kHistMax; k++)
        if ((k % kHistMax) < nh) dd[k] = wave_sum(dd[k]);
    if (threadIdx.x == 0) {
        double* o = partials + int64_t(kRelaxPart) * blockIdx.x;
        o[0] = f2;
        o[1] = mm;
        o[2] = hh;
#pragma unroll
        for (int k = 0; k < 2 * kHistMax; k++) o[3 + k] = dd[k];
    }
}

__global__ __launch_bounds__(kRelaxLanes) void k_relax_step(int64_t n, const double* __restrict__ mats,
                                                            const double* __restrict__ tab,
                                                            const double* __restrict__ hext,
                                                            const double* __restrict__ qdiag,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ M,
                                                            double* __restrict__ F,
                                                            double* __restrict__ Hout,
                                                            double* __restrict__ partials, HistPack hp)
{
    relax_step_block(n, mats, tab, hext, qdiag, y, M, F, Hout, partials, hp.Fp, hp.xp, hp.dFw, hp.dXw, hp.dF,
                     hp.nh);
}

// The Anderson step of the device-driven loop on one wave: the step partials
// reduced across the wave in a fixed order, then the residual and convergence
// test, the restart safeguard, the Gram matrix of the differences in use
// (kept oldest first, the newest row from this step), the regularized
// Cholesky solve of the least-squares problem and the update coefficients,
// and the history extension for the next iteration.  The Gram matrix lives in
// LDS; the factorization runs column by column with the lanes updating the
// trailing entries (the same operations in the same order as a row-by-row
// factorization), and the two triangular solves column by column.
__device__ void anderson_wave(int64_t nb, const double* __restrict__ partials, RelaxDevState* __restrict__ st)
{
    constexpr int H = kHistMax, HP = kHistMax + 1;  // LDS row stride
    __shared__ double sA[H * HP], sL[H * HP], sz[H], sv[kRelaxPart];
    const int lane = threadIdx.x & 63;
    double v[kRelaxPart];
#pragma unroll
    for (int k = 0; k < kRelaxPart; k++) v[k] = 0.0;
    for (int64_t b = lane; b < nb; b += 64) {
        const double* o = partials + int64_t(kRelaxPart) * b;
        v[0] += o[0];
        v[1] = fmax(v[1], o[1]);
        v[2] = fmax(v[2], o[2]);
#pragma unroll
        for (int k = 3; k < kRelaxPart; k++) v[k] += o[k];
    }
    v[0] = wave_sum(v[0]);
    v[1] = wave_max(v[1]);
    v[2] = wave_max(v[2]);
#pragma unroll
    for (int k = 3; k < kRelaxPart; k++) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kRelaxPart; k++) sv[k] = v[k];
    }
    __syncthreads();
    // wave-uniform decisions (every lane reads the same state)
    const int hist = st->hist;
    const int head = st->head;
    const int it = st->it + 1;
    const double misfit = sqrt(v[0] / st->n);
    const bool finite = isfinite(misfit);
    const bool finish = !finite || misfit <= st->prec || it >= st->max_iter;
    const bool restart = !finish && misfit > 1e4 * st->best;
    const double best = fmin(st->best, misfit);
    const double beta = restart ? fmax(0.05, 0.5 * st->beta) : st->beta;
    bool ok = !finish && !restart && hist > 0;
    if (ok) {
        // Gram matrix: live part from the state with the newest row and
        // column from this step; its regularized copy to factor in sL
        for (int e = lane; e < H * H; e += 64) {
            const int a = e / H, c = e - H * (e / H);
            double g = 0.0;
            if (a < hist && c < hist) {
                if (a == hist - 1) g = sv[3 + c];
                else if (c == hist - 1) g = sv[3 + a];
                else g = st->G[a * H + c];
            }
            sA[a * HP + c] = g;
        }
        __syncthreads();
        double tr = 0.0;
        for (int a = 0; a < hist; a++) tr += sA[a * HP + a];
        const double lam = 1e-10 * (tr > 0 ? tr / hist : 1.0);
        for (int e = lane; e < H * H; e += 64) {
            const int a = e / H, c = e - H * (e / H);
            sL[a * HP + c] = sA[a * HP + c] + ((a == c && a < hist) ? lam : 0.0);
        }
        __syncthreads();
        // Cholesky, column k at a time: pivot, the column below it, then the
        // trailing entries (i, j), k < j <= i, lose L_ik L_jk
        for (int k = 0; k < hist && ok; k++) {
            const double dkk = sL[k * HP + k];
            if (!(dkk > 0.0)) {
                ok = false;
                break;
            }
            const double lkk = sqrt(dkk);
            __syncthreads();
            if (lane == k) sL[k * HP + k] = lkk;
            if (lane > k && lane < hist) sL[lane * HP + k] = sL[lane * HP + k] / lkk;
            __syncthreads();
            for (int e = lane; e < H * H; e += 64) {
                const int i = e / H, j = e - H * (e / H);
                if (j > k && i >= j && i < hist) sL[i * HP + j] -= sL[i * HP + k] * sL[j * HP + k];
            }
            __syncthreads();
        }
    }
    if (ok) {
        // L z = r, then L^T gam = z, column by column (gam overwrites z)
        if (lane < hist) sz[lane] = sv[3 + kHistMax + lane];
        __syncthreads();
        for (int k = 0; k < hist; k++) {
            const double zk = sz[k] / sL[k * HP + k];
            __syncthreads();
            if (lane == k) sz[k] = zk;
            if (lane > k && lane < hist) sz[lane] -= sL[lane * HP + k] * zk;
            __syncthreads();
        }
        for (int k = hist - 1; k >= 0; k--) {
            const double gk = sz[k] / sL[k * HP + k];
            __syncthreads();
            if (lane == k) sz[k] = gk;
            if (lane < k) sz[lane] -= sL[k * HP + lane] * gk;
            __syncthreads();
        }
    }
    // state of the next iteration
    if (lane == 0) {
        st->arrive = 0;  // every block of this step has arrived
        st->it = it;
        st->misfit = misfit;
        st->maxM = fmax(1e-17, v[1]);
        st->maxH = fmax(1e-17, v[2]);
        if (!finite) {
            st->fail = 1;
            st->done = 2;
        } else if (misfit <= st->prec && it == 1 && st->keep_if_converged) {
            st->done = 2;
        } else if (finish) {
            st->done = 1;
        } else {
            st->best = best;
            st->beta = beta;
            const int nh = ok ? hist : 0, hd = ok ? head : 0;
            st->cF = beta;
            for (int c = 0; c < H; c++) {
                st->cdX[c] = 0.0;
                st->cdF[c] = 0.0;
            }
            if (ok) {
                const int r0 = (head - hist + 2 * H) % H;  // buffer slot of the oldest difference
                for (int k = 0; k < hist; k++) {
                    const int p = (r0 + k) % H;
                    st->cdX[p] = -sz[k];
                    st->cdF[p] = -beta * sz[k];
                }
            }
            // the next iteration writes a new difference into slot hd
            st->ext = 1;
            st->head = (hd + 1) % H;
            st->hist = nh + 1 < H ? nh + 1 : H;
        }
    }
    // the Gram matrix of the next iteration: the oldest difference drops out
    // when the history is full
    if (ok) {
        const int sh = (hist == H) ? 1 : 0;
        for (int e = lane; e < H * H; e += 64) {
            const int a = e / H, c = e - H * (e / H);
            const int a2 = a + sh, c2 = c + sh;
            st->G[e] = (a2 < H && c2 < H) ? sA[a2 * HP + c2] : 0.0;
        }
    }
}

// The same step in the device-driven loop (history slots from the state); the
// last block to finish runs the Anderson step on the partials of all blocks.
__global__ __launch_bounds__(kRelaxLanes) void k_relax_step_dev(int64_t n, const double* __restrict__ mats,
                                                                const double* __restrict__ tab,
                                                                const double* __restrict__ hext,
                                                                const double* __restrict__ qdiag,
                                                                const double* __restrict__ y,
                                                                const double* __restrict__ M,
                                                                double* __restrict__ F,
                                                                double* __restrict__ Hout,
                                                                double* __restrict__ partials, HistBufs hb,
                                                                RelaxDevState* __restrict__ st)
{
    if (st->done) return;
    const int nh = st->hist, head = st->head;
    const double* dF[kHistMax];
#pragma unroll
    for (int k = 0; k < kHistMax; k++) dF[k] = hb.dF[(head - nh + k + 2 * kHistMax) % kHistMax];
    const int w = (head - 1 + kHistMax) % kHistMax;
    relax_step_block(n, mats, tab, hext, qdiag, y, M, F, Hout, partials, hb.Fp, hb.xp,
                     st->ext ? hb.dF[w] : nullptr, st->ext ? hb.dX[w] : nullptr, dF, nh);
    // release this block's partials, then count it in
    int arrived = 0;
    if (threadIdx.x == 0) {
        __threadfence();
        arrived = atomicAdd(&st->arrive, 1);
    }
    arrived = __shfl(arrived, 0, 64);
    if (arrived != int(gridDim.x) - 1) return;
    __threadfence();  // the other blocks' partials are visible to the last one
    anderson_wave(int64_t(gridDim.x), partials, st);
}

// The update of the device-driven loop (grid-stride over the 3n components).
__global__ __launch_bounds__(256) void k_update_dev(int64_t n3, double* __restrict__ x,
                                                    const double* __restrict__ F, HistBufs hb,
                                                    double* __restrict__ M_out,
                                                    const RelaxDevState* __restrict__ st)
{
    const int done = st->done;
    if (done == 2) return;
    if (done == 1) {
        for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n3; i += int64_t(gridDim.x) * 256)
            M_out[i] = x[i] + F[i];
        return;
    }
    const double cF = st->cF;
    double cX[kHistMax], cD[kHistMax];
#pragma unroll
    for (int b = 0; b < kHistMax; b++) {
        cX[b] = st->cdX[b];
        cD[b] = st->cdF[b];
    }
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n3; i += int64_t(gridDim.x) * 256) {
        double acc = x[i] + cF * F[i];
#pragma unroll
        for (int b = 0; b < kHistMax; b++)
            if (cX[b] != 0.0 || cD[b] != 0.0) acc += cX[b] * hb.dX[b][i] + cD[b] * hb.dF[b][i];
        x[i] = acc;
    }
}

// ---------------------------------------------------------------- currents
__global__ __launch_bounds__(256) void k_current(int64_t ns, const int32_t* __restrict__ kind,
                                                 const double* __restrict__ par,
                                                 const double* __restrict__ cp,
                                                 const double* __restrict__ pts, int64_t np,
                                                 double* __restrict__ B)
{
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= np) return;
    const double x[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    double b[3] = {0, 0, 0};
    for (int64_t s = 0; s < ns; s++) current_add(kind[s], par + 12 * s, cp + 24 * s, x, b);
    for (int k = 0; k < 3; k++) B[3 * p + k] += b[k];
}

}  // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_assemble(const DevModel& m, void* Q, int elem_bytes, int64_t ld, double* qdiag,
                           int64_t r0, int64_t r1, hipStream_t s, bool far32, int planes, int32_t* flag)
{
    if (m.n == 0 || r1 <= r0) return hipSuccess;
    if (planes != 0 && (m.sym || !m.rowt_identity || (planes != 5 && planes != 6) || (planes == 5 && !flag)))
        return hipErrorInvalidValue;
    if (elem_bytes == 4) {
        float* q = static_cast<float*>(Q);
        if (planes == 5) return asm_dispatch<false, false, float, 5>(m, q, ld, qdiag, r0, r1, s, far32, flag);
        if (planes == 6) return asm_dispatch<false, false, float, 6>(m, q, ld, qdiag, r0, r1, s, far32, flag);
        if (!m.sym && m.rowt_identity) return asm_dispatch<false, false>(m, q, ld, qdiag, r0, r1, s, far32, flag);
        if (m.rowt_identity) return asm_dispatch<true, false>(m, q, ld, qdiag, r0, r1, s, far32, flag);
        return asm_dispatch<true, true>(m, q, ld, qdiag, r0, r1, s, far32, flag);
    }
    double* q = static_cast<double*>(Q);
    if (planes == 5) return asm_dispatch<false, false, double, 5>(m, q, ld, qdiag, r0, r1, s, false, flag);
    if (planes == 6) return asm_dispatch<false, false, double, 6>(m, q, ld, qdiag, r0, r1, s, false, flag);
    if (!m.sym && m.rowt_identity) return asm_dispatch<false, false>(m, q, ld, qdiag, r0, r1, s, false, flag);
    if (m.rowt_identity) return asm_dispatch<true, false>(m, q, ld, qdiag, r0, r1, s, false, flag);
    return asm_dispatch<true, true>(m, q, ld, qdiag, r0, r1, s, false, flag);
}

namespace {
// element splits for one kind slot: enough workgroups to fill the chip
int64_t slot_splits(int64_t np, int64_t cnt)
{
    if (cnt == 0) return 0;
    const int64_t blocks = (np + 255) / 256;
    if (blocks >= 2048 || cnt < 64) return 1;
    int64_t sp = (2048 + blocks - 1) / blocks;
    sp = std::min<int64_t>(sp, std::max<int64_t>(1, cnt / 16));
    return std::max<int64_t>(1, sp);
}

template <class R, bool SYM, int KIND>
void field_slot(const DevModel& m, const double* M, const double* pts, int64_t np, int64_t splits,
                int64_t slice0, double* B, double* H, double* part, hipStream_t s)
{
    const int64_t cnt = m.koff[KIND + 1] - m.koff[KIND];
    int64_t js = (cnt + splits - 1) / splits;
    const dim3 grid{unsigned((np + 255) / 256), unsigned(splits), 1u};
    constexpr bool f32 = sizeof(R) == 4;
    if (KIND == 1 && m.charged) {
        js = (js + kTetTile - 1) / kTetTile * kTetTile;
        if constexpr (f32) {
            hipLaunchKernelGGL((k_field_tet<float, SYM, kPassFar>), grid, dim3(256), 0, s, m, pts, np, js,
                               slice0, B, H, part);
            hipLaunchKernelGGL((k_field_tet<double, SYM, kPassNear>), grid, dim3(256), 0, s, m, pts, np, js,
                               slice0, B, H, part);
        } else {
            hipLaunchKernelGGL((k_field_tet<double, SYM, kPassAll>), grid, dim3(256), 0, s, m, pts, np, js,
                               slice0, B, H, part);
        }
        return;
    }
    if constexpr (f32) {
        hipLaunchKernelGGL((k_field<float, SYM, KIND, kPassFar>), grid, dim3(256), 0, s, m, M, pts, np, js,
                           slice0, B, H, part);
        hipLaunchKernelGGL((k_field<double, SYM, KIND, kPassNear>), grid, dim3(256), 0, s, m, M, pts, np, js,
                           slice0, B, H, part);
    } else {
        hipLaunchKernelGGL((k_field<double, SYM, KIND, kPassAll>), grid, dim3(256), 0, s, m, M, pts, np, js,
                           slice0, B, H, part);
    }
}

bool mesh_on(const DevModel& m) { return m.use_mesh && m.mesh.active; }

// slices per kind slot (0 blocks, 1 tetrahedra, 2 general polyhedra); in the
// mesh form slot 1 carries the whole mesh and slot 2 is empty
void slot_plan(const DevModel& m, int64_t np, int64_t sp[3], int64_t work_cap)
{
    sp[0] = slot_splits(np, m.koff[1] - m.koff[0]);
    if (mesh_on(m)) {
        sp[1] = slot_splits(np, m.mesh.ne + m.mesh.nf);
        sp[2] = 0;
    } else {
        sp[1] = slot_splits(np, m.koff[2] - m.koff[1]);
        sp[2] = slot_splits(np, m.koff[3] - m.koff[2]);
    }
    if (work_cap < 0) return;
    // halve the largest split count until the partials fit or every slot is
    // down to one split
    for (;;) {
        const int64_t total = sp[0] + sp[1] + sp[2];
        if (total <= 1 || 6 * np * total <= work_cap) return;
        int big = 0;
        for (int k = 1; k < 3; k++)
            if (sp[k] > sp[big]) big = k;
        if (sp[big] <= 1) return;
        sp[big] = (sp[big] + 1) / 2;
    }
}

template <class R, bool SYM>
hipError_t field_all(const DevModel& m, const double* M, const double* pts, int64_t np, double* B,
                     double* H, double* work, int64_t work_cap, hipStream_t s)
{
    int64_t sp[3];
    slot_plan(m, np, sp, work_cap);
    if (work_cap >= 0 && sp[0] + sp[1] + sp[2] > 1 && 6 * np * (sp[0] + sp[1] + sp[2]) > work_cap)
        return hipErrorInvalidValue;
    const int64_t total = sp[0] + sp[1] + sp[2];
    double* part = (total > 1) ? work : nullptr;
    int64_t slice0 = 0;
    if (sp[0]) field_slot<R, SYM, 0>(m, M, pts, np, sp[0], slice0, B, H, part, s);
    slice0 += sp[0];
    if (sp[1] && mesh_on(m)) {
        const dim3 grid{unsigned((np + 255) / 256), unsigned(sp[1]), 1u};
        hipLaunchKernelGGL((k_field_mesh<R>), grid, dim3(256), 0, s, m, pts, np, sp[1], slice0, B, H,
                           part);
    } else if (sp[1]) {
        field_slot<R, SYM, 1>(m, M, pts, np, sp[1], slice0, B, H, part, s);
    }
    slice0 += sp[1];
    if (sp[2]) field_slot<R, SYM, 2>(m, M, pts, np, sp[2], slice0, B, H, part, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || total <= 1) return e;
    hipLaunchKernelGGL(k_field_reduce, dim3(unsigned((np + 255) / 256)), dim3(256), 0, s, np, total,
                       part, B, H);
    return hipGetLastError();
}
}  // namespace

int64_t field_slices(const DevModel& m, int64_t np, int64_t work_cap)
{
    int64_t sp[3];
    slot_plan(m, np, sp, work_cap);
    return std::max<int64_t>(1, sp[0] + sp[1] + sp[2]);
}

hipError_t launch_field(const DevModel& m0, const double* M, const double* pts, int64_t np,
                        double* B, double* H, double* work, double* charged, bool fp32,
                        hipStream_t s, int64_t work_cap)
{
    if (np == 0) return hipSuccess;
    if (m0.n == 0) {
        if (B) (void)hipMemsetAsync(B, 0, size_t(3 * np) * sizeof(double), s);
        if (H) (void)hipMemsetAsync(H, 0, size_t(3 * np) * sizeof(double), s);
        return hipGetLastError();
    }
    DevModel m = m0;
    const int64_t ntet = m.koff[2] - m.koff[1];
    if (mesh_on(m)) {
        const MeshDev& md = m.mesh;
        if (md.ne)
            hipLaunchKernelGGL(k_mesh_charge_edges, dim3(unsigned((md.ne + 255) / 256)), dim3(256), 0, s,
                               md, M, fp32);
        if (md.nf)
            hipLaunchKernelGGL(k_mesh_charge_faces, dim3(unsigned((md.nf + 255) / 256)), dim3(256), 0, s,
                               md, M, fp32);
        if (md.ni)
            hipLaunchKernelGGL(k_mesh_charge_in, dim3(unsigned((md.ni + 255) / 256)), dim3(256), 0, s, md,
                               M);
    } else if (ntet > 0 && charged) {
        hipLaunchKernelGGL(k_charge_tets, dim3(unsigned((ntet + 255) / 256)), dim3(256), 0, s, m, M,
                           charged);
        m.charged = charged;
    }
    if (fp32) return m.sym ? field_all<float, true>(m, M, pts, np, B, H, work, work_cap, s)
                           : field_all<float, false>(m, M, pts, np, B, H, work, work_cap, s);
    return m.sym ? field_all<double, true>(m, M, pts, np, B, H, work, work_cap, s)
                 : field_all<double, false>(m, M, pts, np, B, H, work, work_cap, s);
}

hipError_t launch_matvec(const void* Q, int elem_bytes, int64_t nrows, int64_t ld, const double* x,
                         const double* add, double* y, hipStream_t s, const int32_t* skip)
{
    if (nrows <= 0) return hipSuccess;
// rows per wave of the matrix pass: on the C-yoke (24,000 x 24,064) 4 rows
// measured 0.668 ms FP64 / 0.358 ms FP32, 2 rows 0.695 / 0.394, 8 rows
// 0.870 / 0.487; unrolling the column loop (RH_MV_UNROLL) gained nothing
#ifndef RH_MV_RW
#define RH_MV_RW 4
#endif
    constexpr int RW = RH_MV_RW;
    const int64_t waves = (nrows + RW - 1) / RW;
    const int64_t blocks = (waves + 3) / 4;
    if (elem_bytes == 4)
        hipLaunchKernelGGL((k_matvec_f32<RW>), dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const float*>(Q), nrows, ld, x, add, y, skip);
    else
        hipLaunchKernelGGL((k_matvec_f64<RW>), dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const double*>(Q), nrows, ld, x, add, y, skip);
    return hipGetLastError();
}

hipError_t launch_matvec_sym(const void* Q, int elem_bytes, int planes, int64_t nel, int64_t ldp,
                             const double* x, const double* add, double* y, hipStream_t s,
                             const int32_t* skip)
{
    if (nel <= 0) return hipSuccess;
    if ((planes != 5 && planes != 6) || ldp % 256 != 0) return hipErrorInvalidValue;
    constexpr int ER = RH_MVS_ER;
    const int64_t waves = (nel + ER - 1) / ER;
    const dim3 grid{unsigned((waves + 3) / 4), 1u, 1u};
    if (elem_bytes == 4) {
        const float* q = static_cast<const float*>(Q);
        if (planes == 5)
            hipLaunchKernelGGL((k_matvec_sym_f32<5, ER>), grid, dim3(256), 0, s, q, nel, ldp, x, add, y, skip);
        else
            hipLaunchKernelGGL((k_matvec_sym_f32<6, ER>), grid, dim3(256), 0, s, q, nel, ldp, x, add, y, skip);
    } else {
        const double* q = static_cast<const double*>(Q);
        if (planes == 5)
            hipLaunchKernelGGL((k_matvec_sym_f64<5, ER>), grid, dim3(256), 0, s, q, nel, ldp, x, add, y, skip);
        else
            hipLaunchKernelGGL((k_matvec_sym_f64<6, ER>), grid, dim3(256), 0, s, q, nel, ldp, x, add, y, skip);
    }
    return hipGetLastError();
}

int64_t relax_blocks(int64_t n) { return (n + kRelaxLanes - 1) / kRelaxLanes; }

hipError_t launch_relax_update(int64_t n, const double* mats, const double* tab, const double* hext,
                               const double* qdiag, const double* y, const double* M, double* F,
                               double* Hout, double* partials, hipStream_t s)
{
    hipLaunchKernelGGL(k_relax_update, dim3(unsigned(relax_blocks(n))), dim3(kRelaxLanes), 0, s, n, mats,
                       tab, hext, qdiag, y, M, F, Hout, partials);
    return hipGetLastError();
}

hipError_t launch_relax_step(int64_t n, const double* mats, const double* tab, const double* hext,
                             const double* qdiag, const double* y, const double* M, double* F,
                             double* Hout, double* partials, const HistPack& hp, hipStream_t s)
{
    if (hp.nh < 0 || hp.nh > kHistMax || (hp.nh > 0 && !hp.dFw)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_relax_step, dim3(unsigned(relax_blocks(n))), dim3(kRelaxLanes), 0, s, n, mats, tab,
                       hext, qdiag, y, M, F, Hout, partials, hp);
    return hipGetLastError();
}

hipError_t launch_relax_step_dev(int64_t n, const double* mats, const double* tab, const double* hext,
                                 const double* qdiag, const double* y, const double* M, double* F,
                                 double* Hout, double* partials, const HistBufs& hb, RelaxDevState* st,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_relax_step_dev, dim3(unsigned(relax_blocks(n))), dim3(kRelaxLanes), 0, s, n, mats, tab,
                       hext, qdiag, y, M, F, Hout, partials, hb, st);
    return hipGetLastError();
}

hipError_t launch_update_dev(int64_t n3, double* x, const double* F, const HistBufs& hb, double* M_out,
                             const RelaxDevState* st, hipStream_t s)
{
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(4096, (n3 + 255) / 256));
    hipLaunchKernelGGL(k_update_dev, dim3(unsigned(blocks)), dim3(256), 0, s, n3, x, F, hb, M_out, st);
    return hipGetLastError();
}

hipError_t launch_refine_defect(int64_t n, const double* H64, const double* rowt, const double* y,
                                const double* qdiag, const double* M, const double* hext,
                                double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_refine_defect, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, H64, rowt,
                       y, qdiag, M, hext, out);
    return hipGetLastError();
}

hipError_t launch_nk_material(int64_t n, const double* mats, const double* tab, const double* hext,
                              const double* qdiag, const double* y, const double* M, double* F, double* D,
                              double* P, double* Hout, double* partials, hipStream_t s)
{
    hipLaunchKernelGGL(k_nk_material, dim3(unsigned(relax_blocks(n))), dim3(kRelaxLanes), 0, s, n, mats, tab, hext,
                       qdiag, y, M, F, D, P, Hout, partials);
    return hipGetLastError();
}

hipError_t launch_nk_prec(int64_t n, const double* P, const double* v, double* z, hipStream_t s)
{
    hipLaunchKernelGGL(k_nk_prec, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, P, v, z);
    return hipGetLastError();
}

hipError_t launch_nk_apply(int64_t n, const double* D, const double* qdiag, const double* z, const double* y,
                           double* w, hipStream_t s)
{
    hipLaunchKernelGGL(k_nk_apply, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, D, qdiag, z, y, w);
    return hipGetLastError();
}

hipError_t launch_mdot(int nv, const double* V, int64_t ldv, const double* w, int64_t len, double* partials,
                       hipStream_t s)
{
    if (nv < 0 || nv > kCoefMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mdot, dim3(unsigned(dot_blocks(len)), unsigned(nv + 1)), dim3(256), 0, s, nv, V, ldv,
                       w, len, partials);
    return hipGetLastError();
}

hipError_t launch_vcomb(int nv, const double* V, int64_t ldv, const double* c, const double* base, double cb,
                        int64_t len, double* out, hipStream_t s)
{
    if (nv < 0 || nv > kCoefMax) return hipErrorInvalidValue;
    Coef64 cc{};
    for (int j = 0; j < nv; j++) cc.c[j] = c[j];
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(4096, (len + 255) / 256));
    hipLaunchKernelGGL(k_vcomb, dim3(unsigned(blocks)), dim3(256), 0, s, nv, V, ldv, cc, base, cb, len, out);
    return hipGetLastError();
}

int64_t dot_blocks(int64_t len) { return std::max<int64_t>(1, std::min<int64_t>(128, (len + 511) / 512)); }

hipError_t launch_dots(int nd, const double* const* a, const double* const* b, int64_t len,
                       double* partials, hipStream_t s)
{
    if (nd < 1 || nd > kDotMax) return hipErrorInvalidValue;
    DotPack pk;
    for (int d = 0; d < nd; d++) {
        pk.a[d] = a[d];
        pk.b[d] = b[d];
    }
    hipLaunchKernelGGL(k_dots, dim3(unsigned(dot_blocks(len))), dim3(256), 0, s, nd, pk, len, partials);
    return hipGetLastError();
}

hipError_t launch_lincomb(int nv, const double* const* v, const double* c, int64_t len, double* y,
                          hipStream_t s)
{
    if (nv < 1 || nv > kCombMax) return hipErrorInvalidValue;
    CombPack pk;
    for (int k = 0; k < nv; k++) {
        pk.v[k] = v[k];
        pk.c[k] = c[k];
    }
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(4096, (len + 255) / 256));
    hipLaunchKernelGGL(k_lincomb, dim3(unsigned(blocks)), dim3(256), 0, s, nv, pk, len, y);
    return hipGetLastError();
}

hipError_t launch_current(int64_t ns, const int32_t* kind, const double* par, const double* cp,
                          const double* pts, int64_t np, double* B, hipStream_t s)
{
    if (np == 0 || ns == 0) return hipSuccess;
    hipLaunchKernelGGL(k_current, dim3(unsigned((np + 255) / 256)), dim3(256), 0, s, ns, kind, par,
                       cp, pts, np, B);
    return hipGetLastError();
}

}  // namespace rh
