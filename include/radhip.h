/*
 * radhip.h -- C ABI of radiacuda_amd, an MI355X (gfx950) implementation of
 * Radia's three GPU-bound paths: interaction-matrix assembly, the relaxation
 * solve, and field evaluation of magnetized blocks and polyhedra.
 *
 * Plain C types only (pointers, sizes, enums); no HIP or torch types cross
 * this boundary.  Streams are passed as opaque `void*` (a hipStream_t, or NULL
 * for the null stream).  Every function returns an rh_status; on failure
 * rh_last_error() returns a thread-local message.
 *
 * Each entry point names the reference interface it replaces
 * (DanielWinklehner/RadiaCUDA, paths relative to the repository root).
 *
 * Units follow Radia: lengths in mm, B, H (as mu0*H) and M in tesla, current
 * density in A/mm^2.
 */
#ifndef RADHIP_H
#define RADHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RH_OK = 0,
    RH_ERR_ARG = 1,         /* invalid argument (shape, index, null pointer)     */
    RH_ERR_HIP = 2,         /* a HIP runtime call failed                          */
    RH_ERR_NOMEM = 3,       /* device allocation failed                           */
    RH_ERR_UNSUPPORTED = 4, /* element / material / option not supported          */
    RH_ERR_GEOMETRY = 5,    /* degenerate or non-convex polyhedron                */
    RH_ERR_NOTCONV = 6      /* relaxation did not reach the requested precision   */
} rh_status;

typedef enum {
    RH_FP64 = 0,  /* FP64 arithmetic, FP64 matrix storage                          */
    RH_MIXED = 1  /* FP32 matrix storage (Radia's own storage) with FP64 self
                   * blocks; far blocks integrated in FP32, near ones in FP64;
                   * field integrals in FP32 with FP64 accumulation            */
} rh_precision;

/* ------------------------------------------------------------------------ */
/* Model: magnetizable elements uploaded once, reused by assembly and field. */
/* ------------------------------------------------------------------------ */

/* Element geometry in each element's own frame plus its transform copies.
 * Copy 0 of element e is its main transform (defines the collocation point and
 * the row transform); further copies come from symmetries.  A copy maps a
 * point x to M*x + V and a field vector h to s*M*h (gmTrans,
 * cpp/src/ext/genmath/gmtrans.h:30-85).  All arrays are host pointers. */
typedef struct {
    int64_t n_elem;
    const int32_t *elem_type;  /* [n]  0 = rectangular block, 1 = polyhedron          */
    const double *rec_center;  /* [3n] block centre (ignored for polyhedra)           */
    const double *rec_dims;    /* [3n] block full dimensions                          */
    const int64_t *vert_off;   /* [n+1] polyhedron vertex ranges                      */
    const double *verts;       /* [3*V] polyhedron vertices                           */
    const int64_t *face_off;   /* [n+1] polyhedron face ranges                        */
    const int32_t *face_len;   /* [F]  vertices per face                              */
    const int32_t *face_idx;   /* [sum face_len] 0-based vertex index within element  */
    const int64_t *copy_off;   /* [n+1] transform-copy ranges                         */
    const double *copy_m;      /* [9*C] copy matrices, row-major                      */
    const double *copy_v;      /* [3*C] copy translations                             */
    const double *copy_s;      /* [C]  field multipliers (+1, or -1 for inversion)   */
} rh_model_desc;

typedef struct rh_model rh_model;

/* Builds face frames, outward normals, edge tables and copy transforms on the
 * host and uploads them.  Replaces the packing half of the reference seam:
 * radGPU_PackGeometryForAsm (cpp/src/core/radgpu_asm.cpp:206-444) and the
 * per-call geometry flattening of radGPU_ComputeField
 * (cpp/src/core/radgpu_fld.cpp:194-536). */
rh_status rh_model_create(const rh_model_desc *desc, int device, void *stream, rh_model **out);
rh_status rh_model_destroy(rh_model *m);
/* Collocation points (lab frame) of the elements, [3n] host doubles:
 * MainTrans(centroid), cpp/src/core/radintrc.cpp:852. */
rh_status rh_model_colloc(const rh_model *m, double *colloc_out);
/* Counts: [0] elements, [1] copies, [2] tetrahedra, [3] hexahedra,
 * [4] other polyhedra, [5] blocks, [6] edges, [7] faces. */
rh_status rh_model_counts(const rh_model *m, int64_t *counts8);
/* Field-sweep form of the polyhedra.  A model without transform copies whose
 * polyhedra have at most 8 faces each is also held as a conforming mesh: the
 * field sweep then evaluates each unique edge and face once with the net
 * weight of the elements sharing it, instead of every element separately (the
 * same sum, regrouped).  form 0: automatic (mesh when held), 1: per element.
 * No counterpart in the reference, which evaluates every polygon of every
 * element (cpp/src/core/radgpu_fld.cu:381-526). */
rh_status rh_model_set_field_form(rh_model *m, int32_t form);
/* The form field sweeps use: 1 per element, 2 mesh. */
rh_status rh_model_field_form(const rh_model *m, int32_t *form);
/* Mesh form sizes: [0] unique edges, [1] unique faces (triangles and
 * quadrilaterals), [2] polyhedra; zeros when the model holds no mesh form. */
rh_status rh_model_mesh_counts(const rh_model *m, int64_t *counts3);

/* ------------------------------------------------------------------------ */
/* Interaction matrix                                                         */
/* ------------------------------------------------------------------------ */
typedef struct rh_matrix rh_matrix;

/* Assembles the dense 3n x 3n interaction matrix (H = Q*M, element frames) on
 * the device and keeps it there.  Replaces radGPU_AssembleMatrix
 * (cpp/src/core/radgpu_asm.cu:889-1158) and the matrix hand-off
 * radGPU_PublishAssembledMatrix / radGPU_AdoptMatrixCache
 * (cpp/src/core/radgpu_asm.cu:40-51, cpp/src/core/radgpurlx.cu:728-746): the
 * handle owns the device matrix, no host copy is made.  A matrix beyond the
 * device budget is handled as the fallback policy says (rh_set_fallback). */
rh_status rh_matrix_assemble(rh_model *m, rh_precision prec, rh_matrix **out);

/* Placement of the matrix rows.  Rows that do not fit in device memory can
 * live in pinned host memory and be streamed through two device tiles (64 MiB,
 * or the rows of 64 elements when larger) in every matrix pass, the upload of
 * one tile overlapping the pass over the other.  The resident part and every
 * tile hold whole groups of 64 elements (the resident count is rounded down
 * to one), so results do not depend on the split: each block is integrated
 * and each row reduced the same way wherever it lives.  Replaces the reference's streaming matrix context
 * and its UtiGpuFallback 'gpu_streaming' policy (cpp/src/core/radgpurlx.cu:258-459,
 * cpp/src/core/radintrc.cpp:796-805). */
typedef struct {
    int64_t device_bytes;     /* device-memory budget for the matrix; <= 0: the free
                               * memory less 10 % and 1 GiB                        */
    double resident_fraction; /* < 0: keep as many element rows resident as the
                               * budget allows; else this fraction of them         */
    int32_t allow_streaming;  /* 0: RH_ERR_NOMEM when the matrix exceeds the
                               * budget; 1: stream the remaining rows;
                               * RH_STREAM_BY_POLICY (-1): as rh_get_fallback()
                               * says ('break' refuses, the others stream)       */
    int32_t fp64_integrals;   /* RH_MIXED only.  0: far blocks (beyond two
                               * circumradii of a block or tetrahedron) are
                               * integrated in FP32, near and self blocks in FP64;
                               * 1: every block integrated in FP64               */
    int32_t dense_rows;       /* 0: a model without transform copies or row
                               * transforms (every block a symmetric tensor) is
                               * stored in the symmetric-block layout (see
                               * rh_matrix_format); 1: always row-major 3x3
                               * blocks                                          */
    int32_t reserved;         /* 0                                               */
} rh_matrix_opts;
#define RH_STREAM_BY_POLICY (-1)
rh_status rh_matrix_assemble_ex(rh_model *m, rh_precision prec, const rh_matrix_opts *opts,
                                rh_matrix **out);
/* Rows held on the device and rows streamed from host memory. */
rh_status rh_matrix_layout(const rh_matrix *q, int64_t *resident_rows, int64_t *streamed_rows);
rh_status rh_matrix_destroy(rh_matrix *q);
/* Copies rows [row0, row0+nrows) of the 3n x 3n matrix to host doubles. */
rh_status rh_matrix_export(const rh_matrix *q, int64_t row0, int64_t nrows, double *out);
/* Raw device pointer and element size (4 or 8 bytes) of the stored matrix
 * (RH_ERR_UNSUPPORTED when part of it is streamed), in the layout that
 * rh_matrix_format reports. */
rh_status rh_matrix_device_ptr(const rh_matrix *q, void **dptr, int32_t *elem_bytes);
/* Storage layout of the matrix (no counterpart in the reference, which stores
 * row-major 3x3 blocks, radintrc.h:126).
 *   planes == 0: row-major, element row i = rows 3i..3i+2 of *ld values
 *     (columns 3j..3j+2 hold block (i, j); padding zero).
 *   planes == 5 or 6: symmetric-block layout, used when no element has
 *     transform copies or a row transform, so that every off-diagonal block
 *     is the symmetric field tensor of source j at collocation point i.
 *     Element row i = planes rows of *ld values: (xx, xy, xz, yy, yz[, zz])
 *     of blocks (i, 0 .. n-1).  With 5 planes zz = -(xx + yy) (the tensor is
 *     traceless outside its source); a model where a collocation point lies
 *     inside another element is stored with 6.
 * The diagonal blocks are held apart in FP64 and read as zeros in both. */
rh_status rh_matrix_format(const rh_matrix *q, int32_t *planes, int64_t *ld);
/* Re-assembles into an existing matrix (same model); the timed unit of work. */
rh_status rh_matrix_reassemble(rh_matrix *q);

/* ------------------------------------------------------------------------ */
/* Relaxation                                                                 */
/* ------------------------------------------------------------------------ */

/* Per-element material record, 16 doubles (radmater.h:164-696):
 *   [0] kind: 0 linear isotropic, 1 linear anisotropic, 2 nonlinear isotropic
 *       (tanh-sum formula MatSatIsoFrm), 3 nonlinear isotropic (table, MatSatIsoTab)
 *   kind 0: [1] ksi
 *   kind 1: [4..12] susceptibility tensor, row-major (element frame)
 *   kind 2: [1] number of terms (1..3), [2..4] ms_i, [5..7] ksi_i
 *   kind 3: [1] first node, [2] node count of a tabulated law (rh_relax_tab)
 *   [13..15] remanent magnetization (element frame) */
#define RH_MAT_STRIDE 16

typedef struct {
    double prec;       /* stop when the RMS magnetization residual (T) <= prec */
    int32_t max_iter;  /* iteration limit                                      */
    int32_t keep_m;    /* 1: start from M_inout ('ZeroM->False'), 0: from M_r  */
    double omega0;     /* mixing factor of the iteration (0 -> 0.5)            */
    int32_t method;    /* 9 (or 0): Anderson-accelerated block Jacobi;         *
                        * 11: Newton-Krylov (GMRES on the analytic Jacobian,    *
                        * radGPU_RelaxNK, cpp/src/core/radgpurlx.cu:1351-2285)  */
    int32_t refine;    /* RH_MIXED matrix only: FP64 refinement passes after   *
                        * convergence (0 -> 1, -1 -> none); each pass takes the *
                        * FP64 field at the collocation points, corrects the   *
                        * external field by its defect against the stored       *
                        * matrix and resumes the iteration                     */
    int32_t krylov_m;  /* method 11: GMRES restart length (0 -> 60, max 63)    */
    int32_t reserved;  /* 0                                                    */
} rh_relax_opts;

typedef struct {
    double misfit;     /* RMS over elements of |M_proposal - M| (method 9)     *
                        * or of |M_mat(H) - M| (method 11), in T               */
    double max_m;      /* max |M| (T)                                          */
    double max_h;      /* max |H| (T)                                          */
    int32_t iters;     /* iterations (method 11: matrix passes, as the          *
                        * reference counts them)                               */
    int32_t matvecs;   /* matrix passes (each streams the matrix)              */
    int32_t converged; /* 1 if misfit <= prec                                  */
    int32_t refined;   /* FP64 refinement passes performed                     */
    int32_t newton;    /* method 11: Newton steps                              */
    int32_t krylov;    /* method 11: GMRES iterations                          */
} rh_relax_result;

/* Relaxes the magnetization to M = M_mat(H_ext + Q*M).  mats: [16n] host
 * records; hext: [3n] external field at the collocation points (element
 * frames); M_inout, H_out: [3n] host doubles.  Replaces radGPU_RelaxAuto
 * (cpp/src/core/radgpurlx.cu:799-1349, header cpp/src/core/radgpurlx.h:69-86)
 * and the method-9 driver radGPU_AutoRelax (cpp/src/core/radrlmet_gpu.cpp:9-53). */
rh_status rh_relax(rh_matrix *q, const double *mats, const double *hext, double *M_inout,
                   double *H_out, const rh_relax_opts *opts, rh_relax_result *res);

/* Same on device-resident vectors (d_mats [16n], d_hext, d_M, d_H [3n] doubles). */
rh_status rh_relax_dev(rh_matrix *q, const double *d_mats, const double *d_hext, double *d_M,
                       double *d_H, const rh_relax_opts *opts, rh_relax_result *res);

/* Relaxation with tabulated isotropic laws (MatSatIsoTab, radmater.h:436-451;
 * GPU counterpart interp_mh_dev, cpp/src/core/radgpurlx.cu:83-100): material
 * kind 3 records hold [1] the first node and [2] the node count of a law in
 * tab, a flat array of ntab doubles = node triples (|H|, |M|, d|M|/d|H|) as
 * the reference builds them (radmater.cpp:460-609).  |M| is the cubic Hermite
 * interpolant between nodes, |M(0)| below the first node and the last slope
 * beyond the last (radmater.cpp:619-646). */
rh_status rh_relax_tab(rh_matrix *q, const double *mats, const double *tab, int64_t ntab,
                       const double *hext, double *M_inout, double *H_out, const rh_relax_opts *opts,
                       rh_relax_result *res);
rh_status rh_relax_tab_dev(rh_matrix *q, const double *d_mats, const double *d_tab, int64_t ntab,
                           const double *d_hext, double *d_M, double *d_H, const rh_relax_opts *opts,
                           rh_relax_result *res);

/* ------------------------------------------------------------------------ */
/* Field evaluation                                                           */
/* ------------------------------------------------------------------------ */

/* B and/or H of the model's elements with magnetizations M ([3n], element
 * frames) at np lab-frame points (B = H + M inside an element).  B_out or
 * H_out may be NULL.  prec = RH_FP64 evaluates the integrals in FP64;
 * RH_MIXED evaluates them in FP32 and accumulates in FP64.  Host arrays.
 * Replaces radGPU_ComputeField (cpp/src/core/radgpu_fld.h:172-184,
 * radgpu_fld.cpp:623-857; use_gpu 1 = FP64, 2 = FP32) and the kernels behind
 * radGPU_Fld{AllocAndCopy,LaunchKernel,RetrieveAndFree}
 * (cpp/src/core/radgpu_fld.cu:381-1023). */
rh_status rh_field(rh_model *m, const double *M, const double *pts, int64_t np, double *B_out,
                   double *H_out, rh_precision prec);
/* Same on device arrays; the library keeps a growable scratch buffer per model. */
rh_status rh_field_dev(rh_model *m, const double *d_M, const double *d_pts, int64_t np,
                       double *d_B, double *d_H, rh_precision prec);

/* No counterpart in the reference.  rh_field_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): same arguments, results and
 * synchronisation, and no hipMalloc/hipFree or other allocation.  Its partial
 * sums use a fixed device array of RH_NOALLOC_SCRATCH_BYTES per GPU, reserved
 * with the library's code on every GPU it runs on until the process ends;
 * the element splits are reduced to fit it.  When even one split per element
 * kind does not fit (6 doubles per point per element kind: more than 699,050
 * points on a model holding two kinds, 466,033 on one holding three) the call
 * returns RH_ERR_ARG before any launch, naming the largest point count that
 * fits.  Calls on one GPU share that array, and the model's charged-record
 * buffers: calls that overlap on one GPU race. */
#define RH_NOALLOC_SCRATCH_BYTES (64u << 20)
rh_status rh_field_dev_noalloc(rh_model *m, const double *d_M, const double *d_pts, int64_t np,
                               double *d_B, double *d_H, rh_precision prec);

/* Current sources.  Record kinds (12 doubles each):
 *   0 block current: centre[3], dims[3], J[3]          (radrec.cpp:49-315, J branch)
 *   1 arc current:   centre[3], rmin, rmax, phimin, phimax, height, j, nseg
 *                    (trapezoid rule, radarccu.cpp:30-234)
 * each with one transform (M [9], V [3], s).  Adds their B (= H outside
 * magnetic material) at np points into B_acc ([3np] host doubles).
 * Replaces the host loop ComputeCoilFieldCPU (cpp/src/core/radgpu_fld.cpp:542-617)
 * and the external-field set-up of radTInteraction::SetupExternFieldArray
 * (cpp/src/core/radintrc.cpp:1208-1236). */
rh_status rh_current_field(int device, void *stream, int64_t nsrc, const int32_t *kind,
                           const double *par, const double *tm, const double *tv,
                           const double *ts, const double *pts, int64_t np, double *B_acc);
/* Same with pts and B_acc in device memory (source records stay on the host). */
rh_status rh_current_field_dev(int device, void *stream, int64_t nsrc, const int32_t *kind,
                               const double *par, const double *tm, const double *tv,
                               const double *ts, const double *d_pts, int64_t np, double *d_B_acc);

/* ------------------------------------------------------------------------ */
/* Adjoint (differentiation)                                                  */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference: the reference has no derivatives of its
 * field or relaxation.  These give the reverse-mode products a gradient-based
 * design loop needs (dL/dM, dL/d(vertex), dL/d(point)) and the adjoint solve
 * of the linear-material relaxation, all in FP64. */

/* Number of polyhedron vertices of the model: the description's vertices in
 * its order (blocks have none); gV below holds 3 doubles per vertex. */
rh_status rh_model_vertex_count(const rh_model *m, int64_t *nvert);

/* Adjoint of rh_field for weights gB = dL/dB (field = 0) or dL/dH (field = 1)
 * at the np points; device arrays of doubles, any output may be null:
 *   gM [3n]      dL/dM, element frames
 *   gV [3 nvert] dL/d(vertex coordinates), element frames (needs d_M)
 *   gX [3 np]    dL/d(point coordinates), lab frame       (needs d_M)
 * gV and gX need a model of polyhedra only (RH_ERR_UNSUPPORTED otherwise).
 * Vertex derivatives treat each face as its fan of triangles, which equals
 * the face for planar faces and stays a closed surface under displacement. */
rh_status rh_field_adjoint_dev(rh_model *m, const double *d_M, const double *d_pts, int64_t np,
                               const double *d_gB, int32_t field, double *d_gM, double *d_gV,
                               double *d_gX);
/* Same on host arrays (M may be null when only gM is requested). */
rh_status rh_field_adjoint(rh_model *m, const double *M, const double *pts, int64_t np,
                           const double *gB, int32_t field, double *gM, double *gV, double *gX);

/* Adjoint of the linear-material relaxation (material kinds 0 and 1):
 * solves (I - Q^T K^T) lam = rhs on the stored matrix, K the susceptibility
 * tensor of each element, by Anderson-accelerated block Jacobi on the
 * transposed system.  With M solving (I - K Q) M = K Hext + M_r and a loss
 * L(M), rhs = dL/dM gives dL/dM_r = lam and dL/dHext = K^T lam.  d_mats
 * [16n], d_rhs [3n], d_lam [3n] device doubles (d_lam is the start value
 * when opts->keep_m, and the result). */
rh_status rh_relax_adjoint_dev(rh_matrix *q, const double *d_mats, const double *d_rhs,
                               double *d_lam, const rh_relax_opts *opts, rh_relax_result *res);
/* Same on host arrays. */
rh_status rh_relax_adjoint(rh_matrix *q, const double *mats, const double *rhs, double *lam,
                           const rh_relax_opts *opts, rh_relax_result *res);

/* ------------------------------------------------------------------------ */
/* Fallback policy and backend report                                         */
/* ------------------------------------------------------------------------ */
/* What rh_matrix_assemble (and rh_matrix_assemble_ex with allow_streaming =
 * RH_STREAM_BY_POLICY) does with a matrix beyond the device budget.  Process
 * wide, default RH_FALLBACK_CPU.  The reference's 'cpu' mode assembles on the
 * host; this library has no host assembly, so 'cpu' and 'gpu_streaming' both
 * keep the rows that fit resident and stream the rest, and 'break' refuses the
 * matrix with RH_ERR_NOMEM.  Replaces RadSetGpuFallbackMode /
 * RadGetGpuFallbackMode (cpp/src/core/radintrc.cpp:45-46, policy values of
 * RadUtiGpuFallback, cpp/src/lib/radentry.cpp:934-944). */
typedef enum {
    RH_FALLBACK_CPU = 0,           /* 'cpu'           */
    RH_FALLBACK_GPU_STREAMING = 1, /* 'gpu_streaming' */
    RH_FALLBACK_BREAK = 2          /* 'break'         */
} rh_fallback;
/* RH_ERR_ARG (policy unchanged) for a mode outside 0..2. */
rh_status rh_set_fallback(int32_t mode);
int32_t rh_get_fallback(void);

/* Backend of the most recent assembly (path 0: rh_matrix_assemble*) or field
 * evaluation (path 1: rh_field*, rh_current_field*) in this process:
 * -1 none yet, 1 device, 2 device with rows streamed from host memory
 * (assembly only); 0 (host) is never reported, as nothing here runs on the
 * host.  -1 for an unknown path.  Replaces RadGetLastAsmBackend
 * (cpp/src/core/radintrc.cpp:44) and RadUtiFldLastBackend
 * (cpp/src/lib/radentry.cpp:1043-1045, 1760-1765). */
int32_t rh_last_backend(int32_t path);

/* Diagnostics */
const char *rh_last_error(void);
int32_t rh_device_count(void);
/* Average duration (ms) and count of the most recent timed launches of the
 * named kernel family ("asm", "field", "matvec"), measured with HIP events on
 * the launch stream when rh_set_timing(1) is on. */
rh_status rh_set_timing(int32_t on);
rh_status rh_kernel_time(const char *family, double *avg_ms, int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* RADHIP_H */
