// rh_api.cpp -- the C ABI (include/radhip.h): device-resident models and
// matrices, the relaxation driver and the field entry points.  Host C++ that
// calls the gfx950 kernels through rh_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/radhip.h"
#include "rh_api_internal.hpp"
#include "rh_geom.hpp"
#include "rh_kernels.hpp"
#include "rh_math.hpp"
#include "rh_mesh.hpp"

namespace {

thread_local std::string g_err;

rh_status fail(rh_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

// Makes `device` current and drops an error left pending on this thread by
// earlier HIP calls outside this library, so that the post-launch checks
// below report this library's launches only.
hipError_t enter_device(int device)
{
    const hipError_t e = hipSetDevice(device);
    (void)hipGetLastError();
    return e;
}

#define RH_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP,           \
                        std::string(#call) + ": " + hipGetErrorString(e_));              \
    } while (0)

// ------------------------------------------------- fallback policy, backends
std::atomic<int32_t> g_fallback{RH_FALLBACK_CPU};
std::atomic<int32_t> g_backend_asm{-1};
std::atomic<int32_t> g_backend_fld{-1};

// ---------------------------------------------------------------- timing
std::atomic<bool> g_timing{false};
struct Family {
    double total_ms = 0.0;
    int64_t count = 0;
};
std::map<std::string, Family> g_fam;
std::mutex g_fam_mu;  // the timing statistics are shared by all threads

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

// Records launches of one kernel family inside one API call and folds their
// durations into the family statistics at the end of the call.
struct Timer {
    std::string fam;
    hipStream_t s;
    std::vector<EventPair> ev;
    Timer(const char* f, hipStream_t st) : fam(f), s(st) {}
    void begin()
    {
        if (!g_timing) return;
        EventPair p;
        (void)hipEventCreate(&p.a);
        (void)hipEventCreate(&p.b);
        (void)hipEventRecord(p.a, s);
        ev.push_back(p);
    }
    void end()
    {
        if (!g_timing || ev.empty()) return;
        (void)hipEventRecord(ev.back().b, s);
    }
    // Keeps the first k recorded launches only (launches that returned at
    // once, after a device-driven loop had finished, are not statistics).
    void keep(size_t k)
    {
        while (ev.size() > k) {
            (void)hipEventSynchronize(ev.back().b);
            (void)hipEventDestroy(ev.back().a);
            (void)hipEventDestroy(ev.back().b);
            ev.pop_back();
        }
    }
    ~Timer()
    {
        if (ev.empty()) return;
        std::lock_guard<std::mutex> lock(g_fam_mu);
        Family& F = g_fam[fam];
        for (EventPair& p : ev) {
            (void)hipEventSynchronize(p.b);
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                F.total_ms += ms;
                F.count++;
            }
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
};

template <class T>
hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

template <class T>
hipError_t upload(T** p, const std::vector<T>& v, hipStream_t s)
{
    hipError_t e = dalloc(p, v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// Anderson-accelerated (type II, depth kHist) under-relaxed block-Jacobi
// iteration on M = G(M), G = per-element implicit material update given the
// field of all other elements (the fixed point of radGPU_RelaxAuto,
// cpp/src/core/radgpurlx.cu:799-1349, with an acceleration of our own).
constexpr int kHist = rh::kHistMax;

// Work arrays of the relaxation: kept with the matrix and reused by every
// solve on it (allocated by the first one).
struct RelaxWork {
    int64_t n3 = 0, ld = 0, nb = 0, db = 0;
    double* x = nullptr;   // [ld] current iterate (padded)
    double* y = nullptr;   // [n3]
    double* F = nullptr;   // [n3]
    double* Fp = nullptr;  // [n3] previous F
    double* xp = nullptr;  // [n3] previous x
    double* dX[kHist] = {nullptr};
    double* dF[kHist] = {nullptr};
    double* part = nullptr;
    double* dots = nullptr;
    rh::RelaxDevState* st = nullptr;   // state of the device-driven loop
    rh::RelaxDevState* hst = nullptr;  // pinned: two read-back slots and an upload slot
    hipEvent_t ev[2] = {nullptr, nullptr};
    double* hbuf = nullptr;            // pinned read-back of the host-driven loops
    double* H64 = nullptr;             // [n3] FP64 refinement: field at the collocation points
    double* hx = nullptr;              // [n3] FP64 refinement: corrected external field
    ~RelaxWork()
    {
        dfree(x);
        dfree(y);
        dfree(F);
        dfree(Fp);
        dfree(xp);
        for (int k = 0; k < kHist; k++) {
            dfree(dX[k]);
            dfree(dF[k]);
        }
        dfree(part);
        dfree(dots);
        dfree(st);
        dfree(H64);
        dfree(hx);
        if (hst) (void)hipHostFree(hst);
        if (hbuf) (void)hipHostFree(hbuf);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    // All arrays of a solve on n elements (3n rows, leading dimension ld).
    hipError_t alloc(int64_t n, int64_t n3_, int64_t ld_)
    {
        n3 = n3_;
        ld = ld_;
        nb = rh::relax_blocks(n);
        db = rh::dot_blocks(n3);
        hipError_t e = dalloc(&x, size_t(ld));
        if (e == hipSuccess) e = dalloc(&y, size_t(n3));
        if (e == hipSuccess) e = dalloc(&F, size_t(n3));
        if (e == hipSuccess) e = dalloc(&Fp, size_t(n3));
        if (e == hipSuccess) e = dalloc(&xp, size_t(n3));
        for (int k = 0; k < kHist && e == hipSuccess; k++) {
            e = dalloc(&dX[k], size_t(n3));
            if (e == hipSuccess) e = dalloc(&dF[k], size_t(n3));
        }
        if (e == hipSuccess) e = dalloc(&part, size_t(rh::kRelaxPart * nb));
        if (e == hipSuccess) e = dalloc(&dots, size_t(rh::kDotMax * db));
        if (e == hipSuccess) e = dalloc(&st, 1);
        if (e == hipSuccess) e = dalloc(&H64, size_t(n3));
        if (e == hipSuccess) e = dalloc(&hx, size_t(n3));
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&hst), 3 * sizeof(rh::RelaxDevState));
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&hbuf), size_t(rh::kRelaxPart * nb + rh::kDotMax * db) * sizeof(double));
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
        return e;
    }
};

}  // namespace

// ---------------------------------------------------------------- handles
struct rh_model {
    int device = 0;
    hipStream_t stream = nullptr;
    rh::PackedModel host;
    rh::DevModel dev;
    int32_t* d_kind = nullptr;
    int64_t* d_off = nullptr;
    double* d_data = nullptr;
    int64_t* d_copy_off = nullptr;
    double* d_copies = nullptr;
    double* d_colloc = nullptr;
    double* d_rowt = nullptr;
    int32_t* d_kidx = nullptr;
    double* d_charged = nullptr;  // field-only tetrahedron records, rebuilt per field call
    // adjoint geometry, uploaded on the first adjoint call
    double* d_adj = nullptr;
    int64_t* d_adj_off = nullptr;
    int64_t* d_vert_off = nullptr;
    int64_t* d_slices = nullptr;
    int64_t n_slices = 0;
    // conforming-mesh field form
    std::vector<void*> mesh_bufs;
    // growable workspace for the field entry points
    double* d_work = nullptr;
    size_t work_cap = 0;
};

struct rh_matrix {
    rh_model* m = nullptr;  // must outlive the matrix
    int device = 0;
    rh_precision prec = RH_MIXED;
    bool far32 = false;  // far blocks integrated in FP32 (RH_MIXED)
    int elem_bytes = 4;
    // Storage of an element row: planes == 0, its 3 matrix rows of ld values;
    // planes == 5 or 6, the symmetric-block planes of ldp values (ld = 3*ldp
    // is then the padded length of the iterate).
    int planes = 0;
    int64_t n3 = 0, ld = 0, ldp = 0;
    int32_t* d_flag = nullptr;  // assembly: a traced block met in the 5-plane layout
    rh_matrix_opts opts{};
    void* d_Q = nullptr;  // rows of elements [0, res_elems)
    double* d_qdiag = nullptr;
    // out-of-core part: rows of elements [res_elems, n) in pinned host memory,
    // streamed through two device tiles of tile_elems elements each
    int64_t res_elems = 0;
    void* h_Q = nullptr;
    void* stage[2] = {nullptr, nullptr};
    int64_t tile_elems = 0;
    hipStream_t copy = nullptr;
    hipEvent_t ev_ready[2] = {nullptr, nullptr};  // tile b holds data for the consumer
    hipEvent_t ev_free[2] = {nullptr, nullptr};   // tile b may be overwritten
    RelaxWork* rw = nullptr;  // relaxation work, kept between solves
    size_t row_bytes() const { return size_t(ld) * size_t(elem_bytes); }
    // bytes of one element row's storage
    size_t elem_stride() const
    {
        return planes ? size_t(planes) * size_t(ldp) * size_t(elem_bytes) : 3 * row_bytes();
    }
    bool streamed() const { return res_elems < n3 / 3; }
};

namespace {

hipError_t grow(rh_model* m, size_t doubles)
{
    if (doubles <= m->work_cap) return hipSuccess;
    dfree(m->d_work);
    m->d_work = nullptr;
    m->work_cap = 0;
    hipError_t e = dalloc(&m->d_work, doubles);
    if (e == hipSuccess) m->work_cap = doubles;
    return e;
}

void free_model(rh_model* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    dfree(m->d_kind);
    dfree(m->d_off);
    dfree(m->d_data);
    dfree(m->d_copy_off);
    dfree(m->d_copies);
    dfree(m->d_colloc);
    dfree(m->d_rowt);
    dfree(m->d_kidx);
    dfree(m->d_charged);
    dfree(m->d_adj);
    dfree(m->d_adj_off);
    dfree(m->d_vert_off);
    dfree(m->d_slices);
    for (void* b : m->mesh_bufs) dfree(b);
    dfree(m->d_work);
    delete m;
}

}  // namespace

extern "C" {

const char* rh_last_error(void) { return g_err.c_str(); }

int32_t rh_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

rh_status rh_set_fallback(int32_t mode)
{
    if (mode < RH_FALLBACK_CPU || mode > RH_FALLBACK_BREAK)
        return fail(RH_ERR_ARG, "fallback mode must be 0 ('cpu'), 1 ('gpu_streaming') or 2 ('break')");
    g_fallback = mode;
    return RH_OK;
}

int32_t rh_get_fallback(void) { return g_fallback; }

int32_t rh_last_backend(int32_t path)
{
    if (path == 0) return g_backend_asm;
    if (path == 1) return g_backend_fld;
    return -1;
}

rh_status rh_set_timing(int32_t on)
{
    std::lock_guard<std::mutex> lock(g_fam_mu);
    g_timing = on != 0;
    g_fam.clear();
    return RH_OK;
}

rh_status rh_kernel_time(const char* family, double* avg_ms, int64_t* count)
{
    if (!family || !avg_ms || !count) return fail(RH_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lock(g_fam_mu);
    auto it = g_fam.find(family);
    if (it == g_fam.end() || it->second.count == 0) {
        *avg_ms = 0.0;
        *count = 0;
        return RH_OK;
    }
    *avg_ms = it->second.total_ms / double(it->second.count);
    *count = it->second.count;
    return RH_OK;
}

rh_status rh_model_create(const rh_model_desc* desc, int device, void* stream, rh_model** out)
{
    if (!desc || !out) return fail(RH_ERR_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(RH_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(RH_ERR_ARG, "device index out of range");
    RH_HIP(enter_device(device));
    rh_model* m = new rh_model();
    m->device = device;
    m->stream = static_cast<hipStream_t>(stream);
    rh_status st = rh::pack_model(desc, &m->host, &g_err);
    if (st != RH_OK) {
        delete m;
        return st;
    }
    const rh::PackedModel& P = m->host;
    hipStream_t s = m->stream;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = upload(&m->d_kind, P.kind, s);
    if (e == hipSuccess) e = upload(&m->d_off, P.off, s);
    if (e == hipSuccess) e = upload(&m->d_data, P.data, s);
    if (e == hipSuccess) e = upload(&m->d_copy_off, P.copy_off, s);
    if (e == hipSuccess) e = upload(&m->d_copies, P.copies, s);
    if (e == hipSuccess) e = upload(&m->d_colloc, P.colloc, s);
    if (e == hipSuccess) e = upload(&m->d_rowt, P.rowt, s);
    if (e == hipSuccess) e = upload(&m->d_kidx, P.kind_idx, s);
    if (e == hipSuccess)
        e = dalloc(&m->d_charged, size_t(rh::kChargedTet * (P.kind_off[2] - P.kind_off[1])));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        std::string msg = std::string("model upload: ") + hipGetErrorString(e);
        free_model(m);
        return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP, msg);
    }
    rh::DevModel& d = m->dev;
    d.n = P.n;
    d.kind = m->d_kind;
    d.off = m->d_off;
    d.data = m->d_data;
    d.copy_off = m->d_copy_off;
    d.copies = m->d_copies;
    d.colloc = m->d_colloc;
    d.rowt = m->d_rowt;
    d.kidx = m->d_kidx;
    for (int k = 0; k < 4; k++) d.koff[k] = P.kind_off[k];
    d.sym = P.sym;
    d.rowt_identity = P.rowt_identity;
    if (P.mesh.active) {
        const rh::MeshForm& F = P.mesh;
        rh::MeshDev& md = d.mesh;
        auto up = [&](auto** dst, const auto& v) {
            using T = typename std::remove_reference<decltype(v)>::type::value_type;
            T* ptr = nullptr;
            hipError_t err = upload(&ptr, v, s);
            m->mesh_bufs.push_back(ptr);
            *dst = ptr;
            return err;
        };
        md.ne = F.n_edges();
        md.nf = F.n_faces();
        md.ni = F.n_in();
        md.ntiles = int64_t(F.tile_box.size() / 6);
        e = up(&md.edge_geo, F.edge_geo);
        if (e == hipSuccess) e = up(&md.edge_inc_off, F.edge_inc_off);
        if (e == hipSuccess) e = up(&md.edge_inc_elem, F.edge_inc_elem);
        if (e == hipSuccess) e = up(&md.edge_inc_w, F.edge_inc_w);
        if (e == hipSuccess) e = up(&md.face_geo, F.face_geo);
        if (e == hipSuccess) e = up(&md.face_inc_off, F.face_inc_off);
        if (e == hipSuccess) e = up(&md.face_inc_elem, F.face_inc_elem);
        if (e == hipSuccess) e = up(&md.face_inc_v, F.face_inc_v);
        if (e == hipSuccess) e = up(&md.in_elem, F.in_elem);
        if (e == hipSuccess) e = up(&md.in_geo, F.in_geo);
        if (e == hipSuccess) e = up(&md.in_F, F.in_F);
        if (e == hipSuccess) e = up(&md.tile_box, F.tile_box);
        if (e == hipSuccess) e = up(&md.edge_cont, F.edge_cont);
        if (e == hipSuccess) e = up(&md.face_cont, F.face_cont);
        if (e == hipSuccess) e = up(&md.face_edge, F.face_edge);
        md.ne_free = F.ne_free;
        md.nf_tri = F.nf_tri;
        double* ch = nullptr;
        if (e == hipSuccess) e = dalloc(&ch, size_t(rh::kMeshEdge * md.ne + rh::kMeshFace * md.nf +
                                                    rh::kMeshIn * md.ni) + 2);
        if (ch) m->mesh_bufs.push_back(ch);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            std::string msg = std::string("mesh upload: ") + hipGetErrorString(e);
            free_model(m);
            return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP, msg);
        }
        md.ch_edge = ch;
        md.ch_face = ch + rh::kMeshEdge * md.ne;
        md.ch_in = md.ch_face + rh::kMeshFace * md.nf;
        md.active = true;
        d.use_mesh = true;
    }
    *out = m;
    return RH_OK;
}

rh_status rh_model_set_field_form(rh_model* m, int32_t form)
{
    if (!m) return fail(RH_ERR_ARG, "null argument");
    if (form != 0 && form != 1) return fail(RH_ERR_ARG, "form must be 0 (automatic) or 1 (per element)");
    m->dev.use_mesh = (form == 0) && m->dev.mesh.active;
    return RH_OK;
}

rh_status rh_model_mesh_counts(const rh_model* m, int64_t* counts3)
{
    if (!m || !counts3) return fail(RH_ERR_ARG, "null argument");
    counts3[0] = m->dev.mesh.ne;
    counts3[1] = m->dev.mesh.nf;
    counts3[2] = m->dev.mesh.ni;
    return RH_OK;
}

rh_status rh_model_field_form(const rh_model* m, int32_t* form)
{
    if (!m || !form) return fail(RH_ERR_ARG, "null argument");
    *form = (m->dev.use_mesh && m->dev.mesh.active) ? 2 : 1;
    return RH_OK;
}

rh_status rh_model_destroy(rh_model* m)
{
    free_model(m);
    return RH_OK;
}

rh_status rh_model_colloc(const rh_model* m, double* colloc_out)
{
    if (!m || !colloc_out) return fail(RH_ERR_ARG, "null argument");
    std::memcpy(colloc_out, m->host.colloc.data(), m->host.colloc.size() * sizeof(double));
    return RH_OK;
}

rh_status rh_model_counts(const rh_model* m, int64_t* counts8)
{
    if (!m || !counts8) return fail(RH_ERR_ARG, "null argument");
    for (int k = 0; k < 8; k++) counts8[k] = m->host.counts[k];
    return RH_OK;
}

// ---------------------------------------------------------------- assembly
// y = Q x over element rows [e0, e1) held at Qp (the first of them).
static hipError_t matvec_rows(const rh_matrix* q, const void* Qp, int64_t e0, int64_t e1, const double* x,
                              double* y, hipStream_t s, const int32_t* skip = nullptr)
{
    if (q->planes)
        return rh::launch_matvec_sym(Qp, q->elem_bytes, q->planes, e1 - e0, q->ldp, x, nullptr, y + 3 * e0, s,
                                     skip);
    return rh::launch_matvec(Qp, q->elem_bytes, 3 * (e1 - e0), q->ld, x, nullptr, y + 3 * e0, s, skip);
}

// Assembles every element row; *traced is set when the 5-plane layout met a
// block with a non-zero trace (the matrix must then be stored with 6 planes).
static rh_status assemble_rows(rh_matrix* q, bool* traced)
{
    rh_model* m = q->m;
    RH_HIP(enter_device(m->device));
    hipStream_t s = m->stream;
    const int64_t n = m->host.n;
    const int64_t lds = q->planes ? q->ldp : q->ld;
    *traced = false;
    if (q->planes == 5) RH_HIP(hipMemsetAsync(q->d_flag, 0, sizeof(int32_t), s));
    Timer t("asm", s);
    t.begin();
    RH_HIP(rh::launch_assemble(m->dev, q->d_Q, q->elem_bytes, lds, q->d_qdiag, 0, q->res_elems, s, q->far32,
                               q->planes, q->d_flag));
    // streamed rows: assemble a tile on the device, copy it to pinned host
    // memory on the copy stream while the next tile is assembled
    int64_t tile = 0;
    for (int64_t e0 = q->res_elems; e0 < n; e0 += q->tile_elems, tile++) {
        const int b = int(tile & 1);
        const int64_t e1 = std::min(n, e0 + q->tile_elems);
        RH_HIP(hipStreamWaitEvent(s, q->ev_free[b], 0));
        RH_HIP(rh::launch_assemble(m->dev, q->stage[b], q->elem_bytes, lds, q->d_qdiag, e0, e1, s, q->far32,
                                   q->planes, q->d_flag));
        RH_HIP(hipEventRecord(q->ev_ready[b], s));
        RH_HIP(hipStreamWaitEvent(q->copy, q->ev_ready[b], 0));
        char* dst = static_cast<char*>(q->h_Q) + size_t(e0 - q->res_elems) * q->elem_stride();
        RH_HIP(hipMemcpyAsync(dst, q->stage[b], size_t(e1 - e0) * q->elem_stride(), hipMemcpyDeviceToHost,
                              q->copy));
        RH_HIP(hipEventRecord(q->ev_free[b], q->copy));
    }
    t.end();
    if (q->copy) RH_HIP(hipStreamSynchronize(q->copy));
    if (q->planes == 5) {
        int32_t f = 0;
        RH_HIP(hipMemcpyAsync(&f, q->d_flag, sizeof(f), hipMemcpyDeviceToHost, s));
        RH_HIP(hipStreamSynchronize(s));
        *traced = f != 0;
    }
    RH_HIP(hipStreamSynchronize(s));
    return RH_OK;
}

// y = Q x over all rows: the resident rows in place, the streamed rows tile by
// tile, the upload of tile t+1 overlapping the pass over tile t.
static hipError_t matrix_pass(rh_matrix* q, const double* x, double* y, hipStream_t s)
{
    const int64_t n = q->m->host.n;
    hipError_t e = matvec_rows(q, q->d_Q, 0, q->res_elems, x, y, s);
    int64_t tile = 0;
    for (int64_t e0 = q->res_elems; e == hipSuccess && e0 < n; e0 += q->tile_elems, tile++) {
        const int b = int(tile & 1);
        const int64_t e1 = std::min(n, e0 + q->tile_elems);
        const char* src = static_cast<const char*>(q->h_Q) + size_t(e0 - q->res_elems) * q->elem_stride();
        e = hipStreamWaitEvent(q->copy, q->ev_free[b], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(q->stage[b], src, size_t(e1 - e0) * q->elem_stride(), hipMemcpyHostToDevice, q->copy);
        if (e == hipSuccess) e = hipEventRecord(q->ev_ready[b], q->copy);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, q->ev_ready[b], 0);
        if (e == hipSuccess) e = matvec_rows(q, q->stage[b], e0, e1, x, y, s);
        if (e == hipSuccess) e = hipEventRecord(q->ev_free[b], s);
    }
    return e;
}

// Frees the matrix storage (device rows, streamed rows, tiles, copy stream).
static void release_storage(rh_matrix* q)
{
    (void)hipSetDevice(q->device);
    if (q->copy) (void)hipStreamSynchronize(q->copy);
    dfree(q->d_Q);
    q->d_Q = nullptr;
    dfree(q->stage[0]);
    dfree(q->stage[1]);
    q->stage[0] = q->stage[1] = nullptr;
    if (q->h_Q) (void)hipHostFree(q->h_Q);
    q->h_Q = nullptr;
    for (int b = 0; b < 2; b++) {
        if (q->ev_ready[b]) (void)hipEventDestroy(q->ev_ready[b]);
        if (q->ev_free[b]) (void)hipEventDestroy(q->ev_free[b]);
        q->ev_ready[b] = q->ev_free[b] = nullptr;
    }
    if (q->copy) (void)hipStreamDestroy(q->copy);
    q->copy = nullptr;
    q->tile_elems = 0;
}

static void free_matrix(rh_matrix* q)
{
    if (!q) return;
    release_storage(q);
    dfree(q->d_qdiag);
    dfree(q->d_flag);
    delete q->rw;
    delete q;
}

// Places and allocates the storage of q for q->planes (q->opts as validated):
// the resident element rows (a whole number of 64-element groups when some
// are streamed) and, beyond the budget, pinned host rows and two tiles.
static rh_status alloc_storage(rh_matrix* q)
{
    rh_model* m = q->m;
    const rh_matrix_opts& o = q->opts;
    const int64_t n = m->host.n;
    if (q->planes) {
        q->ldp = std::max<int64_t>(256, (n + 255) / 256 * 256);
        q->ld = 3 * q->ldp;
    } else {
        q->ldp = 0;
        q->ld = std::max<int64_t>(256, (q->n3 + 255) / 256 * 256);
    }
    const size_t elem_rows = q->elem_stride();  // one element row's storage
    const size_t total = size_t(n) * elem_rows;
    size_t budget = size_t(o.device_bytes);
    if (o.device_bytes <= 0) {
        size_t fr = 0, tot = 0;
        RH_HIP(hipMemGetInfo(&fr, &tot));
        budget = fr - std::min(fr, fr / 10 + (size_t(1) << 30));
    }
    int64_t res = n;
    if (o.resident_fraction >= 0.0) {
        res = std::min<int64_t>(n, int64_t(o.resident_fraction * double(n)));
    } else if (total > budget) {
        if (!o.allow_streaming)
            return fail(RH_ERR_NOMEM, "the " + std::to_string(total) + "-byte matrix exceeds the device budget of " +
                                          std::to_string(budget) + " bytes and streaming is not allowed");
        const size_t stage_bytes = size_t(2) << 26;  // two 64 MiB tiles
        res = (budget > stage_bytes) ? int64_t((budget - stage_bytes) / elem_rows) : 0;
        res = std::min(res, n);
    }
    if (res < n && !o.allow_streaming && o.resident_fraction < 0.0)
        return fail(RH_ERR_NOMEM, "matrix does not fit and streaming is not allowed");
    // the resident rows and every streamed tile start at a multiple of the
    // assembly's 64-row groups, so each block is integrated with the same
    // wave-uniform choices (FP32 or FP64 integrals) wherever its row lives
    constexpr int64_t kRowGroup = 64;
    if (res < n) res = res / kRowGroup * kRowGroup;
    q->res_elems = res;
    const size_t dev_bytes = size_t(res) * elem_rows;
    hipError_t e = hipMalloc(&q->d_Q, std::max<size_t>(dev_bytes, 16));
    // padded columns must read as zero in the matrix pass
    if (e == hipSuccess && dev_bytes) e = hipMemsetAsync(q->d_Q, 0, dev_bytes, m->stream);
    if (e == hipSuccess && res < n) {
        q->tile_elems = int64_t((size_t(1) << 26) / elem_rows) / kRowGroup * kRowGroup;
        q->tile_elems = std::max<int64_t>(kRowGroup, q->tile_elems);
        q->tile_elems = std::min(q->tile_elems, n - res);
        const size_t tb = size_t(q->tile_elems) * elem_rows;
        for (int b = 0; b < 2 && e == hipSuccess; b++) {
            e = hipMalloc(&q->stage[b], tb);
            if (e == hipSuccess) e = hipMemsetAsync(q->stage[b], 0, tb, m->stream);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&q->ev_ready[b], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&q->ev_free[b], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&q->copy, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipHostMalloc(&q->h_Q, size_t(n - res) * elem_rows, hipHostMallocDefault);
        // events start out signalled
        for (int b = 0; b < 2 && e == hipSuccess; b++) {
            e = hipEventRecord(q->ev_free[b], m->stream);
            if (e == hipSuccess) e = hipEventRecord(q->ev_ready[b], m->stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP,
                    std::string("matrix allocation (") + std::to_string(dev_bytes) + " device bytes, " +
                        std::to_string(size_t(n - res) * elem_rows) + " pinned host bytes): " + hipGetErrorString(e));
    return RH_OK;
}

// Assembles q; a 5-plane matrix that meets a traced block is stored again
// with 6 planes.
static rh_status assemble_into(rh_matrix* q)
{
    bool traced = false;
    rh_status st = assemble_rows(q, &traced);
    if (st != RH_OK || !traced) return st;
    release_storage(q);
    q->planes = 6;
    st = alloc_storage(q);
    if (st == RH_OK) st = assemble_rows(q, &traced);
    return st;
}

rh_status rh_matrix_assemble_ex(rh_model* m, rh_precision prec, const rh_matrix_opts* opts,
                                rh_matrix** out)
{
    if (!m || !out) return fail(RH_ERR_ARG, "null argument");
    *out = nullptr;
    if (prec != RH_FP64 && prec != RH_MIXED) return fail(RH_ERR_ARG, "unknown precision");
    rh_matrix_opts o{};
    o.resident_fraction = -1.0;
    if (opts) o = *opts;
    if (o.resident_fraction > 1.0) return fail(RH_ERR_ARG, "resident_fraction above 1");
    if (o.allow_streaming != 0 && o.allow_streaming != 1 && o.allow_streaming != RH_STREAM_BY_POLICY)
        return fail(RH_ERR_ARG, "allow_streaming must be 0, 1 or RH_STREAM_BY_POLICY");
    if (o.allow_streaming == RH_STREAM_BY_POLICY) o.allow_streaming = (g_fallback == RH_FALLBACK_BREAK) ? 0 : 1;
    if (o.fp64_integrals != 0 && o.fp64_integrals != 1) return fail(RH_ERR_ARG, "fp64_integrals must be 0 or 1");
    if (o.dense_rows != 0 && o.dense_rows != 1) return fail(RH_ERR_ARG, "dense_rows must be 0 or 1");
    RH_HIP(enter_device(m->device));
    const int64_t n = m->host.n;
    rh_matrix* q = new rh_matrix();
    q->m = m;
    q->device = m->device;
    q->prec = prec;
    q->opts = o;
    q->far32 = (prec == RH_MIXED) && o.fp64_integrals == 0;
    q->elem_bytes = (prec == RH_FP64) ? 8 : 4;
    q->n3 = 3 * n;
    // lab-frame blocks (no transform copies, no row transforms) are symmetric
    q->planes = (o.dense_rows == 0 && !m->host.sym && m->host.rowt_identity) ? 5 : 0;
    hipError_t e = dalloc(&q->d_qdiag, size_t(9 * n));
    if (e == hipSuccess) e = hipMalloc(&q->d_flag, sizeof(int32_t));
    if (e != hipSuccess) {
        free_matrix(q);
        return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP, hipGetErrorString(e));
    }
    rh_status st = alloc_storage(q);
    if (st == RH_OK) st = assemble_into(q);
    if (st != RH_OK) {
        free_matrix(q);
        return st;
    }
    g_backend_asm = q->streamed() ? 2 : 1;
    *out = q;
    return RH_OK;
}

rh_status rh_matrix_assemble(rh_model* m, rh_precision prec, rh_matrix** out)
{
    rh_matrix_opts o{};
    o.resident_fraction = -1.0;
    o.allow_streaming = RH_STREAM_BY_POLICY;
    return rh_matrix_assemble_ex(m, prec, &o, out);
}

rh_status rh_matrix_layout(const rh_matrix* q, int64_t* resident_rows, int64_t* streamed_rows)
{
    if (!q || !resident_rows || !streamed_rows) return fail(RH_ERR_ARG, "null argument");
    *resident_rows = 3 * q->res_elems;
    *streamed_rows = q->n3 - 3 * q->res_elems;
    return RH_OK;
}

rh_status rh_matrix_reassemble(rh_matrix* q)
{
    if (!q) return fail(RH_ERR_ARG, "null argument");
    return assemble_into(q);
}

rh_status rh_matrix_destroy(rh_matrix* q)
{
    free_matrix(q);
    return RH_OK;
}

rh_status rh_matrix_device_ptr(const rh_matrix* q, void** dptr, int32_t* elem_bytes)
{
    if (!q || !dptr || !elem_bytes) return fail(RH_ERR_ARG, "null argument");
    if (q->streamed()) return fail(RH_ERR_UNSUPPORTED, "the matrix is partly streamed from host memory");
    *dptr = q->d_Q;
    *elem_bytes = q->elem_bytes;
    return RH_OK;
}

rh_status rh_matrix_format(const rh_matrix* q, int32_t* planes, int64_t* ld)
{
    if (!q || !planes || !ld) return fail(RH_ERR_ARG, "null argument");
    *planes = q->planes;
    *ld = q->planes ? q->ldp : q->ld;
    return RH_OK;
}

rh_status rh_matrix_export(const rh_matrix* q, int64_t row0, int64_t nrows, double* out)
{
    if (!q || !out) return fail(RH_ERR_ARG, "null argument");
    if (row0 < 0 || nrows < 0 || row0 + nrows > q->n3) return fail(RH_ERR_ARG, "row range out of bounds");
    rh_model* m = q->m;
    RH_HIP(enter_device(m->device));
    const int64_t n3 = q->n3, n = m->host.n;
    // one element row's storage, read from the device or the pinned rows
    std::vector<char> buf(q->elem_stride());
    int64_t have = -1;
    auto load = [&](int64_t e) -> hipError_t {
        if (e == have) return hipSuccess;
        have = e;
        if (e < q->res_elems)
            return hipMemcpy(buf.data(), static_cast<const char*>(q->d_Q) + size_t(e) * q->elem_stride(),
                             buf.size(), hipMemcpyDeviceToHost);
        std::memcpy(buf.data(), static_cast<const char*>(q->h_Q) + size_t(e - q->res_elems) * q->elem_stride(),
                    buf.size());
        return hipSuccess;
    };
    auto val = [&](int64_t k) -> double {
        return (q->elem_bytes == 4) ? double(reinterpret_cast<const float*>(buf.data())[k])
                                    : reinterpret_cast<const double*>(buf.data())[k];
    };
    // plane of entry (a, b) of a symmetric block
    static const int kPlane[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int64_t r = row0; r < row0 + nrows; r++) {
        const int64_t e = r / 3, a = r % 3;
        RH_HIP(load(e));
        double* o = out + (r - row0) * n3;
        if (!q->planes) {
            for (int64_t c = 0; c < n3; c++) o[c] = val(a * q->ld + c);
            continue;
        }
        for (int64_t j = 0; j < n; j++)
            for (int b = 0; b < 3; b++) {
                const int k = kPlane[3 * a + b];
                o[3 * j + b] = (k < q->planes) ? val(k * q->ldp + j) : -(val(j) + val(3 * q->ldp + j));
            }
    }
    // diagonal blocks are kept in FP64 beside the matrix
    std::vector<double> qd(size_t(9 * n));
    RH_HIP(hipMemcpy(qd.data(), q->d_qdiag, qd.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t r = row0; r < row0 + nrows; r++) {
        const int64_t e = r / 3, a = r % 3;
        for (int b = 0; b < 3; b++) out[(r - row0) * n3 + 3 * e + b] = qd[size_t(9 * e + 3 * a + b)];
    }
    return RH_OK;
}

// ---------------------------------------------------------------- relaxation
namespace {

// Solves (G + lambda I) g = r for a small symmetric positive semi-definite G.
bool small_solve(int m, std::vector<double> G, std::vector<double> r, std::vector<double>& g)
{
    double tr = 0.0;
    for (int i = 0; i < m; i++) tr += G[i * m + i];
    const double lam = 1e-10 * (tr > 0 ? tr / m : 1.0);
    for (int i = 0; i < m; i++) G[i * m + i] += lam;
    // Cholesky
    std::vector<double> L(size_t(m * m), 0.0);
    for (int i = 0; i < m; i++)
        for (int j = 0; j <= i; j++) {
            double s = G[i * m + j];
            for (int k = 0; k < j; k++) s -= L[i * m + k] * L[j * m + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * m + i] = std::sqrt(s);
            } else {
                L[i * m + j] = s / L[j * m + j];
            }
        }
    std::vector<double> z(m);
    for (int i = 0; i < m; i++) {
        double s = r[i];
        for (int k = 0; k < i; k++) s -= L[i * m + k] * z[k];
        z[i] = s / L[i * m + i];
    }
    g.assign(m, 0.0);
    for (int i = m - 1; i >= 0; i--) {
        double s = z[i];
        for (int k = i + 1; k < m; k++) s -= L[k * m + i] * g[k];
        g[i] = s / L[i * m + i];
    }
    return true;
}

}  // namespace

static rh_status field_dev_impl(rh_model* m, const double* d_M, const double* d_pts, int64_t np,
                                double* d_B, double* d_H, bool fp32, const char* family = "field");

namespace {
// Anderson-accelerated block-Jacobi iterations from the iterate in w.x until
// misfit <= prec or max_iter; writes the final magnetization to d_M and the
// element fields to d_H.
// Transposed (adjoint) iteration of a linear-material system: the matrix pass
// becomes y = offdiag(Q)^T K^T lam and the update the transposed local solve.
struct AdjointPass {
    double* mu = nullptr;    // [n3] K^T lam
    double* tpart = nullptr; // [chunks * ld] transposed-pass partials
    int64_t chunks = 0;
};

// Device-driven form of the forward iteration below: the matrix pass, the
// relaxation step (whose last block runs the Anderson step) and the update
// all read the loop state on the device.  For a resident matrix the host enqueues iterations in
// batches without waiting for any of them, copies the state back after each
// batch and looks at it one batch late; iterations enqueued after the loop
// finished return at once.  The batch after the first few is sized from the
// residual's decay.  A streamed matrix (whose passes upload tiles from host
// memory) runs one iteration per batch and is looked at after each, so no
// pass is enqueued beyond the last; its iterates are the resident matrix's,
// bit for bit.
rh_status relax_run_dev(rh_matrix* q, const double* d_mats, const double* d_tab, const double* d_hext,
                        double* d_M, double* d_H, const rh_relax_opts* opts, RelaxWork& w, rh_relax_result* res,
                        bool keep_if_converged)
{
    const bool lagged = !q->streamed();
    rh_model* m = q->m;
    hipStream_t s = m->stream;
    const int64_t n = m->host.n, n3 = q->n3;
    const int max_iter = std::max(1, int(opts->max_iter));
    rh::RelaxDevState& h = w.hst[2];
    std::memset(&h, 0, sizeof(h));
    h.max_iter = max_iter;
    h.keep_if_converged = keep_if_converged ? 1 : 0;
    h.prec = opts->prec;
    h.beta = (opts->omega0 > 0.0) ? opts->omega0 : 0.5;
    h.best = INFINITY;
    h.n = double(n);
    RH_HIP(hipMemcpyAsync(w.st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    rh::HistBufs hb{};
    hb.Fp = w.Fp;
    hb.xp = w.xp;
    for (int k = 0; k < kHist; k++) {
        hb.dF[k] = w.dF[k];
        hb.dX[k] = w.dX[k];
    }
    Timer tm("matvec", s);
    int enq = 0, batch = 0, cur = 0, want = lagged ? 4 : 1;
    int it_prev = 0;
    double mis_prev = 0.0;
    for (;;) {
        const int cnt = std::min(want, max_iter - enq);
        for (int j = 0; j < cnt; j++) {
            tm.begin();
            if (lagged)
                RH_HIP(matvec_rows(q, q->d_Q, 0, n, w.x, w.y, s, &w.st->done));
            else
                RH_HIP(matrix_pass(q, w.x, w.y, s));
            tm.end();
            RH_HIP(rh::launch_relax_step_dev(n, d_mats, d_tab, d_hext, q->d_qdiag, w.y, w.x, w.F, d_H, w.part, hb,
                                             w.st, s));
            RH_HIP(rh::launch_update_dev(n3, w.x, w.F, hb, d_M, w.st, s));
        }
        enq += cnt;
        RH_HIP(hipMemcpyAsync(&w.hst[cur], w.st, sizeof(rh::RelaxDevState), hipMemcpyDeviceToHost, s));
        RH_HIP(hipEventRecord(w.ev[cur], s));
        bool stop = enq >= max_iter;
        if (!lagged) {
            RH_HIP(hipEventSynchronize(w.ev[cur]));
            if (w.hst[cur].done) stop = true;
        } else if (batch > 0) {
            RH_HIP(hipEventSynchronize(w.ev[cur ^ 1]));
            const rh::RelaxDevState& p = w.hst[cur ^ 1];
            if (p.done) {
                stop = true;
            } else if (p.it > it_prev && mis_prev > 0.0 && p.misfit < mis_prev && p.misfit > p.prec) {
                // iterations left at the observed rate, less those in flight
                const double rate = std::log(p.misfit / mis_prev) / double(p.it - it_prev);
                const double left = std::log(p.prec / p.misfit) / rate;
                const int inflight = enq - p.it;
                want = int(std::min(32.0, std::max(1.0, std::ceil(left) - inflight)));
            } else {
                want = 4;
            }
            it_prev = p.it;
            mis_prev = p.misfit;
        }
        if (stop) break;
        batch++;
        cur ^= 1;
    }
    RH_HIP(hipStreamSynchronize(s));
    const rh::RelaxDevState& f = w.hst[cur];
    tm.keep(size_t(std::max(0, f.it)));
    if (f.fail) return fail(RH_ERR_NOTCONV, "relaxation produced a non-finite residual");
    res->misfit = f.misfit;
    res->max_m = f.maxM;
    res->max_h = f.maxH;
    res->iters += f.it;
    res->matvecs += f.it;
    res->converged = f.misfit <= opts->prec ? 1 : 0;
    return RH_OK;
}

rh_status relax_run(rh_matrix* q, const double* d_mats, const double* d_tab, const double* d_hext,
                    double* d_M, double* d_H, const rh_relax_opts* opts, RelaxWork& w, double* hbuf,
                    int64_t nb, int64_t db, rh_relax_result* res,
                    const AdjointPass* adj = nullptr, bool keep_if_converged = false)
{
    if (!adj && !std::getenv("RADHIP_RELAX_HOST_LOOP"))
        return relax_run_dev(q, d_mats, d_tab, d_hext, d_M, d_H, opts, w, res, keep_if_converged);
    rh_model* m = q->m;
    hipStream_t s = m->stream;
    const int64_t n = m->host.n, n3 = q->n3;
    const double prec = opts->prec;
    const int max_iter = std::max(1, int(opts->max_iter));
    double beta = (opts->omega0 > 0.0) ? opts->omega0 : 0.5;
    std::vector<double> G(size_t(kHist * kHist), 0.0);
    int hist = 0, head = 0;  // ring of the last `hist` differences, newest at head-1
    double best = INFINITY;
    Timer tm("matvec", s);
    int it = 0;
    double misfit = INFINITY, maxM = 0.0, maxH = 0.0;
    while (true) {
        it++;
        tm.begin();
        if (adj) {
            RH_HIP(rh::launch_adj_scale(n, d_mats, nullptr, w.x, adj->mu, s));
            if (q->planes)
                RH_HIP(rh::launch_matvec_t_sym(q->d_Q, q->elem_bytes, q->planes, n, q->ldp, adj->mu, adj->tpart,
                                               adj->chunks, w.y, s));
            else
                RH_HIP(rh::launch_matvec_t(q->d_Q, q->elem_bytes, n3, q->ld, adj->mu, adj->tpart,
                                           adj->chunks, w.y, s));
        } else {
            RH_HIP(matrix_pass(q, w.x, w.y, s));
        }
        tm.end();
        // dv: the dot products of the newest difference with every stored
        // one, then of every stored difference with F (oldest first)
        int slot[kHist];
        std::vector<double> dv;
        double f2 = 0.0;
        maxM = 1e-17;
        maxH = 1e-17;
        if (adj) {
            RH_HIP(rh::launch_adj_update(n, d_mats, d_hext, q->d_qdiag, w.y, w.x, w.F, w.part, s));
            // tentatively extend the history with the newest differences
            if (it > 1) {
                const double* v1[2] = {w.F, w.Fp};
                const double c1[2] = {1.0, -1.0};
                RH_HIP(rh::launch_lincomb(2, v1, c1, n3, w.dF[head], s));
                const double* v2[2] = {w.x, w.xp};
                RH_HIP(rh::launch_lincomb(2, v2, c1, n3, w.dX[head], s));
                head = (head + 1) % kHist;
                hist = std::min(hist + 1, kHist);
            }
            RH_HIP(hipMemcpyAsync(w.Fp, w.F, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
            RH_HIP(hipMemcpyAsync(w.xp, w.x, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
            for (int k = 0; k < hist; k++) slot[k] = (head - hist + k + kHist) % kHist;
            int nd = 0;
            if (hist > 0) {
                const double* a[rh::kDotMax];
                const double* b[rh::kDotMax];
                const int newest = slot[hist - 1];
                for (int k = 0; k < hist; k++) {
                    a[nd] = w.dF[newest];
                    b[nd] = w.dF[slot[k]];
                    nd++;
                }
                for (int k = 0; k < hist; k++) {
                    a[nd] = w.dF[slot[k]];
                    b[nd] = w.F;
                    nd++;
                }
                RH_HIP(rh::launch_dots(nd, a, b, n3, w.dots, s));
                RH_HIP(hipMemcpyAsync(hbuf + 3 * nb, w.dots, size_t(nd * db) * sizeof(double),
                                      hipMemcpyDeviceToHost, s));
            }
            RH_HIP(hipMemcpyAsync(hbuf, w.part, size_t(3 * nb) * sizeof(double), hipMemcpyDeviceToHost, s));
            RH_HIP(hipStreamSynchronize(s));
            for (int64_t bb = 0; bb < nb; bb++) {
                f2 += hbuf[3 * bb];
                maxM = std::max(maxM, hbuf[3 * bb + 1]);
                maxH = std::max(maxH, hbuf[3 * bb + 2]);
            }
            dv.assign(size_t(nd), 0.0);
            for (int d = 0; d < nd; d++)
                for (int64_t bb = 0; bb < db; bb++) dv[size_t(d)] += hbuf[3 * nb + d * db + bb];
        } else {
            // the update, the history differences, the copies and the dot
            // products in one pass
            rh::HistPack hp;
            hp.Fp = w.Fp;
            hp.xp = w.xp;
            if (it > 1) {
                hp.dFw = w.dF[head];
                hp.dXw = w.dX[head];
                head = (head + 1) % kHist;
                hist = std::min(hist + 1, kHist);
            }
            for (int k = 0; k < hist; k++) {
                slot[k] = (head - hist + k + kHist) % kHist;
                hp.dF[k] = w.dF[slot[k]];
            }
            hp.nh = hist;
            RH_HIP(rh::launch_relax_step(n, d_mats, d_tab, d_hext, q->d_qdiag, w.y, w.x, w.F, d_H, w.part, hp,
                                         s));
            RH_HIP(hipMemcpyAsync(hbuf, w.part, size_t(rh::kRelaxPart * nb) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            RH_HIP(hipStreamSynchronize(s));
            dv.assign(size_t(2 * hist), 0.0);
            for (int64_t bb = 0; bb < nb; bb++) {
                const double* o = hbuf + rh::kRelaxPart * bb;
                f2 += o[0];
                maxM = std::max(maxM, o[1]);
                maxH = std::max(maxH, o[2]);
                for (int k = 0; k < hist; k++) {
                    dv[size_t(k)] += o[3 + k];
                    dv[size_t(hist + k)] += o[3 + rh::kHistMax + k];
                }
            }
        }
        misfit = std::sqrt(f2 / double(n));
        if (!std::isfinite(misfit)) return fail(RH_ERR_NOTCONV, "relaxation produced a non-finite residual");
        if (misfit <= prec && it == 1 && keep_if_converged) break;  // d_M already holds x
        if (misfit <= prec || it >= max_iter) {
            // final magnetization: the last proposal x + F
            const double* vv[2] = {w.x, w.F};
            const double cc[2] = {1.0, 1.0};
            RH_HIP(rh::launch_lincomb(2, vv, cc, n3, d_M, s));
            break;
        }
        // safeguard: restart the history when the residual grows a lot
        bool plain = (hist == 0);
        if (misfit > 1e4 * best) {
            hist = 0;
            head = 0;
            beta = std::max(0.05, 0.5 * beta);
            plain = true;
        }
        best = std::min(best, misfit);
        std::vector<double> gam;
        if (!plain) {
            const int newest = slot[hist - 1];
            for (int k = 0; k < hist; k++) {
                G[size_t(newest * kHist + slot[k])] = dv[size_t(k)];
                G[size_t(slot[k] * kHist + newest)] = dv[size_t(k)];
            }
            std::vector<double> Gm(static_cast<size_t>(hist) * hist);
            std::vector<double> r(static_cast<size_t>(hist));
            for (int i = 0; i < hist; i++) {
                r[size_t(i)] = dv[size_t(hist + i)];
                for (int j = 0; j < hist; j++) Gm[size_t(i * hist + j)] = G[size_t(slot[i] * kHist + slot[j])];
            }
            if (!small_solve(hist, Gm, r, gam)) {
                hist = 0;
                head = 0;
                plain = true;
            }
        }
        if (plain) {
            const double* v[2] = {w.x, w.F};
            const double c[2] = {1.0, beta};
            RH_HIP(rh::launch_lincomb(2, v, c, n3, w.x, s));
            continue;
        }
        // x <- x + beta F - sum_k gam_k (dX_k + beta dF_k)
        const double* v[rh::kCombMax];
        double c[rh::kCombMax];
        int nv = 0;
        v[nv] = w.x;
        c[nv++] = 1.0;
        v[nv] = w.F;
        c[nv++] = beta;
        for (int k = 0; k < hist; k++) {
            v[nv] = w.dX[slot[k]];
            c[nv++] = -gam[size_t(k)];
            v[nv] = w.dF[slot[k]];
            c[nv++] = -beta * gam[size_t(k)];
        }
        RH_HIP(rh::launch_lincomb(nv, v, c, n3, w.x, s));
    }
    RH_HIP(hipStreamSynchronize(s));
    res->misfit = misfit;
    res->max_m = maxM;
    res->max_h = maxH;
    res->iters += it;
    res->matvecs += it;
    res->converged = misfit <= prec ? 1 : 0;
    return RH_OK;
}
}  // namespace

namespace {
// ---------------------------------------------------------------- Newton-Krylov
// Method 11 (the root finding of radGPU_RelaxNK, cpp/src/core/radgpurlx.cu:1351-2285,
// with a driver of our own): Newton on F(M) = M_mat(H_ext + Q M) - M with the
// analytic Jacobian J = D Q - I; each step solves (I - D Q) delta = F by
// restarted GMRES right-preconditioned with the element blocks
// (I - D_i Q_ii)^{-1} (classical Gram-Schmidt with one re-orthogonalisation,
// Givens rotations on the host), forcing terms after Eisenstat and Walker
// (choice 2), and a backtracking line search on |F|.  A step whose line search
// fails falls back to Anderson iterations.
constexpr int kKrylovDefault = 60;
constexpr int kKrylovPerStep = 400;

struct NKWork {
    double* Mc = nullptr;  // [ld] current magnetization (matrix-pass input)
    double* Mt = nullptr;  // [ld] trial magnetization
    double *F = nullptr, *Ft = nullptr;  // [n3]
    double *D = nullptr, *Dt = nullptr;  // [9n]
    double *P = nullptr, *Pt = nullptr;  // [9n]
    double *H = nullptr, *Ht = nullptr;  // [n3]
    double* y = nullptr;   // [n3] matrix-pass output
    double* z = nullptr;   // [ld] preconditioned vector (matrix-pass input)
    double* w = nullptr;   // [n3]
    double* u = nullptr;   // [n3] GMRES solution in the preconditioned space
    double* V = nullptr;   // [(m + 1) * ldv] Krylov basis
    double* part = nullptr;  // [3 nb]
    double* dots = nullptr;  // [(m + 2) * db]
    double* hpin = nullptr;  // pinned [max(3 nb, (m + 2) db)]
    int64_t ldv = 0;
    int m = 0;
    ~NKWork()
    {
        for (double* p : {Mc, Mt, F, Ft, D, Dt, P, Pt, H, Ht, y, z, w, u, V, part, dots}) dfree(p);
        if (hpin) (void)hipHostFree(hpin);
    }
    hipError_t alloc(int64_t n, int64_t ld, int krylov_m, int64_t nb, int64_t db, hipStream_t s)
    {
        const int64_t n3 = 3 * n;
        m = krylov_m;
        ldv = (n3 + 63) / 64 * 64;
        hipError_t e = dalloc(&Mc, size_t(ld));
        if (e == hipSuccess) e = dalloc(&Mt, size_t(ld));
        if (e == hipSuccess) e = dalloc(&z, size_t(ld));
        for (double** p : {&F, &Ft, &H, &Ht, &y, &w, &u})
            if (e == hipSuccess) e = dalloc(p, size_t(n3));
        for (double** p : {&D, &Dt, &P, &Pt})
            if (e == hipSuccess) e = dalloc(p, size_t(9 * n));
        if (e == hipSuccess) e = dalloc(&V, size_t((m + 1) * ldv));
        if (e == hipSuccess) e = dalloc(&part, size_t(3 * nb));
        if (e == hipSuccess) e = dalloc(&dots, size_t((m + 2) * db));
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&hpin),
                              size_t(std::max<int64_t>(3 * nb, (m + 2) * db)) * sizeof(double));
        // padded tails of the matrix-pass inputs read as zero
        for (double* p : {Mc, Mt, z})
            if (e == hipSuccess) e = hipMemsetAsync(p, 0, size_t(ld) * sizeof(double), s);
        return e;
    }
};

rh_status nk_run(rh_matrix* q, const double* d_mats, const double* d_tab, const double* d_hext, double* d_M,
                 double* d_H, const rh_relax_opts* opts, NKWork& k, RelaxWork& rw, double* hbuf, int64_t nb,
                 int64_t db, rh_relax_result* res)
{
    rh_model* m = q->m;
    hipStream_t s = m->stream;
    const int64_t n = m->host.n, n3 = q->n3;
    const double prec = opts->prec;
    const int max_mv = std::max(1, int(opts->max_iter));
    const double prec2 = prec * std::sqrt(double(n));  // misfit prec as a 2-norm of F
    Timer tm("matvec", s);
    int mv = 0, newton = 0, krylov = 0;
    double maxM = 0.0, maxH = 0.0;
    // F, D, P, H at magnetization M; returns |F|^2
    auto evaluate = [&](double* M, double* F, double* D, double* P, double* H, double& f2) -> rh_status {
        tm.begin();
        RH_HIP(matrix_pass(q, M, k.y, s));
        tm.end();
        RH_HIP(rh::launch_nk_material(n, d_mats, d_tab, d_hext, q->d_qdiag, k.y, M, F, D, P, H, k.part, s));
        RH_HIP(hipMemcpyAsync(k.hpin, k.part, size_t(3 * nb) * sizeof(double), hipMemcpyDeviceToHost, s));
        RH_HIP(hipStreamSynchronize(s));
        mv++;
        f2 = 0.0;
        maxM = 1e-17;
        maxH = 1e-17;
        for (int64_t b = 0; b < nb; b++) {
            f2 += k.hpin[3 * b];
            maxM = std::max(maxM, k.hpin[3 * b + 1]);
            maxH = std::max(maxH, k.hpin[3 * b + 2]);
        }
        if (!std::isfinite(f2)) return fail(RH_ERR_NOTCONV, "Newton-Krylov produced a non-finite residual");
        return RH_OK;
    };
    // dot products of w with V_0..V_{nv-1} and with itself
    auto mdot = [&](int nv, const double* w, std::vector<double>& out) -> rh_status {
        RH_HIP(rh::launch_mdot(nv, k.V, k.ldv, w, n3, k.dots, s));
        RH_HIP(hipMemcpyAsync(k.hpin, k.dots, size_t((nv + 1) * db) * sizeof(double), hipMemcpyDeviceToHost, s));
        RH_HIP(hipStreamSynchronize(s));
        out.assign(size_t(nv + 1), 0.0);
        for (int j = 0; j <= nv; j++)
            for (int64_t b = 0; b < db; b++) out[size_t(j)] += k.hpin[j * db + b];
        return RH_OK;
    };
    // w = (I - D Q) P v
    auto apply = [&](const double* v, double* w) -> rh_status {
        RH_HIP(rh::launch_nk_prec(n, k.P, v, k.z, s));
        tm.begin();
        RH_HIP(matrix_pass(q, k.z, k.y, s));
        tm.end();
        RH_HIP(rh::launch_nk_apply(n, k.D, q->d_qdiag, k.z, k.y, w, s));
        mv++;
        return RH_OK;
    };
    rh_status st;
    // Newton steps from a far start need heavy damping on stiff (high-chi)
    // models, so Anderson iterations run first until their misfit is within
    // 100 x prec (at most RADHIP_NK_PRESMOOTH iterations, default 200);
    // Newton then converges in a few steps.  RADHIP_NK_DEBUG=1: one line
    // per Newton step on stderr.
    const char* env_ps = std::getenv("RADHIP_NK_PRESMOOTH");
    const int presmooth = env_ps ? std::max(0, std::atoi(env_ps)) : 200;
    const bool debug = std::getenv("RADHIP_NK_DEBUG") != nullptr;
    if (presmooth > 0) {
        rh_relax_opts o2 = *opts;
        o2.prec = 100.0 * prec;
        o2.max_iter = std::min(presmooth, max_mv);
        rh_relax_result r2{};
        if ((st = relax_run(q, d_mats, d_tab, d_hext, d_M, d_H, &o2, rw, hbuf, nb, db, &r2)) != RH_OK) return st;
        mv += r2.matvecs;
        RH_HIP(hipMemcpyAsync(k.Mc, d_M, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        RH_HIP(hipMemcpyAsync(k.Mc, rw.x, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    double f2 = 0.0;
    if ((st = evaluate(k.Mc, k.F, k.D, k.P, k.H, f2)) != RH_OK) return st;
    double fn = std::sqrt(f2), fn_prev = 0.0, eta = 0.5;
    const int mk = k.m;
    const size_t um = static_cast<size_t>(mk);
    std::vector<double> Hs((um + 1) * um), cs(um), sn(um), g(um + 1), yv(um), h1, h2, coef(um + 1);
    while (fn > prec2 && mv < max_mv) {
        newton++;
        // forcing term (Eisenstat-Walker choice 2, gamma 0.9, alpha 2, with safeguards)
        if (newton > 1) {
            const double r = fn / fn_prev;
            double e2 = 0.9 * r * r;
            const double sg = 0.9 * eta * eta;
            if (sg > 0.1) e2 = std::max(e2, sg);
            eta = e2;
        }
        eta = std::min(eta, 0.5);
        eta = std::max(eta, std::min(0.5, 0.5 * prec2 / fn));
        const double tol = eta * fn;
        // restarted GMRES on (I - D Q) P u = F, u0 = 0
        RH_HIP(hipMemsetAsync(k.u, 0, size_t(n3) * sizeof(double), s));
        const double* r0 = k.F;
        double beta = fn, resid = fn;
        int its = 0;
        while (true) {
            const double one[1] = {0.0};
            RH_HIP(rh::launch_vcomb(0, k.V, k.ldv, one, r0, 1.0 / beta, n3, k.V, s));
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int j = 0;
            bool stop = false;
            for (; j < mk; j++) {
                if ((st = apply(k.V + j * k.ldv, k.w)) != RH_OK) return st;
                if ((st = mdot(j + 1, k.w, h1)) != RH_OK) return st;
                for (int i = 0; i <= j; i++) coef[size_t(i)] = -h1[size_t(i)];
                RH_HIP(rh::launch_vcomb(j + 1, k.V, k.ldv, coef.data(), k.w, 1.0, n3, k.w, s));
                if ((st = mdot(j + 1, k.w, h2)) != RH_OK) return st;
                double hh2 = 0.0;
                for (int i = 0; i <= j; i++) hh2 += h2[size_t(i)] * h2[size_t(i)];
                const double bnext = std::sqrt(std::max(h2[size_t(j + 1)] - hh2, 0.0));
                double* col = Hs.data() + size_t(j) * (mk + 1);
                for (int i = 0; i <= j; i++) col[i] = h1[size_t(i)] + h2[size_t(i)];
                const bool breakdown = !(bnext > 1e-14 * fn);
                if (!breakdown && j + 1 <= mk) {
                    for (int i = 0; i <= j; i++) coef[size_t(i)] = -h2[size_t(i)] / bnext;
                    RH_HIP(rh::launch_vcomb(j + 1, k.V, k.ldv, coef.data(), k.w, 1.0 / bnext, n3,
                                            k.V + (j + 1) * k.ldv, s));
                }
                // Givens rotations
                for (int i = 0; i < j; i++) {
                    const double a = col[i], b = col[i + 1];
                    col[i] = cs[size_t(i)] * a + sn[size_t(i)] * b;
                    col[i + 1] = -sn[size_t(i)] * a + cs[size_t(i)] * b;
                }
                const double t = std::hypot(col[j], bnext);
                cs[size_t(j)] = (t > 0.0) ? col[j] / t : 1.0;
                sn[size_t(j)] = (t > 0.0) ? bnext / t : 0.0;
                col[j] = t;
                g[size_t(j + 1)] = -sn[size_t(j)] * g[size_t(j)];
                g[size_t(j)] = cs[size_t(j)] * g[size_t(j)];
                resid = std::fabs(g[size_t(j + 1)]);
                its++;
                if (resid <= tol || breakdown || its >= kKrylovPerStep || mv >= max_mv) {
                    stop = true;
                    j++;
                    break;
                }
            }
            // y = R^{-1} g, u += V y
            for (int i = j - 1; i >= 0; i--) {
                double v = g[size_t(i)];
                for (int c = i + 1; c < j; c++) v -= Hs[size_t(c) * (mk + 1) + i] * yv[size_t(c)];
                const double rii = Hs[size_t(i) * (mk + 1) + i];
                yv[size_t(i)] = (rii != 0.0) ? v / rii : 0.0;
            }
            RH_HIP(rh::launch_vcomb(j, k.V, k.ldv, yv.data(), k.u, 1.0, n3, k.u, s));
            if (stop) break;
            // restart from the true residual F - (I - D Q) P u
            if ((st = apply(k.u, k.w)) != RH_OK) return st;
            const double* vv[2] = {k.F, k.w};
            const double cc[2] = {1.0, -1.0};
            RH_HIP(rh::launch_lincomb(2, vv, cc, n3, k.w, s));
            std::vector<double> rr;
            if ((st = mdot(0, k.w, rr)) != RH_OK) return st;
            beta = std::sqrt(rr[0]);
            resid = beta;
            if (beta <= tol || mv >= max_mv) break;
            r0 = k.w;
        }
        krylov += its;
        // delta = P u; backtracking line search on |F|
        RH_HIP(rh::launch_nk_prec(n, k.P, k.u, k.w, s));
        double alpha = 1.0, ft2 = 0.0;
        bool accepted = false;
        for (int ls = 0; ls < 12 && mv < max_mv; ls++) {
            const double* vv[2] = {k.Mc, k.w};
            const double cc[2] = {1.0, alpha};
            RH_HIP(rh::launch_lincomb(2, vv, cc, n3, k.Mt, s));
            if ((st = evaluate(k.Mt, k.Ft, k.Dt, k.Pt, k.Ht, ft2)) != RH_OK) return st;
            if (std::sqrt(ft2) <= (1.0 - 1e-4 * alpha * (1.0 - eta)) * fn) {
                accepted = true;
                break;
            }
            alpha *= 0.5;
        }
        if (debug)
            std::fprintf(stderr, "nk step %d: |F| %.3e eta %.2e gmres %d (resid %.2e) alpha %g %s, matvecs %d\n",
                         newton, fn, eta, its, resid, alpha, accepted ? "accepted" : "rejected", mv);
        fn_prev = fn;
        if (accepted) {
            std::swap(k.Mc, k.Mt);
            std::swap(k.F, k.Ft);
            std::swap(k.D, k.Dt);
            std::swap(k.P, k.Pt);
            std::swap(k.H, k.Ht);
            fn = std::sqrt(ft2);
            continue;
        }
        if (mv >= max_mv) break;
        // no descent along the Newton direction: Anderson iterations from the
        // current magnetization, then Newton again
        rh_relax_opts o2 = *opts;
        o2.max_iter = std::min(50, max_mv - mv);
        rh_relax_result r2{};
        RH_HIP(hipMemcpyAsync(rw.x, k.Mc, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
        if ((st = relax_run(q, d_mats, d_tab, d_hext, d_M, d_H, &o2, rw, hbuf, nb, db, &r2)) != RH_OK) return st;
        mv += r2.matvecs;
        RH_HIP(hipMemcpyAsync(k.Mc, d_M, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
        if ((st = evaluate(k.Mc, k.F, k.D, k.P, k.H, f2)) != RH_OK) return st;
        fn = std::sqrt(f2);
        eta = 0.5;
    }
    RH_HIP(hipMemcpyAsync(d_M, k.Mc, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    RH_HIP(hipMemcpyAsync(d_H, k.H, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    // the next call's start (refinement passes resume from here)
    RH_HIP(hipMemcpyAsync(rw.x, k.Mc, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    RH_HIP(hipStreamSynchronize(s));
    const double misfit = fn / std::sqrt(double(n));
    res->misfit = misfit;
    res->max_m = maxM;
    res->max_h = maxH;
    res->iters += mv;
    res->matvecs += mv;
    res->newton += newton;
    res->krylov += krylov;
    res->converged = misfit <= prec ? 1 : 0;
    return RH_OK;
}
}  // namespace

rh_status rh_relax_dev(rh_matrix* q, const double* d_mats, const double* d_hext, double* d_M,
                       double* d_H, const rh_relax_opts* opts, rh_relax_result* res)
{
    return rh_relax_tab_dev(q, d_mats, nullptr, 0, d_hext, d_M, d_H, opts, res);
}

rh_status rh_relax_tab_dev(rh_matrix* q, const double* d_mats, const double* d_tab, int64_t ntab,
                           const double* d_hext, double* d_M, double* d_H, const rh_relax_opts* opts,
                           rh_relax_result* res)
{
    if (!q || !d_mats || !d_hext || !d_M || !d_H || !opts || !res) return fail(RH_ERR_ARG, "null argument");
    if (ntab < 0 || (ntab > 0 && !d_tab)) return fail(RH_ERR_ARG, "bad table");
    if (opts->method != 9 && opts->method != 0 && opts->method != 11)
        return fail(RH_ERR_UNSUPPORTED, "relaxation method must be 9 or 11");
    if (opts->krylov_m < 0 || opts->krylov_m > rh::kCoefMax - 1)
        return fail(RH_ERR_ARG, "krylov_m must be 0 to " + std::to_string(rh::kCoefMax - 1));
    rh_model* m = q->m;
    RH_HIP(enter_device(m->device));
    hipStream_t s = m->stream;
    const int64_t n = m->host.n, n3 = q->n3;
    std::memset(res, 0, sizeof(*res));
    if (n == 0) {
        res->converged = 1;
        return RH_OK;
    }
    {
        // material kinds and table ranges (tabulated laws index d_tab)
        std::vector<double> recs(size_t(16 * n), 0.0);
        RH_HIP(hipMemcpyAsync(recs.data(), d_mats, recs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        RH_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; i++) {
            const double* r = recs.data() + 16 * i;
            const int kind = int(r[0]);
            if (kind < 0 || kind > 3 || double(kind) != r[0])
                return fail(RH_ERR_UNSUPPORTED, "unknown material kind in record " + std::to_string(i));
            if (kind == 2 && !(r[1] >= 1 && r[1] <= 3)) return fail(RH_ERR_ARG, "tanh law needs 1 to 3 terms");
            if (kind == 3 && !(r[1] >= 0 && r[2] >= 2 && 3 * (r[1] + r[2]) <= double(ntab)))
                return fail(RH_ERR_ARG, "tabulated law of record " + std::to_string(i) + " outside the table");
        }
    }
    if (!q->rw) {
        RelaxWork* rw = new RelaxWork();
        const hipError_t e = rw->alloc(n, n3, q->ld);
        if (e != hipSuccess) {
            delete rw;
            return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP,
                        std::string("relaxation work: ") + hipGetErrorString(e));
        }
        q->rw = rw;
    }
    RelaxWork& w = *q->rw;
    const int64_t nb = w.nb;
    const int64_t db = w.db;
    RH_HIP(hipMemsetAsync(w.x, 0, size_t(q->ld) * sizeof(double), s));
    // d_H starts the per-element nonlinear solves (previous local field); a
    // zero start keeps repeated solves bit-identical
    RH_HIP(hipMemsetAsync(d_H, 0, size_t(n3) * sizeof(double), s));
    if (opts->keep_m) {
        RH_HIP(hipMemcpyAsync(w.x, d_M, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        // start from the remanent magnetization (radintrc.h:385-395)
        for (int k = 0; k < 3; k++)
            RH_HIP(hipMemcpy2DAsync(w.x + k, 3 * sizeof(double), d_mats + 13 + k, 16 * sizeof(double),
                                    sizeof(double), size_t(n), hipMemcpyDeviceToDevice, s));
    }
    // pinned read-back of the host-driven loops: residual partials, then dot partials
    double* hbuf = w.hbuf;
    const bool nk_method = (opts->method == 11);
    NKWork nk;
    if (nk_method) {
        const int km = opts->krylov_m > 0 ? std::max(2, int(opts->krylov_m)) : kKrylovDefault;
        hipError_t e = nk.alloc(n, q->ld, km, nb, db, s);
        if (e != hipSuccess) return fail(RH_ERR_NOMEM, std::string("Newton-Krylov work: ") + hipGetErrorString(e));
    }
    auto run = [&](const double* hx, const rh_relax_opts* o, bool keep) -> rh_status {
        if (nk_method) return nk_run(q, d_mats, d_tab, hx, d_M, d_H, o, nk, w, hbuf, nb, db, res);
        return relax_run(q, d_mats, d_tab, hx, d_M, d_H, o, w, hbuf, nb, db, res, nullptr, keep);
    };
    rh_status st = run(d_hext, opts, false);
    if (st != RH_OK) return st;
    // FP64 refinement of a solve on the FP32-stored matrix: the defect between
    // the FP64 field at the collocation points and the stored-matrix field
    // becomes a correction to the external field; the iteration resumes from
    // the current magnetization unless that already meets prec on the
    // corrected system.
    const int passes = (q->prec == RH_MIXED) ? (opts->refine == 0 ? 1 : std::max(0, int(opts->refine))) : 0;
    if (passes == 0) return RH_OK;
    double* d_H64 = w.H64;
    double* d_hx = w.hx;
    const rh_relax_opts o2 = *opts;
    for (int pss = 0; pss < passes; pss++) {
        st = field_dev_impl(m, d_M, m->d_colloc, n, nullptr, d_H64, false, "refine");
        if (st != RH_OK) return st;
        RH_HIP(hipMemcpyAsync(w.x, d_M, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
        RH_HIP(matrix_pass(q, w.x, w.y, s));
        RH_HIP(rh::launch_refine_defect(n, d_H64, m->host.rowt_identity ? nullptr : m->d_rowt, w.y,
                                        q->d_qdiag, d_M, d_hext, d_hx, s));
        res->matvecs += 1;
        // a magnetization that already meets prec on the corrected system is kept
        st = run(d_hx, &o2, true);
        if (st != RH_OK) return st;
        res->refined += 1;
    }
    return RH_OK;
}

rh_status rh_relax(rh_matrix* q, const double* mats, const double* hext, double* M_inout,
                   double* H_out, const rh_relax_opts* opts, rh_relax_result* res)
{
    return rh_relax_tab(q, mats, nullptr, 0, hext, M_inout, H_out, opts, res);
}

rh_status rh_relax_tab(rh_matrix* q, const double* mats, const double* tab, int64_t ntab,
                       const double* hext, double* M_inout, double* H_out, const rh_relax_opts* opts,
                       rh_relax_result* res)
{
    if (!q || !opts || !res) return fail(RH_ERR_ARG, "null argument");
    if (q->m->host.n == 0) {
        std::memset(res, 0, sizeof(*res));
        res->converged = 1;
        return RH_OK;
    }
    if (!mats || !hext || !M_inout || !H_out) return fail(RH_ERR_ARG, "null argument");
    if (ntab < 0 || (ntab > 0 && !tab)) return fail(RH_ERR_ARG, "bad table");
    rh_model* m = q->m;
    RH_HIP(enter_device(m->device));
    const int64_t n = m->host.n;
    double *dm = nullptr, *dh = nullptr, *dM = nullptr, *dH = nullptr, *dt = nullptr;
    hipStream_t s = m->stream;
    auto cleanup = [&]() {
        dfree(dm);
        dfree(dh);
        dfree(dM);
        dfree(dH);
        dfree(dt);
    };
    hipError_t e = dalloc(&dm, size_t(16 * n));
    if (e == hipSuccess) e = dalloc(&dh, size_t(3 * n));
    if (e == hipSuccess) e = dalloc(&dM, size_t(3 * n));
    if (e == hipSuccess) e = dalloc(&dH, size_t(3 * n));
    if (e == hipSuccess && ntab > 0) e = dalloc(&dt, size_t(ntab));
    if (e == hipSuccess && ntab > 0)
        e = hipMemcpyAsync(dt, tab, size_t(ntab) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dm, mats, size_t(16 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dh, hext, size_t(3 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dM, M_inout, size_t(3 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        cleanup();
        return fail(RH_ERR_HIP, std::string("relax upload: ") + hipGetErrorString(e));
    }
    rh_status st = rh_relax_tab_dev(q, dm, dt, ntab, dh, dM, dH, opts, res);
    if (st == RH_OK) {
        e = hipMemcpyAsync(M_inout, dM, size_t(3 * n) * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(H_out, dH, size_t(3 * n) * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(RH_ERR_HIP, std::string("relax download: ") + hipGetErrorString(e));
    }
    cleanup();
    return st;
}

// ---------------------------------------------------------------- field
static rh_status field_dev_impl(rh_model* m, const double* d_M, const double* d_pts, int64_t np,
                                double* d_B, double* d_H, bool fp32, const char* family)
{
    const int64_t slices = rh::field_slices(m->dev, np);
    double* work = nullptr;
    if (slices > 1) {
        RH_HIP(grow(m, size_t(6 * np * slices)));
        work = m->d_work;
    }
    Timer t(family, m->stream);
    t.begin();
    RH_HIP(rh::launch_field(m->dev, d_M, d_pts, np, d_B, d_H, work, m->d_charged, fp32, m->stream));
    t.end();
    return RH_OK;
}

rh_status rh_field_dev(rh_model* m, const double* d_M, const double* d_pts, int64_t np,
                       double* d_B, double* d_H, rh_precision prec)
{
    if (!m || (!d_B && !d_H) || np < 0) return fail(RH_ERR_ARG, "bad argument");
    if (np > 0 && (!d_M || !d_pts)) return fail(RH_ERR_ARG, "null argument");
    if (prec != RH_FP64 && prec != RH_MIXED) return fail(RH_ERR_ARG, "unknown precision");
    RH_HIP(enter_device(m->device));
    rh_status st = field_dev_impl(m, d_M, d_pts, np, d_B, d_H, prec == RH_MIXED);
    if (st != RH_OK) return st;
    RH_HIP(hipStreamSynchronize(m->stream));
    g_backend_fld = 1;
    return RH_OK;
}

}  // extern "C"

namespace rh_internal {

rh_status fail(rh_status st, const std::string& msg) { return ::fail(st, msg); }

int model_device(const rh_model* m) { return m->device; }

hipError_t enter_device(int device) { return ::enter_device(device); }

rh_status field_dev_into(rh_model* m, const double* d_M, const double* d_pts, int64_t np, double* d_B,
                         double* d_H, rh_precision prec, double* work, int64_t work_cap)
{
    if (!m || (!d_B && !d_H) || np < 0) return fail(RH_ERR_ARG, "bad argument");
    if (np > 0 && (!d_M || !d_pts)) return fail(RH_ERR_ARG, "null argument");
    if (prec != RH_FP64 && prec != RH_MIXED) return fail(RH_ERR_ARG, "unknown precision");
    if (!work || work_cap < 0) return fail(RH_ERR_ARG, "no work buffer");
    const int64_t slices = rh::field_slices(m->dev, np, work_cap);
    if (slices > 1 && 6 * np * slices > work_cap)
        return fail(RH_ERR_ARG, "rh_field_dev_noalloc: " + std::to_string(np) + " points need " +
                                    std::to_string(6 * np * slices * 8) + " bytes of partial sums for this model, "
                                    "beyond the fixed scratch of " + std::to_string(work_cap * 8) +
                                    " bytes; split the points into calls of at most " +
                                    std::to_string(work_cap / (6 * slices)) + " points");
    RH_HIP(::enter_device(m->device));
    RH_HIP(rh::launch_field(m->dev, d_M, d_pts, np, d_B, d_H, work, m->d_charged, prec == RH_MIXED, m->stream,
                            work_cap));
    RH_HIP(hipStreamSynchronize(m->stream));
    g_backend_fld = 1;
    return RH_OK;
}

}  // namespace rh_internal

extern "C" {

rh_status rh_field(rh_model* m, const double* M, const double* pts, int64_t np, double* B_out,
                   double* H_out, rh_precision prec)
{
    if (prec != RH_FP64 && prec != RH_MIXED) return fail(RH_ERR_ARG, "unknown precision");
    if (!m || !pts || np < 0 || (!B_out && !H_out)) return fail(RH_ERR_ARG, "bad argument");
    const int64_t n = m->host.n;
    if (n > 0 && !M) return fail(RH_ERR_ARG, "null magnetization argument");
    RH_HIP(enter_device(m->device));
    if (np == 0) return RH_OK;
    double *dM = nullptr, *dP = nullptr, *dB = nullptr, *dH = nullptr;
    auto cleanup = [&]() {
        dfree(dM);
        dfree(dP);
        dfree(dB);
        dfree(dH);
    };
    hipStream_t s = m->stream;
    hipError_t e = dalloc(&dM, size_t(3 * std::max<int64_t>(n, 1)));
    if (e == hipSuccess) e = dalloc(&dP, size_t(3 * np));
    if (e == hipSuccess) e = dalloc(&dB, size_t(3 * np));
    if (e == hipSuccess) e = dalloc(&dH, size_t(3 * np));
    if (e == hipSuccess && n > 0) e = hipMemcpyAsync(dM, M, size_t(3 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dP, pts, size_t(3 * np) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        cleanup();
        return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP,
                    std::string("field upload: ") + hipGetErrorString(e));
    }
    rh_status st = field_dev_impl(m, dM, dP, np, dB, dH, prec == RH_MIXED);
    if (st == RH_OK) {
        if (B_out) e = hipMemcpyAsync(B_out, dB, size_t(3 * np) * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && H_out)
            e = hipMemcpyAsync(H_out, dH, size_t(3 * np) * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(RH_ERR_HIP, std::string("field download: ") + hipGetErrorString(e));
    }
    cleanup();
    if (st == RH_OK) g_backend_fld = 1;
    return st;
}

// ---------------------------------------------------------------- currents
// ---------------------------------------------------------------- adjoint
rh_status rh_model_vertex_count(const rh_model* m, int64_t* nvert)
{
    if (!m || !nvert) return fail(RH_ERR_ARG, "null argument");
    *nvert = m->host.vert_off.empty() ? 0 : m->host.vert_off.back();
    return RH_OK;
}

static rh_status adj_geometry(rh_model* m)
{
    if (m->d_adj) return RH_OK;
    const rh::PackedModel& P = m->host;
    std::string err;
    std::vector<int64_t> sl = rh::adj_geom_slices(P.vert_off, &err);
    if (sl.empty()) return fail(RH_ERR_UNSUPPORTED, err);
    hipStream_t s = m->stream;
    RH_HIP(upload(&m->d_adj, P.adj, s));
    RH_HIP(upload(&m->d_adj_off, P.adj_off, s));
    RH_HIP(upload(&m->d_vert_off, P.vert_off, s));
    RH_HIP(upload(&m->d_slices, sl, s));
    m->n_slices = int64_t(sl.size()) - 1;
    RH_HIP(hipStreamSynchronize(s));
    return RH_OK;
}

rh_status rh_field_adjoint_dev(rh_model* m, const double* d_M, const double* d_pts, int64_t np,
                               const double* d_gB, int32_t field, double* d_gM, double* d_gV,
                               double* d_gX)
{
    if (!m || np < 0 || (np > 0 && (!d_pts || !d_gB))) return fail(RH_ERR_ARG, "bad argument");
    if ((d_gV || d_gX) && !d_M) return fail(RH_ERR_ARG, "vertex or point gradients need M");
    if (field != 0 && field != 1) return fail(RH_ERR_ARG, "field must be 0 (B) or 1 (H)");
    RH_HIP(enter_device(m->device));
    hipStream_t s = m->stream;
    const rh::PackedModel& P = m->host;
    const int64_t n = P.n;
    const int64_t nvert = P.vert_off.empty() ? 0 : P.vert_off.back();
    if ((d_gV || d_gX) && P.kind_off[1] > 0)
        return fail(RH_ERR_UNSUPPORTED, "vertex and point gradients need a model of polyhedra only");
    if (np == 0 || n == 0) {
        if (d_gM) RH_HIP(hipMemsetAsync(d_gM, 0, size_t(3 * n) * sizeof(double), s));
        if (d_gV) RH_HIP(hipMemsetAsync(d_gV, 0, size_t(3 * nvert) * sizeof(double), s));
        if (d_gX) RH_HIP(hipMemsetAsync(d_gX, 0, size_t(3 * np) * sizeof(double), s));
        RH_HIP(hipStreamSynchronize(s));
        return RH_OK;
    }
    Timer t("adjoint", s);
    if (d_gM) {
        const int64_t groups = rh::adj_point_groups(np, rh::adj_m_slices(n));
        RH_HIP(grow(m, size_t(groups * 3 * n)));
        t.begin();
        RH_HIP(rh::launch_adj_m(m->dev, d_pts, np, d_gB, field == 0, m->d_work, groups, d_gM, s));
        t.end();
    }
    if (d_gV || d_gX) {
        rh_status st = adj_geometry(m);
        if (st != RH_OK) return st;
        const int64_t groups = rh::adj_point_groups(np, m->n_slices);
        const size_t nv_part = size_t(groups * 3 * nvert);
        const size_t nx_part = d_gX ? size_t(m->n_slices * 3 * np) : 0;
        RH_HIP(grow(m, nv_part + nx_part));
        t.begin();
        RH_HIP(rh::launch_adj_geom(m->dev, m->d_adj, m->d_adj_off, m->d_vert_off, m->d_slices,
                                   m->n_slices, nvert, d_M, d_pts, np, d_gB, m->d_work,
                                   m->d_work + nv_part, groups, d_gV, d_gX, s));
        t.end();
    }
    RH_HIP(hipStreamSynchronize(s));
    return RH_OK;
}

rh_status rh_field_adjoint(rh_model* m, const double* M, const double* pts, int64_t np,
                           const double* gB, int32_t field, double* gM, double* gV, double* gX)
{
    if (!m || np < 0 || (np > 0 && (!pts || !gB))) return fail(RH_ERR_ARG, "bad argument");
    RH_HIP(enter_device(m->device));
    hipStream_t s = m->stream;
    const int64_t n = m->host.n;
    const int64_t nvert = m->host.vert_off.empty() ? 0 : m->host.vert_off.back();
    std::vector<double*> bufs;
    auto cleanup = [&]() {
        for (double* b : bufs) dfree(b);
    };
    // device copy of a host array (h may be null: allocation only)
    auto dev = [&](const double* h, size_t cnt, double** d) -> hipError_t {
        hipError_t e = dalloc(d, std::max<size_t>(cnt, 1));
        if (e != hipSuccess) return e;
        bufs.push_back(*d);
        if (h && cnt) e = hipMemcpyAsync(*d, h, cnt * sizeof(double), hipMemcpyHostToDevice, s);
        return e;
    };
    double *dM = nullptr, *dP = nullptr, *dG = nullptr, *dgM = nullptr, *dgV = nullptr, *dgX = nullptr;
    hipError_t e = hipSuccess;
    if (M) e = dev(M, size_t(3 * n), &dM);
    if (e == hipSuccess) e = dev(pts, size_t(3 * np), &dP);
    if (e == hipSuccess) e = dev(gB, size_t(3 * np), &dG);
    if (e == hipSuccess && gM) e = dev(nullptr, size_t(3 * n), &dgM);
    if (e == hipSuccess && gV) e = dev(nullptr, size_t(3 * nvert), &dgV);
    if (e == hipSuccess && gX) e = dev(nullptr, size_t(3 * np), &dgX);
    if (e != hipSuccess) {
        cleanup();
        return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_HIP,
                    std::string("adjoint upload: ") + hipGetErrorString(e));
    }
    rh_status st = rh_field_adjoint_dev(m, dM, dP, np, dG, field, dgM, dgV, dgX);
    if (st == RH_OK) {
        if (gM) e = hipMemcpy(gM, dgM, size_t(3 * n) * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && gV)
            e = hipMemcpy(gV, dgV, size_t(3 * nvert) * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && gX)
            e = hipMemcpy(gX, dgX, size_t(3 * np) * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = fail(RH_ERR_HIP, std::string("adjoint download: ") + hipGetErrorString(e));
    }
    cleanup();
    return st;
}

rh_status rh_relax_adjoint_dev(rh_matrix* q, const double* d_mats, const double* d_rhs, double* d_lam,
                               const rh_relax_opts* opts, rh_relax_result* res)
{
    if (!q || !d_mats || !d_rhs || !d_lam || !opts || !res) return fail(RH_ERR_ARG, "null argument");
    if (q->streamed()) return fail(RH_ERR_UNSUPPORTED, "relaxation adjoint needs a device-resident matrix");
    rh_model* m = q->m;
    RH_HIP(enter_device(m->device));
    hipStream_t s = m->stream;
    const int64_t n = m->host.n, n3 = q->n3;
    std::memset(res, 0, sizeof(*res));
    if (n == 0) {
        res->converged = 1;
        return RH_OK;
    }
    {
        // linear materials only: the transposed local solve needs K
        std::vector<double> kinds(static_cast<size_t>(n), 0.0);
        RH_HIP(hipMemcpy2DAsync(kinds.data(), sizeof(double), d_mats, 16 * sizeof(double),
                                sizeof(double), size_t(n), hipMemcpyDeviceToHost, s));
        RH_HIP(hipStreamSynchronize(s));
        for (double k : kinds)
            if (k != 0.0 && k != 1.0)
                return fail(RH_ERR_UNSUPPORTED, "relaxation adjoint needs linear materials (kinds 0, 1)");
    }
    RelaxWork w;
    RH_HIP(dalloc(&w.x, size_t(q->ld)));
    RH_HIP(dalloc(&w.y, size_t(n3)));
    RH_HIP(dalloc(&w.F, size_t(n3)));
    RH_HIP(dalloc(&w.Fp, size_t(n3)));
    RH_HIP(dalloc(&w.xp, size_t(n3)));
    for (int k = 0; k < kHist; k++) {
        RH_HIP(dalloc(&w.dX[k], size_t(n3)));
        RH_HIP(dalloc(&w.dF[k], size_t(n3)));
    }
    const int64_t nb = rh::relax_blocks(n);
    const int64_t db = rh::dot_blocks(n3);
    RH_HIP(dalloc(&w.part, size_t(rh::kRelaxPart * nb)));
    RH_HIP(dalloc(&w.dots, size_t(rh::kDotMax * db)));
    AdjointPass ap;
    ap.chunks = rh::matvec_t_chunks(n3, q->ld);
    struct Free2 {
        double** a;
        double** b;
        ~Free2()
        {
            dfree(*a);
            dfree(*b);
        }
    } free2{&ap.mu, &ap.tpart};
    RH_HIP(dalloc(&ap.mu, size_t(n3)));
    RH_HIP(dalloc(&ap.tpart, size_t(ap.chunks * q->ld)));
    RH_HIP(hipMemsetAsync(w.x, 0, size_t(q->ld) * sizeof(double), s));
    if (opts->keep_m)
        RH_HIP(hipMemcpyAsync(w.x, d_lam, size_t(n3) * sizeof(double), hipMemcpyDeviceToDevice, s));
    double* hbuf = nullptr;
    RH_HIP(hipHostMalloc(reinterpret_cast<void**>(&hbuf), size_t(rh::kRelaxPart * nb + rh::kDotMax * db) * sizeof(double)));
    struct PinnedFree {
        double* p;
        ~PinnedFree() { (void)hipHostFree(p); }
    } pinned_guard{hbuf};
    return relax_run(q, d_mats, nullptr, d_rhs, d_lam, w.y /* unused field output */, opts, w, hbuf, nb,
                     db, res, &ap);
}

rh_status rh_relax_adjoint(rh_matrix* q, const double* mats, const double* rhs, double* lam,
                           const rh_relax_opts* opts, rh_relax_result* res)
{
    if (!q || !mats || !rhs || !lam || !opts || !res) return fail(RH_ERR_ARG, "null argument");
    rh_model* m = q->m;
    RH_HIP(enter_device(m->device));
    const int64_t n = m->host.n;
    hipStream_t s = m->stream;
    double *dm = nullptr, *dr = nullptr, *dl = nullptr;
    auto cleanup = [&]() {
        dfree(dm);
        dfree(dr);
        dfree(dl);
    };
    hipError_t e = dalloc(&dm, size_t(16 * n) + 1);
    if (e == hipSuccess) e = dalloc(&dr, size_t(3 * n) + 1);
    if (e == hipSuccess) e = dalloc(&dl, size_t(3 * n) + 1);
    if (e == hipSuccess) e = hipMemcpyAsync(dm, mats, size_t(16 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rhs, size_t(3 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, lam, size_t(3 * n) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        cleanup();
        return fail(RH_ERR_HIP, std::string("adjoint relax upload: ") + hipGetErrorString(e));
    }
    rh_status st = rh_relax_adjoint_dev(q, dm, dr, dl, opts, res);
    if (st == RH_OK) {
        e = hipMemcpyAsync(lam, dl, size_t(3 * n) * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(RH_ERR_HIP, std::string("adjoint relax download: ") + hipGetErrorString(e));
    }
    cleanup();
    return st;
}

static rh_status pack_currents(int64_t nsrc, const double* tm, const double* tv, const double* ts,
                               std::vector<double>& cp)
{
    cp.assign(size_t(24 * nsrc), 0.0);
    for (int64_t k = 0; k < nsrc; k++) {
        double* o = cp.data() + 24 * k;
        rh::inv3(tm + 9 * k, o);
        for (int i = 0; i < 3; i++) o[9 + i] = tv[3 * k + i];
        for (int i = 0; i < 9; i++) o[12 + i] = ts[k] * tm[9 * k + i];
        for (int i = 0; i < 9; i++)
            if (!std::isfinite(o[i])) return fail(RH_ERR_ARG, "singular current-source transform");
    }
    return RH_OK;
}

rh_status rh_current_field_dev(int device, void* stream, int64_t nsrc, const int32_t* kind,
                               const double* par, const double* tm, const double* tv,
                               const double* ts, const double* d_pts, int64_t np, double* d_B)
{
    if (nsrc < 0 || np < 0) return fail(RH_ERR_ARG, "negative size");
    if (nsrc == 0 || np == 0) return RH_OK;
    if (!kind || !par || !tm || !tv || !ts || !d_pts || !d_B) return fail(RH_ERR_ARG, "null argument");
    for (int64_t k = 0; k < nsrc; k++) {
        if (kind[k] != 0 && kind[k] != 1) return fail(RH_ERR_UNSUPPORTED, "unknown current-source kind");
        if (kind[k] == 1 && !(par[12 * k + 9] >= 1.0)) return fail(RH_ERR_ARG, "arc needs nseg >= 1");
    }
    std::vector<double> cp;
    rh_status st = pack_currents(nsrc, tm, tv, ts, cp);
    if (st != RH_OK) return st;
    RH_HIP(enter_device(device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int32_t* dk = nullptr;
    double *dp = nullptr, *dc = nullptr;
    auto cleanup = [&]() {
        dfree(dk);
        dfree(dp);
        dfree(dc);
    };
    hipError_t e = dalloc(&dk, size_t(nsrc));
    if (e == hipSuccess) e = dalloc(&dp, size_t(12 * nsrc));
    if (e == hipSuccess) e = dalloc(&dc, size_t(24 * nsrc));
    if (e == hipSuccess) e = hipMemcpyAsync(dk, kind, size_t(nsrc) * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, par, size_t(12 * nsrc) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dc, cp.data(), cp.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = rh::launch_current(nsrc, dk, dp, dc, d_pts, np, d_B, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return fail(RH_ERR_HIP, std::string("current field: ") + hipGetErrorString(e));
    g_backend_fld = 1;
    return RH_OK;
}

rh_status rh_current_field(int device, void* stream, int64_t nsrc, const int32_t* kind,
                           const double* par, const double* tm, const double* tv, const double* ts,
                           const double* pts, int64_t np, double* B_acc)
{
    if (np == 0 || nsrc == 0) return RH_OK;
    if (!pts || !B_acc) return fail(RH_ERR_ARG, "null argument");
    RH_HIP(enter_device(device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *dP = nullptr, *dB = nullptr;
    hipError_t e = dalloc(&dP, size_t(3 * np));
    if (e == hipSuccess) e = dalloc(&dB, size_t(3 * np));
    if (e == hipSuccess) e = hipMemcpyAsync(dP, pts, size_t(3 * np) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dB, B_acc, size_t(3 * np) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        dfree(dP);
        dfree(dB);
        return fail(RH_ERR_HIP, std::string("current upload: ") + hipGetErrorString(e));
    }
    rh_status st = rh_current_field_dev(device, stream, nsrc, kind, par, tm, tv, ts, dP, np, dB);
    if (st == RH_OK) {
        e = hipMemcpyAsync(B_acc, dB, size_t(3 * np) * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(RH_ERR_HIP, std::string("current download: ") + hipGetErrorString(e));
    }
    dfree(dP);
    dfree(dB);
    return st;
}

}  // extern "C"
