// C-ABI of libcapytaine_amd: plain pointers and sizes, int status codes
// (0 = success), no exceptions across the boundary.  Declarations and the
// reference interface each entry point replaces are in
// include/capytaine_amd.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "cpt_internal.hpp"
#include "../../include/capytaine_amd.h"

// Scratch of cpt_lu_solve_noalloc: the inverted 64 x 64 diagonal blocks of
// the factors (a strip of n x 64 complex128 per matrix) for at most
// CPT_LU_SOLVE_NOALLOC_MAX_ROWS rows over the batch: 32 MiB, held by every
// device the library runs on until the process ends.  Its device address is
// taken once per device (scratch_address) and only that address is used.
static __device__ __align__(256) double scratch[2 * 64 * (size_t)CPT_LU_SOLVE_NOALLOC_MAX_ROWS];

namespace cpt {

int tabulate_on_device(const double* r_range, int nr, const double* z_range, int nz, int npts, double* tab,
                       hipStream_t stream);
int launch_table_relayout(const double* src, int nr, int nz, double* dst, hipStream_t stream);
int launch_zgemm(int M, int N, int K, double alpha, const cplx* A, int64_t lda, int64_t sA, const cplx* B,
                 int64_t ldb, int64_t sB, cplx* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream);
int launch_getrf(cplx* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                 cplx* work, const LuSetup* aux, hipStream_t stream);
size_t getrf_workspace_bytes(int N, int batch);
size_t mixed_workspace_bytes(int n, int nrhs, int batch);
size_t mixed_grouped_bytes(const MixedGroup* g, int ng);
int launch_mixed_grouped(const MixedGroup* g, int ng, void* work, const LuSetup* tune, int max_iter, double tol,
                         int* iterations, double* residual, int* info_out);
int launch_getrf_lanes(const LuGroupArgs* g, int ng, const LuSetup* tune, void* work);
int launch_getrs_lanes(const LuGroupArgs* g, int ng, void* work);
size_t getrs_workspace_bytes(int N, int batch);
int launch_getrs(const cplx* LU, int N, int64_t lda, int64_t strideA, const int* ipiv, int64_t strideP, cplx* B,
                 int nrhs, int64_t ldb, int64_t strideB, int batch, cplx* work, hipStream_t stream);
int launch_cgemm(int M, int N, int K, float alpha, const void* A, int64_t lda, int64_t sA, const void* B,
                 int64_t ldb, int64_t sB, void* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream);
size_t cgetrf_workspace_bytes(int N, int batch);
int launch_cgetrf(void* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                  void* work, const LuSetup* aux, hipStream_t stream);
size_t cgetrs_workspace_bytes(int N, int batch);
int launch_cgetrs(const void* LU, int N, int64_t lda, int64_t strideA, const int* ipiv, int64_t strideP, void* B,
                  int nrhs, int64_t ldb, int64_t strideB, int batch, void* work, hipStream_t stream);
size_t gmres_workspace_bytes(int n, int restart, bool single);
int gmres_max_restart();
int run_gmres(bool single, const void* A, int n, int64_t lda, const void* B, int nrhs, int64_t ldb, void* X,
              int64_t ldx, int restart, int maxiter, double rtol, double atol, void* work, hipStream_t stream,
              int* iterations, double* residual, int* not_converged);
int launch_block_transform(int xor_kind, int n, int inverse, int R, int C, int nbatch, const cplx* in, int64_t sbi,
                           int64_t sj, int64_t ldi, cplx* out, int64_t sbo, int64_t sk, int64_t ldo, hipStream_t stream);
int block_transform_max_blocks();

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// Device address of `scratch` per device, looked up once with that device
// current.
constexpr int MAX_DEVICES = 64;
static void* g_scratch_addr[MAX_DEVICES];
static std::mutex g_scratch_mutex;

int scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) {
        set_error("scratch_init: device index beyond " + std::to_string(MAX_DEVICES));
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_addr[device]) return 0;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) prev = device;
    hipError_t e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return -1;
    }
    g_scratch_addr[device] = p;
    return 0;
}

thread_local KernelTimer* active_timer = nullptr;
thread_local int active_gemm_tiles = 0;

int KernelTimer::start(hipStream_t s, int kind, double work, double bytes)
{
    if (!on) return -1;
    if (used == recs.size()) {
        Rec r;
        if (hipEventCreate(&r.a) != hipSuccess) return -1;
        if (hipEventCreate(&r.b) != hipSuccess) {
            (void)hipEventDestroy(r.a);
            return -1;
        }
        recs.push_back(r);
    }
    Rec& r = recs[used];
    r.kind = kind;
    r.work = work;
    r.bytes = bytes;
    if (hipEventRecord(r.a, s) != hipSuccess) return -1;
    return (int)used++;
}

void KernelTimer::stop(int slot, hipStream_t s) { (void)hipEventRecord(recs[slot].b, s); }

void KernelTimer::release()
{
    for (Rec& r : recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    recs.clear();
    used = 0;
    on = false;
}

// Installs a context's timer and product tiling for the calling thread
// during one entry point.
struct CallScope {
    KernelTimer* prev;
    int prev_tiles;
    CallScope(KernelTimer* t, int tiles) : prev(active_timer), prev_tiles(active_gemm_tiles)
    {
        active_timer = (t && t->on) ? t : nullptr;
        active_gemm_tiles = tiles;
    }
    ~CallScope()
    {
        active_timer = prev;
        active_gemm_tiles = prev_tiles;
    }
};

// Growable device buffer owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int reserve(size_t need)
    {
        if (need <= bytes) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, need) != hipSuccess) return -1;
        bytes = need;
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace cpt

struct cpt_context {
    int device = 0;
    // tabulation
    cpt::DevBuf r_range, z_range, table;
    cpt::DevBuf r_range_f, z_range_f, table_f;  // float copies (far field of the mixed form)
    int far_fp32 = 0;                           // mixed-precision assembly
    int out_c64 = 0;                            // S and K stored as complex64
    int nr = 0, nz = 0, grid = 1, npts = 1001;
    double rmax = 0, zmax_abs = 0, dz = 0;
    int index_of_1 = 81;
    bool has_table = false;
    // per-call workspace
    cpt::DevBuf faces, quad, freq, prony;
    cpt::DevBuf bfaces, bquad, bfreq, bargs, btiles;  // multi-body assembly
    cpt::DevBuf lu_work, mixed_work, solve_work;  // solve_work: inverted diagonal blocks of the solves
    cpt::DevBuf gmres_work;                       // Krylov basis and reductions of the GMRES solves
    cpt::DevBuf fg_roots, fg_ws, fg_lists, fg_tab; // FinGreen3D: roots, epsilon tables, pair lists, factor tables
    cpt::LuSetup lu_streams;
    int asm_chunk = 0;  // panels per assembly work-group, 0 = by size
    int gemm_tiles = 0; // 1: complex64 products on 128 x 128 tiles, 0: 128 x 64
    long mixed_retries = 0;    // FP32-input re-factorisations after FP16-input products fell short
    void* ext_work = nullptr;  // caller-owned workspace of the mixed solves
    size_t ext_bytes = 0;
    // lanes of the grouped LU: default-priority streams, a fork event and
    // one join event per lane
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_done;
    hipEvent_t fork = nullptr;
    cpt::KernelTimer timer;
};

using namespace cpt;

#define CPT_CHECK_HIP(expr, what)                                                            \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_error(std::string(what) + ": " + hipGetErrorString(e_));                     \
            return CPT_ERR_HIP;                                                              \
        }                                                                                    \
    } while (0)

namespace {

// Frequency table from device arrays (cpt_build_matrices_dev): prony holds
// nb x (lambda[nexp], a[nexp]).
__global__ void freq_args_kernel(int nb, const double* __restrict__ k, const int* __restrict__ gfs,
                                 const double* __restrict__ prony, int nexp, int rankine_only,
                                 FreqArgs* __restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    FreqArgs f;
    f.k = k[b];
    f.gfs = gfs[b];
    f.wave = (!rankine_only && 0.0 < f.k && isfinite(f.k)) ? 1 : 0;
    f.nexp = nexp;
    f.pad = 0;
    f.prony_lambda = prony ? prony + (size_t)b * 2 * nexp : nullptr;
    f.prony_a = prony ? prony + (size_t)b * 2 * nexp + nexp : nullptr;
    out[b] = f;
}

__global__ void pack_faces_kernel(const double* __restrict__ vertices, int nv, const int32_t* __restrict__ faces,
                                  const double* __restrict__ centers, const double* __restrict__ normals,
                                  const double* __restrict__ areas, const double* __restrict__ radii,
                                  const double* __restrict__ qp, const double* __restrict__ qw, int m, int nq,
                                  Face* __restrict__ out, double* __restrict__ quad)
{
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    if (J >= m) return;
    Face F;
#pragma unroll
    for (int L = 0; L < 4; ++L) {
        // Indices are validated by the host wrapper; clamping keeps a bad
        // index from reading outside the vertex array.
        int v = faces[4 * (size_t)J + L];
        v = (v < 0 || v >= nv) ? 0 : v;
        for (int c = 0; c < 3; ++c) F.v[L][c] = vertices[3 * (size_t)v + c];
    }
    for (int c = 0; c < 3; ++c) {
        F.c[c] = centers[3 * (size_t)J + c];
        F.n[c] = normals[3 * (size_t)J + c];
    }
    F.area = areas[J];
    F.radius = radii[J];
    for (int c = 0; c < 4; ++c) F.pad[c] = 0.0;
    out[J] = F;
    for (int q = 0; q < nq; ++q) {
        const size_t s = (size_t)J * nq + q;
        quad[4 * s + 0] = qp[3 * s + 0];
        quad[4 * s + 1] = qp[3 * s + 1];
        quad[4 * s + 2] = qp[3 * s + 2];
        quad[4 * s + 3] = qw[s];
    }
}

template <bool UNTAB>
__global__ void wave_points_kernel(Table T, int np, const double* __restrict__ x0i, const double* __restrict__ x0j,
                                   double k, int gfs, cplx* __restrict__ g, cplx* __restrict__ ng)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    cplx G, nG[3];
    wave_point<UNTAB>(T, k, gfs, x0i[3 * (size_t)p] - x0j[3 * (size_t)p],
                      x0i[3 * (size_t)p + 1] - x0j[3 * (size_t)p + 1], x0i[3 * (size_t)p + 2] + x0j[3 * (size_t)p + 2],
                      G, nG);
    g[p] = G;
    for (int c = 0; c < 3; ++c) ng[3 * (size_t)p + c] = nG[c];
}

__global__ void to_float_kernel(const double* __restrict__ src, float* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ void zero_int_kernel(int* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Side stream (default priority) and events of the LU look-ahead.
int ensure_lu_streams(cpt_context* h)
{
    if (h->lu_streams.side) return 0;
    if (hipStreamCreateWithFlags(&h->lu_streams.side, hipStreamNonBlocking) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&h->lu_streams.ready, hipEventDisableTiming) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&h->lu_streams.done, hipEventDisableTiming) != hipSuccess) return -1;
    return 0;
}

// Three lane streams: with the caller's stream they take the four hardware
// queues of a process (GPU_MAX_HW_QUEUES); four lanes measured slower on the
// 64-hull fleet (0.30 s against 0.25 s per chunk of about 20 bodies).
constexpr int NUM_LANES = 3;

int ensure_lanes(cpt_context* h)
{
    if (!h->fork && hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess) return -1;
    while ((int)h->lanes.size() < NUM_LANES) {
        hipStream_t s;
        hipEvent_t ev;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -1;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(s);
            return -1;
        }
        h->lanes.push_back(s);
        h->lane_done.push_back(ev);
    }
    return 0;
}

// Lanes wait for the work already on `caller`.
int fork_lanes(cpt_context* h, hipStream_t caller)
{
    if (hipEventRecord(h->fork, caller) != hipSuccess) return -1;
    for (hipStream_t l : h->lanes)
        if (hipStreamWaitEvent(l, h->fork, 0) != hipSuccess) return -1;
    return 0;
}

// `caller` waits for the work enqueued on every lane.
int join_lanes(cpt_context* h, hipStream_t caller)
{
    for (size_t i = 0; i < h->lanes.size(); ++i) {
        if (hipEventRecord(h->lane_done[i], h->lanes[i]) != hipSuccess) return -1;
        if (hipStreamWaitEvent(caller, h->lane_done[i], 0) != hipSuccess) return -1;
    }
    return 0;
}

// Longest-processing-time assignment of groups (cost n^3 batch) to lanes.
std::vector<hipStream_t> assign_lanes(const cpt_context* h, const cpt_lu_group* g, int ng)
{
    std::vector<int> order(ng);
    for (int i = 0; i < ng; ++i) order[i] = i;
    auto cost = [&](int i) { return (double)g[i].n * g[i].n * g[i].n * g[i].batch; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    std::vector<double> load(h->lanes.size(), 0.0);
    std::vector<hipStream_t> out(ng);
    for (int i : order) {
        const size_t l = std::min_element(load.begin(), load.end()) - load.begin();
        out[i] = h->lanes[l];
        load[l] += cost(i);
    }
    return out;
}

bool valid_groups(const cpt_lu_group* g, int ng, bool need_factor, bool need_rhs, bool need_x, const char* who)
{
    for (int i = 0; i < ng; ++i) {
        const cpt_lu_group& G = g[i];
        bool ok = G.n >= 0 && G.batch >= 0 && G.lda >= G.n && (G.batch <= 1 || G.strideA >= G.lda * G.n);
        if (ok && G.n > 0 && G.batch > 0) {
            ok = ok && G.A;
            if (need_factor) ok = ok && G.ipiv && G.info && (G.batch <= 1 || G.strideP >= G.n);
            if (need_rhs)
                ok = ok && G.nrhs >= 0 && G.ldb >= G.n && (G.nrhs == 0 || G.B) &&
                     (G.batch <= 1 || G.strideB >= G.ldb * G.nrhs);
            if (need_x) ok = ok && G.ldx >= G.n && (G.nrhs == 0 || G.X) && (G.batch <= 1 || G.strideX >= G.ldx * G.nrhs);
            if (G.n > 24576) ok = false;
        }
        if (!ok) {
            set_error(std::string(who) + ": invalid sizes, strides or null pointers in group " + std::to_string(i));
            return false;
        }
    }
    return true;
}

// Workspace of a mixed solve: the caller's (cpt_set_workspace) when large
// enough, else the context's own, grown as needed.  nullptr on failure.
void* mixed_workspace(cpt_context* h, size_t bytes)
{
    if (h->ext_work && h->ext_bytes >= bytes) return h->ext_work;
    if (h->mixed_work.reserve(bytes)) return nullptr;
    return h->mixed_work.p;
}

// Header of a caller-owned complex64 factor (cpt_mixed_factor): magic, n,
// batch and the product form that factored it, in the first 256 bytes.
constexpr int FACTOR_MAGIC = 0x4650434d;
constexpr size_t FACTOR_HEADER = 256;

__global__ void factor_header_kernel(int* p, int magic, int n, int batch, int tiles)
{
    if (threadIdx.x == 0) {
        p[0] = magic;
        p[1] = n;
        p[2] = batch;
        p[3] = tiles;
    }
}

// Refinement of one group with its factor F, factored first when `factor`.
// *tiles is the product form of F (set when factoring here).  When the
// FP16-input products (form 4) leave the refinement short of tol, F is
// factored again with FP32-input products and the refinement repeated; the
// sweeps of both are reported.  CPT_ERR_NOT_CONVERGED when the final
// residual exceeds tol or the complex64 factor is exactly singular; X,
// iterations and residual are written in every case.
int mixed_solve_one(cpt_context* h, const MixedGroup& g, const MixedFactor& F, void* fws, void* rws, bool factor,
                    int* tiles, int max_iter, double tol, int* iterations, double* residual, const char* who)
{
    int info = 0, e;
    if (factor) {
        *tiles = h->gemm_tiles;
        if ((e = mixed_factor(g.K, g.n, g.lda, g.strideK, g.batch, F, fws, &h->lu_streams, g.s))) {
            set_error(std::string(who) + ": complex64 factorisation: " + hipGetErrorString((hipError_t)e));
            return CPT_ERR_HIP;
        }
    }
    if ((e = mixed_refine(&g, &F, 1, rws, max_iter, tol, iterations, residual, &info))) {
        set_error(std::string(who) + ": refinement: " + hipGetErrorString((hipError_t)e));
        return CPT_ERR_HIP;
    }
    if (*tiles == 4 && !(*residual <= tol)) {
        const int sweeps = *iterations;
        CallScope fp32_scope(&h->timer, 3);
        if ((e = mixed_factor(g.K, g.n, g.lda, g.strideK, g.batch, F, fws, &h->lu_streams, g.s)) ||
            (e = mixed_refine(&g, &F, 1, rws, max_iter, tol, iterations, residual, &info))) {
            set_error(std::string(who) + ": FP32-input retry: " + hipGetErrorString((hipError_t)e));
            return CPT_ERR_HIP;
        }
        *tiles = 3;
        ++h->mixed_retries;
        *iterations += sweeps;
    }
    if (info < 0) {
        set_error(std::string(who) + ": the complex64 panel factorisation gave up waiting for its work-groups");
        return CPT_ERR_NOT_CONVERGED;
    }
    if (info != 0) {
        set_error(std::string(who) + ": the complex64 factor is exactly singular (pivot " + std::to_string(info) +
                  " is zero); relative residual " + std::to_string(*residual));
        return CPT_ERR_NOT_CONVERGED;
    }
    if (!(*residual <= tol)) {
        char buf[160];
        std::snprintf(buf, sizeof buf, ": the refinement stopped at relative residual %.3e > tol %.3e after %d sweeps",
                      *residual, tol, *iterations);
        set_error(std::string(who) + buf);
        return CPT_ERR_NOT_CONVERGED;
    }
    return CPT_OK;
}

Table table_of(const cpt_context* h)
{
    Table T;
    T.tab = (const double*)h->table.p;
    T.r_range = (const double*)h->r_range.p;
    T.z_range = (const double*)h->z_range.p;
    T.nr = h->nr;
    T.nz = h->nz;
    T.grid = h->grid;
    T.npts = h->npts;
    T.rmax = h->rmax;
    T.zmax_abs = h->zmax_abs;
    T.dz = h->dz;
    T.index_of_1 = h->index_of_1;
    T.tab_f = (const float*)h->table_f.p;
    T.r_range_f = (const float*)h->r_range_f.p;
    T.z_range_f = (const float*)h->z_range_f.p;
    return T;
}

}  // namespace

extern "C" {

const char* cpt_last_error(void) { return g_last_error.c_str(); }

int cpt_version(void) { return CPT_ABI_VERSION; }

int cpt_create(int device, cpt_handle* out)
{
    if (!out) {
        set_error("cpt_create: null output pointer");
        return CPT_ERR_ARG;
    }
    int count = 0;
    CPT_CHECK_HIP(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) {
        set_error("cpt_create: invalid device index " + std::to_string(device));
        return CPT_ERR_ARG;
    }
    if (scratch_init(device)) return CPT_ERR_HIP;
    cpt_context* h = new cpt_context();
    h->device = device;
    // device counter of speculative panels factored again (CPT_TUNE_SPEC_REPAIRS)
    if (hipSetDevice(device) != hipSuccess || hipMalloc((void**)&h->lu_streams.spec_repairs, sizeof(int)) != hipSuccess ||
        hipMemset(h->lu_streams.spec_repairs, 0, sizeof(int)) != hipSuccess) {
        set_error("cpt_create: could not allocate the context's counters");
        if (h->lu_streams.spec_repairs) (void)hipFree(h->lu_streams.spec_repairs);
        delete h;
        return CPT_ERR_HIP;
    }
    *out = h;
    return CPT_OK;
}

int cpt_destroy(cpt_handle h)
{
    if (!h) return CPT_OK;
    (void)hipSetDevice(h->device);
    h->r_range.release();
    h->z_range.release();
    h->table.release();
    h->r_range_f.release();
    h->z_range_f.release();
    h->table_f.release();
    h->faces.release();
    h->quad.release();
    h->bfaces.release();
    h->bquad.release();
    h->bfreq.release();
    h->bargs.release();
    h->btiles.release();
    h->freq.release();
    h->prony.release();
    h->lu_work.release();
    h->gmres_work.release();
    h->mixed_work.release();
    h->solve_work.release();
    if (h->lu_streams.side) (void)hipStreamDestroy(h->lu_streams.side);
    if (h->lu_streams.ready) (void)hipEventDestroy(h->lu_streams.ready);
    if (h->lu_streams.done) (void)hipEventDestroy(h->lu_streams.done);
    if (h->lu_streams.spec_repairs) (void)hipFree(h->lu_streams.spec_repairs);
    h->timer.release();
    for (hipStream_t l : h->lanes) (void)hipStreamDestroy(l);
    for (hipEvent_t ev : h->lane_done) (void)hipEventDestroy(ev);
    if (h->fork) (void)hipEventDestroy(h->fork);
    delete h;
    return CPT_OK;
}

int cpt_tabulate(cpt_handle h, const double* r_range, int nr, const double* z_range, int nz,
                 int nb_integration_points, double* tab, hipStream_t stream)
{
    if (!h || !r_range || !z_range || !tab || nr <= 0 || nz <= 0 || nb_integration_points <= 0) {
        set_error("cpt_tabulate: invalid arguments");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    const int e = tabulate_on_device(r_range, nr, z_range, nz, nb_integration_points, tab, stream);
    return e == 0 ? CPT_OK : (e == -3 ? CPT_ERR_ALLOC : CPT_ERR_HIP);
}

int cpt_set_tabulation(cpt_handle h, const double* r_range, int nr, const double* z_range, int nz,
                       const double* tab, int grid_shape, int nb_integration_points, hipStream_t stream)
{
    if (!h || nr < 0 || nz < 0 || (grid_shape != CPT_LEGACY_GRID && grid_shape != CPT_SCALED_NEMOH3_GRID)) {
        set_error("cpt_set_tabulation: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (nr > 1 && nz > 1 && (!r_range || !z_range || !tab)) {
        set_error("cpt_set_tabulation: null table pointer");
        return CPT_ERR_ARG;
    }
    if (nr <= 1 || nz <= 1) {
        // no table: the wave term integrates over theta for each point
        // (Green_wave.f90:226-228), with nb_integration_points Simpson points
        if (nb_integration_points < 1) {
            set_error("cpt_set_tabulation: nb_integration_points must be positive");
            return CPT_ERR_ARG;
        }
        h->nr = nr;
        h->nz = nz;
        h->grid = grid_shape;
        h->npts = nb_integration_points;
        h->rmax = 0.0;
        h->zmax_abs = 0.0;
        h->dz = 0.0;
        h->index_of_1 = 0;
        h->has_table = true;
        return CPT_OK;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    if (h->r_range.reserve(sizeof(double) * nr) || h->z_range.reserve(sizeof(double) * nz) ||
        h->table.reserve(sizeof(double) * 5 * (size_t)nr * nz)) {
        set_error("cpt_set_tabulation: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    CPT_CHECK_HIP(hipMemcpyAsync(h->r_range.p, r_range, sizeof(double) * nr, hipMemcpyDeviceToDevice, stream),
                  "copy r_range");
    CPT_CHECK_HIP(hipMemcpyAsync(h->z_range.p, z_range, sizeof(double) * nz, hipMemcpyDeviceToDevice, stream),
                  "copy z_range");
    int e = launch_table_relayout(tab, nr, nz, (double*)h->table.p, stream);
    CPT_CHECK_HIP((hipError_t)e, "table relayout launch");
    if (h->r_range_f.reserve(sizeof(float) * nr) || h->z_range_f.reserve(sizeof(float) * nz) ||
        h->table_f.reserve(sizeof(float) * 5 * (size_t)nr * nz)) {
        set_error("cpt_set_tabulation: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    hipLaunchKernelGGL(to_float_kernel, dim3((nr + 255) / 256), dim3(256), 0, stream, (const double*)h->r_range.p,
                       (float*)h->r_range_f.p, (size_t)nr);
    hipLaunchKernelGGL(to_float_kernel, dim3((nz + 255) / 256), dim3(256), 0, stream, (const double*)h->z_range.p,
                       (float*)h->z_range_f.p, (size_t)nz);
    hipLaunchKernelGGL(to_float_kernel, dim3((5 * (size_t)nr * nz + 255) / 256), dim3(256), 0, stream,
                       (const double*)h->table.p, (float*)h->table_f.p, 5 * (size_t)nr * nz);
    CPT_CHECK_HIP(hipGetLastError(), "table conversion launch");
    double last[2];
    CPT_CHECK_HIP(hipMemcpyAsync(&last[0], (const double*)r_range + (nr - 1), sizeof(double),
                                 hipMemcpyDeviceToHost, stream),
                  "read rmax");
    CPT_CHECK_HIP(hipMemcpyAsync(&last[1], (const double*)z_range + (nz - 1), sizeof(double),
                                 hipMemcpyDeviceToHost, stream),
                  "read zmin");
    CPT_CHECK_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    h->nr = nr;
    h->nz = nz;
    h->grid = grid_shape;
    h->npts = nb_integration_points;
    h->rmax = last[0];
    h->zmax_abs = std::fabs(last[1]);
    // Delhommeau_integrals.f90:360 and :324 (the latter in default real)
    h->dz = (std::log10(std::fabs(last[1])) + 10.0) / nz;
    h->index_of_1 = (int)std::lround((float)(nr * 81) / 676.0f);
    h->has_table = true;
    return CPT_OK;
}

static int build_matrices_impl(bool device_params, bool no_wave, cpt_handle h, int nb_collocation_points,
                               const double* collocation_points,
                               const double* dot_product_normals, int nb_vertices, const double* vertices, int nb_faces,
                               const int32_t* faces, const double* centers, const double* normals, const double* areas,
                               const double* radiuses, int nb_quad_points, const double* quad_points,
                               const double* quad_weights, int nb_wavenumbers, const double* wavenumbers,
                               const int* gf_singularities, double depth, int finite_depth_method, int nexp,
                               const double* prony, int adjoint_double_layer, int early_dot_product,
                               int diagonal_term_in_double_layer, double free_surface, void* S, int64_t ldS,
                               int64_t strideS, void* K, int64_t ldK, int64_t strideK, hipStream_t stream)
{
    if (!h) {
        set_error("cpt_build_matrices: null handle");
        return CPT_ERR_ARG;
    }
    const int n = nb_collocation_points, m = nb_faces, nq = nb_quad_points, nb = nb_wavenumbers;
    if (n < 0 || m < 0 || nq < 1 || nq > 16 || nb < 1 || nb_vertices < 0 || ldS < n || ldK < n ||
        (nb > 1 && (strideS < ldS * m || strideK < ldK * m * (early_dot_product ? 1 : 3)))) {
        set_error("cpt_build_matrices: invalid sizes or strides");
        return CPT_ERR_ARG;
    }
    if (n == 0 || m == 0) return CPT_OK;
    if (!collocation_points || !dot_product_normals || !vertices || !faces || !centers || !normals || !areas ||
        !radiuses || !quad_points || !quad_weights || !wavenumbers || !gf_singularities || !S || !K) {
        set_error("cpt_build_matrices: null pointer argument");
        return CPT_ERR_ARG;
    }
    const bool rankine_only = std::isinf(free_surface);
    const bool finite_depth = std::isfinite(depth);
    if (rankine_only && finite_depth) {
        set_error("cpt_build_matrices: free_surface=inf requires an infinite water depth");
        return CPT_ERR_ARG;
    }
    if (finite_depth && finite_depth_method != CPT_LEGACY_FINITE_DEPTH &&
        finite_depth_method != CPT_NEWER_FINITE_DEPTH) {
        set_error(finite_depth_method == CPT_FINGREEN3D_METHOD
                      ? "cpt_build_matrices: FinGreen3D takes its dispersion roots through cpt_build_matrices_fingreen3d"
                      : "cpt_build_matrices: unknown finite_depth_method");
        return CPT_ERR_UNSUPPORTED;
    }
    if (finite_depth && nexp > 0 && !prony) {
        set_error("cpt_build_matrices: finite depth needs the Prony decomposition");
        return CPT_ERR_ARG;
    }
    // no_wave: the Rankine term and its free-surface image only, the wave
    // part being added by another kernel (FinGreen3D)
    bool need_table = device_params && !rankine_only && !no_wave;  // the wavenumbers are read on the device
    for (int b = 0; b < nb && !device_params; ++b) {
        const double k = wavenumbers[b];
        const int g = gf_singularities[b];
        if (g != CPT_HIGH_FREQ && g != CPT_LOW_FREQ && g != CPT_LOW_FREQ_WITH_RANKINE_PART) {
            set_error("cpt_build_matrices: invalid gf_singularities value");
            return CPT_ERR_ARG;
        }
        if (std::isnan(k) || k < 0) {
            set_error("cpt_build_matrices: invalid wavenumber");
            return CPT_ERR_ARG;
        }
        if (!rankine_only && !no_wave && k > 0 && std::isfinite(k)) need_table = true;
    }
    if (need_table && !h->has_table) {
        set_error("cpt_build_matrices: no tabulation set (call cpt_set_tabulation first)");
        return CPT_ERR_STATE;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);

    // Pack panels and quadrature points.
    const int ne = finite_depth ? std::max(nexp, 0) : 0;
    if (h->faces.reserve(sizeof(Face) * (size_t)m) ||
        h->quad.reserve(sizeof(double) * 4 * (size_t)m * nq) || h->freq.reserve(sizeof(FreqArgs) * nb) ||
        h->prony.reserve(sizeof(double) * (2 * (size_t)nb * ne + 1))) {
        set_error("cpt_build_matrices: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    hipLaunchKernelGGL(pack_faces_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, vertices, nb_vertices, faces,
                       centers, normals, areas, radiuses, quad_points, quad_weights, m, nq, (Face*)h->faces.p,
                       (double*)h->quad.p);
    CPT_CHECK_HIP(hipGetLastError(), "pack kernel launch");

    if (device_params) {
        // the frequency table from the device arrays, at run time: nothing
        // from host memory enters the stream, so the call can be captured
        hipLaunchKernelGGL(freq_args_kernel, dim3((nb + 63) / 64), dim3(64), 0, stream, nb, wavenumbers,
                           gf_singularities, ne > 0 ? prony : nullptr, ne, rankine_only ? 1 : 0, (FreqArgs*)h->freq.p);
        CPT_CHECK_HIP(hipGetLastError(), "frequency table kernel launch");
    } else {
        std::vector<FreqArgs> fa(nb);
        const double* prony_dev = (const double*)h->prony.p;
        if (ne > 0)
            CPT_CHECK_HIP(hipMemcpyAsync(h->prony.p, prony, sizeof(double) * 2 * (size_t)nb * ne,
                                         hipMemcpyHostToDevice, stream),
                          "copy prony");
        for (int b = 0; b < nb; ++b) {
            const double k = wavenumbers[b];
            fa[b].k = k;
            fa[b].gfs = gf_singularities[b];
            fa[b].wave = (!rankine_only && !no_wave && 0.0 < k && std::isfinite(k)) ? 1 : 0;
            fa[b].nexp = ne;
            fa[b].pad = 0;
            fa[b].prony_lambda = prony_dev + (size_t)b * 2 * ne;
            fa[b].prony_a = prony_dev + (size_t)b * 2 * ne + ne;
        }
        CPT_CHECK_HIP(hipMemcpyAsync(h->freq.p, fa.data(), sizeof(FreqArgs) * nb, hipMemcpyHostToDevice, stream),
                      "copy frequency table");
    }

    AssemblyArgs A;
    std::memset(&A, 0, sizeof(A));
    A.colloc = collocation_points;
    A.dpn = dot_product_normals;
    A.n = n;
    A.m = m;
    A.nq = nq;
    A.nk = early_dot_product ? 1 : 3;
    A.faces = (const Face*)h->faces.p;
    A.quad = (const double*)h->quad.p;
    A.T = table_of(h);
    A.depth = depth;
    A.finite_depth = finite_depth ? 1 : 0;
    A.fd_method = finite_depth_method;
    A.adjoint = adjoint_double_layer ? 1 : 0;
    A.rankine_only = rankine_only ? 1 : 0;
    A.diag = diagonal_term_in_double_layer ? 1 : 0;
    A.n_diag = std::min(n, m);
    A.free_surface = free_surface;
    A.nb = nb;
    // Enough work-groups to fill 256 CUs several times over.
    const int gx = (n + 255) / 256;
    int chunk = (int)std::max<int64_t>(1, ((int64_t)gx * m) / 4096);
    A.j_chunk = h->asm_chunk ? h->asm_chunk : std::min(chunk, 64);
    A.far_fp32 = (h->far_fp32 && !finite_depth && !rankine_only) ? 1 : 0;
    A.c64 = h->out_c64;
    A.freq = (const FreqArgs*)h->freq.p;
    A.S = (cplx*)S;
    A.ldS = ldS;
    A.strideS = strideS;
    A.K = (cplx*)K;
    A.ldK = ldK;
    A.strideK = strideK;
    int e = launch_assembly(A, stream);
    CPT_CHECK_HIP((hipError_t)e, "assembly kernel launch");
    return CPT_OK;
}

int cpt_build_matrices(cpt_handle h, int nb_collocation_points, const double* collocation_points,
                       const double* dot_product_normals, int nb_vertices, const double* vertices, int nb_faces,
                       const int32_t* faces, const double* centers, const double* normals, const double* areas,
                       const double* radiuses, int nb_quad_points, const double* quad_points,
                       const double* quad_weights, int nb_wavenumbers, const double* wavenumbers,
                       const int* gf_singularities, double depth, int finite_depth_method, int nexp,
                       const double* prony, int adjoint_double_layer, int early_dot_product,
                       int diagonal_term_in_double_layer, double free_surface, void* S, int64_t ldS,
                       int64_t strideS, void* K, int64_t ldK, int64_t strideK, hipStream_t stream)
{
    return build_matrices_impl(false, false, h, nb_collocation_points, collocation_points, dot_product_normals, nb_vertices,
                               vertices, nb_faces, faces, centers, normals, areas, radiuses, nb_quad_points,
                               quad_points, quad_weights, nb_wavenumbers, wavenumbers, gf_singularities, depth,
                               finite_depth_method, nexp, prony, adjoint_double_layer, early_dot_product,
                               diagonal_term_in_double_layer, free_surface, S, ldS, strideS, K, ldK, strideK, stream);
}

int cpt_build_matrices_dev(cpt_handle h, int nb_collocation_points, const double* collocation_points,
                           const double* dot_product_normals, int nb_vertices, const double* vertices, int nb_faces,
                           const int32_t* faces, const double* centers, const double* normals, const double* areas,
                           const double* radiuses, int nb_quad_points, const double* quad_points,
                           const double* quad_weights, int nb_wavenumbers, const double* wavenumbers,
                           const int* gf_singularities, double depth, int finite_depth_method, int nexp,
                           const double* prony, int adjoint_double_layer, int early_dot_product,
                           int diagonal_term_in_double_layer, double free_surface, void* S, int64_t ldS,
                           int64_t strideS, void* K, int64_t ldK, int64_t strideK, hipStream_t stream)
{
    return build_matrices_impl(true, false, h, nb_collocation_points, collocation_points, dot_product_normals, nb_vertices,
                               vertices, nb_faces, faces, centers, normals, areas, radiuses, nb_quad_points,
                               quad_points, quad_weights, nb_wavenumbers, wavenumbers, gf_singularities, depth,
                               finite_depth_method, nexp, prony, adjoint_double_layer, early_dot_product,
                               diagonal_term_in_double_layer, free_surface, S, ldS, strideS, K, ldK, strideK, stream);
}

// The FinGreen3D finite-depth Green function (matrices.f90:95-209 with
// finite_depth_method FINGREEN3D_METHOD; Green_wave.f90:545-616): the
// Rankine term and its free-surface image integrated on the panels by the
// assembly kernel (as in infinite depth at zero frequency, the sign of the
// image following gf_singularities), then the FinGreen3D wave part at the
// quadrature points added by its own kernel.
int cpt_build_matrices_fingreen3d(cpt_handle h, int nb_collocation_points, const double* collocation_points,
                                  const double* dot_product_normals, int nb_vertices, const double* vertices,
                                  int nb_faces, const int32_t* faces, const double* centers, const double* normals,
                                  const double* areas, const double* radiuses, int nb_quad_points,
                                  const double* quad_points, const double* quad_weights, double wavenumber,
                                  double depth, int nb_roots, const double* roots, int gf_singularities,
                                  int adjoint_double_layer, int early_dot_product, int diagonal_term_in_double_layer,
                                  double free_surface, void* S, int64_t ldS, void* K, int64_t ldK, hipStream_t stream)
{
    if (!h) {
        set_error("cpt_build_matrices_fingreen3d: null handle");
        return CPT_ERR_ARG;
    }
    if (!(std::isfinite(depth) && depth > 0.0) || !std::isfinite(free_surface)) {
        set_error("cpt_build_matrices_fingreen3d: FinGreen3D is only implemented for finite depth with a free surface");
        return CPT_ERR_ARG;
    }
    if (!(wavenumber > 0.0 && std::isfinite(wavenumber))) {
        set_error("cpt_build_matrices_fingreen3d: FinGreen3D is only implemented for non-zero and non-infinite "
                  "frequencies");
        return CPT_ERR_ARG;
    }
    // the series read roots up to index 98 (Pidcock at R/h = 0.0005)
    if (nb_roots < 100 || !roots) {
        set_error("cpt_build_matrices_fingreen3d: at least 100 dispersion roots are needed");
        return CPT_ERR_ARG;
    }
    if (h->out_c64) {
        set_error("cpt_build_matrices_fingreen3d: FinGreen3D assembles complex128 matrices only");
        return CPT_ERR_UNSUPPORTED;
    }
    const int n = nb_collocation_points, m = nb_faces, nk = early_dot_product ? 1 : 3;
    const int gfs = gf_singularities;
    // the Rankine parts (validates the remaining arguments, packs the panels)
    int st = build_matrices_impl(false, true, h, n, collocation_points, dot_product_normals, nb_vertices, vertices, m,
                                 faces, centers, normals, areas, radiuses, nb_quad_points, quad_points, quad_weights,
                                 1, &wavenumber, &gfs, INFINITY, CPT_NEWER_FINITE_DEPTH, 0, nullptr,
                                 adjoint_double_layer, early_dot_product, diagonal_term_in_double_layer, free_surface,
                                 S, ldS, ldS * m, K, ldK, ldK * (int64_t)m * nk, stream);
    if (st != CPT_OK || n == 0 || m == 0) return st;
    const int threads = fingreen3d_threads((int64_t)n * m);
    if (h->fg_roots.reserve(sizeof(double) * (size_t)nb_roots) ||
        h->fg_ws.reserve(sizeof(double) * FG_WS_DOUBLES * (size_t)threads) ||
        h->fg_lists.reserve(fingreen3d_list_bytes(n)) ||
        h->fg_tab.reserve(fingreen3d_table_bytes((size_t)n + (size_t)m * nb_quad_points))) {
        set_error("cpt_build_matrices_fingreen3d: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    CPT_CHECK_HIP(hipMemcpyAsync(h->fg_roots.p, roots, sizeof(double) * nb_roots, hipMemcpyHostToDevice, stream),
                  "copy dispersion roots");
    FgArgs A;
    A.colloc = collocation_points;
    A.dpn = dot_product_normals;
    A.quad = (const double*)h->quad.p;
    A.n = n;
    A.m = m;
    A.nq = nb_quad_points;
    A.nk = nk;
    A.adjoint = adjoint_double_layer ? 1 : 0;
    A.roots = (const double*)h->fg_roots.p;
    A.depth = depth;
    A.v = wavenumber * std::tanh(wavenumber * depth);  // omega^2/g (Green_wave.f90:571)
    A.S = (cplx*)S;
    A.ldS = ldS;
    A.K = (cplx*)K;
    A.ldK = ldK;
    A.ws = (double*)h->fg_ws.p;
    A.nthreads = threads;
    A.lists = (unsigned*)h->fg_lists.p;
    A.list_cap = n > FG_CHUNK_PAIRS ? n : FG_CHUNK_PAIRS;
    A.npts = (size_t)n + (size_t)m * nb_quad_points;
    A.tab = (double*)h->fg_tab.p;
    A.root_tab = A.tab + 4 * (size_t)100 * A.npts;
    CallScope call_scope(&h->timer, h->gemm_tiles);
    CPT_CHECK_HIP((hipError_t)launch_fingreen3d_pairs(A, stream), "FinGreen3D kernel launch");
    return CPT_OK;
}

int cpt_fingreen3d_points(cpt_handle h, int np, const double* rr, const double* zf, const double* zp, double v,
                          int nb_roots, const double* roots, double depth, void* grn, hipStream_t stream)
{
    if (!h || np < 0 || (np > 0 && (!rr || !zf || !zp || !roots || !grn))) {
        set_error("cpt_fingreen3d_points: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (nb_roots < 100 || !(std::isfinite(depth) && depth > 0.0) || !(v > 0.0 && std::isfinite(v))) {
        set_error("cpt_fingreen3d_points: needs at least 100 roots, a finite positive depth and omega^2/g");
        return CPT_ERR_ARG;
    }
    if (np == 0) return CPT_OK;
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    const int threads = fingreen3d_threads(np);
    if (h->fg_ws.reserve(sizeof(double) * FG_WS_DOUBLES * (size_t)threads)) {
        set_error("cpt_fingreen3d_points: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    CPT_CHECK_HIP((hipError_t)launch_fingreen3d_points(np, rr, zf, zp, v, roots, depth, (cplx*)grn,
                                                       (double*)h->fg_ws.p, threads, stream),
                  "FinGreen3D points kernel launch");
    return CPT_OK;
}

int cpt_build_matrices_batched(cpt_handle h, int nb_bodies, const cpt_body* bodies, int adjoint_double_layer,
                               int diagonal_term_in_double_layer, double free_surface, hipStream_t stream)
{
    if (!h || nb_bodies < 0 || (nb_bodies > 0 && !bodies)) {
        set_error("cpt_build_matrices_batched: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (std::isinf(free_surface)) {
        set_error("cpt_build_matrices_batched: free_surface=inf is not supported in the batched form");
        return CPT_ERR_UNSUPPORTED;
    }
    size_t tot_m = 0, tot_q = 0, tot_nb = 0;
    bool need_table = false;
    for (int i = 0; i < nb_bodies; ++i) {
        const cpt_body& B = bodies[i];
        const int m = B.nb_faces, nb = B.nb_wavenumbers;
        if (m < 0 || B.nb_vertices < 0 || B.nb_quad_points < 1 || B.nb_quad_points > 16 || nb < 1 || B.ldS < m ||
            B.ldK < m || (nb > 1 && (B.strideS < B.ldS * m || B.strideK < B.ldK * m))) {
            set_error("cpt_build_matrices_batched: invalid sizes or strides of body " + std::to_string(i));
            return CPT_ERR_ARG;
        }
        if (m > 0 && (!B.vertices || !B.faces || !B.centers || !B.normals || !B.areas || !B.radiuses ||
                      !B.quad_points || !B.quad_weights || !B.wavenumbers || !B.gf_singularities || !B.S || !B.K)) {
            set_error("cpt_build_matrices_batched: null pointer in body " + std::to_string(i));
            return CPT_ERR_ARG;
        }
        for (int b = 0; b < nb; ++b) {
            const double k = B.wavenumbers[b];
            const int g = B.gf_singularities[b];
            if (g != CPT_HIGH_FREQ && g != CPT_LOW_FREQ && g != CPT_LOW_FREQ_WITH_RANKINE_PART) {
                set_error("cpt_build_matrices_batched: invalid gf_singularities value");
                return CPT_ERR_ARG;
            }
            if (std::isnan(k) || k < 0) {
                set_error("cpt_build_matrices_batched: invalid wavenumber");
                return CPT_ERR_ARG;
            }
            if (k > 0 && std::isfinite(k)) need_table = true;
        }
        tot_m += m;
        tot_q += (size_t)m * B.nb_quad_points;
        tot_nb += nb;
    }
    if (tot_m == 0) return CPT_OK;
    if (need_table && !h->has_table) {
        set_error("cpt_build_matrices_batched: no tabulation set (call cpt_set_tabulation first)");
        return CPT_ERR_STATE;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    if (h->bfaces.reserve(sizeof(Face) * tot_m) || h->bquad.reserve(sizeof(double) * 4 * tot_q) ||
        h->bfreq.reserve(sizeof(FreqArgs) * tot_nb) || h->bargs.reserve(sizeof(AssemblyArgs) * nb_bodies)) {
        set_error("cpt_build_matrices_batched: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    // Panels per tile: the same for every body, sized for ~8 tiles per
    // resident work-group over the whole work-list.
    size_t rowblock_panels = 0;
    for (int i = 0; i < nb_bodies; ++i)
        rowblock_panels += (size_t)((bodies[i].nb_faces + 255) / 256) * bodies[i].nb_faces;
    const int chunk = h->asm_chunk ? h->asm_chunk
                                   : (int)std::min<size_t>(64, std::max<size_t>(1, rowblock_panels / 8192));
    std::vector<FreqArgs> fa(tot_nb);
    std::vector<AssemblyArgs> args(nb_bodies);
    std::vector<AsmTile> tiles;
    std::vector<double> cost;
    Face* faces_d = (Face*)h->bfaces.p;
    double* quad_d = (double*)h->bquad.p;
    const FreqArgs* freq_d = (const FreqArgs*)h->bfreq.p;
    size_t om = 0, oq = 0, ob = 0;
    double pairs = 0.0;
    const Table T = table_of(h);
    for (int i = 0; i < nb_bodies; ++i) {
        const cpt_body& B = bodies[i];
        const int m = B.nb_faces, nq = B.nb_quad_points, nb = B.nb_wavenumbers;
        if (m > 0) {
            hipLaunchKernelGGL(pack_faces_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, B.vertices,
                               B.nb_vertices, B.faces, B.centers, B.normals, B.areas, B.radiuses, B.quad_points,
                               B.quad_weights, m, nq, faces_d + om, quad_d + 4 * oq);
            CPT_CHECK_HIP(hipGetLastError(), "pack kernel launch");
        }
        for (int b = 0; b < nb; ++b) {
            const double k = B.wavenumbers[b];
            FreqArgs& f = fa[ob + b];
            f.k = k;
            f.gfs = B.gf_singularities[b];
            f.wave = (0.0 < k && std::isfinite(k)) ? 1 : 0;
            f.nexp = 0;
            f.pad = 0;
            f.prony_lambda = nullptr;
            f.prony_a = nullptr;
        }
        AssemblyArgs& A = args[i];
        std::memset(&A, 0, sizeof(A));
        A.colloc = B.centers;
        A.dpn = B.normals;
        A.n = m;
        A.m = m;
        A.nq = nq;
        A.nk = 1;
        A.faces = faces_d + om;
        A.quad = quad_d + 4 * oq;
        A.T = T;
        A.depth = INFINITY;
        A.adjoint = adjoint_double_layer ? 1 : 0;
        A.diag = diagonal_term_in_double_layer ? 1 : 0;
        A.n_diag = m;
        A.free_surface = free_surface;
        A.nb = nb;
        A.j_chunk = chunk;
        A.far_fp32 = h->far_fp32;
        A.c64 = h->out_c64;
        A.freq = freq_d + ob;
        A.S = (cplx*)B.S;
        A.ldS = B.ldS;
        A.strideS = B.strideS;
        A.K = (cplx*)B.K;
        A.ldK = B.ldK;
        A.strideK = B.strideK;
        for (int ib = 0; ib < (m + 255) / 256; ++ib)
            for (int jc = 0; jc * chunk < m; ++jc) {
                tiles.push_back(AsmTile{i, ib, jc, 0});
                cost.push_back((double)nb * (std::min(m, (jc + 1) * chunk) - jc * chunk));
            }
        pairs += (double)m * m * nb;
        om += m;
        oq += (size_t)m * nq;
        ob += nb;
    }
    // Most expensive tiles first so that the grid stride ends evenly.
    std::vector<int> order(tiles.size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = (int)t;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<AsmTile> sorted(tiles.size());
    for (size_t t = 0; t < order.size(); ++t) sorted[t] = tiles[order[t]];
    if (h->btiles.reserve(sizeof(AsmTile) * sorted.size())) {
        set_error("cpt_build_matrices_batched: device allocation failed");
        return CPT_ERR_ALLOC;
    }
    CPT_CHECK_HIP(hipMemcpyAsync(h->bfreq.p, fa.data(), sizeof(FreqArgs) * tot_nb, hipMemcpyHostToDevice, stream),
                  "copy frequency table");
    CPT_CHECK_HIP(hipMemcpyAsync(h->bargs.p, args.data(), sizeof(AssemblyArgs) * nb_bodies, hipMemcpyHostToDevice,
                                 stream),
                  "copy body table");
    CPT_CHECK_HIP(hipMemcpyAsync(h->btiles.p, sorted.data(), sizeof(AsmTile) * sorted.size(), hipMemcpyHostToDevice,
                                 stream),
                  "copy work-list");
    // Persistent grid: two work-groups per CU (the kernel's occupancy) on
    // every CU, or fewer when the work-list is short.
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device);
    const int grid = (int)std::min<size_t>(sorted.size(), (size_t)ncu * 2);
    const int e = launch_assembly_worklist((const AssemblyArgs*)h->bargs.p, (const AsmTile*)h->btiles.p,
                                           (int)sorted.size(), grid, pairs, h->far_fp32, h->nr <= 1 || h->nz <= 1,
                                           stream);
    CPT_CHECK_HIP((hipError_t)e, "batched assembly kernel launch");
    // The host tables above are copied from pageable memory: wait for the
    // copies before they go out of scope.
    CPT_CHECK_HIP(hipStreamSynchronize(stream), "batched assembly upload");
    return CPT_OK;
}

int cpt_wave_part_infinite_depth(cpt_handle h, int np, const double* x0i, const double* x0j, double wavenumber,
                                 int gf_singularities, void* g, void* nablag, hipStream_t stream)
{
    if (!h || np < 0 || (np > 0 && (!x0i || !x0j || !g || !nablag))) {
        set_error("cpt_wave_part_infinite_depth: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (!h->has_table) {
        set_error("cpt_wave_part_infinite_depth: no tabulation set");
        return CPT_ERR_STATE;
    }
    if (np == 0) return CPT_OK;
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    if (h->nr <= 1 || h->nz <= 1)
        hipLaunchKernelGGL(wave_points_kernel<true>, dim3((np + 255) / 256), dim3(256), 0, stream, table_of(h), np,
                           x0i, x0j, wavenumber, gf_singularities, (cplx*)g, (cplx*)nablag);
    else
        hipLaunchKernelGGL(wave_points_kernel<false>, dim3((np + 255) / 256), dim3(256), 0, stream, table_of(h), np,
                           x0i, x0j, wavenumber, gf_singularities, (cplx*)g, (cplx*)nablag);
    CPT_CHECK_HIP(hipGetLastError(), "wave kernel launch");
    return CPT_OK;
}

int cpt_lu_factor(cpt_handle h, void* A, int n, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP,
                  int* info, int batch, hipStream_t stream)
{
    if (!h || n < 0 || batch < 0 || lda < n || (batch > 1 && (strideA < lda * n || strideP < n))) {
        set_error("cpt_lu_factor: invalid sizes or strides");
        return CPT_ERR_ARG;
    }
    if (n == 0 || batch == 0) return CPT_OK;
    if (!A || !ipiv || !info) {
        set_error("cpt_lu_factor: null pointer argument");
        return CPT_ERR_ARG;
    }
    if (n > 24576) {
        set_error("cpt_lu_factor: n > 24576 is not supported");
        return CPT_ERR_UNSUPPORTED;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    hipLaunchKernelGGL(zero_int_kernel, dim3((batch + 255) / 256), dim3(256), 0, stream, info, batch);
    CPT_CHECK_HIP(hipGetLastError(), "zero info launch");
    if (h->lu_work.reserve(getrf_workspace_bytes(n, batch))) {
        set_error("cpt_lu_factor: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lu_streams(h)) {
        set_error("cpt_lu_factor: could not create the look-ahead streams: " + std::string(cpt_last_error()));
        return CPT_ERR_HIP;
    }
    CPT_CHECK_HIP((hipError_t)launch_getrf((cplx*)A, n, lda, strideA, ipiv, strideP, info, batch,
                                           (cplx*)h->lu_work.p, &h->lu_streams, stream),
                  "getrf launch");
    return CPT_OK;
}

int cpt_lu_solve(cpt_handle h, const void* LU, int n, int64_t lda, int64_t strideA, const int* ipiv,
                 int64_t strideP, void* B, int nrhs, int64_t ldb, int64_t strideB, int batch, hipStream_t stream)
{
    if (!h || n < 0 || nrhs < 0 || batch < 0 || lda < n || ldb < n ||
        (batch > 1 && (strideA < lda * n || strideP < n || strideB < ldb * nrhs))) {
        set_error("cpt_lu_solve: invalid sizes or strides");
        return CPT_ERR_ARG;
    }
    if (n == 0 || nrhs == 0 || batch == 0) return CPT_OK;
    if (!LU || !ipiv || !B) {
        set_error("cpt_lu_solve: null pointer argument");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    if (h->solve_work.reserve(getrs_workspace_bytes(n, batch))) {
        set_error("cpt_lu_solve: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    CPT_CHECK_HIP((hipError_t)launch_getrs((const cplx*)LU, n, lda, strideA, ipiv, strideP, (cplx*)B, nrhs, ldb,
                                           strideB, batch, (cplx*)h->solve_work.p, stream),
                  "getrs launch");
    return CPT_OK;
}

int cpt_lu_solve_noalloc(cpt_handle h, const void* LU, int n, int64_t lda, int64_t strideA, const int* ipiv,
                         int64_t strideP, void* B, int nrhs, int64_t ldb, int64_t strideB, int batch,
                         hipStream_t stream)
{
    if (!h || n < 0 || nrhs < 0 || batch < 0 || lda < n || ldb < n ||
        (batch > 1 && (strideA < lda * n || strideP < n || strideB < ldb * nrhs))) {
        set_error("cpt_lu_solve_noalloc: invalid sizes or strides");
        return CPT_ERR_ARG;
    }
    if ((int64_t)n * batch > CPT_LU_SOLVE_NOALLOC_MAX_ROWS) {
        set_error("cpt_lu_solve_noalloc: n x batch = " + std::to_string((int64_t)n * batch) +
                  " exceeds the static scratch (" + std::to_string(CPT_LU_SOLVE_NOALLOC_MAX_ROWS) +
                  " rows); use cpt_lu_solve");
        return CPT_ERR_ARG;
    }
    if (n == 0 || nrhs == 0 || batch == 0) return CPT_OK;
    if (!LU || !ipiv || !B) {
        set_error("cpt_lu_solve_noalloc: null pointer argument");
        return CPT_ERR_ARG;
    }
    void* ptr = (h->device >= 0 && h->device < MAX_DEVICES) ? g_scratch_addr[h->device] : nullptr;
    if (!ptr) {
        set_error("cpt_lu_solve_noalloc: no scratch address for this device");
        return CPT_ERR_STATE;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(nullptr, h->gemm_tiles);  // no timing events
    CPT_CHECK_HIP((hipError_t)launch_getrs((const cplx*)LU, n, lda, strideA, ipiv, strideP, (cplx*)B, nrhs, ldb,
                                           strideB, batch, (cplx*)ptr, stream),
                  "getrs launch");
    return CPT_OK;
}

int cpt_zgemm(cpt_handle h, int M, int N, int K, double alpha, const void* A, int64_t lda, int64_t strideA,
              const void* B, int64_t ldb, int64_t strideB, void* C, int64_t ldc, int64_t strideC, int beta_zero,
              int batch, hipStream_t stream)
{
    if (!h || M < 0 || N < 0 || K < 0 || batch < 0 || lda < M || ldb < K || ldc < M) {
        set_error("cpt_zgemm: invalid sizes");
        return CPT_ERR_ARG;
    }
    if (M == 0 || N == 0 || batch == 0) return CPT_OK;
    if (!A || !B || !C) {
        set_error("cpt_zgemm: null pointer argument");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    CPT_CHECK_HIP((hipError_t)launch_zgemm(M, N, K, alpha, (const cplx*)A, lda, strideA, (const cplx*)B, ldb, strideB,
                                           (cplx*)C, ldc, strideC, beta_zero, batch, stream),
                  "zgemm launch");
    return CPT_OK;
}

// complex64 products and LU run the FP32-input tile products: the FP16-input
// ones (tile setting 4) serve the mixed solve, whose refinement restores the
// digits they lose, and are replaced here by the 128 x 128 FP32 tiles.
static int fp32_tiles(const cpt_context* h) { return h->gemm_tiles == 4 ? 1 : h->gemm_tiles; }

int cpt_cgemm(cpt_handle h, int M, int N, int K, float alpha, const void* A, int64_t lda, int64_t strideA,
              const void* B, int64_t ldb, int64_t strideB, void* C, int64_t ldc, int64_t strideC, int beta_zero,
              int batch, hipStream_t stream)
{
    if (!h || M < 0 || N < 0 || K < 0 || batch < 0 || lda < M || ldb < K || ldc < M) {
        set_error("cpt_cgemm: invalid sizes");
        return CPT_ERR_ARG;
    }
    if (M == 0 || N == 0 || batch == 0) return CPT_OK;
    if (!A || !B || !C) {
        set_error("cpt_cgemm: null pointer argument");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, fp32_tiles(h));
    CPT_CHECK_HIP((hipError_t)launch_cgemm(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, C, ldc, strideC,
                                           beta_zero, batch, stream),
                  "cgemm launch");
    return CPT_OK;
}

int cpt_clu_factor(cpt_handle h, void* A, int n, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP,
                   int* info, int batch, hipStream_t stream)
{
    if (!h || n < 0 || batch < 0 || lda < n || (batch > 1 && (strideA < lda * n || strideP < n))) {
        set_error("cpt_clu_factor: invalid sizes or strides");
        return CPT_ERR_ARG;
    }
    if (n == 0 || batch == 0) return CPT_OK;
    if (!A || !ipiv || !info) {
        set_error("cpt_clu_factor: null pointer argument");
        return CPT_ERR_ARG;
    }
    if (n > 24576) {
        set_error("cpt_clu_factor: n > 24576 is not supported");
        return CPT_ERR_UNSUPPORTED;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, fp32_tiles(h));
    hipLaunchKernelGGL(zero_int_kernel, dim3((batch + 255) / 256), dim3(256), 0, stream, info, batch);
    CPT_CHECK_HIP(hipGetLastError(), "zero info launch");
    if (h->lu_work.reserve(cgetrf_workspace_bytes(n, batch))) {
        set_error("cpt_clu_factor: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lu_streams(h)) {
        set_error("cpt_clu_factor: could not create the look-ahead streams: " + std::string(cpt_last_error()));
        return CPT_ERR_HIP;
    }
    CPT_CHECK_HIP((hipError_t)launch_cgetrf(A, n, lda, strideA, ipiv, strideP, info, batch, h->lu_work.p,
                                            &h->lu_streams, stream),
                  "cgetrf launch");
    return CPT_OK;
}

int cpt_clu_solve(cpt_handle h, const void* LU, int n, int64_t lda, int64_t strideA, const int* ipiv,
                  int64_t strideP, void* B, int nrhs, int64_t ldb, int64_t strideB, int batch, hipStream_t stream)
{
    if (!h || n < 0 || nrhs < 0 || batch < 0 || lda < n || ldb < n ||
        (batch > 1 && (strideA < lda * n || strideP < n || strideB < ldb * nrhs))) {
        set_error("cpt_clu_solve: invalid sizes or strides");
        return CPT_ERR_ARG;
    }
    if (n == 0 || nrhs == 0 || batch == 0) return CPT_OK;
    if (!LU || !ipiv || !B) {
        set_error("cpt_clu_solve: null pointer argument");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, fp32_tiles(h));
    if (h->solve_work.reserve(cgetrs_workspace_bytes(n, batch))) {
        set_error("cpt_clu_solve: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    CPT_CHECK_HIP((hipError_t)launch_cgetrs(LU, n, lda, strideA, ipiv, strideP, B, nrhs, ldb, strideB, batch,
                                            h->solve_work.p, stream),
                  "cgetrs launch");
    return CPT_OK;
}

static int gmres_entry(cpt_handle h, bool single, const char* name, const void* A, int n, int64_t lda, const void* B,
                       int nrhs, int64_t ldb, void* X, int64_t ldx, int restart, int maxiter, double rtol,
                       double atol, int* iterations, double* residual, hipStream_t stream)
{
    const std::string who(name);
    if (!h || n < 0 || nrhs < 0 || lda < n || ldb < n || ldx < n || !(rtol >= 0.0) || !(atol >= 0.0) ||
        restart > gmres_max_restart()) {
        set_error(who + ": invalid arguments");
        return CPT_ERR_ARG;
    }
    if (n == 0 || nrhs == 0) return CPT_OK;
    if (!A || !B || !X || !iterations || !residual) {
        set_error(who + ": null pointer argument");
        return CPT_ERR_ARG;
    }
    // scipy.sparse.linalg.gmres defaults: restart 20 (at most n), 10 n cycles
    if (restart <= 0) restart = 20;
    restart = std::min(restart, n);
    if (maxiter <= 0) maxiter = (int)std::min<int64_t>(10 * (int64_t)n, 1 << 30);
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, fp32_tiles(h));
    if (h->gmres_work.reserve(gmres_workspace_bytes(n, restart, single))) {
        set_error(who + ": workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    int not_converged = 0;
    CPT_CHECK_HIP((hipError_t)run_gmres(single, A, n, lda, B, nrhs, ldb, X, ldx, restart, maxiter, rtol, atol,
                                        h->gmres_work.p, stream, iterations, residual, &not_converged),
                  name);
    if (not_converged) {
        set_error(who + ": " + std::to_string(not_converged) + " right-hand side(s) did not converge");
        return CPT_ERR_NOT_CONVERGED;
    }
    return CPT_OK;
}

int cpt_zgmres(cpt_handle h, const void* A, int n, int64_t lda, const void* B, int nrhs, int64_t ldb, void* X,
               int64_t ldx, int restart, int maxiter, double rtol, double atol, int* iterations, double* residual,
               hipStream_t stream)
{
    return gmres_entry(h, false, "cpt_zgmres", A, n, lda, B, nrhs, ldb, X, ldx, restart, maxiter, rtol, atol,
                       iterations, residual, stream);
}

int cpt_cgmres(cpt_handle h, const void* A, int n, int64_t lda, const void* B, int nrhs, int64_t ldb, void* X,
               int64_t ldx, int restart, int maxiter, double rtol, double atol, int* iterations, double* residual,
               hipStream_t stream)
{
    return gmres_entry(h, true, "cpt_cgmres", A, n, lda, B, nrhs, ldb, X, ldx, restart, maxiter, rtol, atol,
                       iterations, residual, stream);
}

int cpt_block_transform(cpt_handle h, int kind, int n, int inverse, int rows, int cols, int batch, const void* in,
                        int64_t in_batch_stride, int64_t in_block_stride, int64_t ldi, void* out,
                        int64_t out_batch_stride, int64_t out_block_stride, int64_t ldo, hipStream_t stream)
{
    if (!h || (kind != CPT_BLOCKS_CYCLIC && kind != CPT_BLOCKS_XOR) || n < 1 || rows < 0 || cols < 0 || batch < 0 ||
        ldi < rows || ldo < rows || in_block_stride < 0 || out_block_stride < 0 || in_batch_stride < 0 ||
        out_batch_stride < 0) {
        set_error("cpt_block_transform: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (n > block_transform_max_blocks()) {
        set_error("cpt_block_transform: at most " + std::to_string(block_transform_max_blocks()) + " blocks");
        return CPT_ERR_UNSUPPORTED;
    }
    if (kind == CPT_BLOCKS_XOR && (n & (n - 1)) != 0) {
        set_error("cpt_block_transform: the reflection kind needs a power-of-two number of blocks");
        return CPT_ERR_ARG;
    }
    if (rows == 0 || cols == 0 || batch == 0) return CPT_OK;
    if (!in || !out) {
        set_error("cpt_block_transform: null pointer argument");
        return CPT_ERR_ARG;
    }
    if (in == out) {
        set_error("cpt_block_transform: input and output must be distinct buffers");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CPT_CHECK_HIP((hipError_t)launch_block_transform(kind == CPT_BLOCKS_XOR, n, inverse ? 1 : 0, rows, cols, batch,
                                                     (const cplx*)in, in_batch_stride, in_block_stride, ldi,
                                                     (cplx*)out, out_batch_stride, out_block_stride, ldo, stream),
                  "block transform launch");
    return CPT_OK;
}

int cpt_lu_solve_mixed(cpt_handle h, const void* K, int n, int64_t lda, int64_t strideK, int batch, const void* B,
                       int nrhs, int64_t ldb, int64_t strideB, void* X, int64_t ldx, int64_t strideX, int max_iter,
                       double tol, int* iterations, double* residual, hipStream_t stream)
{
    if (!h || n < 0 || nrhs < 0 || batch < 0 || lda < n || ldb < n || ldx < n || max_iter < 0 || !(tol >= 0) ||
        (batch > 1 && (strideK < lda * n || strideB < ldb * nrhs || strideX < ldx * nrhs))) {
        set_error("cpt_lu_solve_mixed: invalid sizes, strides, max_iter or tol");
        return CPT_ERR_ARG;
    }
    if (!iterations || !residual) {
        set_error("cpt_lu_solve_mixed: null output pointer");
        return CPT_ERR_ARG;
    }
    *iterations = 0;
    *residual = 0.0;
    if (n == 0 || nrhs == 0 || batch == 0) return CPT_OK;
    if (!K || !B || !X) {
        set_error("cpt_lu_solve_mixed: null pointer argument");
        return CPT_ERR_ARG;
    }
    if (n > 24576) {
        set_error("cpt_lu_solve_mixed: n > 24576 is not supported");
        return CPT_ERR_UNSUPPORTED;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    const MixedGroup g{(const cplx*)K, n, lda, strideK, batch, (const cplx*)B, nrhs, ldb, strideB, (cplx*)X, ldx,
                       strideX, stream};
    const size_t fb = mixed_factor_bytes(n, batch), wb = mixed_factor_ws_bytes(n, batch);
    void* work = mixed_workspace(h, mixed_workspace_bytes(n, nrhs, batch));
    if (!work) {
        set_error("cpt_lu_solve_mixed: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lu_streams(h)) {
        set_error("cpt_lu_solve_mixed: could not create the look-ahead streams: " + std::string(cpt_last_error()));
        return CPT_ERR_HIP;
    }
    char* base = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    const MixedFactor F = mixed_factor_at(base, n, batch);
    int tiles = 0;
    return mixed_solve_one(h, g, F, base + fb, base + fb + wb, true, &tiles, max_iter, tol, iterations, residual,
                           "cpt_lu_solve_mixed");
}

int cpt_mixed_factor_size(int n, int batch, size_t* bytes)
{
    if (n < 0 || batch < 0 || !bytes) {
        set_error("cpt_mixed_factor_size: invalid arguments");
        return CPT_ERR_ARG;
    }
    *bytes = FACTOR_HEADER + mixed_factor_bytes(n, batch);
    return CPT_OK;
}

int cpt_mixed_factor(cpt_handle h, const void* K, int n, int64_t lda, int64_t strideK, int batch, void* factor,
                     size_t factor_bytes, hipStream_t stream)
{
    if (!h || n < 1 || batch < 1 || lda < n || (batch > 1 && strideK < lda * n) || !K || !factor ||
        ((uintptr_t)factor & 255) != 0) {
        set_error("cpt_mixed_factor: invalid sizes, strides or pointers (the factor must be 256-byte aligned)");
        return CPT_ERR_ARG;
    }
    if (factor_bytes < FACTOR_HEADER + mixed_factor_bytes(n, batch)) {
        set_error("cpt_mixed_factor: factor buffer smaller than cpt_mixed_factor_size");
        return CPT_ERR_ARG;
    }
    if (n > 24576) {
        set_error("cpt_mixed_factor: n > 24576 is not supported");
        return CPT_ERR_UNSUPPORTED;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    if (h->lu_work.reserve(mixed_factor_ws_bytes(n, batch))) {
        set_error("cpt_mixed_factor: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lu_streams(h)) {
        set_error("cpt_mixed_factor: could not create the look-ahead streams: " + std::string(cpt_last_error()));
        return CPT_ERR_HIP;
    }
    const MixedFactor F = mixed_factor_at((char*)factor + FACTOR_HEADER, n, batch);
    hipLaunchKernelGGL(factor_header_kernel, dim3(1), dim3(64), 0, stream, (int*)factor, FACTOR_MAGIC, n, batch,
                       h->gemm_tiles);
    CPT_CHECK_HIP(hipGetLastError(), "factor header");
    CPT_CHECK_HIP((hipError_t)mixed_factor((const cplx*)K, n, lda, strideK, batch, F, h->lu_work.p, &h->lu_streams,
                                           stream),
                  "cpt_mixed_factor");
    return CPT_OK;
}

int cpt_mixed_refine(cpt_handle h, const void* K, int n, int64_t lda, int64_t strideK, int batch, void* factor,
                     size_t factor_bytes, const void* B, int nrhs, int64_t ldb, int64_t strideB, void* X, int64_t ldx,
                     int64_t strideX, int max_iter, double tol, int* iterations, double* residual, hipStream_t stream)
{
    if (!h || n < 1 || batch < 1 || nrhs < 0 || lda < n || ldb < n || ldx < n || max_iter < 0 || !(tol >= 0) ||
        (batch > 1 && (strideK < lda * n || strideB < ldb * nrhs || strideX < ldx * nrhs)) || !K || !factor ||
        ((uintptr_t)factor & 255) != 0) {
        set_error("cpt_mixed_refine: invalid sizes, strides, pointers, max_iter or tol");
        return CPT_ERR_ARG;
    }
    if (!iterations || !residual) {
        set_error("cpt_mixed_refine: null output pointer");
        return CPT_ERR_ARG;
    }
    *iterations = 0;
    *residual = 0.0;
    if (factor_bytes < FACTOR_HEADER + mixed_factor_bytes(n, batch)) {
        set_error("cpt_mixed_refine: factor buffer smaller than cpt_mixed_factor_size");
        return CPT_ERR_ARG;
    }
    if (nrhs == 0) return CPT_OK;
    if (!B || !X) {
        set_error("cpt_mixed_refine: null pointer argument");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    int head[4];
    CPT_CHECK_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    CPT_CHECK_HIP(hipMemcpy(head, factor, sizeof head, hipMemcpyDeviceToHost), "reading the factor header");
    if (head[0] != FACTOR_MAGIC || head[1] != n || head[2] != batch) {
        set_error("cpt_mixed_refine: the factor buffer does not hold a cpt_mixed_factor of this order and batch");
        return CPT_ERR_ARG;
    }
    const MixedGroup g{(const cplx*)K, n, lda, strideK, batch, (const cplx*)B, nrhs, ldb, strideB, (cplx*)X, ldx,
                       strideX, stream};
    const size_t wb = mixed_factor_ws_bytes(n, batch);
    void* work = mixed_workspace(h, 256 + wb + mixed_refine_bytes(&g, 1));
    if (!work) {
        set_error("cpt_mixed_refine: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lu_streams(h)) {
        set_error("cpt_mixed_refine: could not create the look-ahead streams: " + std::string(cpt_last_error()));
        return CPT_ERR_HIP;
    }
    char* base = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    const MixedFactor F = mixed_factor_at((char*)factor + FACTOR_HEADER, n, batch);
    int tiles = head[3];
    const int status = mixed_solve_one(h, g, F, base, base + wb, false, &tiles, max_iter, tol, iterations, residual,
                                       "cpt_mixed_refine");
    if (tiles != head[3]) {
        // factored again with FP32-input products: later refinements start there
        hipLaunchKernelGGL(factor_header_kernel, dim3(1), dim3(64), 0, stream, (int*)factor, FACTOR_MAGIC, n, batch,
                           tiles);
        CPT_CHECK_HIP(hipGetLastError(), "factor header");
    }
    return status;
}

}  // extern "C"

int cpt_profile_begin(cpt_handle h)
{
    if (!h) {
        set_error("cpt_profile_begin: null handle");
        return CPT_ERR_ARG;
    }
    h->timer.used = 0;
    h->timer.on = true;
    return CPT_OK;
}

int cpt_profile_end(cpt_handle h, double* ms, double* work, int64_t* launches, double* bytes)
{
    if (!h || !ms || !work || !launches) {
        set_error("cpt_profile_end: null argument");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    for (int k = 0; k < CPT_TIMED_KERNELS; ++k) {
        ms[k] = 0.0;
        work[k] = 0.0;
        launches[k] = 0;
        if (bytes) bytes[k] = 0.0;
    }
    KernelTimer& t = h->timer;
    t.on = false;
    for (size_t i = 0; i < t.used; ++i) {
        KernelTimer::Rec& r = t.recs[i];
        CPT_CHECK_HIP(hipEventSynchronize(r.b), "hipEventSynchronize");
        float dt = 0.0f;
        CPT_CHECK_HIP(hipEventElapsedTime(&dt, r.a, r.b), "hipEventElapsedTime");
        ms[r.kind] += dt;
        work[r.kind] += r.work;
        launches[r.kind] += 1;
        if (bytes) bytes[r.kind] += r.bytes;
    }
    t.used = 0;
    return CPT_OK;
}

int cpt_set_tuning(cpt_handle h, int param, int value)
{
    if (!h) {
        set_error("cpt_set_tuning: null handle");
        return CPT_ERR_ARG;
    }
    switch (param) {
    case CPT_TUNE_LU_BLOCK:
        if (value != 0 && value != 64 && value != 128 && value != 256 && value != 512 && value != 1024) break;
        h->lu_streams.block = value;
        return CPT_OK;
    case CPT_TUNE_PANEL_LEAF:
        if (value < 1 || value > 64) break;
        h->lu_streams.leaf = value;
        return CPT_OK;
    case CPT_TUNE_LOOKAHEAD:
        if (value == 2) {
            // CU-masked streams: refused by the runtime on MI355X, and a
            // process creating one from a partial mask crashed inside it
            set_error("cpt_set_tuning: look-ahead mode 2 (CU-masked streams) is not supported");
            return CPT_ERR_UNSUPPORTED;
        }
        if (value < 0 || value > 1) break;
        h->lu_streams.lookahead = value;
        return CPT_OK;
    case CPT_TUNE_GEMM_TILES:
        if (value < 0 || value > 4) break;
        h->gemm_tiles = value;
        return CPT_OK;
    case CPT_TUNE_ASSEMBLY_CHUNK:
        if (value < 0 || value > 1024) break;
        h->asm_chunk = value;
        return CPT_OK;
    case CPT_TUNE_PANEL_THREADS:
        if (value != 256 && value != 512) break;
        h->lu_streams.panel_threads = value;
        return CPT_OK;
    case CPT_TUNE_PANEL_WGS:
        if (value != 0 && value != 1) break;
        h->lu_streams.panel_mw = value;
        return CPT_OK;
    case CPT_TUNE_PANEL_SPIN:
        h->lu_streams.panel_spin = value;
        return CPT_OK;
    case CPT_TUNE_PIVOT_SPECULATION:
        if (value != 0 && value != 1) break;
        h->lu_streams.speculate = value;
        return CPT_OK;
    default:
        set_error("cpt_set_tuning: unknown parameter " + std::to_string(param));
        return CPT_ERR_ARG;
    }
    set_error("cpt_set_tuning: value " + std::to_string(value) + " out of range for parameter " +
              std::to_string(param));
    return CPT_ERR_ARG;
}

int cpt_get_tuning(cpt_handle h, int param, int* value)
{
    if (!h || !value) {
        set_error("cpt_get_tuning: null argument");
        return CPT_ERR_ARG;
    }
    switch (param) {
    case CPT_TUNE_LU_BLOCK: *value = h->lu_streams.block; return CPT_OK;
    case CPT_TUNE_PANEL_LEAF: *value = h->lu_streams.leaf; return CPT_OK;
    case CPT_TUNE_LOOKAHEAD: *value = h->lu_streams.lookahead; return CPT_OK;
    case CPT_TUNE_GEMM_TILES: *value = h->gemm_tiles; return CPT_OK;
    case CPT_TUNE_LOOKAHEAD_ACTIVE: *value = h->lu_streams.lookahead; return CPT_OK;
    case CPT_TUNE_ASSEMBLY_CHUNK: *value = h->asm_chunk; return CPT_OK;
    case CPT_TUNE_PANEL_THREADS: *value = h->lu_streams.panel_threads; return CPT_OK;
    case CPT_TUNE_PANEL_WGS: *value = h->lu_streams.panel_mw; return CPT_OK;
    case CPT_TUNE_PANEL_SPIN: *value = h->lu_streams.panel_spin; return CPT_OK;
    case CPT_TUNE_MIXED_RETRIES: *value = (int)std::min<long>(h->mixed_retries, 0x7fffffff); return CPT_OK;
    case CPT_TUNE_PIVOT_SPECULATION: *value = h->lu_streams.speculate; return CPT_OK;
    case CPT_TUNE_SPEC_REPAIRS:
        // the device counter: waits for the device's work
        CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
        CPT_CHECK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
        CPT_CHECK_HIP(hipMemcpy(value, h->lu_streams.spec_repairs, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
        return CPT_OK;
    default:
        set_error("cpt_get_tuning: unknown parameter " + std::to_string(param));
        return CPT_ERR_ARG;
    }
}

int cpt_lu_factor_grouped(cpt_handle h, int ngroups, const cpt_lu_group* groups, hipStream_t stream)
{
    if (!h || ngroups < 0 || (ngroups > 0 && !groups)) {
        set_error("cpt_lu_factor_grouped: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (!valid_groups(groups, ngroups, true, false, false, "cpt_lu_factor_grouped")) return CPT_ERR_ARG;
    if (ngroups == 0) return CPT_OK;
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    size_t ws = 0;
    for (int i = 0; i < ngroups; ++i)
        if (groups[i].n > 0 && groups[i].batch > 0) ws += getrf_workspace_bytes(groups[i].n, groups[i].batch);
    if (h->lu_work.reserve(ws)) {
        set_error("cpt_lu_factor_grouped: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lanes(h) || fork_lanes(h, stream)) {
        set_error("cpt_lu_factor_grouped: could not set up the lane streams");
        return CPT_ERR_HIP;
    }
    const std::vector<hipStream_t> lane = assign_lanes(h, groups, ngroups);
    std::vector<LuGroupArgs> g(ngroups);
    for (int i = 0; i < ngroups; ++i) {
        const cpt_lu_group& G = groups[i];
        g[i] = LuGroupArgs{(cplx*)G.A, G.n, G.batch, G.lda, G.strideA, G.ipiv, G.strideP, G.info, nullptr, 0, 0, 0,
                           lane[i]};
    }
    LuSetup tune = h->lu_streams;
    tune.side = nullptr;  // no look-ahead inside a lane: the lanes overlap instead
    tune.lookahead = 0;
    // up to NUM_LANES multi-work-group leaves at once must fit the device together
    tune.mw_budget = 256 / NUM_LANES;
    const int e = launch_getrf_lanes(g.data(), ngroups, &tune, h->lu_work.p);
    if (join_lanes(h, stream)) {
        set_error("cpt_lu_factor_grouped: could not join the lane streams");
        return CPT_ERR_HIP;
    }
    CPT_CHECK_HIP((hipError_t)e, "grouped getrf launch");
    return CPT_OK;
}

int cpt_lu_solve_grouped(cpt_handle h, int ngroups, const cpt_lu_group* groups, hipStream_t stream)
{
    if (!h || ngroups < 0 || (ngroups > 0 && !groups)) {
        set_error("cpt_lu_solve_grouped: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (!valid_groups(groups, ngroups, true, true, false, "cpt_lu_solve_grouped")) return CPT_ERR_ARG;
    if (ngroups == 0) return CPT_OK;
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    size_t ws = 0;
    for (int i = 0; i < ngroups; ++i)
        if (groups[i].n > 0 && groups[i].batch > 0 && groups[i].nrhs > 0)
            ws += getrs_workspace_bytes(groups[i].n, groups[i].batch);
    if (h->solve_work.reserve(ws)) {
        set_error("cpt_lu_solve_grouped: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    if (ensure_lanes(h) || fork_lanes(h, stream)) {
        set_error("cpt_lu_solve_grouped: could not set up the lane streams");
        return CPT_ERR_HIP;
    }
    const std::vector<hipStream_t> lane = assign_lanes(h, groups, ngroups);
    std::vector<LuGroupArgs> g(ngroups);
    for (int i = 0; i < ngroups; ++i) {
        const cpt_lu_group& G = groups[i];
        g[i] = LuGroupArgs{(cplx*)G.A, G.n, G.batch, G.lda, G.strideA, G.ipiv, G.strideP, G.info, (cplx*)G.B,
                           G.nrhs, G.ldb, G.strideB, lane[i]};
    }
    const int e = launch_getrs_lanes(g.data(), ngroups, h->solve_work.p);
    if (join_lanes(h, stream)) {
        set_error("cpt_lu_solve_grouped: could not join the lane streams");
        return CPT_ERR_HIP;
    }
    CPT_CHECK_HIP((hipError_t)e, "grouped getrs launch");
    return CPT_OK;
}

int cpt_lu_solve_mixed_grouped(cpt_handle h, int ngroups, const cpt_lu_group* groups, int max_iter, double tol,
                               int* iterations, double* residual, hipStream_t stream)
{
    if (!h || ngroups < 0 || (ngroups > 0 && (!groups || !iterations || !residual)) || max_iter < 0 ||
        !(tol >= 0)) {
        set_error("cpt_lu_solve_mixed_grouped: invalid arguments");
        return CPT_ERR_ARG;
    }
    if (!valid_groups(groups, ngroups, false, true, true, "cpt_lu_solve_mixed_grouped")) return CPT_ERR_ARG;
    if (ngroups == 0) return CPT_OK;
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CallScope call_scope(&h->timer, h->gemm_tiles);
    if (ensure_lanes(h) || fork_lanes(h, stream)) {
        set_error("cpt_lu_solve_mixed_grouped: could not set up the lane streams");
        return CPT_ERR_HIP;
    }
    const std::vector<hipStream_t> lane = assign_lanes(h, groups, ngroups);
    std::vector<MixedGroup> g(ngroups);
    for (int i = 0; i < ngroups; ++i) {
        const cpt_lu_group& G = groups[i];
        g[i] = MixedGroup{(const cplx*)G.A, G.n, G.lda, G.strideA, G.batch, (const cplx*)G.B, G.nrhs, G.ldb,
                          G.strideB, (cplx*)G.X, G.ldx, G.strideX, lane[i]};
    }
    void* work = mixed_workspace(h, mixed_grouped_bytes(g.data(), ngroups));
    if (!work) {
        set_error("cpt_lu_solve_mixed_grouped: workspace allocation failed");
        return CPT_ERR_ALLOC;
    }
    LuSetup tune = h->lu_streams;
    tune.side = nullptr;
    tune.lookahead = 0;
    tune.mw_budget = 256 / NUM_LANES;
    std::vector<int> info(ngroups, 0);
    const int e = launch_mixed_grouped(g.data(), ngroups, work, &tune, max_iter, tol, iterations, residual, info.data());
    if (join_lanes(h, stream)) {
        set_error("cpt_lu_solve_mixed_grouped: could not join the lane streams");
        return CPT_ERR_HIP;
    }
    CPT_CHECK_HIP((hipError_t)e, "grouped mixed solve");
    if (h->gemm_tiles == 4) {
        // groups the FP16-input products left short of tol are solved again
        // with FP32-input products; their sweeps add up
        std::vector<MixedGroup> again;
        std::vector<int> idx;
        for (int i = 0; i < ngroups; ++i)
            if (!(residual[i] <= tol)) {
                again.push_back(g[i]);
                idx.push_back(i);
            }
        if (!again.empty()) {
            CallScope fp32_scope(&h->timer, 3);
            std::vector<int> it2(again.size()), info2(again.size());
            std::vector<double> res2(again.size());
            void* work2 = mixed_workspace(h, mixed_grouped_bytes(again.data(), (int)again.size()));
            if (!work2) {
                set_error("cpt_lu_solve_mixed_grouped: workspace allocation failed");
                return CPT_ERR_ALLOC;
            }
            if (fork_lanes(h, stream)) {
                set_error("cpt_lu_solve_mixed_grouped: could not set up the lane streams");
                return CPT_ERR_HIP;
            }
            const int e2 = launch_mixed_grouped(again.data(), (int)again.size(), work2, &tune, max_iter, tol,
                                                it2.data(), res2.data(), info2.data());
            if (join_lanes(h, stream)) {
                set_error("cpt_lu_solve_mixed_grouped: could not join the lane streams");
                return CPT_ERR_HIP;
            }
            CPT_CHECK_HIP((hipError_t)e2, "grouped mixed solve (FP32-input products)");
            for (size_t k = 0; k < again.size(); ++k) {
                iterations[idx[k]] += it2[k];
                residual[idx[k]] = res2[k];
                info[idx[k]] = info2[k];
            }
            h->mixed_retries += (long)again.size();
        }
    }
    for (int i = 0; i < ngroups; ++i) {
        if (info[i] != 0 || !(residual[i] <= tol)) {
            char buf[240];
            if (info[i] != 0)
                std::snprintf(buf, sizeof buf,
                              "cpt_lu_solve_mixed_grouped: group %d (order %d): the complex64 factor is exactly "
                              "singular (pivot %d is zero)",
                              i, g[i].n, info[i]);
            else
                std::snprintf(buf, sizeof buf,
                              "cpt_lu_solve_mixed_grouped: group %d (order %d): the refinement stopped at relative "
                              "residual %.3e > tol %.3e after %d sweeps",
                              i, g[i].n, residual[i], tol, iterations[i]);
            set_error(buf);
            return CPT_ERR_NOT_CONVERGED;
        }
    }
    return CPT_OK;
}

int cpt_release_workspace(cpt_handle h)
{
    if (!h) {
        set_error("cpt_release_workspace: null handle");
        return CPT_ERR_ARG;
    }
    CPT_CHECK_HIP(hipSetDevice(h->device), "hipSetDevice");
    CPT_CHECK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    h->faces.release();
    h->quad.release();
    h->freq.release();
    h->prony.release();
    h->bfaces.release();
    h->bquad.release();
    h->bfreq.release();
    h->bargs.release();
    h->btiles.release();
    h->lu_work.release();
    h->mixed_work.release();
    h->solve_work.release();
    h->gmres_work.release();
    return CPT_OK;
}

int cpt_set_workspace(cpt_handle h, void* ptr, size_t bytes)
{
    if (!h || (!ptr && bytes)) {
        set_error("cpt_set_workspace: invalid arguments");
        return CPT_ERR_ARG;
    }
    h->ext_work = ptr;
    h->ext_bytes = ptr ? bytes : 0;
    return CPT_OK;
}

int cpt_mixed_workspace_size(cpt_handle h, int ngroups, const cpt_lu_group* groups, size_t* bytes)
{
    if (!h || ngroups < 0 || (ngroups > 0 && !groups) || !bytes) {
        set_error("cpt_mixed_workspace_size: invalid arguments");
        return CPT_ERR_ARG;
    }
    std::vector<MixedGroup> g(ngroups);
    for (int i = 0; i < ngroups; ++i)
        g[i] = MixedGroup{nullptr, groups[i].n, 0, 0, groups[i].batch, nullptr, groups[i].nrhs, 0, 0, nullptr, 0, 0,
                          nullptr};
    *bytes = mixed_grouped_bytes(g.data(), ngroups);
    return CPT_OK;
}

int cpt_set_assembly_precision(cpt_handle h, int flags)
{
    if (!h || flags < 0 || flags > (CPT_ASSEMBLY_FAR_FIELD_FP32 | CPT_ASSEMBLY_COMPLEX64)) {
        set_error("cpt_set_assembly_precision: invalid arguments");
        return CPT_ERR_ARG;
    }
    h->far_fp32 = (flags & CPT_ASSEMBLY_FAR_FIELD_FP32) ? 1 : 0;
    h->out_c64 = (flags & CPT_ASSEMBLY_COMPLEX64) ? 1 : 0;
    return CPT_OK;
}
