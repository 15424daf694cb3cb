// Internal types shared by the HIP translation units of libcapytaine_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "green.hpp"

namespace cpt {

// Per-wavenumber parameters of a batched assembly launch.
struct FreqArgs {
    double k;                    // wavenumber
    int gfs;                     // gf_singularities after the host-side overrides
    int wave;                    // 0 < k < inf: add the wave term
    int nexp;                    // number of Prony exponentials (finite depth)
    int pad;
    const double* prony_lambda;  // device, nexp
    const double* prony_a;       // device, nexp
};

struct AssemblyArgs {
    const double* colloc;  // (n,3) row-major collocation points
    const double* dpn;     // (n,3) if adjoint else (m,3): normals of the dot product
    int n, m, nq, nk;
    const Face* faces;     // (m)
    const double* quad;    // (m, nq, 4): x, y, z, weight
    Table T;
    double depth;
    int finite_depth, fd_method, adjoint, rankine_only;
    int diag, n_diag;
    double free_surface;
    int nb, j_chunk;
    int far_fp32;          // FP32 far field (infinite depth, mixed form)
    int c64;               // S and K stored as complex64 (single-precision Green function)
    const FreqArgs* freq;  // device, nb
    cplx* S;               // complex128, or complex64 when c64 (strides in entries)
    int64_t ldS, strideS;
    cplx* K;
    int64_t ldK, strideK;
};

int launch_assembly(const AssemblyArgs& A, hipStream_t stream);

// FinGreen3D wave part added to S and K (fingreen3d.hip).  The pairs (I, J)
// of n collocation points and m panels: S[I + ldS J] += -1/(4 pi) sum_q w_q G,
// and the same for the gradient into K (dot product with dpn when nk == 1).
struct FgArgs {
    const double* colloc;  // (n,3)
    const double* dpn;     // (n,3) if adjoint else (m,3)
    const double* quad;    // (m, nq, 4): x, y, z, weight
    int n, m, nq, nk, adjoint;
    const double* roots;   // dispersion roots, roots[0] the propagating one
    double depth, v;       // water depth, omega^2/g
    cplx* S;
    int64_t ldS;
    cplx* K;
    int64_t ldK;
    double* ws;            // FG_WS_DOUBLES per thread, interleaved by thread
    int nthreads;
    unsigned* lists;       // three lists of list_cap pair indices (by series), then 3 counters
    int list_cap;
    double* tab;           // per-point factor tables (4 x 100 entries per point, interleaved by point)
    double* root_tab;      // per-root factors (3 x 100)
    size_t npts;           // points of the tables: the n collocation points, then the m x nq quadrature points
};
size_t fingreen3d_table_bytes(size_t npts);
constexpr int FG_WS_DOUBLES = 400;
// panel pairs per pass of the sorted kernels (bounds the lists' memory)
constexpr int FG_CHUNK_PAIRS = 1 << 24;
size_t fingreen3d_list_bytes(int n);
int fingreen3d_threads(int64_t items);
int launch_fingreen3d_pairs(const FgArgs& A, hipStream_t stream);
int launch_fingreen3d_points(int np, const double* rr, const double* zf, const double* zp, double v,
                             const double* roots, double depth, cplx* grn, double* ws, int nthreads,
                             hipStream_t stream);

// Work-list entry of the persistent multi-body assembly.
struct AsmTile {
    int body, iblk, jc, pad;
};
int launch_assembly_worklist(const AssemblyArgs* bodies, const AsmTile* tiles, int ntiles, int grid, double pairs,
                             int far_fp32, bool untabulated, hipStream_t stream);

// Tuning of the blocked LU and the side stream / events of its look-ahead
// (owned by a context).
struct LuSetup {
    hipStream_t side = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    int block = 0;      // block width: 64, 128 or 256, 0 = by size (256 from 2048 rows, 128 from 1024)
    int leaf = 16;      // panel columns factored by one work-group per matrix
    int panel_threads = 512;  // threads of a panel leaf's work-group: 512 or 256
    int panel_mw = 1;         // leaves over several work-groups per matrix where they fit
    int panel_spin = 0;       // polls a multi-work-group leaf waits per column, 0: default; < 0: give up at once
    int mw_budget = 256;      // work-groups of one multi-work-group leaf launch over the batch
    int lookahead = 1;  // 1: next panel on the side stream, 0: in order
    int speculate = 1;  // panels factored with the diagonal as pivots, checked (else pivot searches)
    int* spec_repairs = nullptr;  // device counter of panels factored again after a failed check
};

// One group of the grouped (variable-size) LU: `batch` matrices of order n,
// run on stream s.
struct MixedGroup {
    const cplx* K;
    int n;
    int64_t lda, strideK;
    int batch;
    const cplx* B;
    int nrhs;
    int64_t ldb, strideB;
    cplx* X;
    int64_t ldx, strideX;
    hipStream_t s;
};
// Complex64 factor of the mixed solve (complex64 values as interleaved
// float pairs): LU factors (batch n^2, leading dimension n), pivots and info,
// inverted 64 x 64 diagonal blocks (n x 64 per matrix).
struct MixedFactor {
    void* K32;
    int* ipiv;
    int* info;
    void* D;
};
size_t mixed_factor_bytes(int n, int batch);
MixedFactor mixed_factor_at(void* base, int n, int batch);
size_t mixed_factor_ws_bytes(int n, int batch);
size_t mixed_refine_bytes(const MixedGroup* g, int ng);
int mixed_factor(const cplx* K, int n, int64_t lda, int64_t strideK, int batch, const MixedFactor& F, void* ws,
                 const LuSetup* tune, hipStream_t s);
int mixed_refine(const MixedGroup* g, const MixedFactor* F, int ng, void* work, int max_iter, double tol,
                 int* iterations, double* residual, int* info_out);

struct LuGroupArgs {
    cplx* A;
    int n, batch;
    int64_t lda, strideA;
    int* ipiv;
    int64_t strideP;
    int* info;
    cplx* B;
    int nrhs;
    int64_t ldb, strideB;
    hipStream_t s;
};

void set_error(const std::string& msg);

// Live timing of selected kernels: start/stop HIP events recorded on the
// stream each timed launch goes to, with the launch's algorithmic work
// (flops, or panel pairs for the assembly).  Enabled per context; the entry
// points install the context's timer for the calling thread.  Brackets may
// nest: a panel bracket includes the tile products launched inside it, which
// are also timed in their own class.
enum TimedKernel {
    TK_ZGEMM = 0,
    TK_CGEMM = 1,
    TK_PANEL_Z = 2,
    TK_PANEL_C = 3,
    TK_ASSEMBLY = 4,
    TK_SKINNY_Z = 5,
    TK_SKINNY_C = 6,
    TK_CGEMM_PACKED = 7,
    TK_COUNT = 8
};
struct KernelTimer {
    struct Rec {
        hipEvent_t a = nullptr, b = nullptr;
        int kind = 0;
        double work = 0.0;
        double bytes = 0.0;  // algorithmic bytes moved (operands read, results written)
    };
    std::vector<Rec> recs;
    size_t used = 0;
    bool on = false;
    int start(hipStream_t s, int kind, double work, double bytes);  // slot, or -1
    void stop(int slot, hipStream_t s);
    void release();
};
extern thread_local KernelTimer* active_timer;
// Tile width of the tiled complex64 product for the calling thread's current
// entry point (1: 128 x 128 tiles, 0: 128 x 64).
extern thread_local int active_gemm_tiles;

// Brackets one launch with the active timer's events (no-op when off).
struct TimedLaunch {
    int slot;
    hipStream_t s;
    TimedLaunch(hipStream_t stream, int kind, double work, double bytes = 0.0)
        : slot(active_timer ? active_timer->start(stream, kind, work, bytes) : -1), s(stream)
    {
    }
    ~TimedLaunch()
    {
        if (slot >= 0 && active_timer) active_timer->stop(slot, s);
    }
};

}  // namespace cpt
