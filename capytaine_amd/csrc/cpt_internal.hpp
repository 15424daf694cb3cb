// Internal types shared by the HIP translation units of libcapytaine_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
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
    const FreqArgs* freq;  // device, nb
    cplx* S;
    int64_t ldS, strideS;
    cplx* K;
    int64_t ldK, strideK;
};

int launch_assembly(const AssemblyArgs& A, hipStream_t stream);

void set_error(const std::string& msg);

}  // namespace cpt
