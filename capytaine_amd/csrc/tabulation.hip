// Construction of the Delhommeau wave-term table on the device and its
// re-layout for the assembly kernel.
//
// Restates Delhommeau_integrals.f90:43-153 (numerical_integration with the
// float32 exp_e1 of Delhommeau's thesis, p. 367-369) and :190-204
// (construct_tabulation): one thread per (r, z) cell, 1001-point Simpson
// rule in theta.  The reference's table is defined by its operations and by
// the host C library's transcendental functions: double cos, exp, sin and
// the float complex logarithm of exp_e1.  The device repeats every
// arithmetic operation in the reference's order (contraction off, correctly
// rounded division, the scaled C99 complex division that the reference's
// Fortran compiler emits), and the host library supplies the transcendental
// values, computed once per distinct argument:
//   cos(theta_k)                       n values
//   exp(z_j)                           nz values
//   cos, sin of r_i cos(theta_k)        nr x n values each
//   clogf(z_j + i y) for the (z_j, y) pairs of exp_e1's logarithmic branch
//       (float z_j > -0.5, |y| <= 10, y = float(r_i cos(theta_k))), per
//       row z_j over the distinct y values.
// The table is then the reference's own, bit for bit, on a host whose C
// library is the reference's (the grids are built the same way).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>
#include "cpt_internal.hpp"
#include "integrals.hpp"

// The C library's float complex logarithm (C99 clogf).
extern "C" float _Complex clogf(float _Complex);

#pragma clang fp contract(off)

namespace cpt {
namespace {

// Writes the rows `rows` (nrows z indices) of the table in the reference
// layout (nr, nz, 5), column-major.  Host-library values: ct (n), ez (nz,
// exp(z_j)), cs / sn (n, nr: cos / sin of r_i ct_k), yidx (n, nr: index of
// float(r_i ct_k) among the distinct |y| <= 10 values, -1 beyond), clog
// (nrows, nuniq: clogf(float(z_j) + i y_u)), NULL for rows without the
// logarithmic branch (float z_j <= -0.5).
__global__ void __launch_bounds__(256) tabulate_kernel(const double* __restrict__ r_range, int nr,
                                                       const double* __restrict__ z_range, int nz,
                                                       const int* __restrict__ rows, int nrows, int n,
                                                       const double* __restrict__ ct_k,
                                                       const double* __restrict__ ez,
                                                       const double* __restrict__ cs,
                                                       const double* __restrict__ sn,
                                                       const int* __restrict__ yidx,
                                                       const float2* __restrict__ clog, int nuniq,
                                                       double* __restrict__ tab)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nr * nrows) return;
    const int i = c % nr, row = c / nr;
    const int j = rows[row];
    const double r = r_range[i], z = z_range[j];
    const float nm1 = (float)n - 1.0f;
    const double d_end = PI / (double)(3.0f * nm1);
    const double d_even = 4 * PI / (double)(3.0f * nm1);
    const double d_odd = 2 * PI / (double)(3.0f * nm1);
    const bool underflow = z <= -30.0;
    const double e = underflow ? 0.0 : ez[j];
    const float2* lrow = clog ? clog + (size_t)row * nuniq : nullptr;
    const float2 nan2 = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (int k = 1; k <= n; ++k) {
        const double dt = (k == 1 || k == n) ? d_end : ((k % 2 == 0) ? d_even : d_odd);
        const size_t ki = (size_t)(k - 1) * nr + i;
        const double ct = ct_k[k - 1];
        const double zr = z, zi = r * ct;
        double er = 0.0, ei = 0.0;
        if (!underflow) {
            er = e * cs[ki];
            ei = e * sn[ki];
        }
        // the host supplies log(z) exactly where exp_e1 takes its log branch
        const int u = yidx[ki];
        const float2* lz = (lrow && u >= 0) ? lrow + u : &nan2;
        const cf g = exp_e1(zr, zi, [lz](float, float) { return *lz; });
        // ii*pi*exp_zeta = (0, pi) * exp_zeta
        const double pr = 0.0 * er - PI * ei, pi_ = 0.0 * ei + PI * er;
        const double jr = (double)g.re + pr, ji = (double)g.im + pi_;
        const cdd inv = ddiv(cdd{1.0, 0.0}, cdd{zr, zi});
        const double mr = jr - inv.re, mi = ji - inv.im;
        a0 += dt * jr;
        a1 += dt * mr;
        a2 += dt * er;
        a3 -= dt * ct * mi;
        a4 -= dt * ct * ei;
    }
    const size_t plane = (size_t)nr * nz;
    const size_t cell = (size_t)j * nr + i;
    tab[cell] = 2 * (a0 / PI);
    tab[cell + plane] = 2 * (a1 / PI);
    tab[cell + 2 * plane] = 2 * a2;
    tab[cell + 3 * plane] = 2 * (a3 / PI);
    tab[cell + 4 * plane] = 2 * a4;
}

// (nr, nz, 5) column-major  ->  (nz, nr, 5) row-major with the five values
// of one cell contiguous.
__global__ void relayout_kernel(const double* __restrict__ src, int nr, int nz, double* __restrict__ dst)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nr * nz) return;
    const size_t plane = (size_t)nr * nz;
#pragma unroll
    for (int c = 0; c < 5; ++c) dst[(size_t)cell * 5 + c] = src[cell + c * plane];
}

}  // namespace

namespace {

// Runs fn(begin, end) over [0, count) on up to 16 host threads.
template <class F>
void parallel_ranges(size_t count, F fn)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = std::max<size_t>(1, std::min<size_t>({(size_t)(hw ? hw : 1), (size_t)16, count / 4096 + 1}));
    std::vector<std::thread> pool;
    const size_t step = (count + nt - 1) / nt;
    for (size_t t = 0; t < nt; ++t) {
        const size_t b = t * step, e = std::min(count, b + step);
        if (b < e) pool.emplace_back([=] { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}

// Device buffers of one tabulation call, freed on every return path.
struct TmpBufs {
    std::vector<void*> p;
    ~TmpBufs()
    {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    T* alloc(size_t count)
    {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        p.push_back(q);
        return (T*)q;
    }
};

}  // namespace

// Host side of cpt_tabulate: evaluates the host library's transcendental
// values, uploads them and launches the Simpson integration.  Synchronises
// `stream`.  Returns 0, or -2 (HIP error) / -3 (allocation) with the error
// message set.
int tabulate_on_device(const double* r_dev, int nr, const double* z_dev, int nz, int npts, double* tab,
                       hipStream_t stream)
{
    auto hip_fail = [](hipError_t e, const char* what) {
        set_error(std::string("cpt_tabulate: ") + what + ": " + hipGetErrorString(e));
        return -2;
    };
    std::vector<double> r(nr), z(nz);
    hipError_t e;
    if ((e = hipMemcpyAsync(r.data(), r_dev, sizeof(double) * nr, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(z.data(), z_dev, sizeof(double) * nz, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
        return hip_fail(e, "reading the grids");

    const int n = (npts % 2 == 0) ? npts + 1 : npts;
    const float nm1 = (float)n - 1.0f;
    // theta = -pi/2 + pi (k - 1) / (n - 1), the integers promoted to default reals
    std::vector<double> ct(n);
    for (int k = 1; k <= n; ++k) ct[k - 1] = std::cos(-PI / 2 + PI * (double)((float)k - 1.0f) / (double)nm1);
    std::vector<double> ez(nz);
    for (int j = 0; j < nz; ++j) ez[j] = (z[j] <= -30.0) ? 0.0 : std::exp(z[j]);

    const size_t nk = (size_t)n * nr;
    std::vector<double> cs(nk), sn(nk);
    std::vector<float> y(nk);
    parallel_ranges(nk, [&](size_t b, size_t en) {
        for (size_t q = b; q < en; ++q) {
            const double zi = r[q % nr] * ct[q / nr];
            ::sincos(zi, &sn[q], &cs[q]);
            y[q] = (float)zi;
        }
    });
    // distinct imaginary parts of the logarithmic branch
    std::vector<float> uniq;
    uniq.reserve(nk);
    for (size_t q = 0; q < nk; ++q)
        if (std::fabs(y[q]) <= 10.0f) uniq.push_back(y[q]);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end(),
                           [](float a, float b) { return a == b && std::signbit(a) == std::signbit(b); }),
               uniq.end());
    std::vector<int> yidx(nk, -1);
    parallel_ranges(nk, [&](size_t b, size_t en) {
        for (size_t q = b; q < en; ++q)
            if (std::fabs(y[q]) <= 10.0f) {
                auto it = std::lower_bound(uniq.begin(), uniq.end(), y[q]);
                // -0.0 and +0.0 compare equal: pick the entry of the same sign
                while (std::signbit(*it) != std::signbit(y[q])) ++it;
                yidx[q] = (int)(it - uniq.begin());
            }
    });
    const int nuniq = (int)uniq.size();

    // rows with and without the logarithmic branch
    std::vector<int> rows_log, rows_other;
    for (int j = 0; j < nz; ++j) ((float)z[j] > -0.5f ? rows_log : rows_other).push_back(j);

    TmpBufs buf;
    double* ct_d = buf.alloc<double>(n);
    double* ez_d = buf.alloc<double>(nz);
    double* cs_d = buf.alloc<double>(nk);
    double* sn_d = buf.alloc<double>(nk);
    int* yidx_d = buf.alloc<int>(nk);
    int* rows_d = buf.alloc<int>(nz);
    // rows of clogf values per chunk: at most 256 MiB
    const size_t row_bytes = (size_t)std::max(nuniq, 1) * sizeof(float2);
    const int chunk_rows = (int)std::max<size_t>(1, std::min<size_t>(rows_log.size(), ((size_t)256 << 20) / row_bytes));
    float2* clog_d = rows_log.empty() ? nullptr : buf.alloc<float2>((size_t)chunk_rows * std::max(nuniq, 1));
    if (!ct_d || !ez_d || !cs_d || !sn_d || !yidx_d || !rows_d || (!rows_log.empty() && !clog_d)) {
        set_error("cpt_tabulate: device allocation failed");
        return -3;
    }
    std::vector<int> rows_all(rows_log);
    rows_all.insert(rows_all.end(), rows_other.begin(), rows_other.end());
    if ((e = hipMemcpyAsync(ct_d, ct.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(ez_d, ez.data(), sizeof(double) * nz, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(cs_d, cs.data(), sizeof(double) * nk, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(sn_d, sn.data(), sizeof(double) * nk, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(yidx_d, yidx.data(), sizeof(int) * nk, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(rows_d, rows_all.data(), sizeof(int) * nz, hipMemcpyHostToDevice, stream)) !=
            hipSuccess)
        return hip_fail(e, "uploading the host-library values");

    auto launch = [&](int row0, int nrows, const float2* clog) {
        const int cells = nr * nrows;
        hipLaunchKernelGGL(tabulate_kernel, dim3((cells + 255) / 256), dim3(256), 0, stream, r_dev, nr, z_dev, nz,
                           rows_d + row0, nrows, n, ct_d, ez_d, cs_d, sn_d, yidx_d, clog, nuniq, tab);
        return hipGetLastError();
    };
    std::vector<float2> clog((size_t)chunk_rows * std::max(nuniq, 1));
    for (size_t r0 = 0; r0 < rows_log.size(); r0 += chunk_rows) {
        const int nrows = (int)std::min<size_t>(chunk_rows, rows_log.size() - r0);
        // the previous chunk's kernel reads clog_d: wait before overwriting
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "tabulation kernel");
        parallel_ranges((size_t)nrows * nuniq, [&](size_t b, size_t en) {
            for (size_t q = b; q < en; ++q) {
                const float zr = (float)z[rows_log[r0 + q / nuniq]];
                float _Complex w;
                __real__ w = zr;
                __imag__ w = uniq[q % nuniq];
                const float _Complex l = clogf(w);
                clog[q] = make_float2(__real__ l, __imag__ l);
            }
        });
        if ((e = hipMemcpyAsync(clog_d, clog.data(), sizeof(float2) * (size_t)nrows * nuniq, hipMemcpyHostToDevice,
                                stream)) != hipSuccess)
            return hip_fail(e, "uploading the logarithms");
        if ((e = launch((int)r0, nrows, clog_d)) != hipSuccess) return hip_fail(e, "tabulation kernel launch");
    }
    if (!rows_other.empty() && (e = launch((int)rows_log.size(), (int)rows_other.size(), nullptr)) != hipSuccess)
        return hip_fail(e, "tabulation kernel launch");
    // the temporaries are freed on return
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "tabulation kernel");
    return 0;
}

int launch_table_relayout(const double* src, int nr, int nz, double* dst, hipStream_t stream)
{
    const int cells = nr * nz;
    hipLaunchKernelGGL(relayout_kernel, dim3((cells + 255) / 256), dim3(256), 0, stream, src, nr, nz, dst);
    return (int)hipGetLastError();
}

}  // namespace cpt
