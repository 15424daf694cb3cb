// Host build of capytaine_amd/csrc/fingreen3d.hpp (the device FinGreen3D
// functions are __host__ __device__): C entry points for the CPU tests that
// compare the same code with the reference's FinGreen3D, point by point,
// before any GPU run.  Built by tests/test_fingreen3d_host.py; not part of
// the library.
#include <vector>
#include "../capytaine_amd/csrc/fingreen3d.hpp"

extern "C" {

// grn: (np, 3) complex128 row-major
void fgh_points(int np, const double* rr, const double* zf, const double* zp, double v, const double* roots,
                double h, double* grn)
{
    std::vector<double> ws(fg::WS_ARRAYS * fg::LIMES_LEN);
    for (int p = 0; p < np; ++p) {
        const fg::Grn g = fg::point(rr[p], zf[p], zp[p], v, roots, h, ws.data(), 1);
        for (int c = 0; c < 3; ++c) {
            grn[6 * p + 2 * c] = g.g[c].re;
            grn[6 * p + 2 * c + 1] = g.g[c].im;
        }
    }
}

// kind: 0 = K0, K1; 1 = J0, J1, Y0, Y1; 2 = E1; 3 = Ei; 4 = erf; 5 = E_n
void fgh_special(int np, int kind, int n, const double* x, double* out)
{
    for (int p = 0; p < np; ++p) {
        double* o = out + 4 * p;
        o[0] = o[1] = o[2] = o[3] = 0.0;
        switch (kind) {
        case 0: fg::bessel_k01(x[p], o[0], o[1]); break;
        case 1: fg::bessel_jy01(x[p], o[0], o[1], o[2], o[3]); break;
        case 2: o[0] = fg::e1_as(x[p]); break;
        case 3: o[0] = fg::ei_series(x[p]); break;
        case 4: o[0] = fg::erf_zj(x[p]); break;
        default: o[0] = fg::en_rec(n, x[p]); break;
        }
    }
}

double fgh_limes(int n, const double* s)
{
    std::vector<double> a(n + 1), t(s, s + n + 1);
    return fg::limes(n, t.data(), a.data(), 1);
}
}
