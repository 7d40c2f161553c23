// gm_theta_f64.hip - the polar-quadrature apply (gm_theta_apply.hpp), with the
// class selection and launch entry points. It runs in FP64 for both
// precisions: its transforms are FP64 matrix-core work and its banded
// product runs on the VALU, where FP32 (packed FP32 off) issues at the FP64
// rate, so an FP32 form was measured slower (0.165 against 0.145 ms at
// l_max 20, N 16384) and a mixed-precision solver's Krylov applies take the
// FP64 one.
#include "gm_theta_apply.hpp"

namespace gmk {

hipError_t launch_tt(int cls, const ApplyArgs *a, const ApplyArgs *args, const int *off, int M, int total,
                     hipStream_t st)
{
    switch (cls) {
    case 4: return launch_t<4>(a, args, off, M, total, st);
    case 8: return launch_t<8>(a, args, off, M, total, st);
    case 12: return launch_t<12>(a, args, off, M, total, st);
    case 16: return launch_t<16>(a, args, off, M, total, st);
    case 20: return launch_t<20>(a, args, off, M, total, st);
    default: return hipErrorInvalidValue;
    }
}

int tapply_class(int lmax, int /*mixed*/)
{
    if (lmax < 0 || lmax > 20) return 0;
    return lmax <= 4 ? 4 : (lmax + 3) / 4 * 4;
}

int tapply_grid(int N) { return (N + TT - 1) / TT; }

hipError_t launch_tapply(const ApplyArgs &a, hipStream_t st)
{
    const int cls = tapply_class(a.g.lmax, a.mixed);
    if (!cls || !a.g.TA) return hipErrorInvalidValue;
    return launch_tt(cls, &a, nullptr, nullptr, 0, 0, st);
}

size_t tapply_u_doubles(int N, int lmax)
{
    const int cls = tapply_class(lmax, 0);
    if (!cls) return 0;
    const size_t per = (size_t)gmt::theta_rule(lmax, 2 * lmax) * (2 * cls + 1) * 4 * TT;
    return (size_t)tapply_grid(N) * per;
}

hipError_t launch_tapply_u(const gmd::Geometry &g, double *Ut, hipStream_t st)
{
    switch (tapply_class(g.lmax, 0)) {
    case 4: return launch_u<4>(g, Ut, st);
    case 8: return launch_u<8>(g, Ut, st);
    case 12: return launch_u<12>(g, Ut, st);
    case 16: return launch_u<16>(g, Ut, st);
    case 20: return launch_u<20>(g, Ut, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tapply_batch(const ApplyArgs *args, const int *off, int M, int total, int cls, hipStream_t st)
{
    return launch_tt(cls, nullptr, args, off, M, total, st);
}

}  // namespace gmk
