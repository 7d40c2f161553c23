// gm_theta_f64.hip - the polar-quadrature apply in FP64 (gm_theta_apply.hpp),
// with the class selection and launch entry points.
#include "gm_theta_apply.hpp"

namespace gmk {

template hipError_t launch_tt<double>(int, const ApplyArgs *, const ApplyArgs *, const int *, int, int, hipStream_t);

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
    return a.mixed ? launch_tt<float>(cls, &a, nullptr, nullptr, 0, 0, st)
                   : launch_tt<double>(cls, &a, nullptr, nullptr, 0, 0, st);
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

hipError_t launch_tapply_batch(const ApplyArgs *args, const int *off, int M, int total, int cls, int mixed,
                               hipStream_t st)
{
    return mixed ? launch_tt<float>(cls, nullptr, args, off, M, total, st)
                 : launch_tt<double>(cls, nullptr, args, off, M, total, st);
}

}  // namespace gmk
