// Launchers of the y-z plane passes (gpe_plane.hpp) for the plane sizes
// that fit in one workgroup's LDS.
#include "gpe_kernels.hpp"

namespace gpe {
namespace {

template <typename T, int L, bool UH>
void launch_mode(const PassArgs<T>& a, int mode, unsigned planes, hipStream_t s) {
    const dim3 g(planes), b(plane_threads<L>());
    switch (mode) {
    case PL_OPEN: hipLaunchKernelGGL((plane_kernel<T, L, PL_OPEN, UH>), g, b, 0, s, a); break;
    case PL_VSTEP2: hipLaunchKernelGGL((plane_kernel<T, L, PL_VSTEP2, UH>), g, b, 0, s, a); break;
    case PL_VSTEP3_NEXT: hipLaunchKernelGGL((plane_kernel<T, L, PL_VSTEP3_NEXT, UH>), g, b, 0, s, a); break;
    default: hipLaunchKernelGGL((plane_kernel<T, L, PL_VSTEP3_LAST, UH>), g, b, 0, s, a); break;
    }
}

template <typename T, int L>
void launch_len(const PassArgs<T>& a, int mode, unsigned planes, bool uh, hipStream_t s) {
    if (uh) launch_mode<T, L, true>(a, mode, planes, s);
    else launch_mode<T, L, false>(a, mode, planes, s);
}

template <typename T>
int launch_plane_t(const PassArgs<T>& a, int L, int mode, unsigned planes, bool uh, hipStream_t s) {
    switch (L) {
    case 32: launch_len<T, 32>(a, mode, planes, uh, s); return 0;
    case 48: launch_len<T, 48>(a, mode, planes, uh, s); return 0;
    case 64: launch_len<T, 64>(a, mode, planes, uh, s); return 0;
    default: return -1;
    }
}

}  // namespace

bool plane_length_ok(int L) { return L == 32 || L == 48 || L == 64; }

template <>
int launch_plane<float>(const PassArgs<float>& a, int L, int mode, unsigned planes, bool uh, hipStream_t s) {
    return launch_plane_t<float>(a, L, mode, planes, uh, s);
}
template <>
int launch_plane<double>(const PassArgs<double>& a, int L, int mode, unsigned planes, bool uh, hipStream_t s) {
    return launch_plane_t<double>(a, L, mode, planes, uh, s);
}

}  // namespace gpe
