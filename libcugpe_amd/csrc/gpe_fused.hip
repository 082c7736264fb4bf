// Launchers of the fused whole-step kernel (gpe_fused.hpp) for the cube
// lengths whose state fits in one workgroup's LDS.
#include "gpe_kernels.hpp"

namespace gpe {
namespace {

template <typename T, int L>
void launch_len(const FusedArgs<T>& a, unsigned batch, bool uh, hipStream_t s) {
    if (uh)
        hipLaunchKernelGGL((fused_steps_kernel<T, L, true>), dim3(batch), dim3(fused_threads<T, L>()), 0, s, a);
    else
        hipLaunchKernelGGL((fused_steps_kernel<T, L, false>), dim3(batch), dim3(fused_threads<T, L>()), 0, s, a);
}

}  // namespace

template <>
bool fused_length_ok<float>(int L) {
    return L == 8 || L == 12 || L == 16 || L == 24;
}
template <>
bool fused_length_ok<double>(int L) {
    return L == 8 || L == 12 || L == 16;
}

template <>
int launch_fused_steps<float>(const FusedArgs<float>& a, int L, unsigned batch, bool uh, hipStream_t s) {
    switch (L) {
    case 8: launch_len<float, 8>(a, batch, uh, s); return 0;
    case 12: launch_len<float, 12>(a, batch, uh, s); return 0;
    case 16: launch_len<float, 16>(a, batch, uh, s); return 0;
    case 24: launch_len<float, 24>(a, batch, uh, s); return 0;
    default: return -1;
    }
}
template <>
int launch_fused_steps<double>(const FusedArgs<double>& a, int L, unsigned batch, bool uh, hipStream_t s) {
    switch (L) {
    case 8: launch_len<double, 8>(a, batch, uh, s); return 0;
    case 12: launch_len<double, 12>(a, batch, uh, s); return 0;
    case 16: launch_len<double, 16>(a, batch, uh, s); return 0;
    default: return -1;
    }
}

}  // namespace gpe
