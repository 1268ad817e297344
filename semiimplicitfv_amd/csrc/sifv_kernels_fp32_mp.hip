// sifv_kernels_fp32_mp.hip - instantiations of the fused stage kernels for
// two-phase explicit states in the FP32-bulk form (FP32 storage and WENO/flux
// arithmetic, FP64 mixture EOS, stored energy relative to the stiffened-gas
// constant).
#include "sifv_kernels.hpp"

namespace sifv {
namespace {
template <int DIM, int R, int RS>
StageKernels<float> mp_pair() {
    StageKernels<float> k;
    if constexpr (DIM == 3) {
        k.xy = &stage_xy_kernel<3, R, RS, 2, false, MODE_XY, float>;
        k.z = &stage_z_kernel<R, RS, 2, false, float>;
    } else {
        k.xy = &stage_xy_kernel<DIM, R, RS, 2, false, MODE_FULL, float>;
    }
    return k;
}

template <int DIM, int R>
StageKernels<float> mp_by_rs(int rs) {
    if (rs == 0) return mp_pair<DIM, R, 0>();
    if (rs == 1) return mp_pair<DIM, R, 1>();
    return mp_pair<DIM, R, 2>();
}

template <int DIM>
StageKernels<float> mp_by_r(int R, int rs) {
    if (R == 1) return mp_by_rs<DIM, 1>(rs);
    if (R == 2) return mp_by_rs<DIM, 2>(rs);
    return mp_by_rs<DIM, 3>(rs);
}
}  // namespace

StageKernels<float> stage_kernels_fp32_mp(int dim, int R, int rs) {
    if (R < 1 || R > 3 || rs < 0 || rs > 2) return StageKernels<float>();
    if (dim == 1) return mp_by_r<1>(R, rs);
    if (dim == 2) return mp_by_r<2>(R, rs);
    return mp_by_r<3>(R, rs);
}

}  // namespace sifv
