// sifv_kernels_fp32.hip - instantiations of the fused stage kernels (fp32).
#include "sifv_kernels.hpp"

namespace sifv {
StageKernels<float> stage_kernels_fp32(int dim, int R, int rs, bool lean) {
    StageKernels<float> k;
    if (lean && dim == 2) {
        if (R == 1 && rs == 0) k.xy = &stage_xy_kernel<2, 1, 0, 0, false, MODE_FULL, float, false>;
        if (R == 1 && rs == 1) k.xy = &stage_xy_kernel<2, 1, 1, 0, false, MODE_FULL, float, false>;
        if (R == 1 && rs == 2) k.xy = &stage_xy_kernel<2, 1, 2, 0, false, MODE_FULL, float, false>;
        if (R == 2 && rs == 0) k.xy = &stage_xy_kernel<2, 2, 0, 0, false, MODE_FULL, float, false>;
        if (R == 2 && rs == 1) k.xy = &stage_xy_kernel<2, 2, 1, 0, false, MODE_FULL, float, false>;
        if (R == 2 && rs == 2) k.xy = &stage_xy_kernel<2, 2, 2, 0, false, MODE_FULL, float, false>;
        if (R == 3 && rs == 0) k.xy = &stage_xy_kernel<2, 3, 0, 0, false, MODE_FULL, float, false>;
        if (R == 3 && rs == 1) k.xy = &stage_xy_kernel<2, 3, 1, 0, false, MODE_FULL, float, false>;
        if (R == 3 && rs == 2) k.xy = &stage_xy_kernel<2, 3, 2, 0, false, MODE_FULL, float, false>;
        return k;
    }
    if (dim == 1 && R == 1 && rs == 0) k.xy = &stage_xy_kernel<1, 1, 0, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 1 && rs == 1) k.xy = &stage_xy_kernel<1, 1, 1, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 1 && rs == 2) k.xy = &stage_xy_kernel<1, 1, 2, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 2 && rs == 0) k.xy = &stage_xy_kernel<1, 2, 0, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 2 && rs == 1) k.xy = &stage_xy_kernel<1, 2, 1, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 2 && rs == 2) k.xy = &stage_xy_kernel<1, 2, 2, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 3 && rs == 0) k.xy = &stage_xy_kernel<1, 3, 0, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 3 && rs == 1) k.xy = &stage_xy_kernel<1, 3, 1, 0, false, MODE_FULL, float>;
    if (dim == 1 && R == 3 && rs == 2) k.xy = &stage_xy_kernel<1, 3, 2, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 1 && rs == 0) k.xy = &stage_xy_kernel<2, 1, 0, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 1 && rs == 1) k.xy = &stage_xy_kernel<2, 1, 1, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 1 && rs == 2) k.xy = &stage_xy_kernel<2, 1, 2, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 2 && rs == 0) k.xy = &stage_xy_kernel<2, 2, 0, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 2 && rs == 1) k.xy = &stage_xy_kernel<2, 2, 1, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 2 && rs == 2) k.xy = &stage_xy_kernel<2, 2, 2, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 3 && rs == 0) k.xy = &stage_xy_kernel<2, 3, 0, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 3 && rs == 1) k.xy = &stage_xy_kernel<2, 3, 1, 0, false, MODE_FULL, float>;
    if (dim == 2 && R == 3 && rs == 2) k.xy = &stage_xy_kernel<2, 3, 2, 0, false, MODE_FULL, float>;
    if (dim == 3 && R == 1 && rs == 0) {
        k.xy = &stage_xy_kernel<3, 1, 0, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<1, 0, 0, false, float>;
    }
    if (dim == 3 && R == 1 && rs == 1) {
        k.xy = &stage_xy_kernel<3, 1, 1, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<1, 1, 0, false, float>;
    }
    if (dim == 3 && R == 1 && rs == 2) {
        k.xy = &stage_xy_kernel<3, 1, 2, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<1, 2, 0, false, float>;
    }
    if (dim == 3 && R == 2 && rs == 0) {
        k.xy = &stage_xy_kernel<3, 2, 0, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<2, 0, 0, false, float>;
    }
    if (dim == 3 && R == 2 && rs == 1) {
        k.xy = &stage_xy_kernel<3, 2, 1, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<2, 1, 0, false, float>;
    }
    if (dim == 3 && R == 2 && rs == 2) {
        k.xy = &stage_xy_kernel<3, 2, 2, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<2, 2, 0, false, float>;
    }
    if (dim == 3 && R == 3 && rs == 0) {
        k.xy = &stage_xy_kernel<3, 3, 0, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<3, 0, 0, false, float>;
    }
    if (dim == 3 && R == 3 && rs == 1) {
        k.xy = &stage_xy_kernel<3, 3, 1, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<3, 1, 0, false, float>;
    }
    if (dim == 3 && R == 3 && rs == 2) {
        k.xy = &stage_xy_kernel<3, 3, 2, 0, false, MODE_XY, float>;
        k.z = &stage_z_kernel<3, 2, 0, false, float>;
    }
    return k;
}

}  // namespace sifv
