// sifv_kernels_2d.hip - instantiations of the fused stage kernels (2d).
#include "sifv_kernels.hpp"

namespace sifv {
StageKernels<double> stage_kernels_2d(int R, int rs, int nph, bool semi, bool lean) {
    StageKernels<double> k;
    if (lean && nph == 0 && !semi) {
        if (R == 1 && rs == 0) k.xy = &stage_xy_kernel<2, 1, 0, 0, false, MODE_FULL, double, false>;
        if (R == 1 && rs == 1) k.xy = &stage_xy_kernel<2, 1, 1, 0, false, MODE_FULL, double, false>;
        if (R == 1 && rs == 2) k.xy = &stage_xy_kernel<2, 1, 2, 0, false, MODE_FULL, double, false>;
        if (R == 2 && rs == 0) k.xy = &stage_xy_kernel<2, 2, 0, 0, false, MODE_FULL, double, false>;
        if (R == 2 && rs == 1) k.xy = &stage_xy_kernel<2, 2, 1, 0, false, MODE_FULL, double, false>;
        if (R == 2 && rs == 2) k.xy = &stage_xy_kernel<2, 2, 2, 0, false, MODE_FULL, double, false>;
        if (R == 3 && rs == 0) k.xy = &stage_xy_kernel<2, 3, 0, 0, false, MODE_FULL, double, false>;
        if (R == 3 && rs == 1) k.xy = &stage_xy_kernel<2, 3, 1, 0, false, MODE_FULL, double, false>;
        if (R == 3 && rs == 2) k.xy = &stage_xy_kernel<2, 3, 2, 0, false, MODE_FULL, double, false>;
        return k;
    }
    if (R == 1 && rs == 0 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 1, 0, 0, false, MODE_FULL, double>;
    if (R == 1 && rs == 1 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 1, 1, 0, false, MODE_FULL, double>;
    if (R == 1 && rs == 2 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 1, 2, 0, false, MODE_FULL, double>;
    if (R == 2 && rs == 0 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 2, 0, 0, false, MODE_FULL, double>;
    if (R == 2 && rs == 1 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 2, 1, 0, false, MODE_FULL, double>;
    if (R == 2 && rs == 2 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 2, 2, 0, false, MODE_FULL, double>;
    if (R == 3 && rs == 0 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 3, 0, 0, false, MODE_FULL, double>;
    if (R == 3 && rs == 1 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 3, 1, 0, false, MODE_FULL, double>;
    if (R == 3 && rs == 2 && nph == 0 && semi == false) k.xy = &stage_xy_kernel<2, 3, 2, 0, false, MODE_FULL, double>;
    if (R == 1 && rs == 0 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 1, 0, 2, false, MODE_FULL, double>;
    if (R == 1 && rs == 1 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 1, 1, 2, false, MODE_FULL, double>;
    if (R == 1 && rs == 2 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 1, 2, 2, false, MODE_FULL, double>;
    if (R == 2 && rs == 0 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 2, 0, 2, false, MODE_FULL, double>;
    if (R == 2 && rs == 1 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 2, 1, 2, false, MODE_FULL, double>;
    if (R == 2 && rs == 2 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 2, 2, 2, false, MODE_FULL, double>;
    if (R == 3 && rs == 0 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 3, 0, 2, false, MODE_FULL, double>;
    if (R == 3 && rs == 1 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 3, 1, 2, false, MODE_FULL, double>;
    if (R == 3 && rs == 2 && nph == 2 && semi == false) k.xy = &stage_xy_kernel<2, 3, 2, 2, false, MODE_FULL, double>;
    if (R == 1 && rs == 0 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 1, 0, 0, true, MODE_FULL, double>;
    if (R == 1 && rs == 1 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 1, 1, 0, true, MODE_FULL, double>;
    if (R == 1 && rs == 2 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 1, 2, 0, true, MODE_FULL, double>;
    if (R == 2 && rs == 0 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 2, 0, 0, true, MODE_FULL, double>;
    if (R == 2 && rs == 1 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 2, 1, 0, true, MODE_FULL, double>;
    if (R == 2 && rs == 2 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 2, 2, 0, true, MODE_FULL, double>;
    if (R == 3 && rs == 0 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 3, 0, 0, true, MODE_FULL, double>;
    if (R == 3 && rs == 1 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 3, 1, 0, true, MODE_FULL, double>;
    if (R == 3 && rs == 2 && nph == 0 && semi == true) k.xy = &stage_xy_kernel<2, 3, 2, 0, true, MODE_FULL, double>;
    if (R == 1 && rs == 0 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 1, 0, 2, true, MODE_FULL, double>;
    if (R == 1 && rs == 1 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 1, 1, 2, true, MODE_FULL, double>;
    if (R == 1 && rs == 2 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 1, 2, 2, true, MODE_FULL, double>;
    if (R == 2 && rs == 0 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 2, 0, 2, true, MODE_FULL, double>;
    if (R == 2 && rs == 1 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 2, 1, 2, true, MODE_FULL, double>;
    if (R == 2 && rs == 2 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 2, 2, 2, true, MODE_FULL, double>;
    if (R == 3 && rs == 0 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 3, 0, 2, true, MODE_FULL, double>;
    if (R == 3 && rs == 1 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 3, 1, 2, true, MODE_FULL, double>;
    if (R == 3 && rs == 2 && nph == 2 && semi == true) k.xy = &stage_xy_kernel<2, 3, 2, 2, true, MODE_FULL, double>;
    return k;
}

}  // namespace sifv
