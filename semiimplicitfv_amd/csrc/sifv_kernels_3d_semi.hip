// sifv_kernels_3d_semi.hip - instantiations of the fused stage kernels (3dsemi).
#include "sifv_kernels.hpp"

namespace sifv {
StageKernels<double> stage_kernels_3d_semi(int R, int rs, int nph) {
    StageKernels<double> k;
    if (R == 1 && rs == 0 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 1, 0, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<1, 0, 0, true, double>;
    }
    if (R == 1 && rs == 1 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 1, 1, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<1, 1, 0, true, double>;
    }
    if (R == 1 && rs == 2 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 1, 2, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<1, 2, 0, true, double>;
    }
    if (R == 2 && rs == 0 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 2, 0, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<2, 0, 0, true, double>;
    }
    if (R == 2 && rs == 1 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 2, 1, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<2, 1, 0, true, double>;
    }
    if (R == 2 && rs == 2 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 2, 2, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<2, 2, 0, true, double>;
    }
    if (R == 3 && rs == 0 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 3, 0, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<3, 0, 0, true, double>;
    }
    if (R == 3 && rs == 1 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 3, 1, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<3, 1, 0, true, double>;
    }
    if (R == 3 && rs == 2 && nph == 0) {
        k.xy = &stage_xy_kernel<3, 3, 2, 0, true, MODE_XY, double>;
        k.z = &stage_z_kernel<3, 2, 0, true, double>;
    }
    if (R == 1 && rs == 0 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 1, 0, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<1, 0, 2, true, double>;
    }
    if (R == 1 && rs == 1 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 1, 1, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<1, 1, 2, true, double>;
    }
    if (R == 1 && rs == 2 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 1, 2, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<1, 2, 2, true, double>;
    }
    if (R == 2 && rs == 0 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 2, 0, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<2, 0, 2, true, double>;
    }
    if (R == 2 && rs == 1 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 2, 1, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<2, 1, 2, true, double>;
    }
    if (R == 2 && rs == 2 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 2, 2, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<2, 2, 2, true, double>;
    }
    if (R == 3 && rs == 0 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 3, 0, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<3, 0, 2, true, double>;
    }
    if (R == 3 && rs == 1 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 3, 1, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<3, 1, 2, true, double>;
    }
    if (R == 3 && rs == 2 && nph == 2) {
        k.xy = &stage_xy_kernel<3, 3, 2, 2, true, MODE_XY, double>;
        k.z = &stage_z_kernel<3, 2, 2, true, double>;
    }
    return k;
}

}  // namespace sifv
