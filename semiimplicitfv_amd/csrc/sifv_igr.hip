// sifv_igr.hip - instantiations and dispatch of the IGR kernels.
#include <hip/hip_runtime.h>

#include "sifv_igr.hpp"
#include "sifv_kernels.hpp"

namespace sifv {

template <typename T>
IgrPreFn<T> igr_pre_kernel_for(int dim, int nph);

template <>
IgrPreFn<double> igr_pre_kernel_for<double>(int dim, int nph) {
    if (nph == 0) {
        if (dim == 1) return igr_pre_kernel<1, 0, double>;
        if (dim == 2) return igr_pre_kernel<2, 0, double>;
        return igr_pre_kernel<3, 0, double>;
    }
    if (nph == 2) {
        if (dim == 1) return igr_pre_kernel<1, 2, double>;
        if (dim == 2) return igr_pre_kernel<2, 2, double>;
        return igr_pre_kernel<3, 2, double>;
    }
    return nullptr;
}

template <>
IgrPreFn<float> igr_pre_kernel_for<float>(int dim, int nph) {
    if (nph == 2) {
        if (dim == 1) return igr_pre_kernel<1, 2, float>;
        if (dim == 2) return igr_pre_kernel<2, 2, float>;
        return igr_pre_kernel<3, 2, float>;
    }
    if (nph != 0) return nullptr;
    if (dim == 1) return igr_pre_kernel<1, 0, float>;
    if (dim == 2) return igr_pre_kernel<2, 0, float>;
    return igr_pre_kernel<3, 0, float>;
}

IgrSweepFn igr_sweep_kernel_for(int dim) {
    if (dim == 1) return igr_sweep_kernel<1>;
    if (dim == 2) return igr_sweep_kernel<2>;
    return igr_sweep_kernel<3>;
}

IgrStripKernels igr_strip_kernels() {
    return {igr_strip_pack_kernel<2>, igr_strip_kernel<2>, igr_strip_unpack_kernel<2>};
}

}  // namespace sifv
