// sifv_visc.hip - instantiations and dispatch of the viscous stress kernel.
#include <hip/hip_runtime.h>

#include "sifv_kernels.hpp"

namespace sifv {

template <>
ViscFn<double> visc_kernel_for<double>(int dim, int nph) {
    if (nph == 0) {
        if (dim == 1) return visc_kernel<1, 0, double>;
        if (dim == 2) return visc_kernel<2, 0, double>;
        return visc_kernel<3, 0, double>;
    }
    if (nph == 2) {
        if (dim == 1) return visc_kernel<1, 2, double>;
        if (dim == 2) return visc_kernel<2, 2, double>;
        return visc_kernel<3, 2, double>;
    }
    return nullptr;
}

template <>
ViscFn<float> visc_kernel_for<float>(int dim, int nph) {
    if (nph == 0) {
        if (dim == 1) return visc_kernel<1, 0, float>;
        if (dim == 2) return visc_kernel<2, 0, float>;
        return visc_kernel<3, 0, float>;
    }
    if (nph == 2) {
        if (dim == 1) return visc_kernel<1, 2, float>;
        if (dim == 2) return visc_kernel<2, 2, float>;
        return visc_kernel<3, 2, float>;
    }
    return nullptr;
}

}  // namespace sifv
