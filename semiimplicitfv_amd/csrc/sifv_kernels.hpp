// sifv_kernels.hpp - dispatch tables of the instantiated stage kernels.
// Instantiations are spread over several translation units
// (sifv_kernels_*.hip) so they compile in parallel.
#pragma once

#include "sifv_igr.hpp"
#include "sifv_stage.hpp"
#include "sifv_visc.hpp"

namespace sifv {

template <typename T>
using StageFn = void (*)(StageParams<T>);

// 1D/2D: xy is the full stage kernel and z is null; 3D: xy writes the x+y
// part of the right-hand side and z completes the stage.
template <typename T>
struct StageKernels {
    StageFn<T> xy = nullptr;
    StageFn<T> z = nullptr;
};

// R = reconstruction radius (1, 2, 3), rs = Riemann solver (RS_*), nph = 0 or 2.
// lean: no body force and no face stresses, so the 2D single-phase explicit
// tables may hand out their SRC = false instantiations.
StageKernels<double> stage_kernels_1d(int R, int rs, int nph, bool semi);
StageKernels<double> stage_kernels_2d(int R, int rs, int nph, bool semi, bool lean);
StageKernels<double> stage_kernels_3d(int R, int rs, int nph);
StageKernels<double> stage_kernels_3d_semi(int R, int rs, int nph);
StageKernels<float> stage_kernels_fp32(int dim, int R, int rs, bool lean);
StageKernels<float> stage_kernels_fp32_mp(int dim, int R, int rs);

// IGR: pre-pass (rhs, diagonal, rho) per dimension / phase count, and the
// wavefront Gauss-Seidel sweep per dimension.
template <typename T>
using IgrPreFn = void (*)(IgrPreArgs<T>);
using IgrSweepFn = void (*)(IgrSweepArgs);
template <typename T>
IgrPreFn<T> igr_pre_kernel_for(int dim, int nph);
template <>
IgrPreFn<double> igr_pre_kernel_for<double>(int dim, int nph);
template <>
IgrPreFn<float> igr_pre_kernel_for<float>(int dim, int nph);
IgrSweepFn igr_sweep_kernel_for(int dim);
// 2D strip-streaming sweep: pack (coefficients + sigma into skewed records),
// the sweeps, unpack (sigma back into the padded array).
struct IgrStripKernels {
    void (*pack)(IgrStripPackArgs);
    void (*sweep)(IgrStripArgs);
    void (*unpack)(Grid, const double*, double*);
};
IgrStripKernels igr_strip_kernels();

// Viscous stress right-hand side per dimension / phase count.
template <typename T>
using ViscFn = void (*)(ViscArgs<T>);
template <typename T>
ViscFn<T> visc_kernel_for(int dim, int nph);
template <>
ViscFn<double> visc_kernel_for<double>(int dim, int nph);
template <>
ViscFn<float> visc_kernel_for<float>(int dim, int nph);

}  // namespace sifv
