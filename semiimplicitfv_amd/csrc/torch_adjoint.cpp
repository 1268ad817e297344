// torch_adjoint.cpp - thin PyTorch extension over the discrete-adjoint C ABI
// (include/sifv_gpu.h): zero-copy views of the solver's resident state as
// torch tensors and the backward sweep on torch device tensors.  The
// autograd.Function around it lives in semiimplicitfv_amd/torch_adjoint.py.
#include <torch/extension.h>
#include <torch/cuda.h>

#include <stdexcept>
#include <string>

#include "../../include/sifv_gpu.h"

namespace {

sifv_handle* as_handle(int64_t h) { return reinterpret_cast<sifv_handle*>(static_cast<intptr_t>(h)); }

void check(int rc) {
    if (rc != SIFV_OK) throw std::runtime_error(std::string("sifv: ") + sifv_last_error());
}

int64_t total_cells(int64_t h) {
    size_t n = 0;
    check(sifv_total_cells(as_handle(h), &n));
    return static_cast<int64_t>(n);
}

// View (4, totalCells) of the resident conservative state on `like`'s device.
torch::Tensor resident_state(int64_t h, const torch::Tensor& like) {
    double* p = nullptr;
    check(sifv_state_device_ptr(as_handle(h), &p));
    auto opts = torch::TensorOptions().dtype(torch::kFloat64).device(like.device());
    return torch::from_blob(p, {4, total_cells(h)}, opts);
}

// Copies U (4, totalCells) into the resident state.
void load_state(int64_t h, const torch::Tensor& U) {
    TORCH_CHECK(U.is_cuda() && U.dtype() == torch::kFloat64 && U.dim() == 2 && U.size(0) == 4 &&
                    U.size(1) == total_cells(h),
                "state must be a float64 device tensor of shape (4, totalCells)");
    resident_state(h, U).copy_(U);
    torch::cuda::synchronize(U.device().index());
}

// grad = (dU^N/dU^0)^T seed for the run recorded by sifv_adjoint_forward;
// returns (grad, d/dgamma as a 0-d CPU tensor).
std::tuple<torch::Tensor, torch::Tensor> adjoint_backward(int64_t h, const torch::Tensor& seed) {
    TORCH_CHECK(seed.is_cuda() && seed.dtype() == torch::kFloat64 && seed.is_contiguous() && seed.dim() == 2 &&
                    seed.size(0) == 4 && seed.size(1) == total_cells(h),
                "seed must be a contiguous float64 device tensor of shape (4, totalCells)");
    auto grad = torch::empty_like(seed);
    torch::cuda::synchronize(seed.device().index());
    double dg = 0.0;
    check(sifv_adjoint_backward(as_handle(h), seed.data_ptr<double>(), grad.data_ptr<double>(), &dg));
    return {grad, torch::tensor(dg, torch::TensorOptions().dtype(torch::kFloat64))};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("total_cells", &total_cells, "padded cells per field");
    m.def("resident_state", &resident_state, "zero-copy view of the resident conservative state");
    m.def("load_state", &load_state, "copy a (4, totalCells) tensor into the resident state");
    m.def("adjoint_backward", &adjoint_backward, "transpose of the recorded run applied to a seed");
}
