// libtorch autograd function over the differentiated window of libk3dom.so
// (include/k3dom.h, k3dom_window_*).  Python module k3dom_amd._k3window:
//
//   loss = window_loss(theta, window_handle)
//
// theta: float tensor of 6 (ls, sigma_occ, sigma_free, prior_free,
// prior_static, prior_dynamic); window_handle: the k3dom_window* as an int.
// forward runs the window's frames on the GPU and returns the loss (a 0-d
// float64 tensor); backward runs the window's reverse pass and returns
// grad_output * d loss / d theta.  Frame inputs are set through the C ABI
// before the call.
#include <torch/extension.h>

#include <stdexcept>
#include <string>

#include "../../include/k3dom.h"

namespace {

k3dom_window* as_window(int64_t handle) { return reinterpret_cast<k3dom_window*>(static_cast<intptr_t>(handle)); }

void check(k3dom_status s, k3dom_window* w, const char* what) {
    if (s != K3DOM_OK) throw std::runtime_error(std::string(what) + " failed: " + k3dom_window_last_error(w));
}

struct K3WindowLoss : public torch::autograd::Function<K3WindowLoss> {
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor theta, int64_t handle) {
        TORCH_CHECK(theta.numel() == 6, "theta must hold 6 values");
        k3dom_window* w = as_window(handle);
        TORCH_CHECK(w != nullptr, "null window handle");
        torch::Tensor th = theta.detach().to(torch::kCPU, torch::kFloat32).contiguous();
        double loss = 0.0;
        check(k3dom_window_forward(w, th.data_ptr<float>(), &loss), w, "k3dom_window_forward");
        uint64_t id = 0;
        check(k3dom_window_forward_id(w, &id), w, "k3dom_window_forward_id");
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["theta"] = th;
        ctx->saved_data["forward_id"] = static_cast<int64_t>(id);
        ctx->saved_data["dtype"] = theta.scalar_type();
        ctx->saved_data["device"] = theta.device();
        return torch::tensor(loss, torch::dtype(torch::kFloat64));
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out) {
        k3dom_window* w = as_window(ctx->saved_data["handle"].toInt());
        const double g = grad_out[0].defined() ? grad_out[0].to(torch::kCPU, torch::kFloat64).item<double>() : 0.0;
        // the window's records are those of its last forward: when another
        // forward ran since this node's, run this node's forward again first
        uint64_t id = 0;
        check(k3dom_window_forward_id(w, &id), w, "k3dom_window_forward_id");
        if (static_cast<int64_t>(id) != ctx->saved_data["forward_id"].toInt()) {
            torch::Tensor th = ctx->saved_data["theta"].toTensor();
            double loss = 0.0;
            check(k3dom_window_forward(w, th.data_ptr<float>(), &loss), w, "k3dom_window_forward");
        }
        double grad[6];
        check(k3dom_window_backward(w, g, grad, nullptr), w, "k3dom_window_backward");
        torch::Tensor gt = torch::from_blob(grad, {6}, torch::kFloat64).clone();
        gt = gt.to(ctx->saved_data["device"].toDevice(), ctx->saved_data["dtype"].toScalarType());
        return {gt, torch::Tensor()};
    }
};

torch::Tensor window_loss(torch::Tensor theta, int64_t handle) { return K3WindowLoss::apply(theta, handle); }

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "torch.autograd function over the K3DOM differentiated window (libk3dom.so)";
    m.def("window_loss", &window_loss, "Occupancy loss of the window; differentiable in theta");
}
