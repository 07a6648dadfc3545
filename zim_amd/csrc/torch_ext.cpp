// torch_ext.cpp — the differentiable triangulation behind a
// torch::autograd::Function (libtorch C++ API). Forward and backward call the
// C ABI (psgc_adjoint_forward / psgc_adjoint_backward) on torch's stream with
// device pointers; no torch types cross into the HIP library.

#include <torch/extension.h>

#include <cstring>

#include "psgc.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(int rc) {
  TORCH_CHECK(rc == PSGC_OK, "psgc error ", rc, ": ", psgc_last_error());
}

torch::Tensor host_f64(const torch::Tensor& t, int64_t n) {
  auto h = t.detach().to(torch::kCPU, torch::kDouble).contiguous();
  TORCH_CHECK(h.numel() == n, "expected ", n, " values");
  return h;
}

class Triangulate : public torch::autograd::Function<Triangulate> {
 public:
  static tensor_list forward(AutogradContext* actx, torch::Tensor theta, torch::Tensor phL,
                             torch::Tensor phR, torch::Tensor match_col, torch::Tensor nominal,
                             int64_t ctx_handle, int64_t stream) {
    TORCH_CHECK(phL.is_cuda() && phL.scalar_type() == torch::kDouble, "phL: f64 device tensor");
    TORCH_CHECK(phR.sizes() == phL.sizes() && phR.scalar_type() == torch::kDouble, "phR");
    TORCH_CHECK(match_col.scalar_type() == torch::kInt && match_col.sizes() == phL.sizes(),
                "match_col: i32 of the phase size");
    auto* ctx = reinterpret_cast<psgc_ctx*>(ctx_handle);
    check(psgc_set_stream(ctx, reinterpret_cast<void*>(stream)));
    auto th = host_f64(theta, 9);
    auto nm = host_f64(nominal, 5);
    auto L = phL.contiguous(), R = phR.contiguous(), M = match_col.contiguous();
    const int rows = (int)L.size(0), cols = (int)L.size(1);
    auto xyz = torch::empty({rows, cols, 3}, L.options());
    auto valid = torch::empty({rows, cols}, L.options().dtype(torch::kUInt8));
    check(psgc_adjoint_forward(ctx, th.data_ptr<double>(), nm.data_ptr<double>(),
                               L.data_ptr<double>(), R.data_ptr<double>(), M.data_ptr<int32_t>(),
                               rows, cols, xyz.data_ptr<double>(), valid.data_ptr<uint8_t>(),
                               PSGC_MEM_DEVICE));
    actx->save_for_backward({th, L, R, M, nm});
    actx->saved_data["ctx"] = ctx_handle;
    actx->saved_data["stream"] = stream;
    actx->saved_data["theta_device"] = theta.device().str();
    actx->saved_data["theta_dtype"] = (int64_t)theta.scalar_type();
    actx->mark_non_differentiable({valid});
    return {xyz, valid};
  }

  static tensor_list backward(AutogradContext* actx, tensor_list grads) {
    auto saved = actx->get_saved_variables();
    auto th = saved[0], L = saved[1], R = saved[2], M = saved[3], nm = saved[4];
    auto* ctx = reinterpret_cast<psgc_ctx*>(actx->saved_data["ctx"].toInt());
    check(psgc_set_stream(ctx, reinterpret_cast<void*>(actx->saved_data["stream"].toInt())));
    const int rows = (int)L.size(0), cols = (int)L.size(1);
    auto g = grads[0].to(torch::kDouble).contiguous();
    auto gth = torch::empty({9}, L.options());
    const bool needL = actx->needs_input_grad(1), needR = actx->needs_input_grad(2);
    torch::Tensor gl, gr;
    if (needL) gl = torch::empty_like(L);
    if (needR) gr = torch::empty_like(R);
    check(psgc_adjoint_backward(ctx, th.data_ptr<double>(), nm.data_ptr<double>(),
                                L.data_ptr<double>(), R.data_ptr<double>(),
                                M.data_ptr<int32_t>(), rows, cols, g.data_ptr<double>(),
                                gth.data_ptr<double>(), needL ? gl.data_ptr<double>() : nullptr,
                                needR ? gr.data_ptr<double>() : nullptr, PSGC_MEM_DEVICE));
    auto dev = torch::Device(actx->saved_data["theta_device"].toStringRef());
    auto dt = (torch::ScalarType)actx->saved_data["theta_dtype"].toInt();
    return {gth.to(dev, dt), gl, gr, torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor()};
  }
};

// Many views in one launch each way: the phase and match maps of all views
// concatenated (flattened), sizes[v] = (rows, cols), thetas [V, 9],
// nominal [V, 5]; outputs xyz [N, 3] and valid [N] in the same order.
class TriangulateBatch : public torch::autograd::Function<TriangulateBatch> {
 public:
  static std::vector<psgc_adjoint_view> views(const torch::Tensor& th, const torch::Tensor& nm,
                                              const torch::Tensor& sizes, const torch::Tensor& L,
                                              const torch::Tensor& R, const torch::Tensor& M) {
    const int64_t V = sizes.size(0);
    std::vector<psgc_adjoint_view> vs(V);
    int64_t off = 0;
    for (int64_t v = 0; v < V; ++v) {
      psgc_adjoint_view& w = vs[v];
      std::memset(&w, 0, sizeof(w));
      for (int k = 0; k < 9; ++k) w.theta[k] = th[v][k].item<double>();
      for (int k = 0; k < 5; ++k) w.nominal[k] = nm[v][k].item<double>();
      w.rows = (int32_t)sizes[v][0].item<int64_t>();
      w.cols = (int32_t)sizes[v][1].item<int64_t>();
      w.phL = L.data_ptr<double>() + off;
      w.phR = R.data_ptr<double>() + off;
      w.match_col = M.data_ptr<int32_t>() + off;
      off += (int64_t)w.rows * w.cols;
    }
    TORCH_CHECK(off == L.numel(), "sizes cover ", off, " pixels, the maps ", L.numel());
    return vs;
  }

  static tensor_list forward(AutogradContext* actx, torch::Tensor thetas, torch::Tensor phL,
                             torch::Tensor phR, torch::Tensor match_col, torch::Tensor nominal,
                             torch::Tensor sizes, int64_t ctx_handle, int64_t stream) {
    TORCH_CHECK(phL.is_cuda() && phL.scalar_type() == torch::kDouble, "phL: f64 device tensor");
    TORCH_CHECK(phR.numel() == phL.numel() && phR.scalar_type() == torch::kDouble, "phR");
    TORCH_CHECK(match_col.scalar_type() == torch::kInt && match_col.numel() == phL.numel(),
                "match_col: i32 of the phase size");
    auto* ctx = reinterpret_cast<psgc_ctx*>(ctx_handle);
    check(psgc_set_stream(ctx, reinterpret_cast<void*>(stream)));
    const int64_t V = sizes.size(0);
    auto th = host_f64(thetas, V * 9).view({V, 9});
    auto nm = host_f64(nominal, V * 5).view({V, 5});
    auto sz = sizes.to(torch::kCPU, torch::kLong).contiguous();
    auto L = phL.contiguous(), R = phR.contiguous(), M = match_col.contiguous();
    const int64_t N = L.numel();
    auto xyz = torch::empty({N, 3}, L.options());
    auto valid = torch::empty({N}, L.options().dtype(torch::kUInt8));
    auto vs = views(th, nm, sz, L, R, M);
    int64_t off = 0;
    for (auto& w : vs) {
      w.xyz = xyz.data_ptr<double>() + off * 3;
      w.valid = valid.data_ptr<uint8_t>() + off;
      off += (int64_t)w.rows * w.cols;
    }
    check(psgc_adjoint_forward_batch(ctx, vs.data(), (int)V));
    actx->save_for_backward({th, L, R, M, nm, sz});
    actx->saved_data["ctx"] = ctx_handle;
    actx->saved_data["stream"] = stream;
    actx->saved_data["theta_device"] = thetas.device().str();
    actx->saved_data["theta_dtype"] = (int64_t)thetas.scalar_type();
    actx->mark_non_differentiable({valid});
    return {xyz, valid};
  }

  static tensor_list backward(AutogradContext* actx, tensor_list grads) {
    auto saved = actx->get_saved_variables();
    auto th = saved[0], L = saved[1], R = saved[2], M = saved[3], nm = saved[4], sz = saved[5];
    auto* ctx = reinterpret_cast<psgc_ctx*>(actx->saved_data["ctx"].toInt());
    check(psgc_set_stream(ctx, reinterpret_cast<void*>(actx->saved_data["stream"].toInt())));
    const int64_t V = sz.size(0);
    auto g = grads[0].to(torch::kDouble).contiguous();
    auto gth = torch::empty({V, 9}, L.options());
    const bool needL = actx->needs_input_grad(1), needR = actx->needs_input_grad(2);
    torch::Tensor gl, gr;
    if (needL) gl = torch::empty_like(L);
    if (needR) gr = torch::empty_like(R);
    auto vs = views(th, nm, sz, L, R, M);
    int64_t off = 0;
    for (int64_t v = 0; v < V; ++v) {
      auto& w = vs[v];
      w.grad_xyz = g.data_ptr<double>() + off * 3;
      w.grad_theta = gth.data_ptr<double>() + v * 9;
      w.grad_phL = needL ? gl.data_ptr<double>() + off : nullptr;
      w.grad_phR = needR ? gr.data_ptr<double>() + off : nullptr;
      off += (int64_t)w.rows * w.cols;
    }
    check(psgc_adjoint_backward_batch(ctx, vs.data(), (int)V));
    auto dev = torch::Device(actx->saved_data["theta_device"].toStringRef());
    auto dt = (torch::ScalarType)actx->saved_data["theta_dtype"].toInt();
    return {gth.to(dev, dt), gl, gr, torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor(), torch::Tensor()};
  }
};

std::vector<torch::Tensor> triangulate_batch(torch::Tensor thetas, torch::Tensor phL,
                                             torch::Tensor phR, torch::Tensor match_col,
                                             torch::Tensor nominal, torch::Tensor sizes,
                                             int64_t ctx_handle, int64_t stream) {
  return TriangulateBatch::apply(thetas, phL, phR, match_col, nominal, sizes, ctx_handle, stream);
}

std::vector<torch::Tensor> triangulate(torch::Tensor theta, torch::Tensor phL, torch::Tensor phR,
                                       torch::Tensor match_col, torch::Tensor nominal,
                                       int64_t ctx_handle, int64_t stream) {
  return Triangulate::apply(theta, phL, phR, match_col, nominal, ctx_handle, stream);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.def("triangulate", &triangulate,
        "Differentiable rectification + sub-pixel root + Q triangulation: "
        "(theta[9], phL, phR, match_col, nominal[5], ctx, stream) -> (xyz, valid)");
  m.def("triangulate_batch", &triangulate_batch,
        "triangulate over many views in one launch each way: (thetas[V,9], phL, phR, "
        "match_col (all views flattened), nominal[V,5], sizes[V,2], ctx, stream) -> "
        "(xyz[N,3], valid[N])");
}
