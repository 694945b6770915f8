// torch::autograd::Function over the frame loop: forward = frame 1 of the
// C++ wavefront renderer (wf_renderer_render), backward = the path-replay
// adjoint kernel (wf_renderer_adjoint).  Parameters are per-material albedo
// and emission, [num_materials, 3] float32 GPU tensors; the image is the
// linear frame estimate, [height, width, 3].
#include <torch/cuda.h>
#include <torch/extension.h>

#include <stdexcept>
#include <string>

#include "wfpt.h"

namespace {

void check(int status, const char* what) {
    if (status != WF_OK)
        throw std::runtime_error(std::string(what) + ": " + wf_status_string(status) + ": " +
                                 wf_last_error());
}

wf_renderer as_renderer(int64_t handle) {
    if (handle == 0) throw std::invalid_argument("null renderer handle");
    return reinterpret_cast<wf_renderer>(handle);
}

torch::Tensor as_param(const torch::Tensor& t, int64_t nmat, const char* name) {
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
    TORCH_CHECK(t.dim() == 2 && t.size(0) == nmat && t.size(1) == 3, name,
                " must have shape [num_materials, 3] = [", nmat, ", 3]");
    return t.detach().to(torch::kFloat32).contiguous();
}

int64_t material_count(wf_renderer r) {
    wf_scene sc = nullptr;
    check(wf_renderer_get_scene(r, &sc), "wf_renderer_get_scene");
    int n = 0;
    check(wf_scene_get_materials(sc, &n, nullptr, nullptr), "wf_scene_get_materials");
    return n;
}

void set_materials(wf_renderer r, const torch::Tensor& albedo, const torch::Tensor& emission) {
    const int64_t nmat = material_count(r);
    torch::Tensor a = as_param(albedo, nmat, "albedo"), e = as_param(emission, nmat, "emission");
    torch::cuda::synchronize();  // the tensors are written on torch's stream
    wf_scene sc = nullptr;
    check(wf_renderer_get_scene(r, &sc), "wf_renderer_get_scene");
    check(wf_scene_set_materials(sc, a.data_ptr<float>(), e.data_ptr<float>(), (int)nmat, nullptr),
          "wf_scene_set_materials");
}

struct RenderFunction : public torch::autograd::Function<RenderFunction> {
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor albedo,
                                 torch::Tensor emission, int64_t handle) {
        wf_renderer r = as_renderer(handle);
        set_materials(r, albedo, emission);
        check(wf_renderer_reset_frame_index(r), "wf_renderer_reset_frame_index");
        check(wf_renderer_render(r, nullptr), "wf_renderer_render");
        const float* px = nullptr;
        check(wf_renderer_pixels_device(r, 0, &px), "wf_renderer_pixels_device");
        int w = 0, h = 0;
        check(wf_renderer_size(r, &w, &h), "wf_renderer_size");
        auto opts = torch::TensorOptions().dtype(torch::kFloat32).device(albedo.device());
        torch::Tensor frame = torch::from_blob(const_cast<float*>(px), {h, w, 4}, opts);
        torch::Tensor img = frame.slice(2, 0, 3).contiguous();  // copy out of the renderer
        torch::cuda::synchronize();
        ctx->save_for_backward({albedo, emission});
        ctx->saved_data["handle"] = handle;
        return img;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out) {
        auto saved = ctx->get_saved_variables();
        wf_renderer r = as_renderer(ctx->saved_data["handle"].toInt());
        // the renderer's materials may have changed since the forward pass
        set_materials(r, saved[0], saved[1]);
        torch::Tensor g = grad_out[0].to(torch::kFloat32);
        const int64_t h = g.size(0), w = g.size(1), nmat = material_count(r);
        auto opts = torch::TensorOptions().dtype(torch::kFloat32).device(g.device());
        torch::Tensor d = torch::zeros({h, w, 4}, opts);
        d.slice(2, 0, 3).copy_(g);
        torch::Tensor ga = torch::empty({nmat, 3}, opts), ge = torch::empty({nmat, 3}, opts);
        torch::cuda::synchronize();
        check(wf_renderer_adjoint(r, 1, d.data_ptr<float>(), ga.data_ptr<float>(),
                                  ge.data_ptr<float>(), nullptr, nullptr),
              "wf_renderer_adjoint");
        return {ga, ge, torch::Tensor()};
    }
};

torch::Tensor render(torch::Tensor albedo, torch::Tensor emission, int64_t handle) {
    return RenderFunction::apply(albedo, emission, handle);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "torch autograd over the wavefront frame loop (path-replay adjoint backward)";
    m.def("render", &render, "image[h, w, 3] = render(albedo[nmat, 3], emission[nmat, 3], renderer)");
}
