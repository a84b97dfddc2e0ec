// torch.ops.tlbm: the C ABI of libtlbm_hip.so as dispatcher operators.
//
// The reference's callers hand torch tensors to its fused step through the backend step-impl
// layer (src/tensorlbm/backends/triton_backend.py:156) and to the functional ops
// (__init__.py:250-282).  Registering the entry points with torch's dispatcher makes them
// visible to torch.compile, graph capture and the profiler:
//
//   * the CUDA key (GPU tensors on ROCm torch) checks shapes, dtypes and devices and calls the
//     C entry point on torch's current stream of the tensor's device;
//   * the Meta key returns outputs of the right shape without touching a device, so tracing
//     with FakeTensors works;
//   * the CPU key raises, as the reference's fused path does for a non-GPU device
//     (triton_fused.py:742-747).
//
// Status codes map to the reference's exception types: shape, dtype, alias and argument errors
// raise ValueError, combinations that are not built NotImplementedError, HIP failures
// RuntimeError.  Outputs are new contiguous tensors; inputs are never written, except the
// optional force accumulator of `step`, which is added to (the caller zeroes it).
#include <ATen/ATen.h>
#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <string>
#include <vector>

#include "tlbm_hip.h"

namespace {

void check(int rc, const char* what) {
  if (rc == TLBM_OK) return;
  const std::string msg = std::string(what) + ": " + tlbm_last_error();
  TORCH_CHECK_NOT_IMPLEMENTED(rc != TLBM_ERR_UNSUPPORTED, msg);
  TORCH_CHECK(rc != TLBM_ERR_HIP, msg);
  TORCH_CHECK_VALUE(false, msg);
}

tlbm_stream_t stream_of(const at::Tensor& t) {
  return reinterpret_cast<tlbm_stream_t>(c10::hip::getCurrentHIPStream(t.get_device()).stream());
}

// storage code of a population tensor; `deviation`: 16-bit storage holds f_i - w_i
int store_code(const at::Tensor& f, bool deviation, const char* what) {
  switch (f.scalar_type()) {
    case at::kFloat:
      TORCH_CHECK_VALUE(!deviation, what, ": deviation storage is a 16-bit form (float16 / bfloat16)");
      return TLBM_STORE_F32;
    case at::kDouble:
      TORCH_CHECK_VALUE(!deviation, what, ": deviation storage is a 16-bit form (float16 / bfloat16)");
      return TLBM_STORE_F64;
    case at::kHalf:
      return deviation ? TLBM_STORE_F16_DEV : TLBM_STORE_F16;
    case at::kBFloat16:
      return deviation ? TLBM_STORE_BF16_DEV : TLBM_STORE_BF16;
    default:
      TORCH_CHECK_VALUE(false, what, ": Unsupported dtype ", f.scalar_type(),
                        "; use float32, float64, float16, or bfloat16.");
  }
  return -1;
}

// (Q, nz, ny, nx) populations (D2Q9: (9, ny, nx)) with unit x stride, any other strides
struct Pops {
  int q, nz, ny, nx;
  int64_t strides[3];  // q, z, y element strides
};

Pops pops(const at::Tensor& f, const char* what) {
  TORCH_CHECK_VALUE(f.dim() == 3 || f.dim() == 4, what, ": populations must be (9, ny, nx) or (Q, nz, ny, nx), got ",
                    f.sizes());
  Pops p{};
  p.q = int(f.size(0));
  TORCH_CHECK_VALUE((f.dim() == 3 && p.q == 9) || (f.dim() == 4 && (p.q == 19 || p.q == 27)), what,
                    ": populations must be (9, ny, nx), (19, nz, ny, nx) or (27, nz, ny, nx), got ", f.sizes());
  p.nz = f.dim() == 4 ? int(f.size(1)) : 1;
  p.ny = int(f.size(-2));
  p.nx = int(f.size(-1));
  TORCH_CHECK_VALUE(f.numel() > 0, what, ": empty population tensor");
  TORCH_CHECK_VALUE(f.stride(-1) == 1, what, ": the x stride of the populations must be 1");
  p.strides[0] = f.stride(0);
  p.strides[2] = f.stride(-2);
  p.strides[1] = f.dim() == 4 ? f.stride(1) : int64_t(p.ny) * f.stride(-2);
  return p;
}

void on_gpu(const at::Tensor& t, const char* what) {
  TORCH_CHECK(t.is_cuda(), what, " runs on a GPU device; got a tensor on ", t.device());
}

tlbm_collision collision(int64_t kind, double tau, at::ArrayRef<double> rates, int64_t tau_mode, double cs_delta_sq,
                         const c10::optional<at::Tensor>& tau_eff, bool eq_rho_floor, const at::Tensor& f,
                         const Pops& p, const char* what) {
  TORCH_CHECK_VALUE(rates.size() <= 8, what, ": at most 8 rates");
  TORCH_CHECK_NOT_IMPLEMENTED(tau_mode != TLBM_TAU_MLP, what,
                              ": the learned closure needs its packed network; use FusedSolver3D(closure=...)");
  tlbm_collision c{};
  c.kind = int32_t(kind);
  c.tau_mode = int32_t(tau_mode);
  c.eq_rho_floor = eq_rho_floor ? 1 : 0;
  c.tau = tau;
  c.cs_delta_sq = cs_delta_sq;
  for (size_t i = 0; i < rates.size(); ++i) c.rate[i] = rates[i];
  if (tau_mode == TLBM_TAU_FIELD || tau_mode == TLBM_TAU_FIELD64) {
    TORCH_CHECK_VALUE(tau_eff.has_value(), what, ": a per-cell tau mode needs tau_eff");
    const at::Tensor& t = *tau_eff;
    on_gpu(t, what);
    TORCH_CHECK_VALUE(t.device() == f.device(), what, ": tau_eff must be on the populations' device");
    TORCH_CHECK_VALUE(t.scalar_type() == (tau_mode == TLBM_TAU_FIELD ? at::kFloat : at::kDouble), what,
                      ": tau_eff must be float32 (TLBM_TAU_FIELD) or float64 (TLBM_TAU_FIELD64)");
    TORCH_CHECK_VALUE(t.is_contiguous() && t.numel() == int64_t(p.nz) * p.ny * p.nx, what,
                      ": tau_eff must be a contiguous field of the grid's shape");
    c.tau_eff = static_cast<const float*>(t.data_ptr());
  }
  return c;
}

// ---- CUDA (HIP) kernels ------------------------------------------------------

at::Tensor equilibrium_gpu(int64_t q, const at::Tensor& rho, const at::Tensor& ux, const at::Tensor& uy,
                           const c10::optional<at::Tensor>& uz) {
  const char* what = "tlbm::equilibrium";
  on_gpu(rho, what);
  TORCH_CHECK_VALUE(q == 9 || q == 19 || q == 27, what, ": Q must be 9, 19 or 27");
  TORCH_CHECK_VALUE((q == 9) == !uz.has_value(), what, ": uz is given exactly for the 3-D lattices");
  std::vector<at::Tensor> in{rho, ux, uy};
  if (uz) in.push_back(*uz);
  for (const auto& t : in) {
    TORCH_CHECK_VALUE(t.sizes() == rho.sizes(), what, ": shapes must match");
    TORCH_CHECK_VALUE(t.device() == rho.device() && t.scalar_type() == rho.scalar_type(), what,
                      ": fields must share device and dtype");
  }
  const int dt = store_code(rho, false, what);
  for (auto& t : in) t = t.contiguous();
  std::vector<int64_t> shape{q};
  for (auto s : rho.sizes()) shape.push_back(s);
  at::Tensor out = at::empty(shape, rho.options());
  const c10::DeviceGuard guard(rho.device());
  check(tlbm_equilibrium(int(q), dt, in[0].data_ptr(), in[1].data_ptr(), in[2].data_ptr(),
                         uz ? in[3].data_ptr() : nullptr, out.data_ptr(), rho.numel(), stream_of(rho)),
        what);
  return out;
}

std::vector<at::Tensor> macroscopic_gpu(const at::Tensor& f) {
  const char* what = "tlbm::macroscopic";
  on_gpu(f, what);
  const Pops p = pops(f, what);
  const int dt = store_code(f, false, what);
  const at::Tensor fc = f.contiguous();
  std::vector<int64_t> shape(f.sizes().begin() + 1, f.sizes().end());
  const int n = p.q == 9 ? 3 : 4;
  std::vector<at::Tensor> out;
  for (int i = 0; i < n; ++i) out.push_back(at::empty(shape, f.options()));
  const c10::DeviceGuard guard(f.device());
  check(tlbm_macroscopic(p.q, dt, fc.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                         n == 4 ? out[3].data_ptr() : nullptr, int64_t(p.nz) * p.ny * p.nx, stream_of(f)),
        what);
  return out;
}

at::Tensor stream_gpu(const at::Tensor& f) {
  const char* what = "tlbm::stream";
  on_gpu(f, what);
  const Pops p = pops(f, what);
  const int st = store_code(f, false, what);
  at::Tensor out = at::empty(f.sizes(), f.options());
  const Pops po = pops(out, what);
  const c10::DeviceGuard guard(f.device());
  check(tlbm_stream(p.q, st, f.data_ptr(), p.strides, out.data_ptr(), po.strides, p.nz, p.ny, p.nx, stream_of(f)),
        what);
  return out;
}

at::Tensor collide_gpu(const at::Tensor& f, int64_t kind, double tau, at::ArrayRef<double> rates, int64_t tau_mode,
                       double cs_delta_sq, const c10::optional<at::Tensor>& tau_eff) {
  const char* what = "tlbm::collide";
  on_gpu(f, what);
  const Pops p = pops(f, what);
  const int st = store_code(f, false, what);
  const tlbm_collision c = collision(kind, tau, rates, tau_mode, cs_delta_sq, tau_eff, false, f, p, what);
  at::Tensor out = at::empty(f.sizes(), f.options());
  const Pops po = pops(out, what);
  const c10::DeviceGuard guard(f.device());
  check(tlbm_collide(p.q, st, f.data_ptr(), p.strides, out.data_ptr(), po.strides, p.nz, p.ny, p.nx, &c,
                     stream_of(f)),
        what);
  return out;
}

at::Tensor step_gpu(const at::Tensor& f, int64_t kind, double tau, at::ArrayRef<double> rates, int64_t tau_mode,
                    double cs_delta_sq, const c10::optional<at::Tensor>& tau_eff,
                    const c10::optional<at::Tensor>& solid, int64_t bc_kind, at::ArrayRef<double> bc_u,
                    int64_t bc_faces, const c10::optional<at::Tensor>& force, bool deviation, int64_t steps) {
  const char* what = "tlbm::step";
  on_gpu(f, what);
  const Pops p = pops(f, what);
  const int st = store_code(f, deviation, what);
  TORCH_CHECK_VALUE(steps == 1 || steps == 2, what, ": steps must be 1 or 2");
  const tlbm_collision c = collision(kind, tau, rates, tau_mode, cs_delta_sq, tau_eff, true, f, p, what);
  at::Tensor mask;
  if (solid) {
    on_gpu(*solid, what);
    TORCH_CHECK_VALUE(solid->device() == f.device(), what, ": the solid mask must be on the populations' device");
    TORCH_CHECK_VALUE(solid->numel() == int64_t(p.nz) * p.ny * p.nx, what,
                      ": the solid mask must have the grid's shape");
    TORCH_CHECK_VALUE(solid->scalar_type() == at::kBool || solid->scalar_type() == at::kChar, what,
                      ": the solid mask must be bool or int8");
    mask = solid->to(at::kChar).contiguous();
  }
  TORCH_CHECK_VALUE(bc_u.size() == 3, what, ": bc_u takes three velocity components");
  tlbm_boundary bc{};
  bc.kind = int32_t(bc_kind);
  bc.faces = int32_t(bc_faces);
  for (int i = 0; i < 3; ++i) bc.u[i] = bc_u[i];
  double* acc = nullptr;
  if (force) {
    on_gpu(*force, what);
    TORCH_CHECK_VALUE(force->device() == f.device() && force->scalar_type() == at::kDouble &&
                          force->numel() == 3 && force->is_contiguous(),
                      what, ": force must be a contiguous float64 device tensor with 3 entries");
    acc = static_cast<double*>(force->data_ptr());
  }
  at::Tensor out = at::empty(f.sizes(), f.options());
  const Pops po = pops(out, what);
  const c10::DeviceGuard guard(f.device());
  const int8_t* m = mask.defined() ? static_cast<const int8_t*>(mask.data_ptr()) : nullptr;
  const tlbm_boundary* b = bc_kind != TLBM_BC_NONE ? &bc : nullptr;
  if (steps == 2) {
    check(tlbm_step2(p.q, st, f.data_ptr(), p.strides, out.data_ptr(), po.strides, p.nz, p.ny, p.nx, &c, m, b, acc,
                     stream_of(f)),
          what);
  } else {
    check(tlbm_step(p.q, st, f.data_ptr(), p.strides, out.data_ptr(), po.strides, p.nz, p.ny, p.nx, &c, m, b, acc,
                    stream_of(f)),
          what);
  }
  return out;
}

// ---- Meta (shapes only) ------------------------------------------------------

at::Tensor equilibrium_meta(int64_t q, const at::Tensor& rho, const at::Tensor&, const at::Tensor&,
                            const c10::optional<at::Tensor>&) {
  std::vector<int64_t> shape{q};
  for (auto s : rho.sizes()) shape.push_back(s);
  return at::empty(shape, rho.options());
}

std::vector<at::Tensor> macroscopic_meta(const at::Tensor& f) {
  std::vector<int64_t> shape(f.sizes().begin() + 1, f.sizes().end());
  const int n = f.size(0) == 9 ? 3 : 4;
  std::vector<at::Tensor> out;
  for (int i = 0; i < n; ++i) out.push_back(at::empty(shape, f.options()));
  return out;
}

at::Tensor stream_meta(const at::Tensor& f) { return at::empty(f.sizes(), f.options()); }

at::Tensor collide_meta(const at::Tensor& f, int64_t, double, at::ArrayRef<double>, int64_t, double,
                        const c10::optional<at::Tensor>&) {
  return at::empty(f.sizes(), f.options());
}

at::Tensor step_meta(const at::Tensor& f, int64_t, double, at::ArrayRef<double>, int64_t, double,
                     const c10::optional<at::Tensor>&, const c10::optional<at::Tensor>&, int64_t,
                     at::ArrayRef<double>, int64_t, const c10::optional<at::Tensor>&, bool, int64_t) {
  return at::empty(f.sizes(), f.options());
}

// ---- CPU: not on the HIP path -------------------------------------------------

[[noreturn]] void no_cpu(const char* what) {
  TORCH_CHECK(false, what, " runs on a GPU device (the MI355X HIP path); got CPU tensors");
  __builtin_unreachable();
}
at::Tensor equilibrium_cpu(int64_t, const at::Tensor&, const at::Tensor&, const at::Tensor&,
                           const c10::optional<at::Tensor>&) {
  no_cpu("tlbm::equilibrium");
}
std::vector<at::Tensor> macroscopic_cpu(const at::Tensor&) { no_cpu("tlbm::macroscopic"); }
at::Tensor stream_cpu(const at::Tensor&) { no_cpu("tlbm::stream"); }
at::Tensor collide_cpu(const at::Tensor&, int64_t, double, at::ArrayRef<double>, int64_t, double,
                       const c10::optional<at::Tensor>&) {
  no_cpu("tlbm::collide");
}
at::Tensor step_cpu(const at::Tensor&, int64_t, double, at::ArrayRef<double>, int64_t, double,
                    const c10::optional<at::Tensor>&, const c10::optional<at::Tensor>&, int64_t, at::ArrayRef<double>,
                    int64_t, const c10::optional<at::Tensor>&, bool, int64_t) {
  no_cpu("tlbm::step");
}

}  // namespace

TORCH_LIBRARY(tlbm, m) {
  m.def("equilibrium(int Q, Tensor rho, Tensor ux, Tensor uy, Tensor? uz=None) -> Tensor");
  m.def("macroscopic(Tensor f) -> Tensor[]");
  m.def("stream(Tensor f) -> Tensor");
  m.def(
      "collide(Tensor f, int kind, float tau, float[] rates=[], int tau_mode=0, float cs_delta_sq=0., "
      "Tensor? tau_eff=None) -> Tensor");
  m.def(
      "step(Tensor f, int kind, float tau, float[] rates=[], int tau_mode=0, float cs_delta_sq=0., "
      "Tensor? tau_eff=None, Tensor? solid=None, int bc_kind=0, float[] bc_u=[0., 0., 0.], int bc_faces=0, "
      "Tensor(a!)? force=None, bool deviation=False, int steps=1) -> Tensor");
}

TORCH_LIBRARY_IMPL(tlbm, CUDA, m) {
  m.impl("equilibrium", &equilibrium_gpu);
  m.impl("macroscopic", &macroscopic_gpu);
  m.impl("stream", &stream_gpu);
  m.impl("collide", &collide_gpu);
  m.impl("step", &step_gpu);
}

TORCH_LIBRARY_IMPL(tlbm, Meta, m) {
  m.impl("equilibrium", &equilibrium_meta);
  m.impl("macroscopic", &macroscopic_meta);
  m.impl("stream", &stream_meta);
  m.impl("collide", &collide_meta);
  m.impl("step", &step_meta);
}

TORCH_LIBRARY_IMPL(tlbm, CPU, m) {
  m.impl("equilibrium", &equilibrium_cpu);
  m.impl("macroscopic", &macroscopic_cpu);
  m.impl("stream", &stream_cpu);
  m.impl("collide", &collide_cpu);
  m.impl("step", &step_cpu);
}
