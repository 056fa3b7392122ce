// Engine interface between the C-ABI layer and the templated HIP driver.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "ls_layout.h"

struct ls_step_params;
struct ls_stats;

namespace lsamd {

class Engine {
 public:
  virtual ~Engine() {}
  // Uploads the layout; geometry is rounded to the engine's working type.
  // Device arrays from the layout: adopted from D when given (D's pointers
  // are taken over and cleared), uploaded from L otherwise.
  virtual int upload(const Layout& L, DeviceLayout* D, std::string& err) = 0;
  virtual int set_phi(const double* phi_user, std::string& err) = 0;
  // Velocity per original element (E x 3); projected on the device into every
  // copy's local frame. mode (LS_VEL_*): each element's own velocity, the
  // patch's first copy's, or (triangles) the advection's velocity from phi.
  virtual int set_velocity(const double* vel, int64_t E, int mode, std::string& err) = 0;
  // Per-patch timestep and narrowband width (P each); empty vectors restore
  // the step's uniform values.
  virtual int set_patch_params(const std::vector<double>& dt, const std::vector<double>& bw,
                               std::string& err) = 0;
  // Mesh of each patch of a batch (P values in [0, ngroups)); a warm start is
  // accepted or rejected per mesh. ngroups <= 1: one mesh.
  virtual int set_patch_groups(const std::vector<int32_t>& group, int ngroups, std::string& err) = 0;
  virtual int step(const ls_step_params& prm, ls_stats& st, std::string& err) = 0;
  virtual int get_phi(double* phi_user, std::string& err) = 0;
  virtual int get_narrowband(std::vector<int32_t>& ids, std::string& err) = 0;
  virtual int working_bytes() const = 0;  // sizeof working real type
  // Learned warm start (tet and triangle engines).
  virtual int mlp_set_params(const float* prm, std::string& err) = 0;
  virtual int mlp_get_params(float* prm, std::string& err) = 0;
  virtual int mlp_predict(std::string& err) = 0;
  virtual int mlp_collect(int phase, int64_t capacity, std::string& err) = 0;
  virtual int64_t mlp_dataset_size() const = 0;
  virtual int mlp_get_dataset(float* X, float* Y, std::string& err) = 0;
  virtual int mlp_gradient(const float* X, const float* Y, int64_t n, float* grad, double* loss,
                           std::string& err) = 0;
  virtual int mlp_train(int epochs, int batch, double lr, uint64_t seed, double* loss, std::string& err) = 0;
  // mlp_gradient through the library's static scratch: no device allocation.
  virtual int mlp_gradient_noalloc(const float* X, const float* Y, int64_t n, float* grad, double* loss,
                                   std::string& err) = 0;
  virtual int get_distances(double* out_user, std::string& err) = 0;
  virtual int set_warm(const double* warm_user, std::string& err) = 0;  // caller order
  virtual int get_warm(double* out_user, std::string& err) = 0;         // caller order
  // CFL coefficients per original element (E values each; null: defaults).
  virtual int set_cfl_coefficients(const double* ceik, const double* ccurv, int64_t E, std::string& err) = 0;
  virtual int compute_dt(double* dt, double* max_rin, std::string& err) = 0;
  virtual void set_exec(int mode) = 0;
  virtual int set_sweep_mode(int mode, std::string& err) = 0;  // LS_SWEEP_*
  virtual int sweep_tasks() const = 0;                         // 1: task-compacted sweeps in use
};

// dim 3 -> tet engine with FP32 (LS_F32) or FP64 (LS_F64) storage,
// dim 2 -> FP64 triangle engine.
Engine* make_engine(int dim, int dtype, int device, std::string& err);

}  // namespace lsamd
