// lh/liquid.hpp -- C++ facade with the reference's class surface over the C ABI.
//
// Mirrors Scene<3> (Scene.h:14-48), LiquidModel<3> (LiquidModel.h:11-61),
// CGSolver<3> (PressureSolver.h:81-145) and ParticleSystem<3>
// (ParticleSystem.h:24-61) so that a driver written against the reference, e.g.
// main.cpp:31-53, compiles against this header with the namespace changed:
//
//   lh::Scene scene("pure_liquid_3d.json");
//   lh::LiquidModel& model = scene.GetLiquidModel();
//   model.Init();
//   float dt = 0.01f;
//   while (running) { model.PreStep(dt); model.Step(dt); model.PostStep(dt); dt = model.GetCFL(); }
//
// Errors throw lh::Error (a std::runtime_error), as the reference throws
// std::runtime_error / nlohmann::json exceptions.  All state lives in HBM;
// the accessors copy in and out.
#pragma once

#include <array>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../lh_api.h"

namespace lh {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace detail {
inline void check(lh_ctx* ctx, int rc) {
  if (rc != LH_OK) throw Error(rc, ctx ? lh_last_error(ctx) : "lh: null context");
}
}  // namespace detail

// ParticleSystem<3>: host copies of the particles held on the device
class ParticleSystem {
 public:
  explicit ParticleSystem(lh_ctx* ctx) : ctx_(ctx) {}
  int64_t NumParticles() const {
    int64_t n = 0;
    detail::check(ctx_, lh_num_particles(ctx_, &n));
    return n;
  }
  // positions / velocities as x,y,z triples; affine matrices column-major (9 per particle)
  void Get(std::vector<float>& pos, std::vector<float>* vel = nullptr, std::vector<float>* C = nullptr) const {
    const int64_t n = NumParticles();
    pos.resize(3 * n);
    if (vel) vel->resize(3 * n);
    if (C) C->resize(9 * n);
    detail::check(ctx_, lh_get_particles(ctx_, pos.data(), vel ? vel->data() : nullptr, C ? C->data() : nullptr, n));
  }
  void Set(const std::vector<float>& pos, const std::vector<float>* vel = nullptr,
           const std::vector<float>* C = nullptr) {
    const int64_t n = (int64_t)(pos.size() / 3);
    detail::check(ctx_, lh_set_particles(ctx_, pos.data(), vel ? vel->data() : nullptr, C ? C->data() : nullptr, n));
  }
  void WriteToFile(const std::string& filename, bool isAppend = false) const {
    detail::check(ctx_, lh_write_positions(ctx_, filename.c_str(), isAppend ? 1 : 0));
  }
  void LoadFromLine(const std::string& line) { detail::check(ctx_, lh_load_positions_line(ctx_, line.c_str())); }
  void Resize(int64_t size) { detail::check(ctx_, lh_resize_particles(ctx_, size)); }

 private:
  lh_ctx* ctx_;
};

// CGSolver<3>: Solve(liquidDensity, dt) builds the system with that density and writes
// Grid<3>::pressure (PressureSolver.h:104)
class CGSolver {
 public:
  explicit CGSolver(lh_ctx* ctx) : ctx_(ctx) {}
  lh_solve_stats Solve(float liquidDensity, float dt) {
    lh_solve_stats st{};
    int rc = lh_solve_pressure(ctx_, liquidDensity, dt, &st);
    if (rc != LH_OK && rc != LH_ERR_NOT_CONVERGED) detail::check(ctx_, rc);
    return st;
  }
  // learned initial guess (no reference counterpart: the reference starts from x = 0)
  void LoadWarmStart(const std::string& path) { detail::check(ctx_, lh_warmstart_load(ctx_, path.c_str())); }
  void SaveWarmStart(const std::string& path) { detail::check(ctx_, lh_warmstart_save(ctx_, path.c_str())); }
  void InitWarmStart(uint64_t seed) { detail::check(ctx_, lh_warmstart_init(ctx_, seed)); }
  void EnableWarmStart(bool on) { detail::check(ctx_, lh_warmstart_enable(ctx_, on ? 1 : 0)); }
  // one Adam step on the last built system and its solved pressure; returns the loss
  float TrainWarmStart(float lr) {
    float loss = 0.f;
    detail::check(ctx_, lh_warmstart_train(ctx_, lr, nullptr, &loss));
    return loss;
  }

 private:
  lh_ctx* ctx_;
};

class LiquidModel {
 public:
  explicit LiquidModel(lh_ctx* ctx) : ctx_(ctx), ps_(ctx), solver_(ctx) {}
  void Init() { detail::check(ctx_, lh_init(ctx_)); }
  void PreStep(float dt) { detail::check(ctx_, lh_pre_step(ctx_, dt)); }
  void Step(float dt) {
    int rc = lh_step(ctx_, dt);
    if (rc != LH_OK && rc != LH_ERR_NOT_CONVERGED) detail::check(ctx_, rc);
  }
  lh_step_stats PostStep(float dt) {
    lh_step_stats st{};
    detail::check(ctx_, lh_post_step(ctx_, dt, &st));
    return st;
  }
  float GetCFL() const {
    float dt = 0.f;
    detail::check(ctx_, lh_get_cfl(ctx_, &dt));
    return dt;
  }
  // LiquidModel.h:57, with the boundary tree as the JSON text of a config "Boundary" node;
  // ComputeSolidPhi() afterwards (each step for a moving solid)
  void SetBoundary(const std::string& boundaryJson) { detail::check(ctx_, lh_set_boundary(ctx_, boundaryJson.c_str())); }
  // snapshots (no reference counterpart): what the next step depends on
  void SaveState(const std::string& path) const { detail::check(ctx_, lh_save(ctx_, path.c_str())); }
  void LoadState(const std::string& path) { detail::check(ctx_, lh_load(ctx_, path.c_str())); }
  // stage methods (LiquidModel.h:25-47)
  void ComputeLiquidPhi() { detail::check(ctx_, lh_compute_liquid_phi(ctx_)); }
  void ComputeSolidPhi() { detail::check(ctx_, lh_compute_solid_phi(ctx_)); }
  void Particle2Grid() { detail::check(ctx_, lh_p2g(ctx_)); }
  void Grid2Particle() { detail::check(ctx_, lh_g2p(ctx_)); }
  void ApplyGravity(float dt) { detail::check(ctx_, lh_apply_gravity(ctx_, dt)); }
  // LiquidModel<3>::SolvePressure (LiquidModel.cpp:825-890): Solve(scene density, dt), velocity update
  void SolvePressure(float dt) {
    float rho = 0.f;
    detail::check(ctx_, lh_liquid_density(ctx_, &rho));
    solver_.Solve(rho, dt);
    detail::check(ctx_, lh_velocity_update(ctx_, dt));
  }
  void Projection(float dt) { SolvePressure(dt); }
  void AdvectParticle(float dt) { detail::check(ctx_, lh_advect(ctx_, dt)); }
  void Extrapolate() { detail::check(ctx_, lh_extrapolate(ctx_)); }
  float ComputeDivergence() const {
    double mean = 0, linf = 0;
    int cnt = 0;
    detail::check(ctx_, lh_divergence(ctx_, &mean, &linf, &cnt));
    return (float)mean;
  }
  ParticleSystem& GetParticleSystem() { return ps_; }
  CGSolver& GetSolver() { return solver_; }
  // Grid<3> public arrays, Array3 layout (x-major, z fastest)
  std::array<int, 3> FieldShape(lh_field f) const {
    std::array<int, 3> s{};
    detail::check(ctx_, lh_field_shape(ctx_, f, s.data()));
    return s;
  }
  std::vector<float> GetField(lh_field f) const {
    auto s = FieldShape(f);
    std::vector<float> a((size_t)s[0] * s[1] * s[2]);
    detail::check(ctx_, lh_get_field(ctx_, f, a.data(), a.size()));
    return a;
  }

 private:
  lh_ctx* ctx_;
  ParticleSystem ps_;
  CGSolver solver_;
};

// Scene<3>(configFile): parses the JSON config and owns the liquid model
class Scene {
 public:
  explicit Scene(const std::string& configFile) {
    lh_ctx* ctx = nullptr;
    int rc = lh_create(configFile.c_str(), &ctx);
    if (rc != LH_OK) {
      std::string msg = ctx ? lh_last_error(ctx) : "lh_create failed";
      if (ctx) lh_destroy(ctx);
      throw Error(rc, msg);
    }
    ctx_.reset(ctx);
    model_.reset(new LiquidModel(ctx));
  }
  LiquidModel& GetLiquidModel() { return *model_; }
  std::array<int, 3> GetGridDimension() const {
    std::array<int, 3> d{};
    std::array<float, 3> mn{}, dx{};
    detail::check(ctx_.get(), lh_grid_info(ctx_.get(), d.data(), mn.data(), dx.data()));
    return d;
  }
  float GetLiquidDensity() const {  // Scene.h:23
    float rho = 0.f;
    detail::check(ctx_.get(), lh_liquid_density(ctx_.get(), &rho));
    return rho;
  }
  // Scene<3>::PlaySimulation (Scene.cpp:26-45) with the renderer replaced by onFrame: the
  // first line's count resizes the particle system, then every following line goes through
  // LoadFromLine (count lines included, as in the reference).  Returns the lines replayed.
  template <class OnFrame>
  long PlaySimulation(const std::string& filename, OnFrame&& onFrame) {
    std::ifstream reader(filename);
    if (!reader.is_open()) throw Error(LH_ERR_ARG, "Failed to open file " + filename);
    std::string line;
    if (!std::getline(reader, line)) throw Error(LH_ERR_ARG, "empty trajectory file " + filename);
    ParticleSystem& ps = model_->GetParticleSystem();
    ps.Resize((int64_t)std::stoul(line));
    long frames = 0;
    while (std::getline(reader, line)) {
      ps.LoadFromLine(line);
      onFrame(ps);
      ++frames;
    }
    return frames;
  }
  long PlaySimulation(const std::string& filename) {
    return PlaySimulation(filename, [](ParticleSystem&) {});
  }
  lh_ctx* handle() const { return ctx_.get(); }

 private:
  struct Del {
    void operator()(lh_ctx* c) const { lh_destroy(c); }
  };
  std::unique_ptr<lh_ctx, Del> ctx_;
  std::unique_ptr<LiquidModel> model_;
};

}  // namespace lh
