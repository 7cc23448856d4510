// lh_headless -- the reference's simulation loop without a window (main.cpp:31-53,
// with the recorder of main.cpp:36-44): builds the scene from a JSON config, runs
// PreStep / Step / PostStep / GetCFL and optionally appends every frame to a
// trajectory file in the ParticleSystem::WriteToFile format.  With --play it is the
// reference's replay instead (Scene<3>::PlaySimulation, main.cpp:29): the trajectory
// file's first count line resizes the particle system, every following line goes
// through LoadFromLine, and where the reference renders a frame this prints the frame's
// particle count and centroid.
//
//   lh_headless <config.json> [steps] [trajectory.csv]
//   lh_headless <config.json> --play <trajectory.csv>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lh/liquid.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <config.json> [steps] [trajectory.csv]\n       %s <config.json> --play <trajectory.csv>\n",
                 argv[0], argv[0]);
    return 2;
  }
  try {
    lh::Scene scene(argv[1]);
    lh::LiquidModel& model = scene.GetLiquidModel();
    if (argc > 3 && std::strcmp(argv[2], "--play") == 0) {
      long frame = 0;
      long n = scene.PlaySimulation(argv[3], [&frame](lh::ParticleSystem& ps) {
        std::vector<float> pos;
        ps.Get(pos);
        double c[3] = {0.0, 0.0, 0.0};
        const size_t np = pos.size() / 3;
        for (size_t i = 0; i < np; ++i)
          for (int a = 0; a < 3; ++a) c[a] += pos[3 * i + a];
        for (int a = 0; a < 3; ++a) c[a] = np ? c[a] / (double)np : 0.0;
        std::printf("frame %ld: %zu particles, centroid %.17g %.17g %.17g\n", frame++, np, c[0], c[1], c[2]);
      });
      std::printf("replayed %ld lines\n", n);
      return 0;
    }
    const int steps = argc > 2 ? std::atoi(argv[2]) : 100;
    const std::string out = argc > 3 ? argv[3] : "";
    model.Init();
    float dt = 0.01f;  // main.cpp:32
    if (!out.empty()) model.GetParticleSystem().WriteToFile(out, false);
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < steps; ++k) {
      model.PreStep(dt);
      model.Step(dt);
      lh_step_stats st = model.PostStep(dt);
      dt = model.GetCFL();
      if (!out.empty()) model.GetParticleSystem().WriteToFile(out, true);
      std::printf("step %d: pcg iterations %d, divergence mean %.3e linf %.3e, next dt %.9g\n", k,
                  st.solve.iterations, st.divergence_mean, st.divergence_linf, dt);
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%d steps in %.3f s (%.2f steps/s), %lld particles, liquid density %g\n", steps, s, steps / s,
                (long long)model.GetParticleSystem().NumParticles(), scene.GetLiquidDensity());
  } catch (const lh::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
