// Timoshenko cantilever driven from C++ through the host layer: a straight
// rod (L = 3, n = 100, r = 0.25, rho = 5000, E = 1e6, G = E / 1.99, damping
// 0.1) clamped at node 0 with a tip force (0, -15, 0), dt = 0.01 L / n, the
// configuration of examples/timoshenko/TimoshenkoCaseStudy.cpp:27-183 with
// the rod along x.  Steps with the reference-order stepper over the
// op-by-op kernels (--plain) or with the fused FP64 kernels, and prints the
// tip position.  --resume PATH steps the first half with the fused kernels,
// writes a checkpoint to PATH, loads it into a new block and steps the rest.
//
//   timoshenko [steps] [--plain | --resume PATH]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "erod.hpp"

int main(int argc, char** argv) {
  const std::uint64_t steps = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1000;
  const bool plain = argc > 2 && std::strcmp(argv[2], "--plain") == 0;
  const char* resume = argc > 3 && std::strcmp(argv[2], "--resume") == 0 ? argv[3] : nullptr;
  try {
    erod::StraightRod p{100, 0.25, 3.0, {0, 0, 0}, {1, 0, 0}, {0, 0, 1}, 5000.0, 1e6, 1e6 / 1.99, 0.1, 0.1};
    std::vector<erod::DeviceBlock> blocks;
    blocks.emplace_back(std::vector<std::uint64_t>{p.n});
    auto fields = erod::straight_rod_state(p);
    erod::load_rod(blocks[0], 0, fields);
    const auto& x = fields[erod::tags::Position];
    const auto& Q = fields[erod::tags::Director];
    const double x0[3] = {x[0], x[p.n + 1], x[2 * (p.n + 1)]};
    double q0[9];
    for (int m = 0; m < 9; ++m) q0[m] = Q[m * p.n];
    blocks[0].clamp(0, x0, q0);
    const double F[3] = {0.0, -15.0, 0.0};
    blocks[0].end_force(0, F);
    const double dt = 0.01 * (p.length / double(p.n));
    double t = 0.0;
    if (plain) {
      for (std::uint64_t i = 0; i < steps; ++i) erod::PositionVerlet::step(blocks, t, dt);
    } else if (resume) {
      blocks[0].set_dt(dt);
      blocks[0].step_fused(steps / 2, EROD_FP64);
      blocks[0].save(resume);
      blocks.clear();
      blocks.push_back(erod::DeviceBlock::load(resume));
      blocks[0].step_fused(steps - steps / 2, EROD_FP64);
      t = double(steps) * dt;
    } else {
      blocks[0].set_dt(dt);
      blocks[0].step_fused(steps, EROD_FP64);
      t = double(steps) * dt;
    }
    blocks[0].synchronize();
    const auto xe = blocks[0].get(erod::tags::Position, 0);
    std::printf("t %.6f tip %.17g %.17g %.17g\n", t, xe[p.n], xe[(p.n + 1) + p.n], xe[2 * (p.n + 1) + p.n]);
  } catch (const erod::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
