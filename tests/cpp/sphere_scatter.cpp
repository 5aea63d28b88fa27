// The dielectric-sphere scatter of XFDTD_CORE's README (a 0.35 m air cube at
// dl = 7.5 mm, a sphere r = 0.1 m with eps_r = 3 and mu_r = 2, a TFSF3D
// Gaussian plane wave from index 15, six 8-cell PMLs, CFL 0.9, an Ex XZ
// movie) written against include/xfdtd_core_amd.hpp the way a user of the
// reference writes it, plus the rest of the reference's user-facing surface:
// xfdtd::constant, a FieldTimeMonitor and a FieldMonitor written by
// output(), the default visitor and a counting one, and the time axis from
// calculationParam().
//
//   sphere_scatter STEPS OUT_DIR [f32]
//
// Writes OUT_DIR/movie/%05d.npy, OUT_DIR/probe/ex_centre.npy (2 x STEPS) and
// OUT_DIR/plane/ez_xy.npy, and prints the time axis ("time: ..."), dt and
// the constants ("dt: ...") and the visitor's counts ("flags: ...").
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "xfdtd_core_amd.hpp"

namespace {

class CountingVisitor : public xfdtd::SimulationFlagVisitor {
 public:
  auto initStep(xfdtd::SimulationInitFlag flag) -> void override { ++init[static_cast<int>(flag)]; }
  auto iteratorStep(xfdtd::SimulationIteratorFlag flag, xfdtd::Index cur, xfdtd::Index start,
                    xfdtd::Index end) -> void override {
    if (flag == xfdtd::SimulationIteratorFlag::NextStep) {
      ++next;
      last = cur;
      last_end = end;
      monotonic = monotonic && cur > prev && start == 0;
      prev = cur;
    }
  }
  int init[6] = {0, 0, 0, 0, 0, 0};
  int next = 0;
  xfdtd::Index last = 0, last_end = 0, prev = 0;
  bool monotonic = true;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s STEPS OUT_DIR [f32]\n", argv[0]);
    return 2;
  }
  const xfdtd::Index steps = std::strtoull(argv[1], nullptr, 10);
  const std::string out = argv[2];
  const bool f32 = argc > 3 && std::string(argv[3]) == "f32";

  xfdtd::Real dl{7.5e-3};
  auto domain{std::make_shared<xfdtd::Object>(
      "domain", std::make_unique<xfdtd::Cube>(xfdtd::Vector{-0.175, -0.175, -0.175}, xfdtd::Vector{0.35, 0.35, 0.35}),
      xfdtd::Material::createAir())};
  auto sphere{std::make_shared<xfdtd::Object>(
      "dielectric_sphere", std::make_unique<xfdtd::Sphere>(xfdtd::Vector{0, 0, 0}, 0.1),
      std::make_unique<xfdtd::Material>("a", xfdtd::ElectroMagneticProperty{3, 2, 0, 0}))};

  // tau = 20 dl / 6e8 s, the same as 20 dl / (2 c) up to c ~ 3e8
  const xfdtd::Real tau{dl * 20 / 6e8};
  auto tfsf{std::make_shared<xfdtd::TFSF3D>(15, 15, 15, 0.0, 0.0, 0.0, xfdtd::Waveform::gaussian(tau, 4.5 * tau))};

  auto movie{std::make_shared<xfdtd::MovieMonitor>(
      std::make_unique<xfdtd::FieldMonitor>(
          std::make_unique<xfdtd::Cube>(xfdtd::Vector{-0.175, 0, -0.175}, xfdtd::Vector{0.35, dl, 0.35}),
          xfdtd::EMF::Field::EX),  // name and output directory left at their defaults
      10, "movie", out + "/movie")};
  auto probe{std::make_shared<xfdtd::FieldTimeMonitor>(
      "ex_centre", std::make_unique<xfdtd::Cube>(xfdtd::Vector{0, 0, 0}, xfdtd::Vector{dl, dl, dl}), out + "/probe",
      xfdtd::EMF::Field::EX)};
  auto plane{std::make_shared<xfdtd::FieldMonitor>(
      std::make_unique<xfdtd::Cube>(xfdtd::Vector{-0.175, -0.175, 0}, xfdtd::Vector{0.35, 0.35, dl}),
      xfdtd::EMF::Field::EZ, "ez_xy", out + "/plane")};

  auto s{xfdtd::Simulation{dl, dl, dl, 0.9}};
  s.addObject(domain);
  s.addObject(sphere);
  s.addWaveformSource(tfsf);
  for (auto d : {xfdtd::Axis::Direction::XN, xfdtd::Axis::Direction::XP, xfdtd::Axis::Direction::YN,
                 xfdtd::Axis::Direction::YP, xfdtd::Axis::Direction::ZN, xfdtd::Axis::Direction::ZP})
    s.addBoundary(std::make_shared<xfdtd::PML>(8, d));
  s.addMonitor(movie);
  s.addMonitor(probe);
  s.addMonitor(plane);
  s.addDefaultVisitor();
  auto counter = std::make_shared<CountingVisitor>();
  s.addVisitor(counter);
  if (f32) s.setPrecision(xfdtd::Precision::FP32);
  s.run(steps);

  probe->output();
  plane->output();
  movie->output();  // writes nothing: the frames were written by the run

  // the time axis of the run and the reference's constants
  const auto& tp = s.calculationParam()->timeParam();
  const auto& h = tp->hTime();
  const auto& e = tp->eTime();
  std::printf("time: %zu %.17g %.17g %.17g %.17g\n", h.size(), h(0), h(h.size() - 1), e(0), e(e.size() - 1));
  const xfdtd::Real dt_check = xfdtd::TimeParam::calculateDt(0.9, dl, dl, dl);
  std::printf("dt: %.17g %.17g c0: %.17g z0: %.17g\n", tp->dt(), dt_check, xfdtd::constant::C_0,
              xfdtd::constant::Z_0);
  std::printf("flags: sim_start %d sim_end %d update_start %d update_end %d next %d last %zu end %zu monotonic %d "
              "current %zu\n",
              counter->init[0], counter->init[1], counter->init[4], counter->init[5], counter->next, counter->last,
              counter->last_end, counter->monotonic ? 1 : 0, tp->currentTimeStep());
  return 0;
}
