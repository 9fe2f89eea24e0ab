// cavity_facade.cpp - drives the GPU path through the HO_2D-shaped C++ facade the way the
// reference's hybrid driver drives HO_2D: arrays in, process_mesh_input, solveto_d,
// getallvorts_d.  Usage: cavity_facade n K nsteps dt [out.bin [sample_points.dat]]
// out.bin holds, as doubles: vorticity (D), stream function (D), velocity (2D), solution points
// (2D), solution areas (D), hoquad weights (K*K).  With a sample-point file, trigger_write(0) also
// writes the reference's output files into the working directory.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../csrc/HO_2D_GPU.hpp"

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 8;
  const int K = argc > 2 ? std::atoi(argv[2]) : 4;
  const int nsteps = argc > 3 ? std::atoi(argv[3]) : 5;
  const double dt = argc > 4 ? std::atof(argv[4]) : 1e-4;
  const char* out = argc > 5 ? argv[5] : nullptr;
  const char* samples = argc > 6 ? argv[6] : nullptr;
  // unit-square n x n bilinear quads: nodes row-major, elements CCW from the SW corner
  std::vector<double> xy;
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) { xy.push_back((double)i / n); xy.push_back((double)j / n); }
  auto id = [&](int i, int j) { return j * (n + 1) + i; };
  std::vector<int32_t> el;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) { el.push_back(id(i, j)); el.push_back(id(i + 1, j)); el.push_back(id(i + 1, j + 1)); el.push_back(id(i, j + 1)); }
  std::vector<int32_t> wall, top;
  for (int i = 0; i < n; ++i) { wall.push_back(id(i, 0)); wall.push_back(id(i + 1, 0)); }
  for (int j = 0; j < n; ++j) { wall.push_back(id(n, j)); wall.push_back(id(n, j + 1)); }
  for (int i = n; i > 0; --i) { top.push_back(id(i, n)); top.push_back(id(i - 1, n)); }
  for (int j = n; j > 0; --j) { wall.push_back(id(0, j)); wall.push_back(id(0, j - 1)); }
  try {
    HO_2D_GPU ho;
    ho.set_defaults();
    ho.set_elemorder(K);
    ho.set_problem(1, 100.0, dt, 4, 1);
    ho.set_linear_solver(1e-14, 500);
    ho.load_mesh_arrays_d(1, (int32_t)xy.size(), xy.data(), (int32_t)el.size(), el.data(), (int32_t)wall.size(),
                          wall.data(), 0, nullptr);
    ho.add_boundary("top", (int32_t)top.size(), top.data());
    ho.process_mesh_input();
    if (samples) ho.read_process_sample_points(samples);
    ho.solveto_d(nsteps * dt, nsteps, 4, 100.0);
    const size_t D = (size_t)n * n * K * K;
    std::vector<double> w(D), psi(D), uv(2 * D), xyp(2 * D), area(D), hq((size_t)K * K);
    ho.getallvorts_d((int32_t)D, w.data());
    ho.getstreamfunction_d((int32_t)D, psi.data());
    ho.getvelocity_d((int32_t)(2 * D), uv.data());
    ho.getsolnpts_d(ho.getsolnptlen(), xyp.data());
    ho.getsolnareas_d((int32_t)D, area.data());
    ho.get_hoquad_weights_d(K * K, hq.data());
    double s = 0.;
    for (double v : w) s += v * v;
    std::printf("time %.6g |omega| %.12e\n", ho.current_time(), std::sqrt(s));
    if (samples) ho.trigger_write(0);
    if (out) {
      std::FILE* f = std::fopen(out, "wb");
      if (!f) { std::printf("cannot open %s\n", out); return 2; }
      for (const auto* v : {&w, &psi, &uv, &xyp, &area, &hq}) std::fwrite(v->data(), sizeof(double), v->size(), f);
      std::fclose(f);
    }
  } catch (const hoc_gpu_error& e) {
    std::printf("hoc_gpu_error %d: %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
