// Test driver: reads a Gmsh 4.1 file with the library's host reader (hoc::read_gmsh41)
// and prints the arrays it produced, one section per line group, for comparison in Python.
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "hoc_host.hpp"

int main(int argc, char** argv) {
  if (argc < 3) { std::fprintf(stderr, "usage: gmsh_dump file.msh occ\n"); return 2; }
  hoc::Mesh m;
  try {
    hoc::read_gmsh41(argv[1], std::atoi(argv[2]), m);
  } catch (const std::exception& e) {
    std::printf("ERROR %s\n", e.what());
    return 1;
  }
  std::printf("lnod %d\nnodes %d\n", m.lnod, m.n_nodes);
  for (int i = 0; i < m.n_nodes; ++i) std::printf("%.17g %.17g\n", m.xy[2 * i], m.xy[2 * i + 1]);
  std::printf("elements %d\n", m.n_elem);
  const int L2 = m.lnod * m.lnod;
  for (int e = 0; e < m.n_elem; ++e) {
    for (int q = 0; q < L2; ++q) std::printf("%d ", m.elem[(size_t)e * L2 + q]);
    std::printf("\n");
  }
  std::printf("bedges %d\n", m.n_bedges);
  for (int e = 0; e < m.n_bedges; ++e) {
    std::printf("%d", m.bedge_bdry[e]);
    for (int v : m.edges[e]) std::printf(" %d", v);
    std::printf("\n");
  }
  std::printf("bnames %d\n", (int)m.bnames.size());
  for (auto& s : m.bnames) std::printf("%s\n", s.c_str());
  return 0;
}
