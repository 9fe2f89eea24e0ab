// ref_mesh_shim.cpp - C entry points around the reference's own Mesh class
// (src/Mesh.hpp, src/Mesh.cpp, compiled unmodified from the reference tree) so the
// tests can compare its connectivity with the oracle's and the product's.
// TEST INFRASTRUCTURE ONLY.  Mesh.hpp grants its friend class HO_2D access to the
// private connectivity; this file's HO_2D is only that accessor.
#include "Mesh.hpp"

#include <unistd.h>  // dup/dup2 (fcntl.h's open() clashes with the reference's enum value "open")

#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

class HO_2D {
 public:
  static Mesh* load(const char* path, int occ) {
    Mesh* m = new Mesh();
    m->input_msh_file = path;
    m->OCC_format = occ;
    // the reader echoes the whole file to stdout: silence it for the duration
    std::cout.flush();
    std::fflush(stdout);
    const int saved = dup(1);
    std::FILE* nul = std::fopen("/dev/null", "w");
    if (nul) { dup2(fileno(nul), 1); std::fclose(nul); }
    char rc = m->read_msh_file();
    std::cout.flush();
    std::fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
    if (rc) { delete m; return nullptr; }
    return m;
  }
  static int counts(Mesh* m, int* out) {
    out[0] = m->N_el; out[1] = m->N_edges_boundary; out[2] = m->Lnod; out[3] = (int)m->boundaries.size();
    out[4] = m->N_nodes;
    return 0;
  }
  static int connectivity(Mesh* m, int* nbr, int* nbr_side, int* is_bnd, int* bnd_index, int* be_elem, int* be_side) {
    for (int el = 0; el < m->N_el; ++el)
      for (int s = 0; s < 4; ++s) {
        const auto& en = m->elem_neighbor[el];
        is_bnd[el * 4 + s] = en.is_on_boundary[s];
        nbr[el * 4 + s] = en.is_on_boundary[s] ? -1 : (int)en.neighbor[s];
        nbr_side[el * 4 + s] = en.is_on_boundary[s] ? -1 : en.neighbor_common_side[s];
        bnd_index[el * 4 + s] = en.is_on_boundary[s] ? (int)en.boundary_index[s] : -1;
      }
    for (int e = 0; e < m->N_edges_boundary; ++e) { be_elem[e] = m->boundary_edges[e].element_index; be_side[e] = m->boundary_edges[e].side; }
    return 0;
  }
  static int edge_nodes(Mesh* m, int* out) {  // boundary edges after process_mesh reorientation
    for (int e = 0; e < m->N_edges_boundary; ++e)
      for (int j = 0; j < m->Lnod; ++j) out[e * m->Lnod + j] = m->edges[e].nodes[j];
    return 0;
  }
  static int elements(Mesh* m, int* out) {
    const int L2 = m->Lnod * m->Lnod;
    for (int el = 0; el < m->N_el; ++el)
      for (int j = 0; j < L2; ++j) out[el * L2 + j] = m->elements[el].nodes[j];
    return 0;
  }
  static int nodes(Mesh* m, double* xy) {
    for (int i = 0; i < m->N_nodes; ++i) { xy[2 * i] = m->nodes[i].coor.x; xy[2 * i + 1] = m->nodes[i].coor.y; }
    return 0;
  }
  static const char* bname(Mesh* m, int b) { return m->boundaries[b].name.c_str(); }
  static int bedges(Mesh* m, int b, int* out) {
    for (size_t i = 0; i < m->boundaries[b].edges.size(); ++i) out[i] = m->boundaries[b].edges[i];
    return (int)m->boundaries[b].edges.size();
  }
};

extern "C" {
void* ref_mesh_load(const char* path, int occ) { return HO_2D::load(path, occ); }
void ref_mesh_free(void* m) { delete static_cast<Mesh*>(m); }
int ref_mesh_counts(void* m, int* out) { return HO_2D::counts(static_cast<Mesh*>(m), out); }
int ref_mesh_connectivity(void* m, int* a, int* b, int* c, int* d, int* e, int* f) {
  return HO_2D::connectivity(static_cast<Mesh*>(m), a, b, c, d, e, f);
}
int ref_mesh_edge_nodes(void* m, int* out) { return HO_2D::edge_nodes(static_cast<Mesh*>(m), out); }
int ref_mesh_elements(void* m, int* out) { return HO_2D::elements(static_cast<Mesh*>(m), out); }
int ref_mesh_nodes(void* m, double* xy) { return HO_2D::nodes(static_cast<Mesh*>(m), xy); }
const char* ref_mesh_bname(void* m, int b) { return HO_2D::bname(static_cast<Mesh*>(m), b); }
int ref_mesh_bedges(void* m, int b, int* out) { return HO_2D::bedges(static_cast<Mesh*>(m), b, out); }
}
