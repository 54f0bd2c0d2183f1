// refinery.cpp -- the reference's conformal mesh refinement and DOF numbering for quadrilateral
// and hexahedral meshes, so that the host setup hands the solver the index arrays a FEAT
// application would build (sparsity, transfers, Dirichlet lists in FEAT's numbering).
//
// Restated from the reference (kernel/geometry):
//  * StandardRefinery: a fine mesh's vertices are the coarse vertices, then one per coarse edge
//    (midpoint), one per coarse quad (centre) and one per coarse hexa (centre), in entity order
//    (intern/standard_vertex_refiner.hpp:65-250, the wrapper recursion vertex -> edge -> quad ->
//    hexa); fine edges are two per coarse edge, then four per coarse quad, then six per coarse
//    hexa; fine quads four per coarse quad, then twelve per coarse hexa; fine hexas eight per
//    coarse hexa (intern/standard_index_refiner.hpp:1341-2700, offsets of
//    intern/entity_counter.hpp).  Each fine entity's local vertex tuple is fixed by the coarse
//    entity's reference cell; here it is written as positions on the coarse cell's reference
//    lattice {0,1,2}^d (0, 2: the cell's corners, 1: midpoints), which name the coarse vertex,
//    edge, face or cell whose refinement vertex sits there.
//  * The face index sets of the fine mesh (edges-at-quad, edges-at-hexa, quads-at-hexa) list,
//    for every local face of FaceIndexMapping (intern/face_index_mapping.hpp), the fine entity
//    with those vertices; the reference computes the same indices from congruency codes
//    (intern/sub_index_mapping.hpp); in a conformal mesh an entity is unique for its vertex
//    set, so looking it up by its vertices gives them directly.
//  * UnitCubeFactory (geometry/common_factories.hpp:70-238): vertex i at ((i>>0)&1, (i>>1)&1,
//    (i>>2)&1), index sets from FaceIndexMapping.
//  * Lagrange-1 DOFs are the vertices; Lagrange-2 DOFs are vertices, then edges, then cells
//    (space/dof_mapping_common.hpp:193-280, DofMappingUniform over the entity dimensions), which
//    is the vertex numbering of the next refinement.
//
// Tested against the refined meshes the reference's own tests hold
// (geometry/test_aux/standard_{quad,hexa}.cpp, tetris_{quad,hexa}.cpp).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "feat3_hip.h"

namespace
{
  typedef std::uint32_t u32;
  typedef std::uint64_t u64;

  // FaceIndexMapping of the reference cells (intern/face_index_mapping.hpp)
  const int kQuadEdge[4][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}};
  const int kHexaEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  const int kHexaQuad[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 2, 4, 6}, {1, 3, 5, 7}};

  struct Mesh
  {
    int dim = 2;
    u64 nv = 0, ne = 0, nq = 0, nh = 0;
    std::vector<double> vtx;               // nv * dim
    std::vector<u32> v_e, v_q, e_q;        // 2 ne, 4 nq, 4 nq
    std::vector<u32> v_c, e_c, q_c;        // 8 nh, 12 nh, 6 nh (hexa meshes)
  };

  // fine vertex at reference-lattice position p of a coarse quad (p in {0,1,2}^2): corner,
  // midpoint of a local edge, or the centre
  inline u32 quad_point(const Mesh& c, u64 q, int px, int py, u64 ioe, u64 ioq)
  {
    if(!(px & 1) && !(py & 1)) return c.v_q[4 * q + (px >> 1) + 2 * (py >> 1)];
    if((px & 1) && (py & 1)) return u32(ioq + q);
    // local edges: 0 = v0v1 (y = 0), 1 = v2v3 (y = 2), 2 = v0v2 (x = 0), 3 = v1v3 (x = 2)
    const int le = (px & 1) ? (py >> 1) : 2 + (px >> 1);
    return u32(ioe + c.e_q[4 * q + le]);
  }

  // fine vertex at reference-lattice position p of a coarse hexa (p in {0,1,2}^3)
  inline u32 hexa_point(const Mesh& c, u64 h, int px, int py, int pz, u64 ioe, u64 ioq, u64 ioc)
  {
    const int odd = (px & 1) + (py & 1) + (pz & 1);
    if(odd == 0) return c.v_c[8 * h + (px >> 1) + 2 * (py >> 1) + 4 * (pz >> 1)];
    if(odd == 3) return u32(ioc + h);
    if(odd == 1)
    {
      // local edges 0-3 along x, 4-7 along y, 8-11 along z, ordered by the other two coordinates
      int le;
      if(px & 1) le = (py >> 1) + 2 * (pz >> 1);
      else if(py & 1) le = 4 + (px >> 1) + 2 * (pz >> 1);
      else le = 8 + (px >> 1) + 2 * (py >> 1);
      return u32(ioe + c.e_c[12 * h + le]);
    }
    // local faces: 0/1 z = 0/2, 2/3 y = 0/2, 4/5 x = 0/2
    int lf;
    if(!(pz & 1)) lf = pz >> 1;
    else if(!(py & 1)) lf = 2 + (py >> 1);
    else lf = 4 + (px >> 1);
    return u32(ioq + c.q_c[6 * h + lf]);
  }

  // incidence lists (entities at each vertex, by their smallest vertex) for lookups by vertices
  struct Incidence
  {
    std::vector<u32> ptr, ent;
    void build(u64 nv, u64 n, int k, const std::vector<u32>& tup)
    {
      ptr.assign(nv + 1, 0);
      for(u64 i = 0; i < n; ++i) ++ptr[*std::min_element(&tup[k * i], &tup[k * i] + k) + 1];
      for(u64 v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
      ent.assign(n, 0);
      std::vector<u32> fill(ptr.begin(), ptr.end() - 1);
      for(u64 i = 0; i < n; ++i) ent[fill[*std::min_element(&tup[k * i], &tup[k * i] + k)]++] = u32(i);
    }
  };

  // index of the entity (k vertices, tuples tup) whose vertex set is {vs[0..k)}; -1 if none
  long long find_entity(const Incidence& inc, int k, const std::vector<u32>& tup, const u32* vs)
  {
    u32 key[4];
    std::copy(vs, vs + k, key);
    std::sort(key, key + k);
    for(u32 a = inc.ptr[key[0]]; a < inc.ptr[key[0] + 1]; ++a)
    {
      const u32 e = inc.ent[a];
      u32 t[4];
      std::copy(&tup[u64(k) * e], &tup[u64(k) * e] + k, t);
      std::sort(t, t + k);
      if(std::equal(t, t + k, key)) return e;
    }
    return -1;
  }

  // one refinement step; with index_sets == false only the vertex coordinates are produced
  // (the last step of a numbering, whose edges and cells are not needed)
  bool refine(const Mesh& c, Mesh& f, bool index_sets)
  {
    const int d = c.dim;
    f.dim = d;
    const u64 iov = 0, ioe = c.nv, ioq = c.nv + c.ne, ioc = c.nv + c.ne + c.nq;
    f.nv = c.nv + c.ne + c.nq + c.nh;
    f.ne = 2 * c.ne + 4 * c.nq + 6 * c.nh;
    f.nq = 4 * c.nq + 12 * c.nh;
    f.nh = 8 * c.nh;
    // vertices: zero, then add scale * v_k over the entity's vertices in order
    // (standard_vertex_refiner.hpp:65-110)
    f.vtx.assign(f.nv * d, 0.0);
    std::copy(c.vtx.begin(), c.vtx.end(), f.vtx.begin());
    auto centre = [&](u64 at, const u32* vs, int k) {
      const double s = 1.0 / double(k);
      for(int j = 0; j < k; ++j)
        for(int x = 0; x < d; ++x) f.vtx[at * d + x] += s * c.vtx[u64(vs[j]) * d + x];
    };
    for(u64 e = 0; e < c.ne; ++e) centre(ioe + e, &c.v_e[2 * e], 2);
    for(u64 q = 0; q < c.nq; ++q) centre(ioq + q, &c.v_q[4 * q], 4);
    for(u64 h = 0; h < c.nh; ++h) centre(ioc + h, &c.v_c[8 * h], 8);
    if(!index_sets) return true;

    // edges
    f.v_e.assign(2 * f.ne, 0);
    u64 k = 0;
    for(u64 e = 0; e < c.ne; ++e)
    {
      f.v_e[2 * k] = u32(iov + c.v_e[2 * e]); f.v_e[2 * k + 1] = u32(ioe + e); ++k;
      f.v_e[2 * k] = u32(ioe + e); f.v_e[2 * k + 1] = u32(iov + c.v_e[2 * e + 1]); ++k;
    }
    // quad-interior edges: along the second axis (lower, upper half), then the first
    static const int qe[4][2][2] = {{{1, 0}, {1, 1}}, {{1, 1}, {1, 2}}, {{0, 1}, {1, 1}}, {{1, 1}, {2, 1}}};
    for(u64 q = 0; q < c.nq; ++q)
      for(int j = 0; j < 4; ++j, ++k)
        for(int a = 0; a < 2; ++a) f.v_e[2 * k + a] = quad_point(c, q, qe[j][a][0], qe[j][a][1], ioe, ioq);
    // hexa-interior edges: along z, y, x (lower, upper half each)
    static const int he[6][2][3] = {{{1, 1, 0}, {1, 1, 1}}, {{1, 1, 1}, {1, 1, 2}}, {{1, 0, 1}, {1, 1, 1}},
                                    {{1, 1, 1}, {1, 2, 1}}, {{0, 1, 1}, {1, 1, 1}}, {{1, 1, 1}, {2, 1, 1}}};
    for(u64 h = 0; h < c.nh; ++h)
      for(int j = 0; j < 6; ++j, ++k)
        for(int a = 0; a < 2; ++a)
          f.v_e[2 * k + a] = hexa_point(c, h, he[j][a][0], he[j][a][1], he[j][a][2], ioe, ioq, ioc);

    // quads: children of coarse quads (child bits in the quad's axes), then the twelve
    // hexa-interior quads (planes x = 1, y = 1, z = 1; children in the plane's two axes)
    f.v_q.assign(4 * f.nq, 0);
    k = 0;
    for(u64 q = 0; q < c.nq; ++q)
      for(int ch = 0; ch < 4; ++ch, ++k)
        for(int v = 0; v < 4; ++v)
          f.v_q[4 * k + v] = quad_point(c, q, (ch & 1) + (v & 1), (ch >> 1) + (v >> 1), ioe, ioq);
    for(u64 h = 0; h < c.nh; ++h)
      for(int plane = 0; plane < 3; ++plane)
        for(int ch = 0; ch < 4; ++ch, ++k)
          for(int v = 0; v < 4; ++v)
          {
            const int a = (ch & 1) + (v & 1), b = (ch >> 1) + (v >> 1);
            int p[3];
            if(plane == 0) { p[0] = 1; p[1] = a; p[2] = b; }
            else if(plane == 1) { p[0] = a; p[1] = 1; p[2] = b; }
            else { p[0] = a; p[1] = b; p[2] = 1; }
            f.v_q[4 * k + v] = hexa_point(c, h, p[0], p[1], p[2], ioe, ioq, ioc);
          }
    // hexas: eight children (child bits x, y, z)
    f.v_c.assign(8 * f.nh, 0);
    for(u64 h = 0; h < c.nh; ++h)
      for(int ch = 0; ch < 8; ++ch)
        for(int v = 0; v < 8; ++v)
          f.v_c[8 * (8 * h + ch) + v] = hexa_point(c, h, (ch & 1) + (v & 1), ((ch >> 1) & 1) + ((v >> 1) & 1),
                                                   (ch >> 2) + (v >> 2), ioe, ioq, ioc);

    // face index sets of the fine mesh by vertex lookup
    Incidence ie;
    ie.build(f.nv, f.ne, 2, f.v_e);
    f.e_q.assign(4 * f.nq, 0);
    for(u64 q = 0; q < f.nq; ++q)
      for(int j = 0; j < 4; ++j)
      {
        const u32 vs[2] = {f.v_q[4 * q + kQuadEdge[j][0]], f.v_q[4 * q + kQuadEdge[j][1]]};
        const long long e = find_entity(ie, 2, f.v_e, vs);
        if(e < 0) return false;
        f.e_q[4 * q + j] = u32(e);
      }
    if(f.nh)
    {
      Incidence iq;
      iq.build(f.nv, f.nq, 4, f.v_q);
      f.e_c.assign(12 * f.nh, 0);
      f.q_c.assign(6 * f.nh, 0);
      for(u64 h = 0; h < f.nh; ++h)
      {
        for(int j = 0; j < 12; ++j)
        {
          const u32 vs[2] = {f.v_c[8 * h + kHexaEdge[j][0]], f.v_c[8 * h + kHexaEdge[j][1]]};
          const long long e = find_entity(ie, 2, f.v_e, vs);
          if(e < 0) return false;
          f.e_c[12 * h + j] = u32(e);
        }
        for(int j = 0; j < 6; ++j)
        {
          u32 vs[4];
          for(int v = 0; v < 4; ++v) vs[v] = f.v_c[8 * h + kHexaQuad[j][v]];
          const long long q = find_entity(iq, 4, f.v_q, vs);
          if(q < 0) return false;
          f.q_c[6 * h + j] = u32(q);
        }
      }
    }
    return true;
  }

  // UnitCubeFactory: the unit square (one quad) or unit cube (one hexa)
  Mesh unit_cube(int dim)
  {
    Mesh m;
    m.dim = dim;
    m.nv = u64(1) << dim;
    m.vtx.assign(m.nv * dim, 0.0);
    for(u64 i = 0; i < m.nv; ++i)
      for(int x = 0; x < dim; ++x) m.vtx[i * dim + x] = double((i >> x) & 1);
    if(dim == 2)
    {
      m.ne = 4; m.nq = 1;
      for(int e = 0; e < 4; ++e) { m.v_e.push_back(u32(kQuadEdge[e][0])); m.v_e.push_back(u32(kQuadEdge[e][1])); }
      m.v_q = {0, 1, 2, 3};
      m.e_q = {0, 1, 2, 3};
    }
    else
    {
      m.ne = 12; m.nq = 6; m.nh = 1;
      for(int e = 0; e < 12; ++e) { m.v_e.push_back(u32(kHexaEdge[e][0])); m.v_e.push_back(u32(kHexaEdge[e][1])); }
      for(int q = 0; q < 6; ++q)
        for(int v = 0; v < 4; ++v) m.v_q.push_back(u32(kHexaQuad[q][v]));
      // edges-at-quad of the reference hexa's faces (FaceIndexMapping<Hypercube<3>, 2, 1>)
      static const u32 eq[6][4] = {{0, 1, 4, 5}, {2, 3, 6, 7}, {0, 2, 8, 9}, {1, 3, 10, 11}, {4, 6, 8, 10}, {5, 7, 9, 11}};
      for(int q = 0; q < 6; ++q)
        for(int v = 0; v < 4; ++v) m.e_q.push_back(eq[q][v]);
      for(u32 v = 0; v < 8; ++v) m.v_c.push_back(v);
      for(u32 e = 0; e < 12; ++e) m.e_c.push_back(e);
      for(u32 q = 0; q < 6; ++q) m.q_c.push_back(q);
    }
    return m;
  }
} // namespace

extern "C" {

int feat3_mesh_refine_sizes(int dim, const uint64_t counts[4], uint64_t out_counts[4])
{
  if((dim != 2 && dim != 3) || !counts || !out_counts) return FEAT3_HIP_ERR_ARG;
  const u64 nh = dim == 3 ? counts[3] : 0;
  out_counts[0] = counts[0] + counts[1] + counts[2] + nh;
  out_counts[1] = 2 * counts[1] + 4 * counts[2] + 6 * nh;
  out_counts[2] = 4 * counts[2] + 12 * nh;
  out_counts[3] = 8 * nh;
  return FEAT3_HIP_OK;
}

int feat3_mesh_refine(int dim, const uint64_t counts[4], const double* vtx, const uint32_t* v_e, const uint32_t* v_q,
                      const uint32_t* e_q, const uint32_t* v_c, const uint32_t* e_c, const uint32_t* q_c, double* f_vtx,
                      uint32_t* f_v_e, uint32_t* f_v_q, uint32_t* f_e_q, uint32_t* f_v_c, uint32_t* f_e_c,
                      uint32_t* f_q_c)
{
  if((dim != 2 && dim != 3) || !counts || !vtx || !v_e || !v_q || !e_q) return FEAT3_HIP_ERR_ARG;
  Mesh c;
  c.dim = dim;
  c.nv = counts[0]; c.ne = counts[1]; c.nq = counts[2]; c.nh = dim == 3 ? counts[3] : 0;
  if(c.nh && (!v_c || !e_c || !q_c)) return FEAT3_HIP_ERR_ARG;
  if(c.nv + c.ne + c.nq + c.nh >= (u64(1) << 32)) return FEAT3_HIP_ERR_ARG;
  c.vtx.assign(vtx, vtx + c.nv * dim);
  c.v_e.assign(v_e, v_e + 2 * c.ne);
  c.v_q.assign(v_q, v_q + 4 * c.nq);
  c.e_q.assign(e_q, e_q + 4 * c.nq);
  for(u32 v : c.v_e) if(v >= c.nv) return FEAT3_HIP_ERR_ARG;
  for(u32 v : c.v_q) if(v >= c.nv) return FEAT3_HIP_ERR_ARG;
  for(u32 e : c.e_q) if(e >= c.ne) return FEAT3_HIP_ERR_ARG;
  if(c.nh)
  {
    c.v_c.assign(v_c, v_c + 8 * c.nh);
    c.e_c.assign(e_c, e_c + 12 * c.nh);
    c.q_c.assign(q_c, q_c + 6 * c.nh);
    for(u32 v : c.v_c) if(v >= c.nv) return FEAT3_HIP_ERR_ARG;
    for(u32 e : c.e_c) if(e >= c.ne) return FEAT3_HIP_ERR_ARG;
    for(u32 q : c.q_c) if(q >= c.nq) return FEAT3_HIP_ERR_ARG;
  }
  Mesh f;
  if(!refine(c, f, true)) return FEAT3_HIP_ERR_ARG;   // not a conformal mesh
  std::copy(f.vtx.begin(), f.vtx.end(), f_vtx);
  std::copy(f.v_e.begin(), f.v_e.end(), f_v_e);
  std::copy(f.v_q.begin(), f.v_q.end(), f_v_q);
  std::copy(f.e_q.begin(), f.e_q.end(), f_e_q);
  if(f.nh)
  {
    if(!f_v_c || !f_e_c || !f_q_c) return FEAT3_HIP_ERR_ARG;
    std::copy(f.v_c.begin(), f.v_c.end(), f_v_c);
    std::copy(f.e_c.begin(), f.e_c.end(), f_e_c);
    std::copy(f.q_c.begin(), f.q_c.end(), f_q_c);
  }
  return FEAT3_HIP_OK;
}

// FEAT's DOF numbering of the Lagrange-1/-2 space on the unit square / cube refined `level` times
// from UnitCubeFactory: lex_of_feat[i] = lexicographic index (x fastest) on the tensor lattice
// of DOF i (the numbering of the host setup feat3_setup_laplace).
int feat3_setup_feat_order(int dim, int degree, int level, uint32_t* lex_of_feat)
{
  if(dim < 2 || dim > 3 || degree < 1 || degree > 2 || level < 0 || level > 12 || !lex_of_feat) return FEAT3_HIP_ERR_ARG;
  // Lagrange-2 DOFs on level L are the vertices of level L + 1
  const int steps = level + (degree == 2 ? 1 : 0);
  u64 m = (u64(1) << steps) + 1, n = 1;
  for(int x = 0; x < dim; ++x) n *= m;
  if(n >= (u64(1) << 32)) return FEAT3_HIP_ERR_ARG;
  Mesh a = unit_cube(dim), b;
  for(int s = 0; s < steps; ++s)
  {
    if(!refine(a, b, s + 1 < steps)) return FEAT3_HIP_ERR_ARG;
    std::swap(a, b);
  }
  if(a.nv != n) return FEAT3_HIP_ERR_ARG;
  // coordinates are dyadic (exact): lattice index = coordinate * (m - 1)
  const double scale = double(m - 1);
  std::vector<char> seen(n, 0);
  for(u64 i = 0; i < n; ++i)
  {
    u64 lex = 0;
    for(int x = dim - 1; x >= 0; --x)
    {
      const double t = a.vtx[i * dim + x] * scale;
      const u64 c = u64(t + 0.5);
      if(double(c) != t || c >= m) return FEAT3_HIP_ERR_ARG;
      lex = lex * m + c;
    }
    if(seen[lex]) return FEAT3_HIP_ERR_ARG;
    seen[lex] = 1;
    lex_of_feat[i] = u32(lex);
  }
  return FEAT3_HIP_OK;
}

// Position of every node in the lexicographic order of its coordinates (last coordinate slowest):
// order[i] = rank of node i.  Nodes of a tensor-product lattice then get their lexicographic
// lattice index.  Coordinates must be distinct.
int feat3_setup_lattice_order(int dim, uint64_t n, const double* coords, uint32_t* order)
{
  if(dim < 1 || dim > 3 || (n && (!coords || !order)) || n >= (u64(1) << 32)) return FEAT3_HIP_ERR_ARG;
  std::vector<u32> idx(n);
  for(u64 i = 0; i < n; ++i) idx[i] = u32(i);
  auto less = [&](u32 a, u32 b) {
    for(int x = dim - 1; x >= 0; --x)
    {
      const double ca = coords[u64(a) * dim + x], cb = coords[u64(b) * dim + x];
      if(ca < cb) return true;
      if(cb < ca) return false;
    }
    return false;
  };
  std::sort(idx.begin(), idx.end(), less);
  for(u64 r = 0; r + 1 < n; ++r)
    if(!less(idx[r], idx[r + 1])) return FEAT3_HIP_ERR_ARG;   // repeated or unordered (NaN) coordinates
  for(u64 r = 0; r < n; ++r) order[idx[r]] = u32(r);
  return FEAT3_HIP_OK;
}

// Symmetric or two-sided permutation of a CSR matrix: old row i becomes row row_order[i], old
// column j becomes column col_order[j] (NULL: unchanged), columns ascending in every new row.
// The values move with their entries (no arithmetic).
int feat3_setup_permute_csr(uint64_t rows, uint64_t cols, const uint32_t* rp, const uint32_t* col, const double* val,
                            const uint32_t* row_order, const uint32_t* col_order, uint32_t* out_rp, uint32_t* out_col,
                            double* out_val)
{
  if(!rp || !out_rp || rows >= (u64(1) << 32) || cols >= (u64(1) << 32)) return FEAT3_HIP_ERR_ARG;
  const u64 nnz = rp[rows];
  if(nnz && (!col || !val || !out_col || !out_val)) return FEAT3_HIP_ERR_ARG;
  std::vector<u32> old_of_new(rows);
  std::vector<char> hit(rows, 0);
  for(u64 i = 0; i < rows; ++i)
  {
    const u32 r = row_order ? row_order[i] : u32(i);
    if(r >= rows || hit[r]) return FEAT3_HIP_ERR_ARG;
    hit[r] = 1;
    old_of_new[r] = u32(i);
  }
  if(col_order)
  {
    std::vector<char> chit(cols, 0);
    for(u64 j = 0; j < cols; ++j)
    {
      if(col_order[j] >= cols || chit[col_order[j]]) return FEAT3_HIP_ERR_ARG;
      chit[col_order[j]] = 1;
    }
  }
  out_rp[0] = 0;
  for(u64 r = 0; r < rows; ++r) out_rp[r + 1] = out_rp[r] + (rp[old_of_new[r] + 1] - rp[old_of_new[r]]);
  std::vector<std::pair<u32, double>> row;
  for(u64 r = 0; r < rows; ++r)
  {
    const u32 o = old_of_new[r];
    row.clear();
    for(u32 k = rp[o]; k < rp[o + 1]; ++k)
    {
      if(col[k] >= cols) return FEAT3_HIP_ERR_ARG;
      row.emplace_back(col_order ? col_order[col[k]] : col[k], val[k]);
    }
    std::stable_sort(row.begin(), row.end(), [](const std::pair<u32, double>& a, const std::pair<u32, double>& b) {
      return a.first < b.first;
    });
    u64 k = out_rp[r];
    for(const auto& e : row) { out_col[k] = e.first; out_val[k] = e.second; ++k; }
  }
  return FEAT3_HIP_OK;
}

} // extern "C"
