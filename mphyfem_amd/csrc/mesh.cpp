// mesh.cpp -- seeded synthetic meshes (Hex8 blocks, Kuhn-split Tet10 cubes) and face helpers.
//
// The reference reads Gmsh 2.2 meshes (femBase.cpp:121-364); the benchmark workloads here are
// generated instead.  Node orders follow the reference's element bases: Hex8 per
// LagrangeBasisFunsHexa (BasisFunctionsLagrange.cpp:1202-1279), Tet10 per LagrangeBasisFunsTetra
// (BasisFunctionsLagrange.cpp:870-956, edges 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(2,3) 9:(1,3)).
// Random numbers: SplitMix64, u = (x >> 11) * 2^-53.

#include <cmath>
#include <cstdint>

#include "../../include/mpf.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline int64_t hex_node(int i, int j, int k, int nx, int ny) { return i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k); }

}  // namespace

extern "C" {

void mpf_mesh_hex_sizes(int32_t nx, int32_t ny, int32_t nz, int32_t* nNode, int32_t* nElem) {
  *nNode = (nx + 1) * (ny + 1) * (nz + 1);
  *nElem = nx * ny * nz;
}

void mpf_mesh_hex(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz, double jitter, uint64_t seed,
                  double* X, int32_t* conn) {
  const double h[3] = {Lx / nx, Ly / ny, Lz / nz};
  const int nn[3] = {nx, ny, nz};
  SplitMix64 rng(seed);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const int64_t n = hex_node(i, j, k, nx, ny);
        const int idx[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          const double r = rng.uniform();
          double x = idx[d] * h[d];
          if (jitter > 0.0 && idx[d] > 0 && idx[d] < nn[d]) x += (2.0 * r - 1.0) * jitter * h[d];
          X[3 * n + d] = x;
        }
      }
  int64_t e = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++e) {
        int32_t* c = conn + 8 * e;
        c[0] = (int32_t)hex_node(i, j, k, nx, ny);
        c[1] = (int32_t)hex_node(i + 1, j, k, nx, ny);
        c[2] = (int32_t)hex_node(i + 1, j + 1, k, nx, ny);
        c[3] = (int32_t)hex_node(i, j + 1, k, nx, ny);
        c[4] = (int32_t)hex_node(i, j, k + 1, nx, ny);
        c[5] = (int32_t)hex_node(i + 1, j, k + 1, nx, ny);
        c[6] = (int32_t)hex_node(i + 1, j + 1, k + 1, nx, ny);
        c[7] = (int32_t)hex_node(i, j + 1, k + 1, nx, ny);
      }
}

void mpf_mesh_tet10_sizes(int32_t n, int32_t* nNode, int32_t* nElem) {
  const int64_t m = 2 * (int64_t)n + 1;
  *nNode = (int32_t)(m * m * m);
  *nElem = 6 * n * n * n;
}

// Kuhn split: the six tets of a cube share the diagonal (0,0,0)-(1,1,1); tet p walks from
// (0,0,0) along the axes in the order of permutation p.  Vertex nodes live on even indices of the
// (2n+1)^3 node grid, mid-edge nodes at the midpoints (odd indices).
void mpf_mesh_tet10(int32_t n, double L, double jitter, uint64_t seed, double* X, int32_t* conn) {
  const int64_t m = 2 * (int64_t)n + 1;
  const double h = L / n;
  auto id = [m](int64_t I, int64_t J, int64_t K) { return I + m * (J + m * K); };
  SplitMix64 rng(seed);
  // vertices (even grid points) in node order, three draws each
  for (int64_t K = 0; K < m; K += 2)
    for (int64_t J = 0; J < m; J += 2)
      for (int64_t I = 0; I < m; I += 2) {
        const int64_t v = id(I, J, K);
        const int64_t idx[3] = {I / 2, J / 2, K / 2};
        for (int d = 0; d < 3; ++d) {
          const double r = rng.uniform();
          double x = idx[d] * h;
          if (jitter > 0.0 && idx[d] > 0 && idx[d] < n) x += (2.0 * r - 1.0) * jitter * h;
          X[3 * v + d] = x;
        }
      }
  // mid-edge nodes at the midpoint of their two vertices: a node with odd indices on the
  // doubled grid lies between (I-oI, J-oJ, K-oK) and (I+oI, J+oJ, K+oK) for its odd axes.
  for (int64_t K = 0; K < m; ++K)
    for (int64_t J = 0; J < m; ++J)
      for (int64_t I = 0; I < m; ++I) {
        const int oi = I & 1, oj = J & 1, ok = K & 1;
        if (!(oi | oj | ok)) continue;
        const int64_t a = id(I - oi, J - oj, K - ok), b = id(I + oi, J + oj, K + ok);
        const int64_t v = id(I, J, K);
        for (int d = 0; d < 3; ++d) X[3 * v + d] = 0.5 * (X[3 * a + d] + X[3 * b + d]);
      }
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int64_t e = 0;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int p = 0; p < 6; ++p, ++e) {
          int64_t P[4][3];
          P[0][0] = 2 * i; P[0][1] = 2 * j; P[0][2] = 2 * k;
          for (int s = 1; s < 4; ++s) {
            for (int d = 0; d < 3; ++d) P[s][d] = P[s - 1][d];
            P[s][perms[p][s - 1]] += 2;
          }
          // odd permutations have negative orientation: swap vertices 1 and 2
          const bool odd = (p == 1 || p == 2 || p == 5);
          if (odd)
            for (int d = 0; d < 3; ++d) { const int64_t t = P[1][d]; P[1][d] = P[2][d]; P[2][d] = t; }
          int32_t* c = conn + 10 * e;
          for (int s = 0; s < 4; ++s) c[s] = (int32_t)id(P[s][0], P[s][1], P[s][2]);
          static const int edges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
          for (int q = 0; q < 6; ++q) {
            const int a = edges[q][0], b = edges[q][1];
            c[4 + q] = (int32_t)id((P[a][0] + P[b][0]) / 2, (P[a][1] + P[b][1]) / 2, (P[a][2] + P[b][2]) / 2);
          }
        }
}

// Nodes of a face of a structured block.  elem_type MPF_HEX8: (nx+1)(ny+1)(nz+1) grid;
// MPF_TET10P1, MPF_TET10: (2n+1)^3 grid with nx = ny = nz = n.
int32_t mpf_mesh_face_nodes(int32_t elem_type, int32_t nx, int32_t ny, int32_t nz, int32_t face, int32_t* nodes) {
  int64_t mx = nx + 1, my = ny + 1, mz = nz + 1;
  if (elem_type == MPF_TET10P1 || elem_type == MPF_TET10) { mx = 2 * (int64_t)nx + 1; my = 2 * (int64_t)ny + 1; mz = 2 * (int64_t)nz + 1; }
  const int axis = face / 2;
  const bool hi = face % 2;
  int32_t cnt = 0;
  for (int64_t k = 0; k < mz; ++k)
    for (int64_t j = 0; j < my; ++j)
      for (int64_t i = 0; i < mx; ++i) {
        const int64_t idx[3] = {i, j, k};
        const int64_t lim[3] = {mx - 1, my - 1, mz - 1};
        if (idx[axis] != (hi ? lim[axis] : 0)) continue;
        if (nodes) nodes[cnt] = (int32_t)(i + mx * (j + my * k));
        ++cnt;
      }
  return cnt;
}

// Quad4 faces of a Hex8 block face with the reference's normal n = (dx/deta x dx/dxi) pointing
// outward (femSolidmechanics1.cpp:1026-1062 with LagrangeBasisFunsFace).
int32_t mpf_mesh_hex_face_quads(int32_t nx, int32_t ny, int32_t nz, int32_t face, int32_t* quads) {
  const int axis = face / 2;
  const bool hi = face % 2;
  const int u = (axis + 1) % 3, v = (axis + 2) % 3;
  const int nn[3] = {nx, ny, nz};
  int32_t cnt = 0;
  for (int b = 0; b < nn[v]; ++b)
    for (int a = 0; a < nn[u]; ++a) {
      int c[4][3];
      const int corners[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
      for (int q = 0; q < 4; ++q) {
        c[q][axis] = hi ? nn[axis] : 0;
        c[q][u] = a + corners[q][0];
        c[q][v] = b + corners[q][1];
      }
      // (u, v, axis) is right-handed, so dxi along u, deta along v gives deta x dxi = -axis;
      // the high face needs +axis: reverse the loop there.
      int order[4] = {0, 1, 2, 3};
      if (hi) { order[1] = 3; order[3] = 1; }
      if (quads)
        for (int q = 0; q < 4; ++q) {
          const int* p = c[order[q]];
          quads[4 * cnt + q] = (int32_t)hex_node(p[0], p[1], p[2], nx, ny);
        }
      ++cnt;
    }
  return cnt;
}

}  // extern "C"
