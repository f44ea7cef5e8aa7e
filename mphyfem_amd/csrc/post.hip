// post.hip -- post-processing of a converged state: Cauchy stress and von Mises at the Gauss
// points, their element means and nodal means, and the .vtu writer.
//
// Reference (chennachaos/mphyfem, src/):
//   Elem_Solid_3D_Disp::projectStress          Elem_Solid_3D_Disp.cpp:480-557  (per Gauss point:
//       reference-configuration gradients, F, computeStressAndTangent, sigma[varindex] for
//       varindex < 9, von Mises for 9, the pressure for 10)
//   Elem_Solid_3D_Disp::elementContourplot     Elem_Solid_3D_Disp.cpp:378-421  (element value =
//       mean over the Gauss points)
//   femSolidmechanics::projectStresses         femSolidmechanics3.cpp:918-960   (nodal value =
//       sum of the element values at the node / number of elements at the node)
//   femSolidmechanics::postProcess             femSolidmechanics3.cpp:267-776   (the .vtu file:
//       points at X + u, "disp", "velocity", cell "ElemID"/"MatlID", field "TIME", the pressure)
//
// The reference extrapolates Gauss-point values to the nodes only for its Bernstein elements
// (Elem_Solid_3D_Disp.cpp:453-467 throws for the Lagrange shapes); here every node of an element
// takes the element's Gauss-point mean, and the nodal mean runs over the adjacent elements in
// increasing element id, so the result is deterministic.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "fem_device.hpp"
#include "mpf_internal.hpp"

namespace mpf {
namespace {
using namespace dev;

constexpr int kNC = MPF_STRESS_NCOMP;  // sigma (9, column-major map), von Mises, tr(sigma)/3

// One lane per (element, Gauss point).
template <int NPE, int NGP>
__global__ void __launch_bounds__(256) k_stress_gauss(ElemArgs A, int nElem, int mixedP1, int p0,
                                                      double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nElem * NGP) return;
  const int e = (int)(t / NGP), gp = (int)(t - (int64_t)e * NGP);
  const int32_t* cn = A.conn + (int64_t)e * NPE;
  double d[3][NPE];
  if constexpr (NPE == 8) {
    hex8_dN(gp, d);
  } else {
    double x1, x2, x3;
    tet_gp(gp, x1, x2, x3);
    tet10_dN(x1, x2, x3, d[0], d[1], d[2]);
  }
  double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Bi[3][3];
  for (int a = 0; a < NPE; ++a) {
    const int32_t nd = cn[a];
    for (int c = 0; c < 3; ++c) {
      const double xc = A.X[3 * (int64_t)nd + c];
      B[0][c] += xc * d[0][a];
      B[1][c] += xc * d[1][a];
      B[2][c] += xc * d[2][a];
    }
  }
  inv3(B, Bi);
  double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int a = 0; a < NPE; ++a) {
    const int32_t nd = cn[a];
    const double gx = d[0][a] * Bi[0][0] + d[1][a] * Bi[0][1] + d[2][a] * Bi[0][2];
    const double gy = d[0][a] * Bi[1][0] + d[1][a] * Bi[1][1] + d[2][a] * Bi[1][2];
    const double gz = d[0][a] * Bi[2][0] + d[1][a] * Bi[2][1] + d[2][a] * Bi[2][2];
    for (int i = 0; i < 3; ++i) {
      const double ui = A.u[3 * (int64_t)nd + i];
      F[i][0] += ui * gx;
      F[i][1] += ui * gy;
      F[i][2] += ui * gz;
    }
  }
  double pres = 0.0;
  if (mixedP1) {
    double Np[4];
    tet_Np(gp, Np);
    for (int q = 0; q < 4; ++q) pres += Np[q] * A.pres[cn[q]];
  } else if (p0) {
    pres = A.pres[e];
  }
  double mu, K;
  elem_material(A, e, mu, K);
  NHPoint P;
  neohooke_point(F, det3(F), mu, K, A.Utype, mixedP1 || p0, pres, 1.0, false, P);
  double* o = out + t * kNC;
  // column-major 9-index map (utilitiesmaterialHyperelastic.cpp:191)
  const double s[9] = {P.sig[0][0], P.sig[1][0], P.sig[2][0], P.sig[0][1], P.sig[1][1],
                       P.sig[2][1], P.sig[0][2], P.sig[1][2], P.sig[2][2]};
  for (int k = 0; k < 9; ++k) o[k] = s[k];
  const double a = s[0] - s[4], b = s[4] - s[8], c = s[8] - s[0];
  o[9] = sqrt((a * a + b * b + c * c + 6.0 * (s[1] * s[1] + s[2] * s[2] + s[5] * s[5])) / 2.0);
  o[10] = (s[0] + s[4] + s[8]) / 3.0;
}

// Element means over the Gauss points, in Gauss-point order.
__global__ void k_stress_elem(int nElem, int ngp, const double* __restrict__ g, double* __restrict__ ev) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nElem * kNC) return;
  const int64_t e = t / kNC;
  const int k = (int)(t - e * kNC);
  double s = 0.0;
  for (int gp = 0; gp < ngp; ++gp) s += g[(e * ngp + gp) * kNC + k];
  ev[t] = s / ngp;
}

// Nodal means over the adjacent elements, visited in increasing element id (the adjacency list
// itself is unordered).
__global__ void k_stress_nodal(int nNode, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nel,
                               const double* __restrict__ ev, double* __restrict__ nv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nNode) return;
  const int a = nptr[n], z = nptr[n + 1];
  double s[kNC];
  for (int k = 0; k < kNC; ++k) s[k] = 0.0;
  int last = -1;
  for (int cnt = 0; cnt < z - a; ++cnt) {
    int next = 0x7fffffff;
    for (int j = a; j < z; ++j)
      if (nel[j] > last && nel[j] < next) next = nel[j];
    last = next;
    for (int k = 0; k < kNC; ++k) s[k] += ev[(int64_t)next * kNC + k];
  }
  const int m = z - a;
  for (int k = 0; k < kNC; ++k) nv[(int64_t)n * kNC + k] = m ? s[k] / m : 0.0;
}

int ngp_of(const mpf_problem* p) { return p->npe == 8 ? 8 : 4; }

ElemArgs post_args(mpf_problem* p) {
  ElemArgs a{};
  a.X = p->X.ptr;
  a.u = p->disp.ptr;
  a.uprev = p->dispPrev.ptr;
  a.pres = p->pres.ptr;
  a.conn = p->conn.ptr;
  a.mu = p->m.mu;
  a.Kinv = p->m.Kinv;
  a.Utype = p->m.Utype;
  return batch_elem_args(p, a);
}

// Gauss-point values into g [nElem][nGP][kNC] (device).
mpf_status enqueue_gauss(mpf_problem* p, double* g) {
  const ElemArgs A = post_args(p);
  const int ngp = ngp_of(p);
  const int64_t tot = (int64_t)p->nElem * ngp;
  const unsigned grid = (unsigned)((tot + 255) / 256);
  const int mixedP1 = p->etype == MPF_TET10P1, p0 = p->p0 ? 1 : 0;
  if (p->npe == 8)
    hipLaunchKernelGGL((k_stress_gauss<8, 8>), dim3(grid), dim3(256), 0, p->stream, A, p->nElem, 0, p0, g);
  else
    hipLaunchKernelGGL((k_stress_gauss<10, 4>), dim3(grid), dim3(256), 0, p->stream, A, p->nElem, mixedP1, 0, g);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status project(mpf_problem* p, std::vector<double>* ev_h, std::vector<double>* nv_h) {
  const int ngp = ngp_of(p);
  DevBuf<double> g, ev, nv;
  MPF_TRY(g.alloc((size_t)p->nElem * ngp * kNC));
  MPF_TRY(ev.alloc((size_t)p->nElem * kNC));
  MPF_TRY(enqueue_gauss(p, g.ptr));
  const int64_t te = (int64_t)p->nElem * kNC;
  hipLaunchKernelGGL(k_stress_elem, dim3((unsigned)((te + 255) / 256)), dim3(256), 0, p->stream, p->nElem, ngp, g.ptr,
                     ev.ptr);
  if (nv_h) {
    MPF_TRY(nv.alloc((size_t)p->nNode * kNC));
    hipLaunchKernelGGL(k_stress_nodal, dim3((unsigned)((p->nNode + 255) / 256)), dim3(256), 0, p->stream, p->nNode,
                       p->pat.nptr.ptr, p->pat.nel.ptr, ev.ptr, nv.ptr);
  }
  MPF_HIP(hipGetLastError());
  if (ev_h) {
    ev_h->resize((size_t)p->nElem * kNC);
    MPF_TRY(ev.download(ev_h->data(), ev_h->size(), p->stream));
  }
  if (nv_h) {
    nv_h->resize((size_t)p->nNode * kNC);
    MPF_TRY(nv.download(nv_h->data(), nv_h->size(), p->stream));
  }
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

// ---- .vtu (VTK XML UnstructuredGrid) ----
struct VtuArray {
  std::string name;
  int ncomp;
  bool is_int;  // Int64 (connectivity, offsets), UInt8 (types), else Float64
  bool is_u8;
  std::vector<double> f;
  std::vector<int64_t> i;
  std::vector<uint8_t> u;
  size_t bytes() const { return is_u8 ? u.size() : (is_int ? i.size() * 8 : f.size() * 8); }
  const void* data() const { return is_u8 ? (const void*)u.data() : (is_int ? (const void*)i.data() : (const void*)f.data()); }
  const char* type() const { return is_u8 ? "UInt8" : (is_int ? "Int64" : "Float64"); }
};

VtuArray farr(const std::string& name, int ncomp, std::vector<double> v) {
  VtuArray a{name, ncomp, false, false, std::move(v), {}, {}};
  return a;
}

void write_array_header(std::FILE* f, const VtuArray& a, bool binary, uint64_t& offset, size_t ntuples) {
  std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\"", a.type(), a.name.c_str());
  if (a.ncomp > 1) std::fprintf(f, " NumberOfComponents=\"%d\"", a.ncomp);
  (void)ntuples;
  if (binary) {
    std::fprintf(f, " format=\"appended\" offset=\"%llu\"/>\n", (unsigned long long)offset);
    offset += 8 + a.bytes();
    return;
  }
  std::fprintf(f, " format=\"ascii\">\n");
  const size_t n = a.is_u8 ? a.u.size() : (a.is_int ? a.i.size() : a.f.size());
  for (size_t k = 0; k < n; ++k) {
    if (a.is_u8) std::fprintf(f, "%u", (unsigned)a.u[k]);
    else if (a.is_int) std::fprintf(f, "%lld", (long long)a.i[k]);
    else std::fprintf(f, "%.17g", a.f[k]);
    std::fputc((k + 1) % 9 == 0 || k + 1 == n ? '\n' : ' ', f);
  }
  std::fprintf(f, "        </DataArray>\n");
}

}  // namespace
}  // namespace mpf

using namespace mpf;

extern "C" {

mpf_status mpf_stress_gauss(mpf_problem* p, double* out) {
  if (!p || p->etype == 0 || !out) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  DevBuf<double> g;
  MPF_TRY(g.alloc((size_t)p->nElem * ngp_of(p) * kNC));
  MPF_TRY(enqueue_gauss(p, g.ptr));
  MPF_TRY(g.download(out, g.n, p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

mpf_status mpf_stress_project(mpf_problem* p, double* elem, double* nodal) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  std::vector<double> ev, nv;
  MPF_TRY(project(p, elem ? &ev : nullptr, nodal ? &nv : nullptr));
  if (elem) std::copy(ev.begin(), ev.end(), elem);
  if (nodal) std::copy(nv.begin(), nv.end(), nodal);
  return MPF_OK;
}

mpf_status mpf_write_vtu(mpf_problem* p, const char* path, double time, int32_t binary) {
  if (!p || p->etype == 0 || !path) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  const int64_t nN = p->nNode, nE = p->nElem, npe = p->npe;
  std::vector<double> u((size_t)nN * 3), pres((size_t)p->npres);
  std::vector<int32_t> conn((size_t)nE * npe);
  MPF_TRY(p->disp.download(u.data(), u.size(), p->stream));
  MPF_TRY(p->pres.download(pres.data(), pres.size(), p->stream));
  MPF_TRY(p->conn.download(conn.data(), conn.size(), p->stream));
  std::vector<double> ev, nv;
  MPF_TRY(project(p, &ev, &nv));  // synchronises the stream
  std::vector<int32_t> emat;
  if (p->batch) {
    std::vector<int32_t> off;
    if (batch_elem_offsets(p, off) == MPF_OK && (int64_t)off.back() == nE) {
      emat.assign((size_t)nE, 0);
      for (size_t m = 0; m + 1 < off.size(); ++m)
        for (int32_t e = off[m]; e < off[m + 1]; ++e) emat[e] = (int32_t)m;
    }
  }
  // points at the current configuration, NodePosNew = NodePosOrig + disp (femSolidmechanics1.cpp:1183)
  std::vector<double> pts((size_t)nN * 3);
  for (size_t k = 0; k < pts.size(); ++k) pts[k] = p->X_h[k] + u[k];
  std::vector<VtuArray> pd, cd;
  pd.push_back(farr("disp", 3, u));
  pd.push_back(farr("velocity", 3, std::vector<double>((size_t)nN * 3, 0.0)));  // static analysis
  if (p->etype == MPF_TET10P1) {
    // P1 pressure at the vertices; a mid-edge node takes the mean of its edge's two vertices
    // (the midnodeData rule of the quadratic 2D mixed elements, femSolidmechanics3.cpp:614-626)
    std::vector<double> pn(pres.begin(), pres.begin() + nN);
    static const int edge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
    for (int64_t e = 0; e < nE; ++e) {
      const int32_t* c = &conn[(size_t)e * 10];
      for (int k = 0; k < 6; ++k) pn[c[4 + k]] = 0.5 * pres[c[edge[k][0]]] + 0.5 * pres[c[edge[k][1]]];
    }
    pd.push_back(farr("pressure", 1, pn));
  }
  std::vector<double> sn((size_t)nN * 9), vmn((size_t)nN), se((size_t)nE * 9), vme((size_t)nE);
  for (int64_t n = 0; n < nN; ++n) {
    for (int k = 0; k < 9; ++k) sn[n * 9 + k] = nv[n * kNC + k];
    vmn[n] = nv[n * kNC + 9];
  }
  for (int64_t e = 0; e < nE; ++e) {
    for (int k = 0; k < 9; ++k) se[e * 9 + k] = ev[e * kNC + k];
    vme[e] = ev[e * kNC + 9];
  }
  pd.push_back(farr("stress", 9, sn));
  pd.push_back(farr("vonMises", 1, vmn));
  cd.push_back(farr("ElemID", 1, std::vector<double>((size_t)nE, 1.0)));
  std::vector<double> mat((size_t)nE, 1.0);
  for (size_t e = 0; e < emat.size(); ++e) mat[e] = 1.0 + emat[e];
  cd.push_back(farr("MatlID", 1, mat));
  if (p->p0) cd.push_back(farr("pressure", 1, std::vector<double>(pres.begin(), pres.begin() + nE)));
  cd.push_back(farr("stress", 9, se));
  cd.push_back(farr("vonMises", 1, vme));
  // cells: VTK_HEXAHEDRON (12) with the reference's node order, VTK_QUADRATIC_TETRA (24) with
  // nodes 8 and 9 exchanged (tetra10nodemap, femSolidmechanics3.cpp:324-325)
  VtuArray vc{"connectivity", 1, true, false, {}, {}, {}}, vo{"offsets", 1, true, false, {}, {}, {}},
      vt{"types", 1, false, true, {}, {}, {}};
  vc.i.resize((size_t)nE * npe);
  vo.i.resize((size_t)nE);
  vt.u.assign((size_t)nE, npe == 8 ? 12 : 24);
  for (int64_t e = 0; e < nE; ++e) {
    for (int a = 0; a < npe; ++a) {
      int slot = a;
      if (npe == 10 && a >= 8) slot = a == 8 ? 9 : 8;
      vc.i[(size_t)e * npe + slot] = conn[(size_t)e * npe + a];
    }
    vo.i[e] = (e + 1) * npe;
  }
  VtuArray vp = farr("Points", 3, pts);
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return MPF_ERR_ARG;
  const bool bin = binary != 0;
  uint64_t off = 0;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
                  "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <UnstructuredGrid>\n");
  std::fprintf(f, "    <FieldData>\n      <DataArray type=\"Float64\" Name=\"TIME\" NumberOfTuples=\"1\" "
                  "format=\"ascii\">%.17g</DataArray>\n    </FieldData>\n", time);
  std::fprintf(f, "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)nN, (long long)nE);
  std::fprintf(f, "      <PointData Vectors=\"disp\"%s>\n", p->etype == MPF_TET10P1 ? " Scalars=\"pressure\"" : "");
  for (auto& a : pd) write_array_header(f, a, bin, off, (size_t)nN);
  std::fprintf(f, "      </PointData>\n      <CellData>\n");
  for (auto& a : cd) write_array_header(f, a, bin, off, (size_t)nE);
  std::fprintf(f, "      </CellData>\n      <Points>\n");
  write_array_header(f, vp, bin, off, (size_t)nN);
  std::fprintf(f, "      </Points>\n      <Cells>\n");
  write_array_header(f, vc, bin, off, (size_t)nE);
  write_array_header(f, vo, bin, off, (size_t)nE);
  write_array_header(f, vt, bin, off, (size_t)nE);
  std::fprintf(f, "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n");
  if (bin) {
    std::fprintf(f, "  <AppendedData encoding=\"raw\">\n   _");
    auto put = [&](const VtuArray& a) {
      const uint64_t nb = a.bytes();
      std::fwrite(&nb, 8, 1, f);
      std::fwrite(a.data(), 1, nb, f);
    };
    for (auto& a : pd) put(a);
    for (auto& a : cd) put(a);
    put(vp);
    put(vc);
    put(vo);
    put(vt);
    std::fprintf(f, "\n  </AppendedData>\n");
  }
  std::fprintf(f, "</VTKFile>\n");
  const bool ok = std::fclose(f) == 0;
  return ok ? MPF_OK : MPF_ERR_ARG;
}

}  // extern "C"
