// deck.cpp -- the reference's input decks: the block-structured `config` file, the Gmsh 2.2 ASCII
// mesh it names, and petsc_options.dat, turned into a problem on the device and run with the
// reference's load stepping.  Host code only.
//
// Reference (chennachaos/mphyfem, src/):
//   femBase::readInput / read*            utilities/femBase.cpp:370-1200   (the config blocks)
//   femBase::readInputGMSH                utilities/femBase.cpp:121-364    (mesh, physical names,
//                                                                           domains and patches)
//   BoundaryPatch::addElement/processData utilities/BoundaryPatch.cpp:33-53
//   isActiveLine                          utilities/util.h:185-188
//   TimeFunction::addTimeBlock / update   utilities/TimeFunction.cpp:25-68
//   MyTime::set / update / cut / stck     utilities/MyTime.cpp:10-90
//   setBoundaryConditions                 femSolidmechanics/femSolidmechanics1.cpp:813-915
//   solveWithNewtonRaphson                femSolidmechanics/femSolidmechanics2.cpp:730-851
//   converged / diverging                 femSolidmechanics/femSolidmechanics1.cpp:616-636
//   SolverPetsc defaults / KSP options    mySolvers/SolverPetsc.cpp:117-140, petsc_options.dat
//
// What a deck may use here: 3D, one domain of Hex8 (Gmsh type 5) or Tet10 (type 11) elements,
// Matl_NeoHookean, ELEM_SOLID_3D_DISP / _MIXED0 / _MIXED1, 'fixed' and 'specified' patches with
// constant values, traction patches of Quad4 (type 3) or Tria6 (type 9) faces, time functions,
// solvertype newton.  All loads that change in time must follow one time function (the problem
// scales its prescribed values and tractions by one load factor); anything else is refused with
// MPF_ERR_ARG and a message on stderr, never silently changed.
#include <algorithm>
#include <array>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mpf.h"

#define MPF_TRY(call)            \
  do {                           \
    mpf_status s_ = (call);      \
    if (s_ != MPF_OK) return s_; \
  } while (0)

namespace {

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// isActiveLine (util.h:185-188)
bool active(const std::string& l) { return l.size() > 1 && !(l[0] == '!' || l[0] == '#' || l[0] == '%'); }

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

// "key : value" (the reference splits on ':'); a line without ':' is read as "key value" (the
// reference's own sample Solver block is written that way).
bool key_value(const std::string& line, std::string& k, std::string& v) {
  const size_t c = line.find(':');
  if (c != std::string::npos) {
    k = trim(line.substr(0, c));
    v = trim(line.substr(c + 1));
    return !k.empty();
  }
  const auto w = split_ws(line);
  if (w.size() < 2) return false;
  k = w[0];
  v = trim(line.substr(line.find(w[0]) + w[0].size()));
  return true;
}

struct TimeFn {  // TimeFunction.cpp:25-68
  std::vector<double> t0, t1;
  std::vector<std::array<double, 8>> p;
  double value(double t) const {
    const double tol = 1e-12;
    double f = 0.0;
    for (size_t i = 0; i < p.size(); ++i) {
      if (t < t0[i] - tol || t > t1[i] - tol) continue;
      const auto& q = p[i];
      const double d = t - t0[i];
      f += q[0] + q[1] * d + q[2] * std::sin(q[3] * d + q[4]) + q[5] * std::cos(q[6] * d + q[7]);
    }
    return f;
  }
};

struct Patch {
  std::string name;
  int tag = 0, dim = 0;
  std::vector<std::vector<int32_t>> elems;  // 0-based node ids of each boundary element
  std::vector<int32_t> nodes;               // sorted unique (BoundaryPatch::processData)
};

struct BC {
  std::string label, type;
  int dof = -1;
  std::string expr;
  int tf = -1;
};

struct Trac {
  std::string label, type;
  double v[3] = {0, 0, 0};
  int tf = -1;
};

struct Matl {
  int id = 0;
  std::string name;
  double mu = 0.0, Kinv = 0.0;
  int Utype = 3;
  bool has_dev = false, has_vol = false;
};

int dof_of(const std::string& s) {
  if (s == "UX") return 0;
  if (s == "UY") return 1;
  if (s == "UZ") return 2;
  return -1;
}

bool parse_double(const std::string& s, double& v) {
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && end != s.c_str() && trim(std::string(end)).empty();
}

}  // namespace

struct mpf_deck {
  std::string config_dir, mesh_name, error;
  int ndim = 3;
  // mesh
  std::vector<double> X;
  std::vector<int32_t> conn;
  int npe = 0, nNode = 0, nElem = 0, gmsh_type = 0;
  std::vector<Patch> patches;
  std::vector<std::string> domains;
  std::map<std::string, std::pair<int, int>> domain_data;  // label -> (material id, element id)
  // config
  std::vector<Matl> mats;
  std::map<int, std::string> elem_types;
  std::vector<BC> bcs;
  std::vector<Trac> tracs;
  std::vector<TimeFn> tfs;
  std::string solvertype = "newton";
  double dt = 1.0, dtMin = 1.0, dtMax = 1.0, finalTime = 1.0, tol = 1e-9;
  int maxSteps = 1 << 30, maxIters = 10;
  bool body_force_nonzero = false, nodal_forces = false;
  mpf_krylov_opts lin{};
  bool lin_from_file = false;
  // problem
  int32_t etype = 0;
  mpf_problem* p = nullptr;
  int load_tf = -1;
};

namespace {

bool fail(mpf_deck* d, const std::string& msg) {
  d->error = msg;
  std::fprintf(stderr, "libmpf deck: %s\n", msg.c_str());
  return false;
}

// femBase::readInputGMSH (femBase.cpp:121-364)
bool read_gmsh(mpf_deck* d, const std::string& path) {
  std::ifstream in(path);
  if (!in) return fail(d, "could not open the mesh file " + path);
  std::string line;
  std::vector<std::string> w;
  // $MeshFormat, version, $EndMeshFormat
  for (int i = 0; i < 3; ++i) std::getline(in, line);
  std::getline(in, line);  // $PhysicalNames
  if (trim(line) != "$PhysicalNames") return fail(d, "mesh: $PhysicalNames expected after the format block");
  std::getline(in, line);
  const int nphys = std::atoi(trim(line).c_str());
  std::vector<int> pdim(nphys), ptag(nphys);
  std::vector<std::string> pname(nphys);
  for (int i = 0; i < nphys; ++i) {
    std::getline(in, line);
    w = split_ws(line);
    if (w.size() < 3) return fail(d, "mesh: bad physical name line: " + line);
    pdim[i] = std::atoi(w[0].c_str());
    ptag[i] = std::atoi(w[1].c_str());
    std::string nm = w[2];
    for (size_t k = 3; k < w.size(); ++k) nm += " " + w[k];
    if (nm.size() >= 2 && nm.front() == '"' && nm.back() == '"') nm = nm.substr(1, nm.size() - 2);
    pname[i] = nm;
  }
  std::getline(in, line);  // $EndPhysicalNames
  int ndim = nphys ? *std::max_element(pdim.begin(), pdim.end()) : 3;
  if (ndim == 1) ndim = 2;
  d->ndim = ndim;
  std::map<int, int> patch_of_tag;
  std::vector<int> domain_tags;
  for (int i = 0; i < nphys; ++i) {
    if (pdim[i] < ndim) {
      Patch pt;
      pt.name = pname[i];
      pt.tag = ptag[i];
      pt.dim = pdim[i];
      patch_of_tag[ptag[i]] = (int)d->patches.size();
      d->patches.push_back(pt);
    } else {
      domain_tags.push_back(ptag[i]);
      d->domains.push_back(pname[i]);
    }
  }
  std::getline(in, line);  // $Nodes
  if (trim(line) != "$Nodes") return fail(d, "mesh: $Nodes expected");
  std::getline(in, line);
  d->nNode = std::atoi(trim(line).c_str());
  d->X.assign(3 * (size_t)d->nNode, 0.0);
  for (int i = 0; i < d->nNode; ++i) {
    std::getline(in, line);
    w = split_ws(line);
    if (w.size() < 4) return fail(d, "mesh: bad node line: " + line);
    for (int c = 0; c < 3; ++c) d->X[3 * (size_t)i + c] = std::strtod(w[1 + c].c_str(), nullptr);
  }
  std::getline(in, line);  // $EndNodes
  std::getline(in, line);  // $Elements
  if (trim(line) != "$Elements") return fail(d, "mesh: $Elements expected");
  std::getline(in, line);
  const int count = std::atoi(trim(line).c_str());
  d->nElem = 0;
  for (int e = 0; e < count; ++e) {
    std::getline(in, line);
    w = split_ws(line);
    if (w.size() < 4) return fail(d, "mesh: bad element line: " + line);
    // elm-number elm-type number-of-tags <tags> node-number-list (nodes from index 3 + ntags)
    const int type = std::atoi(w[1].c_str()), ntags = std::atoi(w[2].c_str());
    const int tag = std::atoi(w[3].c_str());
    const size_t first = 3 + (size_t)ntags;
    std::vector<int32_t> nodes;
    for (size_t k = first; k < w.size(); ++k) nodes.push_back(std::atoi(w[k].c_str()) - 1);
    if (std::count(domain_tags.begin(), domain_tags.end(), tag)) {
      if (d->nElem == 0) {
        d->gmsh_type = type;
        d->npe = (int)nodes.size();
      } else if (type != d->gmsh_type) {
        return fail(d, "mesh: more than one domain element type");
      }
      d->conn.insert(d->conn.end(), nodes.begin(), nodes.end());
      ++d->nElem;
    } else {
      auto it = patch_of_tag.find(tag);
      if (it == patch_of_tag.end()) return fail(d, "mesh: element tag without a physical name");
      Patch& pt = d->patches[it->second];
      pt.elems.push_back(nodes);
      pt.nodes.insert(pt.nodes.end(), nodes.begin(), nodes.end());
    }
  }
  for (auto& pt : d->patches) {
    std::sort(pt.nodes.begin(), pt.nodes.end());
    pt.nodes.erase(std::unique(pt.nodes.begin(), pt.nodes.end()), pt.nodes.end());
  }
  for (int32_t v : d->conn)
    if (v < 0 || v >= d->nNode) return fail(d, "mesh: element node id out of range");
  return true;
}

// a "{ ... }" block's active lines, the braces consumed
bool block_lines(std::istream& in, std::vector<std::string>& out) {
  std::string line;
  out.clear();
  if (!std::getline(in, line) || trim(line) != "{") return false;
  int depth = 1;
  while (std::getline(in, line)) {
    const std::string t = trim(line);
    if (t == "{") {
      ++depth;
      out.push_back(t);
      continue;
    }
    if (!t.empty() && t[0] == '}') {
      if (--depth == 0) return true;
      out.push_back("}");
      continue;
    }
    if (active(t)) out.push_back(t);
  }
  return false;
}

// labelled sub-blocks "label { key : value ... }" of Boundary Conditions / Tractions / Domains
bool sub_blocks(const std::vector<std::string>& lines,
                std::vector<std::pair<std::string, std::vector<std::string>>>& out) {
  for (size_t i = 0; i < lines.size(); ++i) {
    if (i + 1 >= lines.size() || lines[i + 1] != "{") return false;
    std::pair<std::string, std::vector<std::string>> b;
    b.first = lines[i];
    size_t j = i + 2;
    while (j < lines.size() && lines[j] != "}") b.second.push_back(lines[j++]);
    if (j >= lines.size()) return false;
    out.push_back(b);
    i = j;
  }
  return true;
}

// -ksp_type / -pc_type / -ksp_rtol / -ksp_max_it / -ksp_gmres_restart (petsc_options.dat)
void read_petsc_options(mpf_deck* d, const std::string& path) {
  std::ifstream in(path);
  if (!in) return;
  d->lin_from_file = true;
  std::string ksp = "cg", pc = "ilu", line;
  while (std::getline(in, line)) {
    const auto w = split_ws(line);
    if (w.size() < 2) continue;
    if (w[0] == "-ksp_type") ksp = w[1];
    else if (w[0] == "-pc_type") pc = w[1];
    else if (w[0] == "-ksp_rtol") d->lin.rtol = std::strtod(w[1].c_str(), nullptr);
    else if (w[0] == "-ksp_max_it") d->lin.maxit = std::atoi(w[1].c_str());
    else if (w[0] == "-ksp_gmres_restart") d->lin.restart = std::atoi(w[1].c_str());
  }
  const bool direct = ksp == "preonly" || pc == "lu" || pc == "cholesky";
  if (direct) d->lin.ksp = MPF_KSP_BAND;
  else if (ksp == "gmres" || ksp == "fgmres" || ksp == "lgmres") d->lin.ksp = MPF_KSP_GMRES;
  else d->lin.ksp = MPF_KSP_CG;
  if (pc == "none") d->lin.pc = MPF_PC_NONE;
  else if (pc == "jacobi") d->lin.pc = MPF_PC_JACOBI;
  else if (pc == "mg" || pc == "gamg" || pc == "ml" || pc == "hypre" || pc == "fieldsplit")
    d->lin.pc = MPF_PC_PMG;  // multigrid / field split: the two-level p-multigrid (Tet10 meshes)
  else d->lin.pc = MPF_PC_BJACOBI3;  // bjacobi / ilu / sor / asm: the 3x3 node-block Jacobi
}

bool read_config(mpf_deck* d, const std::string& path) {
  std::ifstream in(path);
  if (!in) return fail(d, "could not open the config file " + path);
  std::string line;
  std::vector<std::string> L;
  while (std::getline(in, line)) {
    std::string key = trim(line);
    if (!active(key)) continue;
    // block names as femBase::readInput spells them; the first letter's case is not significant
    // (the reference's own sample deck writes "files")
    key[0] = (char)std::toupper((unsigned char)key[0]);
    if (!block_lines(in, L)) return fail(d, "config: block '" + key + "' has no matching braces");
    std::string k, v;
    if (key == "Files") {
      for (auto& l : L)
        if (key_value(l, k, v) && k == "mesh") d->mesh_name = v;
    } else if (key == "Model Type") {
      for (auto& l : L)
        if (key_value(l, k, v) && k == "dimension") d->ndim = std::atoi(v.c_str());
    } else if (key == "Domains") {
      std::vector<std::pair<std::string, std::vector<std::string>>> sb;
      if (!sub_blocks(L, sb)) return fail(d, "config: bad Domains block");
      for (auto& b : sb) {
        int m = -1, el = -1;
        for (auto& l : b.second)
          if (key_value(l, k, v)) {
            if (k == "Material") m = std::atoi(v.c_str());
            else if (k == "Element") el = std::atoi(v.c_str());
          }
        d->domain_data[b.first] = {m, el};
      }
    } else if (key == "Material") {  // femBase::readMaterialProps, MaterialBase::readInput
      Matl m;
      for (auto& l : L)
        if (key_value(l, k, v)) {
          const auto w = split_ws(v);
          if (k == "id") m.id = std::atoi(v.c_str());
          else if (k == "name") m.name = v;
          else if (k == "data_deviatoric" && !w.empty()) { m.mu = std::strtod(w[0].c_str(), nullptr); m.has_dev = true; }
          else if (k == "data_volumetric" && w.size() >= 2) {
            m.Utype = std::atoi(w[0].c_str());
            m.Kinv = std::strtod(w[1].c_str(), nullptr);
            m.has_vol = true;
          }
        }
      d->mats.push_back(m);
    } else if (key == "Element") {
      int id = 0;
      std::string type;
      for (auto& l : L)
        if (key_value(l, k, v)) {
          if (k == "id") id = std::atoi(v.c_str());
          else if (k == "type") type = v;
        }
      d->elem_types[id] = type;
    } else if (key == "Body Force") {
      for (auto& l : L)
        if (key_value(l, k, v) && k == "value")
          for (auto& t : split_ws(v))
            if (std::strtod(t.c_str(), nullptr) != 0.0) d->body_force_nonzero = true;
    } else if (key == "Boundary Conditions" || key == "Tractions") {
      std::vector<std::pair<std::string, std::vector<std::string>>> sb;
      if (!sub_blocks(L, sb)) return fail(d, "config: bad " + key + " block");
      for (auto& b : sb) {
        if (key == "Boundary Conditions") {
          BC bc;
          bc.label = b.first;
          for (auto& l : b.second)
            if (key_value(l, k, v)) {
              if (k == "type") bc.type = v;
              else if (k == "dof") bc.dof = dof_of(v);
              else if (k == "value") bc.expr = v;
              else if (k == "timefunction") bc.tf = std::atoi(v.c_str()) - 1;
              else return fail(d, "config: option '" + k + "' not available in a boundary condition");
            }
          d->bcs.push_back(bc);
        } else {
          Trac tr;
          tr.label = b.first;
          for (auto& l : b.second)
            if (key_value(l, k, v)) {
              if (k == "type") tr.type = v;
              else if (k == "value") {
                const auto w = split_ws(v);
                for (size_t c = 0; c < w.size() && c < 3; ++c) tr.v[c] = std::strtod(w[c].c_str(), nullptr);
              } else if (k == "timefunction") tr.tf = std::atoi(v.c_str()) - 1;
              else return fail(d, "config: option '" + k + "' not available in a traction");
            }
          d->tracs.push_back(tr);
        }
      }
    } else if (key == "Time Functions") {  // femBase::readTimeFunctions
      for (auto& l : L) {
        const auto w = split_ws(l);
        if (w.size() < 11) return fail(d, "config: a time function needs 11 numbers: " + l);
        const int id = std::atoi(w[0].c_str()) - 1;
        if (id < 0) return fail(d, "config: time function id 0 is not admissible");
        if (id >= (int)d->tfs.size()) d->tfs.resize(id + 1);
        TimeFn& f = d->tfs[id];
        f.t0.push_back(std::strtod(w[1].c_str(), nullptr));
        f.t1.push_back(std::strtod(w[2].c_str(), nullptr));
        std::array<double, 8> q{};
        for (int c = 0; c < 8; ++c) q[c] = std::strtod(w[3 + c].c_str(), nullptr);
        f.p.push_back(q);
      }
    } else if (key == "Solver") {  // femBase::readSolverDetails (femBase.cpp:994-1130)
      for (auto& l : L) {
        if (!key_value(l, k, v)) return fail(d, "config: input error in the Solver block: " + l);
        const auto w = split_ws(v);
        if (k == "solvertype") d->solvertype = v;
        else if (k == "timeStep" && !w.empty()) {
          const double a = std::strtod(w[0].c_str(), nullptr);
          const double b = w.size() > 1 ? std::strtod(w[1].c_str(), nullptr) : a;
          const double c = w.size() > 2 ? std::strtod(w[2].c_str(), nullptr) : (w.size() > 1 ? a : a);
          d->dt = a;
          d->dtMin = b;
          d->dtMax = c;
        } else if (k == "maximumIterations") d->maxIters = std::atoi(v.c_str());
        else if (k == "tolerance") d->tol = std::strtod(v.c_str(), nullptr);
        else if (k == "finalTime") d->finalTime = std::strtod(v.c_str(), nullptr);
        else if (k == "maximumSteps") d->maxSteps = std::atoi(v.c_str());
        // timescheme, spectralRadius, outputFrequency, masslumping, debug, library: no effect on a
        // static (STEADY) Newton run
      }
      if (d->maxIters < 0 || d->tol < 0.0) return fail(d, "config: negative iterations or tolerance");
    } else if (key == "Nodal Forces") {
      d->nodal_forces = true;
    } else if (key == "Output" || key == "Nodal Data Output" || key == "Initial Conditions") {
      // output selections: the .vtu writer writes its fixed set
    } else {
      return fail(d, "config: key not found: " + key);  // femBase::readInput throws the same way
    }
  }
  return true;
}

const Patch* find_patch(const mpf_deck* d, const std::string& name) {
  for (auto& p : d->patches)
    if (p.name == name) return &p;
  return nullptr;
}

// validate what the path supports and build the problem (prepareMatrixPattern)
mpf_status build(mpf_deck* d) {
  if (d->ndim != 3) return fail(d, "only 3D decks run on this path"), MPF_ERR_ARG;
  if (d->solvertype != "newton" && d->solvertype != "arclength")
    return fail(d, "solvertype '" + d->solvertype + "'"), MPF_ERR_ARG;
  if (d->body_force_nonzero) return fail(d, "a nonzero body force is not supported"), MPF_ERR_ARG;
  if (d->nodal_forces) return fail(d, "Nodal Forces are not supported"), MPF_ERR_ARG;
  if (d->domains.size() != 1) return fail(d, "exactly one domain is supported"), MPF_ERR_ARG;
  auto dd = d->domain_data.find(d->domains[0]);
  if (dd == d->domain_data.end()) return fail(d, "the domain '" + d->domains[0] + "' has no Domains entry"), MPF_ERR_ARG;
  const Matl* m = nullptr;
  for (auto& mm : d->mats)
    if (mm.id == dd->second.first) m = &mm;
  if (!m || m->name != "Matl_NeoHookean" || !m->has_dev || !m->has_vol)
    return fail(d, "the domain's material must be Matl_NeoHookean with data_deviatoric and data_volumetric"),
           MPF_ERR_ARG;
  auto et = d->elem_types.find(dd->second.second);
  if (et == d->elem_types.end()) return fail(d, "the domain's element type is not defined"), MPF_ERR_ARG;
  const std::string& en = et->second;
  if (d->gmsh_type == 5 && d->npe == 8) {
    if (en == "ELEM_SOLID_3D_DISP") d->etype = MPF_HEX8;
    else if (en == "ELEM_SOLID_3D_MIXED0") d->etype = MPF_HEX8P0;
  } else if (d->gmsh_type == 11 && d->npe == 10) {
    if (en == "ELEM_SOLID_3D_DISP") d->etype = MPF_TET10;
    else if (en == "ELEM_SOLID_3D_MIXED1") d->etype = MPF_TET10P1;
  }
  if (!d->etype) return fail(d, "element " + en + " on Gmsh type " + std::to_string(d->gmsh_type) + " is not supported"),
                 MPF_ERR_ARG;
  // one load time function
  int tfn = -2;
  auto use_tf = [&](int t) -> bool {
    if (tfn == -2 || tfn == t) { tfn = t; return true; }
    return false;
  };
  for (auto& bc : d->bcs) {
    if (bc.type == "fixed") continue;
    if (bc.type != "specified") return fail(d, "boundary condition type '" + bc.type + "'"), MPF_ERR_ARG;
    if (bc.dof < 0) return fail(d, "boundary condition '" + bc.label + "': dof UX, UY or UZ"), MPF_ERR_ARG;
    double v = 0.0;
    if (!parse_double(bc.expr, v))
      return fail(d, "boundary condition '" + bc.label + "': only constant values (got '" + bc.expr + "')"), MPF_ERR_ARG;
    if (v != 0.0 && !use_tf(bc.tf)) return fail(d, "loads follow more than one time function"), MPF_ERR_ARG;
  }
  for (auto& tr : d->tracs) {
    if (tr.type != "traction") return fail(d, "traction type '" + tr.type + "'"), MPF_ERR_ARG;
    if (tr.v[2] != 0.0) return fail(d, "traction '" + tr.label + "': the third value must be 0 (v0 normal, v1 tangent)"),
                        MPF_ERR_ARG;
    if (!use_tf(tr.tf)) return fail(d, "loads follow more than one time function"), MPF_ERR_ARG;
  }
  d->load_tf = tfn == -2 ? -1 : tfn;
  if (d->load_tf >= (int)d->tfs.size()) return fail(d, "undefined time function"), MPF_ERR_ARG;
  // arc-length: the reference forms Fext once, at time 0 (femSolidmechanics2.cpp:868), and the
  // continuation scales it; here Fext is the tractions at load factor 1, so their time function
  // must be 1 at time 0
  if (d->solvertype == "arclength" && d->load_tf >= 0 && d->tfs[d->load_tf].value(0.0) != 1.0)
    return fail(d, "arclength: the loads' time function must be 1 at time 0"), MPF_ERR_ARG;
  // the linear solver: petsc_options.dat, else the reference's code default CG + ILU
  // (SolverPetsc.cpp:117-140), here CG + 3x3 node-block Jacobi; an indefinite mixed system runs
  // GMRES, and the direct solve (displacement systems only) becomes GMRES at 1e-10 for it
  const bool mixed = d->etype == MPF_TET10P1;
  if (d->lin.pc == MPF_PC_PMG) {  // needs the Tet10 mesh, and changes between applications
    if (d->etype != MPF_TET10P1 && d->etype != MPF_TET10) d->lin.pc = MPF_PC_BJACOBI3;
    else if (d->lin.ksp != MPF_KSP_BAND) d->lin.ksp = MPF_KSP_GMRES;
  }
  if (mixed && (d->lin.ksp == MPF_KSP_CG || d->lin.ksp == MPF_KSP_BAND)) {
    if (d->lin.ksp == MPF_KSP_BAND) d->lin.rtol = 1e-10;
    d->lin.ksp = MPF_KSP_GMRES;
    d->lin.pc = MPF_PC_BJACOBI3;
  }
  mpf_matl mt{};
  mt.mu = m->mu;
  mt.Kinv = m->Kinv;
  mt.Utype = m->Utype;
  mt.mixed = (d->etype == MPF_TET10P1 || d->etype == MPF_HEX8P0) ? 1 : 0;
  mt.tangent = MPF_TANGENT_REFERENCE;  // the reference's element tangent
  MPF_TRY(mpf_problem_create(&d->p, d->etype, d->nNode, d->nElem, d->X.data(), d->conn.data(), &mt));
  for (auto& bc : d->bcs) {  // setBoundaryConditions: later patches win per dof
    const Patch* pt = find_patch(d, bc.label);
    if (!pt) return fail(d, "boundary condition on unknown patch '" + bc.label + "'"), MPF_ERR_ARG;
    if (bc.type == "fixed") {
      MPF_TRY(mpf_add_dirichlet(d->p, (int32_t)pt->nodes.size(), pt->nodes.data(), -1, nullptr));
    } else {
      double v = 0.0;
      parse_double(bc.expr, v);
      std::vector<double> val(pt->nodes.size(), v);
      MPF_TRY(mpf_add_dirichlet(d->p, (int32_t)pt->nodes.size(), pt->nodes.data(), bc.dof, val.data()));
    }
  }
  for (auto& tr : d->tracs) {
    const Patch* pt = find_patch(d, tr.label);
    if (!pt || pt->elems.empty()) return fail(d, "traction on unknown or empty patch '" + tr.label + "'"), MPF_ERR_ARG;
    const size_t npf = pt->elems[0].size();
    std::vector<int32_t> faces;
    for (auto& f : pt->elems) {
      if (f.size() != npf) return fail(d, "traction patch '" + tr.label + "' mixes face types"), MPF_ERR_ARG;
      faces.insert(faces.end(), f.begin(), f.end());
    }
    if (npf == 4) MPF_TRY(mpf_add_traction(d->p, (int32_t)pt->elems.size(), faces.data(), tr.v[0], tr.v[1]));
    else if (npf == 6) MPF_TRY(mpf_add_traction_tri6(d->p, (int32_t)pt->elems.size(), faces.data(), tr.v[0], tr.v[1]));
    else return fail(d, "traction faces must be Quad4 or Tria6"), MPF_ERR_ARG;
  }
  MPF_TRY(mpf_prepare_pattern(d->p));
  return MPF_OK;
}

}  // namespace

extern "C" {

mpf_status mpf_deck_load(const char* config_path, mpf_deck** out) {
  if (!config_path || !out) return MPF_ERR_ARG;
  auto* d = new mpf_deck();
  d->lin.ksp = MPF_KSP_CG;
  d->lin.pc = MPF_PC_BJACOBI3;
  d->lin.restart = 50;
  d->lin.maxit = 20000;
  d->lin.rtol = 1e-8;
  const std::string cfg(config_path);
  const size_t slash = cfg.find_last_of('/');
  d->config_dir = slash == std::string::npos ? "." : cfg.substr(0, slash);
  if (!read_config(d, cfg)) {
    delete d;
    return MPF_ERR_ARG;
  }
  if (d->mesh_name.empty()) {
    fail(d, "config: no mesh in the Files block");
    delete d;
    return MPF_ERR_ARG;
  }
  // the reference opens "./inputs/<mesh>.msh"; here the mesh sits beside the config file
  std::string mesh = d->config_dir + "/" + d->mesh_name;
  if (mesh.size() < 4 || mesh.substr(mesh.size() - 4) != ".msh") mesh += ".msh";
  if (!read_gmsh(d, mesh)) {
    delete d;
    return MPF_ERR_ARG;
  }
  read_petsc_options(d, d->config_dir + "/petsc_options.dat");
  *out = d;
  return MPF_OK;
}

mpf_status mpf_deck_destroy(mpf_deck* d) {
  if (!d) return MPF_OK;
  if (d->p) mpf_problem_destroy(d->p);
  delete d;
  return MPF_OK;
}

mpf_status mpf_deck_get_info(const mpf_deck* d, mpf_deck_info* info) {
  if (!d || !info) return MPF_ERR_ARG;
  std::memset(info, 0, sizeof(*info));
  info->ndim = d->ndim;
  info->nNode = d->nNode;
  info->nElem = d->nElem;
  info->npe = d->npe;
  info->gmsh_type = d->gmsh_type;
  info->elem_type = d->etype;
  info->npatch = (int32_t)d->patches.size();
  info->nbc = (int32_t)d->bcs.size();
  info->ntrac = (int32_t)d->tracs.size();
  info->nmat = (int32_t)d->mats.size();
  info->ntf = (int32_t)d->tfs.size();
  info->arclength = d->solvertype == "arclength" ? 1 : 0;
  info->dt = d->dt;
  info->dt_min = d->dtMin;
  info->dt_max = d->dtMax;
  info->final_time = d->finalTime;
  info->max_steps = d->maxSteps;
  info->max_iters = d->maxIters;
  info->tol = d->tol;
  info->lin = d->lin;
  info->lin_from_file = d->lin_from_file ? 1 : 0;
  return MPF_OK;
}

mpf_status mpf_deck_patch(const mpf_deck* d, int32_t index, char* name, int32_t name_len, int32_t* dim,
                          int32_t* nnodes, int32_t* nodes) {
  if (!d || index < 0 || index >= (int32_t)d->patches.size()) return MPF_ERR_ARG;
  const Patch& p = d->patches[index];
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s", p.name.c_str());
  }
  if (dim) *dim = p.dim;
  if (nnodes) *nnodes = (int32_t)p.nodes.size();
  if (nodes) std::copy(p.nodes.begin(), p.nodes.end(), nodes);
  return MPF_OK;
}

mpf_status mpf_deck_setup(mpf_deck* d, mpf_problem** p) {
  if (!d) return MPF_ERR_ARG;
  if (!d->p) MPF_TRY(build(d));
  if (p) *p = d->p;
  return MPF_OK;
}

// solveWithNewtonRaphson with timeUpdate and MyTime (femSolidmechanics2.cpp:730-851; MyTime.cpp):
// per step cur += dt, load factor lam = timefunction(cur), Newton with the deck's absolute tolerance
// on ||R|| (converged, femSolidmechanics1.cpp:616-622) and the divergence test (:627-636); a failed
// step restores the last converged state and halves dt (cut), a converged one pops the dt stack
// (stck).  vtu_prefix (may be NULL): one .vtu per converged step, "<prefix>-NNNNNN.vtu".
mpf_status mpf_deck_run(mpf_deck* d, const char* vtu_prefix, mpf_deck_report* rep) {
  if (!d || !rep) return MPF_ERR_ARG;
  std::memset(rep, 0, sizeof(*rep));
  MPF_TRY(mpf_deck_setup(d, nullptr));
  mpf_problem* p = d->p;
  if (d->solvertype == "arclength") {  // solveWithArclength (femSolidmechanics2.cpp:855-1039)
    mpf_arclength_opts ao{};
    ao.dt = d->dt;
    ao.final_time = d->finalTime;
    ao.max_steps = d->maxSteps < (1 << 20) ? d->maxSteps : (1 << 20);
    ao.max_iters = d->maxIters;
    ao.tol = d->tol;
    ao.lin = d->lin;
    mpf_arclength_report ar{};
    MPF_TRY(mpf_arclength_solve(p, &ao, nullptr, nullptr, &ar));
    rep->steps_done = ar.steps_done;
    rep->newton_iters = ar.newton_iters;
    rep->krylov_iters = ar.krylov_iters;
    rep->final_lam = ar.final_lam;
    rep->converged = ar.steps_done > 0 ? 1 : 0;
    rep->seconds = ar.seconds;
    if (vtu_prefix) {
      char fn[4096];
      std::snprintf(fn, sizeof fn, "%s-%06d.vtu", vtu_prefix, ar.steps_done);
      MPF_TRY(mpf_write_vtu(p, fn, ar.final_lam, 1));
    }
    return MPF_OK;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const double EPS = 1e-10;  // MathBasic.h:17
  const size_t n3 = 3 * (size_t)d->nNode;
  const size_t npres = d->etype == MPF_HEX8P0 ? (size_t)d->nElem : (size_t)d->nNode;
  std::vector<double> u(n3), pr(npres);
  double cur = 0.0, dt = d->dt;
  std::vector<double> stack{dt};
  int steps = 1, filecount = 0;
  if (vtu_prefix) {
    char fn[4096];
    std::snprintf(fn, sizeof fn, "%s-%06d.vtu", vtu_prefix, filecount);
    MPF_TRY(mpf_write_vtu(p, fn, cur, 1));
  }
  while (cur <= d->finalTime - EPS && steps <= d->maxSteps) {
    cur += dt;
    const double lam = d->load_tf >= 0 ? d->tfs[d->load_tf].value(cur) : 1.0;
    // saveSolution (SolutionDataSolid.cpp:170-184): the converged state is the previous one
    MPF_TRY(mpf_get_state(p, u.data(), pr.data()));
    MPF_TRY(mpf_set_state(p, nullptr, u.data(), nullptr, pr.data()));
    MPF_TRY(mpf_set_bc(p, lam));
    bool conv = false;
    double rn = -1.0, rnPrev = -1.0;
    int iter = 1;
    for (; iter <= d->maxIters; ++iter) {
      const mpf_status st = mpf_assemble(p, iter == 1, lam);
      if (st == MPF_ERR_NEG_JACOBIAN) break;  // "Negative Jacobian encountered": the step fails
      MPF_TRY(st);
      rnPrev = rn;
      MPF_TRY(mpf_rhs_norm(p, &rn));
      rep->newton_iters += 1;
      if (rn < d->tol) {
        conv = true;
        break;
      }
      if (std::isnan(rn) || (iter > 3 && rnPrev > 0.0 && rn / rnPrev > 1.0e7)) break;
      mpf_krylov_report kr{};
      const mpf_status ss = mpf_solve(p, &d->lin, &kr);
      rep->krylov_iters += kr.iterations;
      if (ss != MPF_OK && ss != MPF_ERR_NOT_CONVERGED) {
        if (ss == MPF_ERR_DIVERGED) break;
        return ss;
      }
      MPF_TRY(mpf_update(p, iter == 1));
    }
    rep->final_norm = rn;
    if (conv) {
      ++steps;
      ++filecount;
      ++rep->steps_done;
      rep->final_time = cur;
      rep->final_lam = lam;
      if (vtu_prefix) {
        char fn[4096];
        std::snprintf(fn, sizeof fn, "%s-%06d.vtu", vtu_prefix, filecount);
        MPF_TRY(mpf_write_vtu(p, fn, cur, 1));
      }
      const size_t m = stack.size();
      if (m == 1) dt = stack[0];
      else { dt = stack[m - 1]; stack.pop_back(); }
    } else {
      ++rep->cuts;
      MPF_TRY(mpf_set_state(p, u.data(), nullptr, pr.data(), nullptr));  // reset
      cur -= dt;
      dt *= 0.5;
      if (dt < stack[0] * 0.01) {  // MyTime::cut: "Time step size is less than the minimum value"
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return MPF_ERR_NOT_CONVERGED;
      }
      stack.push_back(dt);
    }
  }
  rep->converged = 1;
  rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return MPF_OK;
}

}  // extern "C"
