// TACSMeshLoader.cpp -- Nastran bulk-data reader for the shell path and the
// assembler set-up behind it.
//
// Cards (the subset TACSMeshLoader::scanBDFFile reads for shells,
// src/io/TACSMeshLoader.cpp:569-1098):
//   GRID    short field (8-column) or free field (comma separated)
//   GRID*   long field (16-column) over two lines
//   CQUAD4 / CQUADR (and the * long-field variants), with continuation
//           lines starting with ' ' or '*'
//   SPC     short field: SID G C D  (C = digits of the constrained dofs)
//   SPC*    long field
// Lines starting with '$' are comments; the bulk section starts after
// BEGIN BULK (the whole file if absent) and ends at ENDDATA / END BULK.
// Numbers accept the Nastran exponent forms 1.5-3, 2.+4, 1.0D-2.
//
// Nodes and elements are ordered by their file ids; CQUAD4 corners
// (counter-clockwise in the file) are stored in the tensor order
// [0, 1, 3, 2] of the shell basis.  createTACS renumbers the nodes by first
// appearance in element order, as TACSCreator does on one rank
// (src/TACSCreator.cpp:1141-1202), so the BSR pattern matches the
// reference's.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "TACSShellGPU.h"

namespace shellgpu {

namespace {

std::string field(const std::string &line, size_t start, size_t width) {
  if (start >= line.size()) return std::string();
  return line.substr(start, std::min(width, line.size() - start));
}

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// Nastran real: plain, E/D exponent, or the implied-exponent form
// "1.5-3" = 1.5e-3 (a sign after the first character starts the exponent)
double nastran_real(const std::string &raw) {
  std::string s = trim(raw);
  if (s.empty()) return 0.0;
  for (char &c : s)
    if (c == 'd' || c == 'D') c = 'e';
  bool has_e = s.find('e') != std::string::npos || s.find('E') != std::string::npos;
  if (!has_e) {
    for (size_t i = 1; i < s.size(); i++) {
      if ((s[i] == '-' || s[i] == '+') && s[i - 1] != 'e' && s[i - 1] != 'E') {
        s.insert(i, 1, 'e');
        break;
      }
    }
  }
  return std::strtod(s.c_str(), nullptr);
}

int to_int(const std::string &s) { return std::atoi(trim(s).c_str()); }

std::vector<std::string> split_free(const std::string &line) {
  std::vector<std::string> out;
  std::stringstream ss(line);
  std::string item;
  while (std::getline(ss, item, ',')) out.push_back(trim(item));
  return out;
}

bool starts(const std::string &line, const char *tag) { return line.compare(0, strlen(tag), tag) == 0; }

}  // namespace

TACSMeshLoader::TACSMeshLoader() : num_components(0) {}

TACSMeshLoader::~TACSMeshLoader() {
  for (TACSElement *e : comp_element)
    if (e) e->decref();
}

int TACSMeshLoader::scanBDFFile(const char *file_name) {
  std::ifstream in(file_name);
  if (!in) {
    fprintf(stderr, "TACSMeshLoader: Unable to open file %s\n", file_name);
    return 1;
  }
  std::vector<std::string> lines;
  std::string line;
  bool has_bulk = false;
  size_t bulk_start = 0;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (starts(line, "BEGIN BULK")) {
      has_bulk = true;
      bulk_start = lines.size() + 1;
    }
    lines.push_back(line);
  }
  if (!has_bulk) bulk_start = 0;

  struct Node {
    int id;
    double x[3];
  };
  struct Elem {
    int id, comp;
    int nodes[4];
  };
  struct Spc {
    int node, dof;
    double val;
  };
  std::vector<Node> nodes;
  std::vector<Elem> elems;
  std::vector<Spc> spcs;
  std::map<int, std::string> comp_type;
  int fail = 0;

  for (size_t i = bulk_start; i < lines.size(); i++) {
    const std::string &l = lines[i];
    if (l.empty() || l[0] == '$') continue;
    if (starts(l, "ENDDATA") || starts(l, "END BULK")) break;
    if (starts(l, "GRID*")) {
      // GRID* id cp x y / * z : fields of 16 columns after the 8-column name
      const std::string l2 = (i + 1 < lines.size()) ? lines[++i] : std::string();
      Node n;
      n.id = to_int(field(l, 8, 16));
      n.x[0] = nastran_real(field(l, 40, 16));
      n.x[1] = nastran_real(field(l, 56, 16));
      n.x[2] = nastran_real(field(l2, 8, 16));
      nodes.push_back(n);
    } else if (starts(l, "GRID")) {
      Node n;
      if (l.find(',') != std::string::npos) {
        std::vector<std::string> f = split_free(l);
        n.id = f.size() > 1 ? to_int(f[1]) : 0;
        for (int k = 0; k < 3; k++) n.x[k] = f.size() > (size_t)(3 + k) ? nastran_real(f[3 + k]) : 0.0;
      } else {
        n.id = to_int(field(l, 8, 8));
        for (int k = 0; k < 3; k++) n.x[k] = nastran_real(field(l, 24 + 8 * k, 8));
      }
      nodes.push_back(n);
    } else if (starts(l, "SPC*")) {
      // SPC* sid g c d : 16-column fields
      Spc s;
      s.node = to_int(field(l, 24, 16));
      s.val = nastran_real(field(l, 56, 16));
      const std::string c = field(l, 40, 16);
      for (char ch : c)
        if (ch >= '1' && ch <= '6') spcs.push_back({s.node, ch - '1', s.val});
    } else if (starts(l, "SPC")) {
      Spc s;
      s.node = to_int(field(l, 16, 8));
      s.val = nastran_real(field(l, 32, 8));
      const std::string c = field(l, 24, 8);
      for (char ch : c)
        if (ch >= '1' && ch <= '6') spcs.push_back({s.node, ch - '1', s.val});
    } else if (starts(l, "CQUAD4") || starts(l, "CQUADR")) {
      const int width = (l.size() > 6 && l[6] == '*') ? 16 : 8;
      Elem e;
      e.id = to_int(field(l, 8, width));
      e.comp = to_int(field(l, 8 + width, width));
      if (e.id <= 0 || e.comp <= 0) {
        fprintf(stderr, "TACSMeshLoader: bad element card\n %s\n", l.c_str());
        fail = 1;
        break;
      }
      // corner ids: the rest of this line, then continuation lines
      std::vector<int> ids;
      size_t pos = 8 + 2 * width;
      std::string cur = l;
      while ((int)ids.size() < 4) {
        for (; pos < cur.size() && (int)ids.size() < 4; pos += width) {
          const std::string f = trim(field(cur, pos, width));
          if (!f.empty()) ids.push_back(std::atoi(f.c_str()));
        }
        if ((int)ids.size() >= 4) break;
        if (i + 1 < lines.size() && !lines[i + 1].empty() && (lines[i + 1][0] == ' ' || lines[i + 1][0] == '*')) {
          cur = lines[++i];
          pos = 8;
        } else {
          break;
        }
      }
      if (ids.size() < 4) {
        fprintf(stderr, "TACSMeshLoader: CQUAD4 %d has %zu nodes\n", e.id, ids.size());
        fail = 1;
        break;
      }
      // counter-clockwise corners -> tensor order
      e.nodes[0] = ids[0];
      e.nodes[1] = ids[1];
      e.nodes[2] = ids[3];
      e.nodes[3] = ids[2];
      elems.push_back(e);
      if (!comp_type.count(e.comp)) comp_type[e.comp] = starts(l, "CQUADR") ? "CQUADR" : "CQUAD4";
    } else {
      fprintf(stderr, "TACSMeshLoader: Element not recognized. Line\n %s\n", l.c_str());
    }
  }
  if (fail) return fail;

  std::stable_sort(nodes.begin(), nodes.end(), [](const Node &a, const Node &b) { return a.id < b.id; });
  std::stable_sort(elems.begin(), elems.end(), [](const Elem &a, const Elem &b) { return a.id < b.id; });
  std::map<int, int> index;
  for (size_t k = 0; k < nodes.size(); k++) index[nodes[k].id] = (int)k;

  Xpts.resize(3 * nodes.size());
  for (size_t k = 0; k < nodes.size(); k++)
    for (int j = 0; j < 3; j++) Xpts[3 * k + j] = nodes[k].x[j];
  conn.resize(4 * elems.size());
  elem_comp.resize(elems.size());
  num_components = 0;
  for (size_t k = 0; k < elems.size(); k++) {
    for (int j = 0; j < 4; j++) {
      auto it = index.find(elems[k].nodes[j]);
      if (it == index.end()) {
        fprintf(stderr, "TACSMeshLoader: element %d references missing node %d\n", elems[k].id, elems[k].nodes[j]);
        return 1;
      }
      conn[4 * k + j] = it->second;
    }
    elem_comp[k] = elems[k].comp - 1;
    num_components = std::max(num_components, elems[k].comp);
  }
  bc_node.clear();
  bc_dof.clear();
  bc_val.clear();
  for (const Spc &s : spcs) {
    auto it = index.find(s.node);
    if (it == index.end()) {
      fprintf(stderr, "TACSMeshLoader: SPC references missing node %d\n", s.node);
      return 1;
    }
    bc_node.push_back(it->second);
    bc_dof.push_back(s.dof);
    bc_val.push_back(s.val);
  }
  comp_elems.assign(num_components, "");
  comp_descript.assign(num_components, "");
  for (auto &kv : comp_type) comp_elems[kv.first - 1] = kv.second;
  for (TACSElement *e : comp_element)
    if (e) e->decref();
  comp_element.assign(num_components, nullptr);
  return 0;
}

const char *TACSMeshLoader::getElementDescript(int comp) {
  if (comp < 0 || comp >= num_components) return NULL;
  return comp_elems[comp].c_str();
}

void TACSMeshLoader::setElement(int comp, TACSElement *elem) {
  if (comp < 0 || comp >= num_components || !elem) return;
  elem->incref();
  if (comp_element[comp]) comp_element[comp]->decref();
  comp_element[comp] = elem;
}

void TACSMeshLoader::getBCs(std::vector<int> &nodes, std::vector<int> &masks, std::vector<TacsScalar> &vals) const {
  // one entry per constrained node: OR of the SPC dofs, values per dof
  std::map<int, size_t> slot;
  nodes.clear();
  masks.clear();
  vals.clear();
  for (size_t k = 0; k < bc_node.size(); k++) {
    auto it = slot.find(bc_node[k]);
    size_t s;
    if (it == slot.end()) {
      s = nodes.size();
      slot[bc_node[k]] = s;
      nodes.push_back(bc_node[k]);
      masks.push_back(0);
      vals.insert(vals.end(), 6, 0.0);
    } else {
      s = it->second;
    }
    masks[s] |= 1 << bc_dof[k];
    vals[6 * s + bc_dof[k]] = bc_val[k];
  }
}

TACSAssembler *TACSMeshLoader::createTACS(int vars_per_node, int device) {
  if (vars_per_node != 6) {
    fprintf(stderr, "TACSMeshLoader::createTACS: shells have 6 variables per node\n");
    return NULL;
  }
  const int nn = getNumNodes(), ne = getNumElements();
  for (int c = 0; c < num_components; c++) {
    if (!comp_element[c]) {
      fprintf(stderr, "TACSMeshLoader::createTACS: component %d has no element\n", c);
      return NULL;
    }
  }
  // first-appearance renumbering in element order; unreferenced nodes last
  std::vector<int> newnum(nn, -1);
  int next = 0;
  for (int k = 0; k < 4 * ne; k++)
    if (newnum[conn[k]] < 0) newnum[conn[k]] = next++;
  for (int k = 0; k < nn; k++)
    if (newnum[k] < 0) newnum[k] = next++;
  std::vector<TacsScalar> X(3 * nn);
  for (int k = 0; k < nn; k++)
    for (int j = 0; j < 3; j++) X[3 * newnum[k] + j] = Xpts[3 * k + j];
  std::vector<int> c2(4 * ne);
  for (int k = 0; k < 4 * ne; k++) c2[k] = newnum[conn[k]];

  TACSAssembler *tacs = new TACSAssembler(nn, ne, c2.data(), X.data(), device);
  std::vector<TACSElement *> elems(ne);
  for (int e = 0; e < ne; e++) elems[e] = comp_element[elem_comp[e]];
  if (tacs->setElements(elems.data())) {
    delete tacs;
    return NULL;
  }
  tacs->setComponents(num_components, elem_comp.data());
  std::vector<int> bn, bm;
  std::vector<TacsScalar> bv;
  getBCs(bn, bm, bv);
  for (int &n : bn) n = newnum[n];
  if (!bn.empty() && tacs->addBCs((int)bn.size(), bn.data(), bm.data(), bv.data())) {
    delete tacs;
    return NULL;
  }
  return tacs;
}

}  // namespace shellgpu
