// Readers for the reference's ASCII mesh formats (Mesh<Type>::Read*,
// ucs/mesh.tcc).  Each reader turns a file into volume elements and tagged
// boundary faces in the local numbering of mesh.cpp's topology tables, then
// hands them to pcfd_mesh_build_from, which winds the boundary faces outward
// and builds edges, half-edges and the median-dual metrics.
//
//   GMSH 2.2 ASCII           ReadGMSH2_Ascii   ucs/mesh.tcc:3599-3925
//   GMSH 4.0 ASCII           ReadGMSH40_Ascii  ucs/mesh.tcc:3926-4503
//   GMSH 4.1 ASCII           ReadGMSH41_Ascii  ucs/mesh.tcc:4505-5242
//   UGRID ASCII              ReadUGRID_Ascii   ucs/mesh.tcc:6740-6932
//   FieldView ("CRUNCH")     ReadCRUNCH_Ascii  ucs/mesh.tcc:6933-7300
//   SU2 ASCII                ReadSU2_Ascii     ucs/mesh.tcc:3197-3548
//
// Kept as in the reference: 1-based node ids on disk; UGRID and FieldView
// face tags taken as |tag| and shifted up by one when a tag 0 occurs (0 is
// the parallel-boundary tag); SU2 faces tagged by marker index + 1; GMSH
// faces tagged by the physical group (the first physical tag of their
// surface entity in 4.x); nodes that no volume element uses are dropped with
// the remaining ids kept in order (EliminateOrphanedNodes).  GMSH point and
// line elements are ignored.  A GMSH 4.x surface with no physical tag gives
// its faces the entity tag.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pcfd_mesh.h"

void pcfd_mesh_set_error(const std::string& msg);  // mesh.cpp

namespace {

struct Raw {
  std::vector<double> xyz;
  std::vector<int> et, ep{0}, en;       // volume elements
  std::vector<int> bt, bp{0}, bn, bg;   // boundary faces and their tags
  void elem(int t, const int* n, int k) {
    et.push_back(t);
    en.insert(en.end(), n, n + k);
    ep.push_back((int)en.size());
  }
  void face(int t, const int* n, int k, int tag) {
    bt.push_back(t);
    bn.insert(bn.end(), n, n + k);
    bp.push_back((int)bn.size());
    bg.push_back(tag);
  }
};

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::string slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Fail("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// whitespace tokenizer over a string it owns (sections are temporaries)
struct Tok {
  std::string buf;
  const char* p;
  const char* e;
  explicit Tok(std::string s) : buf(std::move(s)), p(buf.data()), e(buf.data() + buf.size()) {}
  Tok(const Tok&) = delete;
  Tok& operator=(const Tok&) = delete;
  void skip() {
    while (p < e && std::isspace((unsigned char)*p)) ++p;
  }
  bool done() {
    skip();
    return p >= e;
  }
  std::string word() {
    skip();
    const char* b = p;
    while (p < e && !std::isspace((unsigned char)*p)) ++p;
    if (b == p) throw Fail("unexpected end of file");
    return std::string(b, p);
  }
  long long integer() {
    std::string w = word();
    char* end = nullptr;
    long long v = std::strtoll(w.c_str(), &end, 10);
    if (*end) throw Fail("expected an integer, found '" + w + "'");
    return v;
  }
  double real() {
    std::string w = word();
    char* end = nullptr;
    double v = std::strtod(w.c_str(), &end);
    if (*end) throw Fail("expected a number, found '" + w + "'");
    return v;
  }
};

std::vector<std::string> lines_of(const std::string& s) {
  std::vector<std::string> out;
  size_t b = 0;
  while (b <= s.size()) {
    size_t e = s.find('\n', b);
    if (e == std::string::npos) e = s.size();
    std::string l = s.substr(b, e - b);
    while (!l.empty() && std::isspace((unsigned char)l.back())) l.pop_back();
    size_t f = 0;
    while (f < l.size() && std::isspace((unsigned char)l[f])) ++f;
    out.push_back(l.substr(f));
    b = e + 1;
  }
  return out;
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

int nodes_of(int t) {
  switch (t) {
    case PCFD_TRI: return 3;
    case PCFD_QUAD: return 4;
    case PCFD_TET: return 4;
    case PCFD_PYRAMID: return 5;
    case PCFD_PRISM: return 6;
    case PCFD_HEX: return 8;
  }
  return 0;
}

// local = file[map[k]] for k = 0..n-1
void permute(int* n, const int* map, int k) {
  int t[8];
  for (int i = 0; i < k; ++i) t[i] = n[map[i]];
  std::memcpy(n, t, sizeof(int) * k);
}

// ---------------------------------------------------------------------------
// GMSH.  Element codes 2 tri, 3 quad, 4 tet, 5 hex, 6 prism, 7 pyramid use
// the local node order as they are; 1 (line) and 15 (point) are ignored.
int gmsh_type(long long code) {
  switch (code) {
    case 2: return PCFD_TRI;
    case 3: return PCFD_QUAD;
    case 4: return PCFD_TET;
    case 5: return PCFD_HEX;
    case 6: return PCFD_PRISM;
    case 7: return PCFD_PYRAMID;
    case 1:
    case 15: return -1;
  }
  throw Fail("GMSH element type " + std::to_string(code) + " is not supported (first-order elements only)");
}

struct NodeMap {
  std::map<long long, int> id;  // file node tag -> 0-based position
  int at(long long tag) const {
    auto it = id.find(tag);
    if (it == id.end()) throw Fail("element refers to undefined node " + std::to_string(tag));
    return it->second;
  }
};

void gmsh_add(Raw& r, const NodeMap& nm, int t, const long long* tags, int tag) {
  int k = nodes_of(t), n[8];
  for (int i = 0; i < k; ++i) n[i] = nm.at(tags[i]);
  if (t == PCFD_TRI || t == PCFD_QUAD) r.face(t, n, k, tag);
  else r.elem(t, n, k);
}

// section body: tokens between "$Name" and "$EndName"
std::string section(const std::string& s, const std::string& name, bool required) {
  const std::string b = "$" + name, e = "$End" + name;
  size_t p = s.find(b);
  while (p != std::string::npos && p + b.size() < s.size() && !std::isspace((unsigned char)s[p + b.size()]))
    p = s.find(b, p + 1);
  if (p == std::string::npos) {
    if (required) throw Fail("GMSH file has no " + b + " section");
    return std::string();
  }
  size_t q = s.find(e, p);
  if (q == std::string::npos) throw Fail("GMSH file has no " + e);
  return s.substr(p + b.size(), q - p - b.size());
}

void read_gmsh2(const std::string& s, Raw& r) {
  NodeMap nm;
  {
    Tok t(section(s, "Nodes", true));
    long long nn = t.integer();
    for (long long i = 0; i < nn; ++i) {
      long long tag = t.integer();
      nm.id[tag] = (int)i;
      for (int d = 0; d < 3; ++d) r.xyz.push_back(t.real());
    }
  }
  Tok t(section(s, "Elements", true));
  long long ne = t.integer();
  for (long long i = 0; i < ne; ++i) {
    t.integer();  // element id
    long long code = t.integer();
    long long ntag = t.integer();
    if (ntag < 0) throw Fail("GMSH 2 element with a negative tag count");
    std::vector<long long> tags(ntag);
    for (auto& v : tags) v = t.integer();
    int ty = gmsh_type(code);
    int k = ty < 0 ? (code == 1 ? 2 : 1) : nodes_of(ty);
    long long n[8];
    for (int j = 0; j < k; ++j) n[j] = t.integer();
    if (ty < 0) continue;
    gmsh_add(r, nm, ty, n, ntag > 0 ? (int)tags[0] : 0);
  }
}

// $Entities of 4.0 / 4.1 -> physical tag of each surface (entity tag when none)
std::map<long long, int> gmsh4_surface_tags(const std::string& s, bool v41) {
  std::map<long long, int> out;
  std::string body = section(s, "Entities", false);
  if (body.empty()) return out;
  Tok t(body);
  long long cnt[4];
  for (auto& c : cnt) c = t.integer();
  for (int dim = 0; dim < 4; ++dim) {
    for (long long i = 0; i < cnt[dim]; ++i) {
      long long tag = t.integer();
      const int nbox = (dim == 0 && v41) ? 3 : 6;
      for (int k = 0; k < nbox; ++k) t.real();
      long long nphys = t.integer();
      long long phys0 = tag;
      for (long long k = 0; k < nphys; ++k) {
        long long v = t.integer();
        if (k == 0) phys0 = v;
      }
      if (dim > 0) {
        long long nb = t.integer();
        for (long long k = 0; k < nb; ++k) t.integer();
      }
      if (dim == 2) out[tag] = (int)(phys0 < 0 ? -phys0 : phys0);
    }
  }
  return out;
}

void read_gmsh4(const std::string& s, Raw& r, bool v41) {
  auto stag = gmsh4_surface_tags(s, v41);
  NodeMap nm;
  {
    Tok t(section(s, "Nodes", true));
    long long nblk = t.integer();
    long long nn = t.integer();
    if (v41) {
      t.integer();
      t.integer();
    }
    int pos = 0;
    for (long long b = 0; b < nblk; ++b) {
      long long a = t.integer(), c = t.integer();
      (void)a;
      (void)c;
      long long param = t.integer();
      long long cnt = t.integer();
      if (param) throw Fail("GMSH 4 parametric nodes are not supported");
      if (v41) {
        std::vector<long long> tags(cnt);
        for (auto& v : tags) v = t.integer();
        for (long long i = 0; i < cnt; ++i) {
          nm.id[tags[i]] = pos++;
          for (int d = 0; d < 3; ++d) r.xyz.push_back(t.real());
        }
      } else {
        for (long long i = 0; i < cnt; ++i) {
          nm.id[t.integer()] = pos++;
          for (int d = 0; d < 3; ++d) r.xyz.push_back(t.real());
        }
      }
    }
    if (pos != nn) throw Fail("GMSH 4 node count does not match its header");
  }
  Tok t(section(s, "Elements", true));
  long long nblk = t.integer();
  t.integer();  // number of elements
  if (v41) {
    t.integer();
    t.integer();
  }
  for (long long b = 0; b < nblk; ++b) {
    long long etag, edim;
    if (v41) {
      edim = t.integer();
      etag = t.integer();
    } else {
      etag = t.integer();
      edim = t.integer();
    }
    long long code = t.integer();
    long long cnt = t.integer();
    int ty = gmsh_type(code);
    int k = ty < 0 ? (code == 1 ? 2 : 1) : nodes_of(ty);
    int tag = (int)etag;
    if (edim == 2) {
      auto it = stag.find(etag);
      if (it != stag.end()) tag = it->second;
    }
    for (long long i = 0; i < cnt; ++i) {
      t.integer();  // element tag
      long long n[8];
      for (int j = 0; j < k; ++j) n[j] = t.integer();
      if (ty >= 0) gmsh_add(r, nm, ty, n, tag);
    }
  }
}

void read_gmsh(const std::string& s, Raw& r) {
  Tok t(section(s, "MeshFormat", true));
  std::string ver = t.word();
  long long ftype = t.integer();
  if (ftype != 0) throw Fail("binary GMSH files are not supported (file-type " + std::to_string(ftype) + ")");
  if (ver[0] == '2') read_gmsh2(s, r);
  else if (ver == "4" || ver == "4.0") read_gmsh4(s, r, false);
  else if (ver == "4.1") read_gmsh4(s, r, true);
  else throw Fail("GMSH format version " + ver + " is not supported");
}

// ---------------------------------------------------------------------------
// UGRID: header, coordinates, tri and quad boundary faces, their tags, then
// tets, pyramids, prisms, hexes.  Pyramids are listed with the apex third:
// local (base 0-3, apex 4) = file (0, 3, 4, 1, 2) (ucs/mesh.tcc:6753-6760).
void read_ugrid(const std::string& s, Raw& r) {
  Tok t(s);
  long long c[7];
  for (auto& v : c) v = t.integer();
  const long long nn = c[0];
  for (long long i = 0; i < 3 * nn; ++i) r.xyz.push_back(t.real());
  auto ids = [&](int* n, int k) {
    for (int j = 0; j < k; ++j) n[j] = (int)t.integer() - 1;
  };
  std::vector<int> fn;
  for (long long i = 0; i < c[1]; ++i) {
    int n[3];
    ids(n, 3);
    fn.insert(fn.end(), n, n + 3);
  }
  for (long long i = 0; i < c[2]; ++i) {
    int n[4];
    ids(n, 4);
    fn.insert(fn.end(), n, n + 4);
  }
  std::vector<long long> tag(c[1] + c[2]);
  bool shift = false;
  for (auto& v : tag) {
    v = t.integer();
    if (v < 0) v = -v;
    if (v == 0) shift = true;
  }
  size_t off = 0;
  for (long long i = 0; i < c[1] + c[2]; ++i) {
    const int ty = i < c[1] ? PCFD_TRI : PCFD_QUAD, k = nodes_of(ty);
    r.face(ty, &fn[off], k, (int)(tag[i] + (shift ? 1 : 0)));
    off += k;
  }
  static const int pyr[5] = {0, 3, 4, 1, 2};
  const int types[4] = {PCFD_TET, PCFD_PYRAMID, PCFD_PRISM, PCFD_HEX};
  for (int g = 0; g < 4; ++g) {
    for (long long i = 0; i < c[3 + g]; ++i) {
      int n[8], k = nodes_of(types[g]);
      ids(n, k);
      if (types[g] == PCFD_PYRAMID) permute(n, pyr, 5);
      r.elem(types[g], n, k);
    }
  }
}

// ---------------------------------------------------------------------------
// FieldView unstructured ASCII ("CRUNCH"): "Boundary Table", "Nodes",
// "Boundary Faces" (tag, count, nodes), "Elements" (type, subtype, nodes)
// with types 1 tet, 2 hex, 3 prism, 4 pyramid.  Local numbering from the
// file's (ucs/mesh.tcc:6977-6985): tet (0,1,3,2), prism (0,3,5,1,2,4),
// hex (0,1,3,2,4,5,7,6), pyramid as is.
void read_fieldview(const std::string& s, Raw& r) {
  auto L = lines_of(s);
  size_t i = 0;
  auto find = [&](const std::string& key) {
    for (; i < L.size(); ++i) {
      std::string l = lower(L[i]);
      if (l.compare(0, key.size(), key) == 0) return;
    }
    throw Fail("FieldView file has no '" + key + "' header");
  };
  auto count_after = [&](const std::string& key) {
    std::istringstream ss(L[i].substr(key.size()));
    long long v = -1;
    ss >> v;
    if (v < 0) {
      ++i;
      if (i >= L.size()) throw Fail("FieldView '" + key + "' count missing");
      v = std::atoll(L[i].c_str());
    }
    ++i;
    return v;
  };
  find("nodes");
  long long nn = count_after("nodes");
  for (long long k = 0; k < nn; ++k, ++i) {
    if (i >= L.size()) throw Fail("FieldView node list is short");
    std::istringstream ss(L[i]);
    double x, y, z;
    if (!(ss >> x >> y >> z)) throw Fail("FieldView node line " + std::to_string(i + 1) + " is malformed");
    r.xyz.push_back(x);
    r.xyz.push_back(y);
    r.xyz.push_back(z);
  }
  find("boundary faces");
  long long nf = count_after("boundary faces");
  bool shift = false;
  std::vector<std::vector<long long>> faces;
  for (long long k = 0; k < nf; ++k, ++i) {
    if (i >= L.size()) throw Fail("FieldView boundary-face list is short");
    std::istringstream ss(L[i]);
    long long tag, np;
    ss >> tag >> np;
    if (np != 3 && np != 4) throw Fail("FieldView boundary face with " + std::to_string(np) + " nodes");
    std::vector<long long> f{tag < 0 ? -tag : tag, np};
    for (long long j = 0; j < np; ++j) {
      long long v;
      if (!(ss >> v)) throw Fail("FieldView boundary face line is malformed");
      f.push_back(v - 1);
    }
    if (f[0] == 0) shift = true;
    faces.push_back(f);
  }
  for (auto& f : faces) {
    int n[4];
    for (int j = 0; j < (int)f[1]; ++j) n[j] = (int)f[2 + j];
    r.face(f[1] == 3 ? PCFD_TRI : PCFD_QUAD, n, (int)f[1], (int)(f[0] + (shift ? 1 : 0)));
  }
  find("elements");
  ++i;
  static const int tet[4] = {0, 1, 3, 2}, pri[6] = {0, 3, 5, 1, 2, 4}, hex[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  for (; i < L.size(); ++i) {
    if (L[i].empty() || !std::isdigit((unsigned char)L[i][0])) break;
    std::istringstream ss(L[i]);
    long long code, sub;
    ss >> code >> sub;
    int ty, k;
    const int* map = nullptr;
    switch (code) {
      case 1: ty = PCFD_TET; map = tet; break;
      case 2: ty = PCFD_HEX; map = hex; break;
      case 3: ty = PCFD_PRISM; map = pri; break;
      case 4: ty = PCFD_PYRAMID; break;
      default: throw Fail("FieldView element type " + std::to_string(code) + " is not supported");
    }
    k = nodes_of(ty);
    int n[8];
    for (int j = 0; j < k; ++j) {
      long long v;
      if (!(ss >> v)) throw Fail("FieldView element line " + std::to_string(i + 1) + " is malformed");
      n[j] = (int)v - 1;
    }
    if (map) permute(n, map, k);
    r.elem(ty, n, k);
  }
}

// ---------------------------------------------------------------------------
// SU2 ASCII: NDIME= 3, NELEM= with (VTK type, nodes, index) lines, NPOIN=
// with (x, y, z[, index]) lines, NMARK= sections of MARKER_TAG= /
// MARKER_ELEMS=.  VTK types 5 tri, 9 quad, 10 tet, 12 hex, 13 prism,
// 14 pyramid use the local node order as they are.
int su2_type(long long code) {
  switch (code) {
    case 5: return PCFD_TRI;
    case 9: return PCFD_QUAD;
    case 10: return PCFD_TET;
    case 12: return PCFD_HEX;
    case 13: return PCFD_PRISM;
    case 14: return PCFD_PYRAMID;
  }
  throw Fail("SU2 element type " + std::to_string(code) + " is not supported");
}

void read_su2(const std::string& s, Raw& r) {
  auto L = lines_of(s);
  size_t i = 0;
  auto next = [&]() -> std::string& {
    while (i < L.size() && (L[i].empty() || L[i][0] == '%')) ++i;
    if (i >= L.size()) throw Fail("SU2 file ends early");
    return L[i++];
  };
  auto keyed = [&](const std::string& key) {
    std::string& l = next();
    size_t p = l.find(key);
    if (p == std::string::npos) throw Fail("SU2 file: expected '" + key + "' on line " + std::to_string(i));
    return l.substr(p + key.size());
  };
  {
    std::istringstream ss(keyed("NDIME="));
    int nd = 0;
    ss >> nd;
    if (nd != 3) throw Fail("SU2 file: only 3-D meshes are supported");
  }
  long long ne = std::atoll(keyed("NELEM=").c_str());
  for (long long k = 0; k < ne; ++k) {
    std::istringstream ss(next());
    long long code;
    ss >> code;
    int ty = su2_type(code), n[8], kk = nodes_of(ty);
    for (int j = 0; j < kk; ++j) {
      long long v;
      if (!(ss >> v)) throw Fail("SU2 element line is malformed");
      n[j] = (int)v;
    }
    if (ty == PCFD_TRI || ty == PCFD_QUAD) throw Fail("SU2 file: surface element in the volume list");
    r.elem(ty, n, kk);
  }
  long long np = std::atoll(keyed("NPOIN=").c_str());
  for (long long k = 0; k < np; ++k) {
    std::istringstream ss(next());
    double x, y, z;
    if (!(ss >> x >> y >> z)) throw Fail("SU2 point line is malformed");
    r.xyz.push_back(x);
    r.xyz.push_back(y);
    r.xyz.push_back(z);
  }
  long long nm = std::atoll(keyed("NMARK=").c_str());
  for (long long m = 0; m < nm; ++m) {
    keyed("MARKER_TAG=");
    long long nf = std::atoll(keyed("MARKER_ELEMS=").c_str());
    for (long long k = 0; k < nf; ++k) {
      std::istringstream ss(next());
      long long code;
      ss >> code;
      int ty = su2_type(code), n[4], kk = nodes_of(ty);
      if (ty != PCFD_TRI && ty != PCFD_QUAD) throw Fail("SU2 marker element is not a tri/quad");
      for (int j = 0; j < kk; ++j) {
        long long v;
        if (!(ss >> v)) throw Fail("SU2 marker line is malformed");
        n[j] = (int)v;
      }
      r.face(ty, n, kk, (int)(m + 1));
    }
  }
}

// EliminateOrphanedNodes: keep the nodes that volume elements use, in order
void drop_orphans(Raw& r) {
  const int nn = (int)(r.xyz.size() / 3);
  for (int v : r.en)
    if (v < 0 || v >= nn) throw Fail("element node id out of range");
  for (int v : r.bn)
    if (v < 0 || v >= nn) throw Fail("boundary face node id out of range");
  std::vector<int> used(nn, 0);
  for (int v : r.en) used[v] = 1;
  std::vector<int> nid(nn, -1);
  int k = 0;
  for (int i = 0; i < nn; ++i)
    if (used[i]) nid[i] = k++;
  if (k == nn) return;
  std::vector<double> x2;
  x2.reserve(3 * (size_t)k);
  for (int i = 0; i < nn; ++i)
    if (used[i]) x2.insert(x2.end(), &r.xyz[3 * (size_t)i], &r.xyz[3 * (size_t)i] + 3);
  r.xyz.swap(x2);
  for (int& v : r.en) v = nid[v];
  for (int& v : r.bn) {
    if (nid[v] < 0) throw Fail("a boundary face uses a node that no volume element uses");
    v = nid[v];
  }
}

bool ends_with(const std::string& s, const std::string& e) {
  return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

}  // namespace

extern "C" int pcfd_mesh_read(const char* path, int format, int rcm, pcfd_mesh** out) {
  if (!path || !out) {
    pcfd_mesh_set_error("pcfd_mesh_read: null argument");
    return -1;
  }
  *out = nullptr;
  const std::string p(path), lp = lower(p);
  int fmt = format;
  if (fmt == PCFD_MESH_FMT_AUTO) {
    if (ends_with(lp, ".msh")) fmt = PCFD_MESH_FMT_GMSH;
    else if (ends_with(lp, ".ugrid")) fmt = PCFD_MESH_FMT_UGRID;
    else if (ends_with(lp, ".crunch") || ends_with(lp, ".fv") || ends_with(lp, ".uns")) fmt = PCFD_MESH_FMT_FIELDVIEW;
    else if (ends_with(lp, ".su2")) fmt = PCFD_MESH_FMT_SU2;
    else {
      pcfd_mesh_set_error("pcfd_mesh_read: cannot tell the format of " + p + " from its extension");
      return -1;
    }
  }
  Raw r;
  try {
    const std::string s = slurp(p);
    switch (fmt) {
      case PCFD_MESH_FMT_GMSH: read_gmsh(s, r); break;
      case PCFD_MESH_FMT_UGRID:
        if (ends_with(lp, ".b8.ugrid") || ends_with(lp, ".lb8.ugrid") || ends_with(lp, ".r8.ugrid"))
          throw Fail("binary UGRID files are not supported");
        read_ugrid(s, r);
        break;
      case PCFD_MESH_FMT_FIELDVIEW: read_fieldview(s, r); break;
      case PCFD_MESH_FMT_SU2: read_su2(s, r); break;
      default: throw Fail("unknown format id " + std::to_string(fmt));
    }
    if (r.et.empty()) throw Fail("no volume elements");
    drop_orphans(r);
  } catch (const std::exception& e) {
    pcfd_mesh_set_error(std::string("pcfd_mesh_read: ") + p + ": " + e.what());
    return -1;
  }
  const int nn = (int)(r.xyz.size() / 3);
  return pcfd_mesh_build_from(nn, r.xyz.data(), (int)r.et.size(), r.et.data(), r.ep.data(), r.en.data(),
                              (int)r.bt.size(), r.bt.data(), r.bp.data(), r.bn.data(), r.bg.data(), rcm, out);
}
