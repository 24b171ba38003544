// Host-side synthetic mesh generator and median-dual metrics (C ABI in
// include/pcfd_mesh.h).  Plain C++17, built with g++; no GPU code.
//
// Reference behaviour followed (ucs/ = reference solver sources):
//  * edges n[0]<n[1], one interior edge per node pair, numbered in the
//    order of the point-to-point map                      ucs/mesh.tcc:517-660,1077-1091
//  * boundary half-edges: one per (boundary face, node), phantom id nnode+k
//                                                          ucs/mesh.tcc:1102-1133
//  * dual facet of edge (p,q) inside an element: the loop edge-midpoint,
//    face centroids of the two faces sharing the edge, element centroid
//                                                          ucs/mesh.tcc:1859-1975
//  * CV volume = cones from the node over its dual facets ucs/mesh.tcc:1936-1962
//  * boundary half-edge area: the node's share of the boundary face
//                                                          ucs/mesh.tcc:2140-2185
//  * closure check, then a[3]=|a| and a[0..2] normalised  ucs/mesh.tcc:2273-2333
//  * reverse Cuthill-McKee numbering                      ucs/mesh.tcc:2411-2535
#include "../../include/pcfd_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct Vec3 {
  double x, y, z;
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(double s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double det3(Vec3 a, Vec3 b, Vec3 c) { return dot(a, cross(b, c)); }

// counter-based RNG: splitmix64 finaliser of (seed, counter)
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double urand_pm1(uint64_t seed, uint64_t ctr) {
  uint64_t h = mix64(mix64(seed) ^ (ctr * 0xD1B54A32D192ED03ull));
  double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
  return 2.0 * u - 1.0;
}

// Local topology tables.  Faces are wound outward for a positively oriented
// element (tet: det[x1-x0,x2-x0,x3-x0]>0; prism: bottom 0,1,2 counter-clockwise
// seen from the top face 3,4,5, node k+3 above node k).
struct ElemTopo {
  int nnode, nface, nedge;
  int face_n[6];
  int face[6][4];
  int edge[12][2];
};
const ElemTopo kTet = {4, 4, 6, {3, 3, 3, 3, 0, 0},
                       {{0, 2, 1, 0}, {0, 1, 3, 0}, {0, 3, 2, 0}, {1, 2, 3, 0}},
                       {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}}};
const ElemTopo kPrism = {6, 5, 9, {3, 3, 4, 4, 4, 0},
                         {{0, 2, 1, 0}, {3, 4, 5, 0}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}},
                         {{0, 1}, {1, 2}, {2, 0}, {3, 4}, {4, 5}, {5, 3}, {0, 3}, {1, 4}, {2, 5}}};
// pyramid: base 0,1,2,3 counter-clockwise seen from the apex 4
const ElemTopo kPyramid = {5, 5, 8, {4, 3, 3, 3, 3, 0},
                           {{0, 3, 2, 1}, {0, 1, 4, 0}, {1, 2, 4, 0}, {2, 3, 4, 0}, {3, 0, 4, 0}},
                           {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 4}, {2, 4}, {3, 4}}};
// hexahedron: bottom 0,1,2,3 counter-clockwise seen from the top, node k+4 above k
const ElemTopo kHex = {8, 6, 12, {4, 4, 4, 4, 4, 4},
                       {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}},
                       {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}}};

struct Work {
  int kind = 0;
  std::vector<double> xyz;
  std::vector<int> etype, eptr, enodes;
  std::vector<int> btype, bptr, bnodes, btag;
};

// ----------------------------------------------------------------------------
// structured generators
// ----------------------------------------------------------------------------
struct Grid {
  int nx, ny, nz;
  int id(int i, int j, int k) const { return (k * (ny + 1) + j) * (nx + 1) + i; }
};

void add_elem(Work& w, int type, const int* n, int nn) {
  w.etype.push_back(type);
  w.enodes.insert(w.enodes.end(), n, n + nn);
  w.eptr.push_back((int)w.enodes.size());
}
void add_bface(Work& w, int type, const int* n, int nn, int tag) {
  w.btype.push_back(type);
  w.bnodes.insert(w.bnodes.end(), n, n + nn);
  w.bptr.push_back((int)w.bnodes.size());
  w.btag.push_back(tag);
}

Vec3 P(const Work& w, int i) { return {w.xyz[3 * i], w.xyz[3 * i + 1], w.xyz[3 * i + 2]}; }

// Kuhn split of hex (i,j,k): 6 tets along the (0,0,0)-(1,1,1) diagonal.  The
// permutation parity fixes the orientation independent of node jitter.
void kuhn_hex(Work& w, const Grid& g, int i, int j, int k) {
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const int parity[6] = {+1, -1, -1, +1, +1, -1};
  for (int p = 0; p < 6; ++p) {
    int c[3] = {0, 0, 0};
    int n[4];
    n[0] = g.id(i, j, k);
    for (int s = 0; s < 3; ++s) {
      c[perms[p][s]] = 1;
      n[s + 1] = g.id(i + c[0], j + c[1], k + c[2]);
    }
    // identity permutation (x,y,z) gives det>0 on an undistorted grid
    if (parity[p] < 0) std::swap(n[2], n[3]);
    add_elem(w, PCFD_TET, n, 4);
  }
}

// Two prisms of hex (i,j,k): the xy quad is split along (i,j)-(i+1,j+1), the
// same diagonal the Kuhn tets use on z-faces, so layers and core conform.
void prism_hex(Work& w, const Grid& g, int i, int j, int k) {
  int a = g.id(i, j, k), b = g.id(i + 1, j, k), c = g.id(i + 1, j + 1, k), d = g.id(i, j + 1, k);
  int A = g.id(i, j, k + 1), B = g.id(i + 1, j, k + 1), C = g.id(i + 1, j + 1, k + 1),
      D = g.id(i, j + 1, k + 1);
  int p1[6] = {a, b, c, A, B, C};
  int p2[6] = {a, c, d, A, C, D};
  add_elem(w, PCFD_PRISM, p1, 6);
  add_elem(w, PCFD_PRISM, p2, 6);
}

// boundary quad (low corner L, corners in cyclic order c0..c3 with c0=L and
// c2=H) -> two triangles sharing the L-H diagonal, or one quad; oriented so
// that the normal points along `out`.
void bquad(Work& w, int c0, int c1, int c2, int c3, Vec3 out, int tag, bool as_quad) {
  Vec3 n = cross(P(w, c2) - P(w, c0), P(w, c3) - P(w, c1));
  bool flip = dot(n, out) < 0.0;
  if (as_quad) {
    int q[4] = {c0, c1, c2, c3};
    if (flip) std::swap(q[1], q[3]);
    add_bface(w, PCFD_QUAD, q, 4, tag);
  } else {
    int t1[3] = {c0, c1, c2}, t2[3] = {c0, c2, c3};
    if (flip) {
      std::swap(t1[1], t1[2]);
      std::swap(t2[1], t2[2]);
    }
    add_bface(w, PCFD_TRI, t1, 3, tag);
    add_bface(w, PCFD_TRI, t2, 3, tag);
  }
}

int generate(const pcfd_mesh_spec& s, Work& w) {
  if (s.nx < 1 || s.ny < 1 || s.nz < 1) {
    g_err = "pcfd_mesh_build: nx, ny, nz must be >= 1";
    return -1;
  }
  Grid g{s.nx, s.ny, s.nz};
  const int nn = (s.nx + 1) * (s.ny + 1) * (s.nz + 1);
  w.kind = s.kind;
  w.xyz.assign(3 * (size_t)nn, 0.0);
  const double pi = 3.141592653589793;
  // vertical coordinate distribution for the plate mesh
  std::vector<double> zc(s.nz + 1, 0.0);
  int nlay = 0;
  if (s.kind == PCFD_MESH_PLATE) {
    nlay = std::max(0, std::min(s.nlayers, s.nz - 1));
    double z = 0.0, h = s.first_h;
    for (int k = 0; k <= nlay; ++k) {
      zc[k] = z;
      z += h;
      h *= s.growth;
    }
    double ztop = zc[nlay];
    if (ztop >= s.lz) {
      g_err = "pcfd_mesh_build: prism layers exceed the domain height";
      return -1;
    }
    for (int k = nlay + 1; k <= s.nz; ++k)
      zc[k] = ztop + (s.lz - ztop) * (double)(k - nlay) / (double)(s.nz - nlay);
  }
  for (int k = 0; k <= s.nz; ++k)
    for (int j = 0; j <= s.ny; ++j)
      for (int i = 0; i <= s.nx; ++i) {
        int id = g.id(i, j, k);
        double xi = (double)i / s.nx, eta = (double)j / s.ny, zeta = (double)k / s.nz;
        bool interior = i > 0 && i < s.nx && j > 0 && j < s.ny && k > 0 && k < s.nz;
        if (interior && s.jitter > 0.0 && !(s.kind == PCFD_MESH_PLATE && k <= nlay + 1)) {
          uint64_t c = 3ull * (uint64_t)id;
          xi += s.jitter * urand_pm1(s.seed, c + 0) / s.nx;
          eta += s.jitter * urand_pm1(s.seed, c + 1) / s.ny;
          zeta += s.jitter * urand_pm1(s.seed, c + 2) / s.nz;
        }
        double x = 0, y = 0, z = 0;
        if (s.kind == PCFD_MESH_WEDGE) {
          x = s.lx * xi;
          double yw = x > s.x0 ? std::tan(s.angle_deg * pi / 180.0) * (x - s.x0) : 0.0;
          y = yw + eta * (s.ly - yw);
          z = s.lz * zeta;
        } else if (s.kind == PCFD_MESH_BUMP) {
          x = s.lx * xi;
          y = s.ly * eta;
          double sx = std::sin(pi * x / s.lx), sy = std::sin(pi * y / s.ly);
          double zf = s.bump_h * sx * sx * sy * sy;
          z = zf + zeta * (s.lz - zf);
        } else if (s.kind == PCFD_MESH_BOX) {
          x = s.lx * xi;
          y = s.ly * eta;
          z = s.lz * zeta;
        } else if (s.kind == PCFD_MESH_PLATE) {
          x = s.lx * xi;
          y = s.ly * eta;
          if (interior && s.jitter > 0.0 && k > nlay + 1) {
            z = zc[k] + s.jitter * urand_pm1(s.seed, 3ull * id + 2) * (zc[k + 1] - zc[k]);
          } else {
            z = zc[k];
          }
        } else {
          g_err = "pcfd_mesh_build: unknown mesh kind";
          return -1;
        }
        w.xyz[3 * (size_t)id + 0] = x;
        w.xyz[3 * (size_t)id + 1] = y;
        w.xyz[3 * (size_t)id + 2] = z;
      }
  w.eptr.assign(1, 0);
  w.bptr.assign(1, 0);
  for (int k = 0; k < s.nz; ++k)
    for (int j = 0; j < s.ny; ++j)
      for (int i = 0; i < s.nx; ++i) {
        if (k < nlay)
          prism_hex(w, g, i, j, k);
        else
          kuhn_hex(w, g, i, j, k);
      }
  // boundary faces, grouped by tag (ascending) so inflow (tag 1) comes first
  auto quad_is_prism = [&](int k) { return k < nlay; };
  for (int tag = 1; tag <= 6; ++tag) {
    if (tag == 1 || tag == 2) {
      int i = (tag == 1) ? 0 : s.nx;
      Vec3 out = {tag == 1 ? -1.0 : 1.0, 0, 0};
      for (int k = 0; k < s.nz; ++k)
        for (int j = 0; j < s.ny; ++j)
          bquad(w, g.id(i, j, k), g.id(i, j + 1, k), g.id(i, j + 1, k + 1), g.id(i, j, k + 1), out,
                tag, quad_is_prism(k));
    } else if (tag == 3 || tag == 4) {
      int j = (tag == 3) ? 0 : s.ny;
      Vec3 out = {0, tag == 3 ? -1.0 : 1.0, 0};
      for (int k = 0; k < s.nz; ++k)
        for (int i = 0; i < s.nx; ++i)
          bquad(w, g.id(i, j, k), g.id(i + 1, j, k), g.id(i + 1, j, k + 1), g.id(i, j, k + 1), out,
                tag, quad_is_prism(k));
    } else {
      int k = (tag == 5) ? 0 : s.nz;
      Vec3 out = {0, 0, tag == 5 ? -1.0 : 1.0};
      for (int j = 0; j < s.ny; ++j)
        for (int i = 0; i < s.nx; ++i)
          bquad(w, g.id(i, j, k), g.id(i + 1, j, k), g.id(i + 1, j + 1, k), g.id(i, j + 1, k), out,
                tag, false);
    }
  }
  return 0;
}

// ----------------------------------------------------------------------------
// graph utilities
// ----------------------------------------------------------------------------
const ElemTopo& topo(int t) {
  switch (t) {
    case PCFD_PRISM: return kPrism;
    case PCFD_PYRAMID: return kPyramid;
    case PCFD_HEX: return kHex;
    default: return kTet;
  }
}

// element volume through its outward faces (divergence theorem, apex at the centroid)
double elem_volume(const Work& w, int type, const int* n) {
  const ElemTopo& T = topo(type);
  Vec3 x[8];
  Vec3 g = {0, 0, 0};
  for (int k = 0; k < T.nnode; ++k) {
    x[k] = P(w, n[k]);
    g = g + x[k];
  }
  g = (1.0 / T.nnode) * g;
  double ev = 0.0;
  for (int f = 0; f < T.nface; ++f) {
    int nf = T.face_n[f];
    Vec3 c = {0, 0, 0};
    for (int k = 0; k < nf; ++k) c = c + x[T.face[f][k]];
    c = (1.0 / nf) * c;
    for (int k = 0; k < nf; ++k) ev += det3(x[T.face[f][k]] - g, x[T.face[f][(k + 1) % nf]] - g, c - g) / 6.0;
  }
  return ev;
}

// reflect an element's local numbering (reverses its orientation)
void reflect(int type, int* n) {
  switch (type) {
    case PCFD_TET: std::swap(n[1], n[2]); break;
    case PCFD_PYRAMID: std::swap(n[1], n[3]); break;
    case PCFD_PRISM: std::swap(n[1], n[2]); std::swap(n[4], n[5]); break;
    case PCFD_HEX: std::swap(n[1], n[3]); std::swap(n[5], n[7]); break;
    default: break;
  }
}

// sorted unique edge keys (n0<<32 | n1) from the element edges
std::vector<uint64_t> edge_keys(const Work& w) {
  std::vector<uint64_t> keys;
  size_t ne = w.etype.size();
  size_t guess = 0;
  for (size_t e = 0; e < ne; ++e) guess += topo(w.etype[e]).nedge;
  keys.reserve(guess);
  for (size_t e = 0; e < ne; ++e) {
    const ElemTopo& T = topo(w.etype[e]);
    const int* n = &w.enodes[w.eptr[e]];
    for (int k = 0; k < T.nedge; ++k) {
      uint32_t a = n[T.edge[k][0]], b = n[T.edge[k][1]];
      if (a > b) std::swap(a, b);
      keys.push_back(((uint64_t)a << 32) | b);
    }
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  return keys;
}

// Neighbour order of each local node of an element, in the order the
// reference's point-to-point map visits them (the connection table of
// Mesh::BuildPsp, ucs/mesh.tcc:538-545), indexed by element type
// (PCFD_TRI .. PCFD_HEX) and local node; -1 ends a list.
const int kPspOrder[6][8][5] = {
    {{1, 2, -1}, {2, 0, -1}, {0, 1, -1}},
    {{1, 3, -1}, {2, 0, -1}, {3, 1, -1}, {0, 2, -1}},
    {{1, 2, 3, -1}, {2, 0, 3, -1}, {0, 1, 3, -1}, {0, 1, 2, -1}},
    {{1, 3, 4, -1}, {2, 0, 4, -1}, {3, 1, 4, -1}, {0, 2, 4, -1}, {0, 1, 2, 3, -1}},
    {{1, 2, 3, -1}, {2, 0, 4, -1}, {0, 1, 5, -1}, {4, 5, 0, -1}, {5, 3, 1, -1}, {3, 4, 2, -1}},
    {{1, 3, 4, -1}, {2, 0, 5, -1}, {3, 1, 6, -1}, {0, 2, 7, -1}, {5, 7, 0, -1}, {6, 4, 1, -1}, {7, 5, 2, -1},
     {4, 6, 3, -1}}};

// Edges in the order ucs.x numbers them (Mesh::BuildPsp + BuildEdges,
// ucs/mesh.tcc:517-660,1077-1091): node i's neighbours come from the
// elements holding i in element-list order -- boundary faces first, then
// volume elements, the order a partition file stores them -- each adding the
// neighbours of i's local node in the table's order, first occurrence kept;
// then edge (i, j) for every neighbour j > i, in that order, i ascending.
// Returns the edges as (n0 << 32 | n1) keys in that order.
std::vector<uint64_t> edges_in_psp_order(const Work& w, int nn) {
  const size_t nbf = w.btype.size(), ne = w.etype.size();
  std::vector<int> cnt(nn + 1, 0);
  auto nodes_of = [&](size_t el, const int*& n, int& type) -> int {
    if (el < nbf) {
      n = &w.bnodes[w.bptr[el]];
      type = w.btype[el];
      return w.bptr[el + 1] - w.bptr[el];
    }
    const size_t e = el - nbf;
    n = &w.enodes[w.eptr[e]];
    type = w.etype[e];
    return w.eptr[e + 1] - w.eptr[e];
  };
  for (size_t el = 0; el < nbf + ne; ++el) {
    const int* n;
    int t;
    const int k = nodes_of(el, n, t);
    for (int j = 0; j < k; ++j) cnt[n[j] + 1]++;
  }
  for (int i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> elsp(cnt[nn]), fill(cnt.begin(), cnt.end() - 1);
  for (size_t el = 0; el < nbf + ne; ++el) {
    const int* n;
    int t;
    const int k = nodes_of(el, n, t);
    for (int j = 0; j < k; ++j) elsp[fill[n[j]]++] = (int)el;
  }
  std::vector<uint64_t> out;
  std::vector<int> stamp(nn, -1);
  for (int i = 0; i < nn; ++i) {
    for (int p = cnt[i]; p < cnt[i + 1]; ++p) {
      const int* n;
      int t;
      const int k = nodes_of((size_t)elsp[p], n, t);
      int ln = 0;
      while (ln < k && n[ln] != i) ++ln;
      for (const int* c = kPspOrder[t][ln]; *c >= 0; ++c) {
        const int j = n[*c];
        if (stamp[j] == i) continue;
        stamp[j] = i;
        if (j > i) out.push_back(((uint64_t)i << 32) | (uint32_t)j);
      }
    }
  }
  return out;
}

// reverse Cuthill-McKee: BFS from a pseudo-peripheral minimum-degree node,
// unvisited neighbours appended in ascending degree, then the order reversed.
std::vector<int> rcm_order(int nn, const std::vector<uint64_t>& keys) {
  std::vector<int> deg(nn, 0);
  for (uint64_t k : keys) {
    deg[k >> 32]++;
    deg[k & 0xffffffffu]++;
  }
  std::vector<int> ptr(nn + 1, 0);
  for (int i = 0; i < nn; ++i) ptr[i + 1] = ptr[i] + deg[i];
  std::vector<int> adj(ptr[nn]), fill(ptr.begin(), ptr.end() - 1);
  for (uint64_t k : keys) {
    int a = (int)(k >> 32), b = (int)(k & 0xffffffffu);
    adj[fill[a]++] = b;
    adj[fill[b]++] = a;
  }
  std::vector<int> order;
  order.reserve(nn);
  std::vector<int> level(nn, -1);
  auto bfs = [&](int start, std::vector<int>& out) -> int {
    // returns the last node visited (a far node); out gets the CM order
    out.clear();
    std::fill(level.begin(), level.end(), -1);
    out.push_back(start);
    level[start] = 0;
    size_t head = 0;
    std::vector<int> nb;
    while (head < out.size()) {
      int z = out[head++];
      nb.clear();
      for (int p = ptr[z]; p < ptr[z + 1]; ++p)
        if (level[adj[p]] < 0) {
          level[adj[p]] = level[z] + 1;
          nb.push_back(adj[p]);
        }
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
      out.insert(out.end(), nb.begin(), nb.end());
    }
    return out.back();
  };
  std::vector<int> seen(nn, 0);
  std::vector<int> comp;
  for (int s0 = 0; s0 < nn; ++s0) {
    if (seen[s0]) continue;
    // pseudo-peripheral start: two BFS sweeps
    int far = bfs(s0, comp);
    int far2 = bfs(far, comp);
    bfs(far2, comp);
    for (int v : comp) seen[v] = 1;
    order.insert(order.end(), comp.begin(), comp.end());
  }
  std::reverse(order.begin(), order.end());
  return order;  // order[new] = old
}

}  // namespace

struct pcfd_mesh {
  pcfd_mesh_data d;
  std::vector<double> xyz, vol, edge_a, bedge_a;
  std::vector<int> edges, bedges, bfactag, etype, eptr, enodes, btype, bptr, bnodes, btag, perm;
};

// used by the file readers (mesh_io.cpp)
void pcfd_mesh_set_error(const std::string& msg) { g_err = msg; }

extern "C" {

const char* pcfd_mesh_error(void) { return g_err.c_str(); }

void pcfd_mesh_default_spec(int kind, pcfd_mesh_spec* s) {
  std::memset(s, 0, sizeof(*s));
  s->kind = kind;
  s->seed = 1;
  s->rcm = 0;
  s->growth = 1.0;
  if (kind == PCFD_MESH_WEDGE) {
    // 2.5D supersonic wedge channel: [0,3]x[0,1]x[0,0.1], 15 deg ramp after x=0.5
    s->nx = 60; s->ny = 30; s->nz = 2;
    s->lx = 3.0; s->ly = 1.0; s->lz = 0.1;
    s->angle_deg = 15.0; s->x0 = 0.5; s->jitter = 0.05; s->seed = 1;
  } else if (kind == PCFD_MESH_BUMP) {
    // unit cube with a sin^2 bump of height 0.1 on the floor
    s->nx = s->ny = s->nz = 100;
    s->lx = s->ly = s->lz = 1.0;
    s->bump_h = 0.1; s->jitter = 0.03; s->seed = 2; s->rcm = 1;
  } else if (kind == PCFD_MESH_PLATE) {
    s->nx = 100; s->ny = 20; s->nz = 40;
    s->lx = 1.0; s->ly = 0.2; s->lz = 0.5;
    s->nlayers = 20; s->first_h = 1e-4; s->growth = 1.15; s->jitter = 0.0; s->seed = 3; s->rcm = 1;
  } else {
    s->nx = s->ny = s->nz = 4;
    s->lx = s->ly = s->lz = 1.0;
  }
}

static int finish(Work& w, int rcm, pcfd_mesh** out);

int pcfd_mesh_build(const pcfd_mesh_spec* spec, pcfd_mesh** out) {
  if (!spec || !out) {
    g_err = "pcfd_mesh_build: null argument";
    return -1;
  }
  *out = nullptr;
  Work w;
  if (generate(*spec, w) != 0) return -1;
  return finish(w, spec->rcm, out);
}

int pcfd_mesh_build_from(int nnode, const double* xyz, int nelem, const int* elem_type, const int* elem_ptr,
                         const int* elem_nodes, int nbface, const int* bface_type, const int* bface_ptr,
                         const int* bface_nodes, const int* bface_tag, int rcm, pcfd_mesh** out) {
  if (!out || nnode <= 0 || !xyz || nelem <= 0 || !elem_type || !elem_ptr || !elem_nodes || nbface < 0 ||
      (nbface > 0 && (!bface_type || !bface_ptr || !bface_nodes || !bface_tag))) {
    g_err = "pcfd_mesh_build_from: bad arguments";
    return -1;
  }
  *out = nullptr;
  Work w;
  w.xyz.assign(xyz, xyz + 3 * (size_t)nnode);
  w.eptr.assign(1, 0);
  w.bptr.assign(1, 0);
  for (int e = 0; e < nelem; ++e) {
    int t = elem_type[e];
    int nn = elem_ptr[e + 1] - elem_ptr[e];
    if ((t != PCFD_TET && t != PCFD_PYRAMID && t != PCFD_PRISM && t != PCFD_HEX) || nn != topo(t).nnode) {
      g_err = "pcfd_mesh_build_from: element " + std::to_string(e) + " has an unsupported type/size";
      return -1;
    }
    int n[8];
    for (int k = 0; k < nn; ++k) {
      n[k] = elem_nodes[elem_ptr[e] + k];
      if (n[k] < 0 || n[k] >= nnode) {
        g_err = "pcfd_mesh_build_from: element node out of range";
        return -1;
      }
    }
    if (elem_volume(w, t, n) < 0.0) reflect(t, n);
    add_elem(w, t, n, nn);
  }
  // element faces keyed by their sorted node set -> owning element, to wind
  // the boundary faces outward (FixBoundaryWindings, ucs/mesh.tcc:2336-2410)
  std::vector<std::pair<std::vector<int>, int>> fmap;
  for (int e = 0; e < nelem; ++e) {
    const ElemTopo& T = topo(w.etype[e]);
    const int* n = &w.enodes[w.eptr[e]];
    for (int f = 0; f < T.nface; ++f) {
      std::vector<int> key;
      for (int k = 0; k < T.face_n[f]; ++k) key.push_back(n[T.face[f][k]]);
      std::sort(key.begin(), key.end());
      fmap.push_back({key, e});
    }
  }
  std::sort(fmap.begin(), fmap.end());
  for (int f = 0; f < nbface; ++f) {
    int t = bface_type[f];
    int nf = bface_ptr[f + 1] - bface_ptr[f];
    if ((t != PCFD_TRI || nf != 3) && (t != PCFD_QUAD || nf != 4)) {
      g_err = "pcfd_mesh_build_from: boundary face " + std::to_string(f) + " is not a tri/quad";
      return -1;
    }
    int n[4];
    for (int k = 0; k < nf; ++k) n[k] = bface_nodes[bface_ptr[f] + k];
    std::vector<int> key(n, n + nf);
    std::sort(key.begin(), key.end());
    auto it = std::lower_bound(fmap.begin(), fmap.end(), std::make_pair(key, -1));
    if (it == fmap.end() || it->first != key) {
      g_err = "pcfd_mesh_build_from: boundary face " + std::to_string(f) + " matches no element face";
      return -1;
    }
    int e = it->second;
    const ElemTopo& T = topo(w.etype[e]);
    Vec3 g = {0, 0, 0};
    for (int k = 0; k < T.nnode; ++k) g = g + P(w, w.enodes[w.eptr[e] + k]);
    g = (1.0 / T.nnode) * g;
    Vec3 c = {0, 0, 0};
    for (int k = 0; k < nf; ++k) c = c + P(w, n[k]);
    c = (1.0 / nf) * c;
    Vec3 a = {0, 0, 0};
    for (int k = 0; k < nf; ++k) a = a + cross(P(w, n[k]) - c, P(w, n[(k + 1) % nf]) - c);
    if (dot(a, c - g) < 0.0) std::reverse(n + 1, n + nf);
    add_bface(w, t, n, nf, bface_tag[f]);
  }
  return finish(w, rcm, out);
}

static int finish(Work& w, int rcm, pcfd_mesh** out) {
  const int nn = (int)(w.xyz.size() / 3);

  // optional RCM renumbering of nodes
  std::vector<int> perm(nn);
  std::iota(perm.begin(), perm.end(), 0);
  if (rcm) {
    std::vector<uint64_t> keys = edge_keys(w);
    perm = rcm_order(nn, keys);
    std::vector<int> inv(nn);
    for (int i = 0; i < nn; ++i) inv[perm[i]] = i;
    std::vector<double> x2(w.xyz.size());
    for (int i = 0; i < nn; ++i)
      for (int d = 0; d < 3; ++d) x2[3 * (size_t)i + d] = w.xyz[3 * (size_t)perm[i] + d];
    w.xyz.swap(x2);
    for (int& v : w.enodes) v = inv[v];
    for (int& v : w.bnodes) v = inv[v];
  }

  pcfd_mesh* m = new pcfd_mesh();
  const std::vector<uint64_t> order = edges_in_psp_order(w, nn);
  const size_t ned = order.size();
  m->edges.resize(2 * ned);
  for (size_t e = 0; e < ned; ++e) {
    m->edges[2 * e] = (int)(order[e] >> 32);
    m->edges[2 * e + 1] = (int)(order[e] & 0xffffffffu);
  }
  // sorted keys -> edge number, for the metric accumulation below
  std::vector<std::pair<uint64_t, int>> keys(ned);
  for (size_t e = 0; e < ned; ++e) keys[e] = {order[e], (int)e};
  std::sort(keys.begin(), keys.end());
  for (size_t e = 1; e < ned; ++e)
    if (keys[e].first == keys[e - 1].first) {
      g_err = "mesh: duplicate edge in the point-to-point map";
      delete m;
      return -1;
    }
  std::vector<double> A(3 * ned, 0.0);
  m->vol.assign(nn, 0.0);
  double elem_vol = 0.0;
  const size_t ne = w.etype.size();
  for (size_t e = 0; e < ne; ++e) {
    const ElemTopo& T = topo(w.etype[e]);
    const int* n = &w.enodes[w.eptr[e]];
    Vec3 x[8];
    Vec3 g = {0, 0, 0};
    for (int k = 0; k < T.nnode; ++k) {
      x[k] = P(w, n[k]);
      g = g + x[k];
    }
    g = (1.0 / T.nnode) * g;
    Vec3 fc[6];
    for (int f = 0; f < T.nface; ++f) {
      Vec3 c = {0, 0, 0};
      for (int k = 0; k < T.face_n[f]; ++k) c = c + x[T.face[f][k]];
      fc[f] = (1.0 / T.face_n[f]) * c;
    }
    // element volume from its outward faces (divergence theorem, centroid apex)
    double ev = 0.0;
    for (int f = 0; f < T.nface; ++f) {
      int nf = T.face_n[f];
      for (int k = 0; k < nf; ++k) {
        Vec3 a = x[T.face[f][k]], b = x[T.face[f][(k + 1) % nf]];
        ev += det3(a - g, b - g, fc[f] - g) / 6.0;
      }
    }
    elem_vol += ev;
    for (int k = 0; k < T.nedge; ++k) {
      int lp = T.edge[k][0], lq = T.edge[k][1];
      // f1: face whose outward winding contains lp->lq, f2: the one with lq->lp
      int f1 = -1, f2 = -1;
      for (int f = 0; f < T.nface; ++f) {
        int nf = T.face_n[f];
        for (int s = 0; s < nf; ++s) {
          int u = T.face[f][s], v = T.face[f][(s + 1) % nf];
          if (u == lp && v == lq) f1 = f;
          if (u == lq && v == lp) f2 = f;
        }
      }
      Vec3 xp = x[lp], xq = x[lq];
      Vec3 mid = 0.5 * (xp + xq);
      // dual facet loop mid -> f2 -> g -> f1, vector area points p -> q
      Vec3 Apq = 0.5 * cross(g - mid, fc[f1] - fc[f2]);
      double vp = (det3(mid - xp, fc[f2] - xp, g - xp) + det3(mid - xp, g - xp, fc[f1] - xp)) / 6.0;
      double vq = -(det3(mid - xq, fc[f2] - xq, g - xq) + det3(mid - xq, g - xq, fc[f1] - xq)) / 6.0;
      int gp = n[lp], gq = n[lq];
      m->vol[gp] += vp;
      m->vol[gq] += vq;
      uint32_t a = gp, b = gq;
      double sgn = 1.0;
      if (a > b) {
        std::swap(a, b);
        sgn = -1.0;
      }
      uint64_t key = ((uint64_t)a << 32) | b;
      const size_t idx = (size_t)std::lower_bound(keys.begin(), keys.end(), std::make_pair(key, -1))->second;
      A[3 * idx + 0] += sgn * Apq.x;
      A[3 * idx + 1] += sgn * Apq.y;
      A[3 * idx + 2] += sgn * Apq.z;
    }
  }
  // boundary half-edges, in boundary-face order, one per face node
  const size_t nbf = w.btype.size();
  size_t nhe = 0;
  for (size_t f = 0; f < nbf; ++f) nhe += w.bptr[f + 1] - w.bptr[f];
  m->bedges.resize(2 * nhe);
  m->bfactag.resize(nhe);
  std::vector<double> B(3 * nhe, 0.0);
  size_t h = 0;
  for (size_t f = 0; f < nbf; ++f) {
    int nf = w.bptr[f + 1] - w.bptr[f];
    const int* n = &w.bnodes[w.bptr[f]];
    Vec3 c = {0, 0, 0};
    for (int k = 0; k < nf; ++k) c = c + P(w, n[k]);
    c = (1.0 / nf) * c;
    for (int k = 0; k < nf; ++k) {
      Vec3 xp = P(w, n[k]);
      Vec3 mn = 0.5 * (xp + P(w, n[(k + 1) % nf]));
      Vec3 mp = 0.5 * (xp + P(w, n[(k + nf - 1) % nf]));
      // polygon p -> mid(p,next) -> centroid -> mid(p,prev), outward
      Vec3 a = 0.5 * cross(c - xp, mp - mn);
      m->bedges[2 * h] = n[k];
      m->bedges[2 * h + 1] = nn + (int)h;
      m->bfactag[h] = w.btag[f];
      B[3 * h] = a.x;
      B[3 * h + 1] = a.y;
      B[3 * h + 2] = a.z;
      ++h;
    }
  }
  // closure check: sum of outward dual areas of every CV
  std::vector<double> close(3 * (size_t)nn, 0.0);
  for (size_t e = 0; e < ned; ++e) {
    int a = m->edges[2 * e], b = m->edges[2 * e + 1];
    for (int d = 0; d < 3; ++d) {
      close[3 * (size_t)a + d] += A[3 * e + d];
      close[3 * (size_t)b + d] -= A[3 * e + d];
    }
  }
  for (size_t k = 0; k < nhe; ++k) {
    int a = m->bedges[2 * k];
    for (int d = 0; d < 3; ++d) close[3 * (size_t)a + d] += B[3 * k + d];
  }
  double cmax = 0.0;
  for (int i = 0; i < nn; ++i) {
    double s2 = close[3 * (size_t)i] * close[3 * (size_t)i] + close[3 * (size_t)i + 1] * close[3 * (size_t)i + 1] +
                close[3 * (size_t)i + 2] * close[3 * (size_t)i + 2];
    cmax = std::max(cmax, std::sqrt(s2));
  }
  // a[3] = |a|, a[0..2] = a/|a|
  m->edge_a.resize(4 * ned);
  for (size_t e = 0; e < ned; ++e) {
    double ax = A[3 * e], ay = A[3 * e + 1], az = A[3 * e + 2];
    double mag = std::sqrt(ax * ax + ay * ay + az * az);
    m->edge_a[4 * e] = ax / mag;
    m->edge_a[4 * e + 1] = ay / mag;
    m->edge_a[4 * e + 2] = az / mag;
    m->edge_a[4 * e + 3] = mag;
  }
  m->bedge_a.resize(4 * nhe);
  for (size_t k = 0; k < nhe; ++k) {
    double ax = B[3 * k], ay = B[3 * k + 1], az = B[3 * k + 2];
    double mag = std::sqrt(ax * ax + ay * ay + az * az);
    m->bedge_a[4 * k] = ax / mag;
    m->bedge_a[4 * k + 1] = ay / mag;
    m->bedge_a[4 * k + 2] = az / mag;
    m->bedge_a[4 * k + 3] = mag;
  }
  double tv = 0.0;
  for (int i = 0; i < nn; ++i) tv += m->vol[i];

  m->xyz.swap(w.xyz);
  m->etype.swap(w.etype);
  m->eptr.swap(w.eptr);
  m->enodes.swap(w.enodes);
  m->btype.swap(w.btype);
  m->bptr.swap(w.bptr);
  m->bnodes.swap(w.bnodes);
  m->btag.swap(w.btag);
  m->perm.swap(perm);

  pcfd_mesh_data& d = m->d;
  d.nnode = nn;
  d.nelem = (int)ne;
  d.nedge = (int)ned;
  d.nbface = (int)nbf;
  d.nbedge = (int)nhe;
  int maxtag = 0;
  for (int t : m->btag) maxtag = std::max(maxtag, t);
  d.nfactag = std::max(PCFD_NFACTAG, maxtag + 1);
  d.xyz = m->xyz.data();
  d.vol = m->vol.data();
  d.edges = m->edges.data();
  d.edge_a = m->edge_a.data();
  d.bedges = m->bedges.data();
  d.bedge_a = m->bedge_a.data();
  d.bfactag = m->bfactag.data();
  d.elem_type = m->etype.data();
  d.elem_ptr = m->eptr.data();
  d.elem_nodes = m->enodes.data();
  d.bface_type = m->btype.data();
  d.bface_ptr = m->bptr.data();
  d.bface_nodes = m->bnodes.data();
  d.bface_tag = m->btag.data();
  d.perm = m->perm.data();
  d.closure_max = cmax;
  d.total_volume = tv;
  d.elem_volume = elem_vol;
  *out = m;
  return 0;
}

const pcfd_mesh_data* pcfd_mesh_get(const pcfd_mesh* m) { return m ? &m->d : nullptr; }

void pcfd_mesh_free(pcfd_mesh* m) { delete m; }

int pcfd_color_nodes(int nnode, int nedge, const int* edges, int* colour) {
  if (nnode < 0 || nedge < 0 || (nedge > 0 && !edges) || (nnode > 0 && !colour)) {
    g_err = "pcfd_color_nodes: bad arguments";
    return -1;
  }
  std::vector<int> deg(nnode + 1, 0);
  for (int e = 0; e < nedge; ++e) {
    int a = edges[2 * e], b = edges[2 * e + 1];
    if (a < 0 || b < 0 || a >= nnode || b >= nnode) {
      g_err = "pcfd_color_nodes: edge node out of range";
      return -1;
    }
    deg[a + 1]++;
    deg[b + 1]++;
  }
  for (int i = 0; i < nnode; ++i) deg[i + 1] += deg[i];
  std::vector<int> adj(deg[nnode]), fill(deg.begin(), deg.end() - 1);
  for (int e = 0; e < nedge; ++e) {
    int a = edges[2 * e], b = edges[2 * e + 1];
    adj[fill[a]++] = b;
    adj[fill[b]++] = a;
  }
  if (nnode == 0) return 0;
  // greedy colouring in a given visiting order, smallest free colour
  std::vector<int> mark;
  auto greedy = [&](const std::vector<int>& order) {
    for (int i = 0; i < nnode; ++i) colour[i] = -1;
    int ncol = 0;
    for (int v : order) {
      mark.assign(ncol + 1, 0);
      for (int p = deg[v]; p < deg[v + 1]; ++p) {
        int c = colour[adj[p]];
        if (c >= 0) mark[c] = 1;
      }
      int c = 0;
      while (c < ncol && mark[c]) ++c;
      colour[v] = c;
      if (c == ncol) ++ncol;
    }
    return ncol;
  };
  // smallest-last order (Matula-Beck): repeatedly remove a node of least
  // remaining degree (ties: lowest id), colour in reverse removal order;
  // uses at most degeneracy + 1 colours
  std::vector<int> d(nnode);
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<std::pair<int, int>>> pq;
  for (int i = 0; i < nnode; ++i) {
    d[i] = deg[i + 1] - deg[i];
    pq.push({d[i], i});
  }
  std::vector<char> gone(nnode, 0);
  std::vector<int> order;
  order.reserve(nnode);
  while (!pq.empty()) {
    const auto [dv, v] = pq.top();
    pq.pop();
    if (gone[v] || d[v] != dv) continue;  // stale entry
    gone[v] = 1;
    order.push_back(v);
    for (int p = deg[v]; p < deg[v + 1]; ++p) {
      int u = adj[p];
      if (!gone[u]) pq.push({--d[u], u});
    }
  }
  std::reverse(order.begin(), order.end());
  int ncol = greedy(order);
  // iterated greedy (Culberson): revisit whole colour classes, largest
  // first; a class-contiguous order never needs more colours, and small
  // classes tend to empty
  for (int round = 0; round < 8; ++round) {
    std::vector<int> cnt(ncol, 0);
    for (int i = 0; i < nnode; ++i) cnt[colour[i]]++;
    std::vector<int> cls(ncol);
    std::iota(cls.begin(), cls.end(), 0);
    std::stable_sort(cls.begin(), cls.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
    std::vector<int> rank(ncol);
    for (int k = 0; k < ncol; ++k) rank[cls[k]] = k;
    std::vector<int> start(ncol + 1, 0);
    for (int c = 0; c < ncol; ++c) start[rank[c] + 1] = cnt[c];
    for (int k = 0; k < ncol; ++k) start[k + 1] += start[k];
    std::vector<int> o2(nnode);
    for (int v : order) o2[start[rank[colour[v]]]++] = v;
    order.swap(o2);
    int nc2 = greedy(order);
    if (nc2 > ncol) {
      g_err = "pcfd_color_nodes: internal error (iterated greedy grew)";
      return -1;
    }
    ncol = nc2;
  }
  // relabel: the multicolour SGS skips the pass over its last colour at each
  // turn of a symmetric sweep and over colour 0 at the start of each later
  // forward sweep, so the largest class becomes colour 0, the second largest
  // the last colour, the rest (largest first) in between
  std::vector<int> cnt(ncol, 0);
  for (int i = 0; i < nnode; ++i) cnt[colour[i]]++;
  std::vector<int> cls(ncol);
  std::iota(cls.begin(), cls.end(), 0);
  std::stable_sort(cls.begin(), cls.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
  std::vector<int> label(ncol);
  for (int k = 0; k < ncol; ++k) {
    int pos = k == 0 ? 0 : (k == 1 ? ncol - 1 : k - 1);
    label[cls[k]] = pos;
  }
  for (int i = 0; i < nnode; ++i) colour[i] = label[colour[i]];
  return ncol;
}

}  // extern "C"
