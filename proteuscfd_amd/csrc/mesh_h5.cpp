// Partitioned HDF5 mesh files, the on-disk mesh format ucs.x runs from:
// one file <case>.<rank>.h5 per MPI rank.
//
//   writer  udecomp.x (ucs/decomp.cpp:281-535) and Mesh<Type>::WriteParallelMesh
//           (ucs/mesh.tcc:2750-2818)
//   reader  Mesh<Type>::ReadPartedMesh (ucs/mesh.tcc:2820-2990)
//
// Layout kept alike so either side reads the other's files.  Root group:
// scalars "Number Of Processors", "Reordered", "Rescaled", "Global Number Of
// Nodes", "Global Number Of Elements", "Number Of Factags".  Group /Mesh/:
// per-type element counts ("Number Of Triangles" ... "Number Of Hexahedron"),
// "Number Of Local Nodes", "Number Of Ghost Nodes", and the arrays "Ghost
// Nodes Owning Process", "Ghost Nodes Local Id", "Nodal Coordinates" (local
// then ghost nodes, xyz interleaved), "Element Factags" and "Element Data"
// (per element its type id followed by its nodes in the rank's local
// numbering, ghosts numbered after the local nodes).  Scalars are 1-element
// 1-D datasets and arrays (n, 1) 2-D datasets of native int / double, as the
// reference's HDF_WriteScalar / HDF_WriteArray make them (ucs/h5layer.tcc).
//
// Writing: a rank's local nodes are the nodes the partition vector gives it,
// in ascending global id; an element goes to every rank that owns one of its
// nodes, first the elements wholly inside the rank, then the split ones, each
// group in element order (boundary faces before volume elements); the ghost
// nodes of a rank are the other ranks' nodes its split elements touch, grouped
// by owning rank and ascending within a group.
//
// Reading: every rank's file is read and the partitions are joined back into
// one mesh for the single device: rank r's local node j becomes global node
// offset[r] + j (offset = local node counts of the ranks before r), ghosts map
// to their owner's copy, and an element stored by several ranks is kept once
// (first occurrence, in rank order).  The joined element list then goes
// through the same median-dual build as the ASCII readers.
#include <hdf5.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcfd_mesh_h5.h"

void pcfd_mesh_set_error(const std::string& msg);  // mesh.cpp

namespace {

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

const int kNodesOf[6] = {3, 4, 4, 5, 6, 8};  // TRI QUAD TET PYRAMID PRISM HEX

const char* kCountNames[6] = {"Number Of Triangles", "Number Of Quadrilaterals", "Number Of Tetrahedron",
                              "Number Of Pyramids",  "Number Of Prisms",        "Number Of Hexahedron"};

struct File {
  hid_t id = -1;
  std::string name;
  ~File() {
    if (id >= 0) H5Fclose(id);
  }
};

struct Group {
  hid_t id = -1;
  ~Group() {
    if (id >= 0) H5Gclose(id);
  }
};

hid_t open_or_create_group(hid_t file, const char* path) {
  if (std::string(path) == "/") return H5Gopen2(file, "/", H5P_DEFAULT);
  htri_t e = H5Lexists(file, path, H5P_DEFAULT);
  if (e > 0) return H5Gopen2(file, path, H5P_DEFAULT);
  return H5Gcreate2(file, path, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

template <class T>
hid_t native();
template <>
hid_t native<int>() {
  return H5T_NATIVE_INT;
}
template <>
hid_t native<double>() {
  return H5T_NATIVE_DOUBLE;
}

template <class T>
void write_data(hid_t file, const char* group, const char* name, const T* data, hsize_t n, bool scalar) {
  Group g;
  g.id = open_or_create_group(file, group);
  if (g.id < 0) throw Fail(std::string("cannot create group ") + group);
  hsize_t dims[2] = {n, 1};
  hid_t space = H5Screate_simple(scalar ? 1 : 2, dims, nullptr);
  hid_t dset = H5Dcreate2(g.id, name, native<T>(), space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t rc = dset < 0 ? -1 : (n ? H5Dwrite(dset, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) : 0);
  if (dset >= 0) H5Dclose(dset);
  H5Sclose(space);
  if (rc < 0) throw Fail(std::string("cannot write dataset ") + group + name);
}

template <class T>
std::vector<T> read_data(hid_t file, const File& f, const char* group, const char* name) {
  std::string path = std::string(group) + name;
  hid_t dset = H5Dopen2(file, path.c_str(), H5P_DEFAULT);
  if (dset < 0) throw Fail(f.name + ": no dataset " + path);
  hid_t space = H5Dget_space(dset);
  hssize_t n = H5Sget_simple_extent_npoints(space);
  std::vector<T> v(n > 0 ? (size_t)n : 0);
  herr_t rc = n > 0 ? H5Dread(dset, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) : 0;
  H5Sclose(space);
  H5Dclose(dset);
  if (rc < 0) throw Fail(f.name + ": cannot read " + path);
  return v;
}

template <class T>
T read_scalar(hid_t file, const File& f, const char* group, const char* name) {
  std::vector<T> v = read_data<T>(file, f, group, name);
  if (v.size() != 1) throw Fail(f.name + ": " + group + name + " is not a scalar");
  return v[0];
}

std::string rank_file(const char* casename, int r) { return std::string(casename) + "." + std::to_string(r) + ".h5"; }

// one element (boundary face or volume element) of the global list
struct Elem {
  int type, tag;
  const int* nodes;
};

}  // namespace

extern "C" int pcfd_mesh_write_h5(const char* casename, int np, const int* part, int nnode, const double* xyz,
                                  int nelem, const int* elem_type, const int* elem_ptr, const int* elem_nodes,
                                  int nbface, const int* bface_type, const int* bface_ptr, const int* bface_nodes,
                                  const int* bface_tag, int reordered) {
  try {
    if (!casename || np < 1 || nnode < 1 || !xyz || nelem < 1) throw Fail("pcfd_mesh_write_h5: bad arguments");
    if (np > 1 && !part) throw Fail("pcfd_mesh_write_h5: np > 1 needs a partition vector");
    std::vector<int> owner(nnode, 0);
    for (int i = 0; i < nnode && np > 1; ++i) {
      if (part[i] < 0 || part[i] >= np) throw Fail("pcfd_mesh_write_h5: partition id out of range");
      owner[i] = part[i];
    }
    // global element list: boundary faces first, then volume elements
    std::vector<Elem> el;
    el.reserve((size_t)nbface + nelem);
    int nfactag = 0;
    for (int f = 0; f < nbface; ++f) {
      if (bface_type[f] != PCFD_TRI && bface_type[f] != PCFD_QUAD) throw Fail("boundary face type must be TRI/QUAD");
      el.push_back({bface_type[f], bface_tag[f], bface_nodes + bface_ptr[f]});
      nfactag = std::max(nfactag, bface_tag[f]);
    }
    for (int e = 0; e < nelem; ++e) {
      if (elem_type[e] < PCFD_TET || elem_type[e] > PCFD_HEX) throw Fail("volume element type out of range");
      el.push_back({elem_type[e], 0, elem_nodes + elem_ptr[e]});
    }
    for (const Elem& x : el)
      for (int j = 0; j < kNodesOf[x.type]; ++j)
        if (x.nodes[j] < 0 || x.nodes[j] >= nnode) throw Fail("element node id out of range");
    // local numbering per rank
    std::vector<int> local_id(nnode), count(np, 0);
    for (int i = 0; i < nnode; ++i) local_id[i] = count[owner[i]]++;
    const int gnelem = (int)el.size();
    for (int r = 0; r < np; ++r) {
      std::vector<int> local_el, split_el;
      for (int e = 0; e < gnelem; ++e) {
        const Elem& x = el[e];
        bool any = false, all = true;
        for (int j = 0; j < kNodesOf[x.type]; ++j) {
          bool mine = owner[x.nodes[j]] == r;
          any |= mine;
          all &= mine;
        }
        if (all) local_el.push_back(e);
        else if (any) split_el.push_back(e);
      }
      std::vector<int> ghosts;
      for (int e : split_el)
        for (int j = 0; j < kNodesOf[el[e].type]; ++j)
          if (owner[el[e].nodes[j]] != r) ghosts.push_back(el[e].nodes[j]);
      std::sort(ghosts.begin(), ghosts.end(), [&](int a, int b) {
        return owner[a] != owner[b] ? owner[a] < owner[b] : a < b;
      });
      ghosts.erase(std::unique(ghosts.begin(), ghosts.end()), ghosts.end());
      std::map<int, int> ghost_slot;
      for (size_t k = 0; k < ghosts.size(); ++k) ghost_slot[ghosts[k]] = count[r] + (int)k;
      std::vector<int> gown(ghosts.size()), glid(ghosts.size());
      std::vector<double> coords;
      coords.reserve(3 * ((size_t)count[r] + ghosts.size()));
      for (int i = 0; i < nnode; ++i)
        if (owner[i] == r) coords.insert(coords.end(), xyz + 3 * (size_t)i, xyz + 3 * (size_t)i + 3);
      for (size_t k = 0; k < ghosts.size(); ++k) {
        gown[k] = owner[ghosts[k]];
        glid[k] = local_id[ghosts[k]];
        coords.insert(coords.end(), xyz + 3 * (size_t)ghosts[k], xyz + 3 * (size_t)ghosts[k] + 3);
      }
      std::vector<int> tags, data;
      int type_count[6] = {0, 0, 0, 0, 0, 0};
      for (int pass = 0; pass < 2; ++pass)
        for (int e : (pass == 0 ? local_el : split_el)) {
          const Elem& x = el[e];
          type_count[x.type]++;
          tags.push_back(x.tag);
          data.push_back(x.type);
          for (int j = 0; j < kNodesOf[x.type]; ++j) {
            int n = x.nodes[j];
            data.push_back(owner[n] == r ? local_id[n] : ghost_slot.at(n));
          }
        }
      File f;
      f.name = rank_file(casename, r);
      f.id = H5Fcreate(f.name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
      if (f.id < 0) throw Fail("cannot create " + f.name);
      int rescaled = 0, reord = reordered ? 1 : 0, gn = nnode, lnode = count[r], gnode = (int)ghosts.size();
      write_data(f.id, "/", "Number Of Processors", &np, 1, true);
      write_data(f.id, "/", "Reordered", &reord, 1, true);
      write_data(f.id, "/", "Global Number Of Nodes", &gn, 1, true);
      write_data(f.id, "/", "Global Number Of Elements", &gnelem, 1, true);
      write_data(f.id, "/", "Number Of Factags", &nfactag, 1, true);
      write_data(f.id, "/", "Rescaled", &rescaled, 1, true);
      write_data(f.id, "/Mesh/", "Number Of Local Nodes", &lnode, 1, true);
      write_data(f.id, "/Mesh/", "Number Of Ghost Nodes", &gnode, 1, true);
      for (int t = 0; t < 6; ++t) write_data(f.id, "/Mesh/", kCountNames[t], &type_count[t], 1, true);
      write_data(f.id, "/Mesh/", "Ghost Nodes Owning Process", gown.data(), gown.size(), false);
      write_data(f.id, "/Mesh/", "Ghost Nodes Local Id", glid.data(), glid.size(), false);
      write_data(f.id, "/Mesh/", "Nodal Coordinates", coords.data(), coords.size(), false);
      write_data(f.id, "/Mesh/", "Element Factags", tags.data(), tags.size(), false);
      write_data(f.id, "/Mesh/", "Element Data", data.data(), data.size(), false);
    }
    return 0;
  } catch (const std::exception& e) {
    pcfd_mesh_set_error(e.what());
    return -1;
  }
}

extern "C" int pcfd_mesh_read_h5(const char* casename, int rcm, pcfd_mesh** out) {
  try {
    if (!casename || !out) throw Fail("pcfd_mesh_read_h5: bad arguments");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    int np = 0;
    std::vector<int> nloc, gnnode_seen;
    struct Rank {
      int nnode, gnode;
      std::vector<int> gown, glid, tags, data;
      std::vector<double> xyz;
      int counts[6];
    };
    std::vector<Rank> ranks;
    for (int r = 0; r == 0 || r < np; ++r) {
      File f;
      f.name = rank_file(casename, r);
      f.id = H5Fopen(f.name.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
      if (f.id < 0) throw Fail("cannot open " + f.name);
      int npr = read_scalar<int>(f.id, f, "/", "Number Of Processors");
      if (r == 0) np = npr;
      if (npr != np || np < 1) throw Fail(f.name + ": Number Of Processors disagrees with rank 0's file");
      gnnode_seen.push_back(read_scalar<int>(f.id, f, "/", "Global Number Of Nodes"));
      Rank k;
      k.nnode = read_scalar<int>(f.id, f, "/Mesh/", "Number Of Local Nodes");
      k.gnode = read_scalar<int>(f.id, f, "/Mesh/", "Number Of Ghost Nodes");
      for (int t = 0; t < 6; ++t) k.counts[t] = read_scalar<int>(f.id, f, "/Mesh/", kCountNames[t]);
      k.gown = read_data<int>(f.id, f, "/Mesh/", "Ghost Nodes Owning Process");
      k.glid = read_data<int>(f.id, f, "/Mesh/", "Ghost Nodes Local Id");
      k.xyz = read_data<double>(f.id, f, "/Mesh/", "Nodal Coordinates");
      k.tags = read_data<int>(f.id, f, "/Mesh/", "Element Factags");
      k.data = read_data<int>(f.id, f, "/Mesh/", "Element Data");
      if (k.nnode < 0 || k.gnode < 0 || (int)k.gown.size() != k.gnode || (int)k.glid.size() != k.gnode ||
          k.xyz.size() != 3 * (size_t)(k.nnode + k.gnode))
        throw Fail(f.name + ": node arrays do not match the node counts");
      ranks.push_back(std::move(k));
    }
    std::vector<int> offset(np + 1, 0);
    for (int r = 0; r < np; ++r) offset[r + 1] = offset[r] + ranks[r].nnode;
    const int nnode = offset[np];
    for (int g : gnnode_seen)
      if (g != nnode) throw Fail("Global Number Of Nodes disagrees with the sum of the local node counts");
    std::vector<double> xyz(3 * (size_t)nnode);
    for (int r = 0; r < np; ++r)
      std::copy(ranks[r].xyz.begin(), ranks[r].xyz.begin() + 3 * (size_t)ranks[r].nnode,
                xyz.begin() + 3 * (size_t)offset[r]);
    std::vector<int> et, ep{0}, en, bt, bp{0}, bn, bg;
    std::map<std::pair<int, std::vector<int>>, int> seen;
    for (int r = 0; r < np; ++r) {
      const Rank& k = ranks[r];
      auto global = [&](int l) -> int {
        if (l < 0 || l >= k.nnode + k.gnode) throw Fail("element node outside the rank's node list");
        if (l < k.nnode) return offset[r] + l;
        int g = l - k.nnode, o = k.gown[g], id = k.glid[g];
        if (o < 0 || o >= np || id < 0 || id >= ranks[o].nnode) throw Fail("ghost node with a bad owner / local id");
        return offset[o] + id;
      };
      size_t pos = 0;
      int nel = 0;
      for (int t = 0; t < 6; ++t) nel += k.counts[t];
      if ((int)k.tags.size() != nel) throw Fail("Element Factags does not match the element counts");
      for (int e = 0; e < nel; ++e) {
        if (pos >= k.data.size()) throw Fail("Element Data ends early");
        int t = k.data[pos++];
        if (t < 0 || t > 5 || pos + kNodesOf[t] > k.data.size()) throw Fail("Element Data is malformed");
        std::vector<int> nodes(kNodesOf[t]);
        for (int j = 0; j < kNodesOf[t]; ++j) nodes[j] = global(k.data[pos++]);
        std::vector<int> key = nodes;
        std::sort(key.begin(), key.end());
        if (!seen.emplace(std::make_pair(t, key), 1).second) continue;
        if (t <= PCFD_QUAD) {
          bt.push_back(t);
          bn.insert(bn.end(), nodes.begin(), nodes.end());
          bp.push_back((int)bn.size());
          bg.push_back(k.tags[e]);
        } else {
          et.push_back(t);
          en.insert(en.end(), nodes.begin(), nodes.end());
          ep.push_back((int)en.size());
        }
      }
      if (pos != k.data.size()) throw Fail("Element Data has trailing entries");
    }
    if (et.empty()) throw Fail("the partition files hold no volume elements");
    return pcfd_mesh_build_from(nnode, xyz.data(), (int)et.size(), et.data(), ep.data(), en.data(), (int)bt.size(),
                                bt.data(), bp.data(), bn.data(), bg.data(), rcm, out);
  } catch (const std::exception& e) {
    pcfd_mesh_set_error(e.what());
    return -1;
  }
}
