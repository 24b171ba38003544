/*
 * pcfd_mesh_h5.h -- partitioned HDF5 mesh files <casename>.<rank>.h5, the
 * format ucs.x reads.  Exported by libpcfd_h5.so only, which is built where
 * HDF5 is installed; libpcfd.so and libpcfd_mesh.so do not need HDF5.
 * The mesh object is the one of pcfd_mesh.h (pcfd_mesh_get, pcfd_mesh_free).
 */
#ifndef PCFD_MESH_H5_H
#define PCFD_MESH_H5_H

#include "pcfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pcfd_mesh_write_h5 replaces udecomp.x's writer (ucs/decomp.cpp:281-535) and
 * Mesh<Type>::WriteParallelMesh (ucs/mesh.tcc:2750-2818): np files from the
 * element and boundary-face lists of a mesh (pcfd_mesh_data's elem_* and
 * bface_* arrays) and a partition vector part[nnode] (NULL when np == 1).
 * pcfd_mesh_read_h5 replaces Mesh<Type>::ReadPartedMesh
 * (ucs/mesh.tcc:2820-2990): it reads every rank's file, joins the partitions
 * into one mesh (rank r's local node j is node offset[r] + j; elements stored
 * by several ranks are kept once) and builds it as pcfd_mesh_build_from. */
int  pcfd_mesh_write_h5(const char* casename, int np, const int* part, int nnode, const double* xyz, int nelem,
                        const int* elem_type, const int* elem_ptr, const int* elem_nodes, int nbface,
                        const int* bface_type, const int* bface_ptr, const int* bface_nodes, const int* bface_tag,
                        int reordered);
int  pcfd_mesh_read_h5(const char* casename, int rcm, pcfd_mesh** out);

#ifdef __cplusplus
}
#endif
#endif
