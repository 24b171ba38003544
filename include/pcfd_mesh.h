/*
 * pcfd_mesh.h -- host-side (CPU, C++ implementation, C ABI) synthetic mesh
 * generator and median-dual metric builder for the edge-based finite-volume
 * hot path.
 *
 * What it replaces in the reference (ucs.x set-up, run once per mesh):
 *   Mesh<Type>::BuildEdges            ucs/mesh.tcc:1055-1167  (edges n[0]<n[1],
 *                                     boundary half-edges with phantom ids
 *                                     nnode+k, one per boundary-face node)
 *   Mesh<Type>::CalcAreasVolumes      ucs/mesh.tcc:1652-2243  (median-dual
 *                                     edge area vectors and CV volumes)
 *   Mesh<Type>::CheckForClosure       ucs/mesh.tcc:2273-2333  (closure check,
 *                                     then a[3]=|a|, a[0..2] normalised)
 *   Mesh<Type>::ReorderMeshCuthillMcKee ucs/mesh.tcc:2411-2535 (RCM numbering)
 *   Edges<Type>/HalfEdges<Type>       ucs/uns_base.h:12-37     (array layouts)
 *
 * It generates the synthetic benchmark meshes (hex grids split into 6 Kuhn
 * tetrahedra, or prism layers under a Kuhn-tet core) from a seed with a
 * counter-based RNG, so the CPU oracle and the GPU path see bit-identical
 * inputs, and reads the reference's ASCII mesh files (pcfd_mesh_read).
 *
 * All arrays are owned by the pcfd_mesh object and stay valid until
 * pcfd_mesh_free().  Return codes: 0 ok, <0 error (pcfd_mesh_error()).
 */
#ifndef PCFD_MESH_H
#define PCFD_MESH_H

#ifdef __cplusplus
extern "C" {
#endif

/* mesh families */
#define PCFD_MESH_WEDGE     1  /* [0,L]x[0,H]x[0,W] channel, lower wall ramp of angle_deg after x0 */
#define PCFD_MESH_BUMP      2  /* unit box, floor z = h*sin^2(pi x)*sin^2(pi y) */
#define PCFD_MESH_BOX       3  /* plain box (tests) */
#define PCFD_MESH_PLATE     4  /* flat-plate channel: prism layers on the floor over a Kuhn-tet core */

/* element types (same numbering as the reference etypes: ucs/etypes.h) */
#define PCFD_TRI     0
#define PCFD_QUAD    1
#define PCFD_TET     2
#define PCFD_PYRAMID 3
#define PCFD_PRISM   4
#define PCFD_HEX     5

typedef struct {
  int kind;                /* PCFD_MESH_* */
  int nx, ny, nz;          /* hex cells per direction */
  unsigned long long seed; /* jitter seed */
  double jitter;           /* fraction of local spacing, uniform +-jitter, interior nodes only */
  int rcm;                 /* 1: reverse Cuthill-McKee renumbering of the nodes */
  double lx, ly, lz;       /* box extents */
  double angle_deg;        /* WEDGE: ramp angle */
  double x0;               /* WEDGE: ramp start */
  double bump_h;           /* BUMP: bump height */
  int nlayers;             /* PLATE: number of prism layers on the floor */
  double first_h;          /* PLATE: first layer height */
  double growth;           /* PLATE: geometric growth ratio of the layers */
} pcfd_mesh_spec;

/* Face tags of the box sides: 1 xmin, 2 xmax, 3 ymin, 4 ymax, 5 zmin, 6 zmax.
 * Tag 0 is the reference's parallel boundary (bc_defines.h:8), never used here.
 * Boundary half-edges are ordered by ascending face tag, then by face. */
#define PCFD_NFACTAG 7

typedef struct {
  int nnode;        /* interior (owned) nodes */
  int nelem;        /* volume elements */
  int nedge;        /* interior edges */
  int nbface;       /* boundary triangles/quads */
  int nbedge;       /* boundary half-edges == phantom nodes (nbnode) */
  int nfactag;      /* PCFD_NFACTAG */
  double* xyz;      /* [nnode*3] */
  double* vol;      /* [nnode] median-dual CV volume */
  int* edges;       /* [nedge*2] n0<n1, sorted ascending (n0,n1) */
  double* edge_a;   /* [nedge*4] unit normal n0->n1, a[3]=|area| */
  int* bedges;      /* [nbedge*2] (node, phantom id nnode+k) */
  double* bedge_a;  /* [nbedge*4] outward unit normal, a[3]=|area| */
  int* bfactag;     /* [nbedge] */
  int* elem_type;   /* [nelem] PCFD_TET / PCFD_PRISM */
  int* elem_ptr;    /* [nelem+1] offsets into elem_nodes */
  int* elem_nodes;  /* element connectivity (positively oriented) */
  int* bface_type;  /* [nbface] PCFD_TRI / PCFD_QUAD */
  int* bface_ptr;   /* [nbface+1] */
  int* bface_nodes; /* outward-wound boundary faces */
  int* bface_tag;   /* [nbface] */
  int* perm;        /* [nnode] new id -> generator id (identity when rcm=0) */
  double closure_max; /* max_i |sum of signed dual areas of CV i| (CheckForClosure) */
  double total_volume;/* sum of CV volumes */
  double elem_volume; /* sum of element volumes (independent of the dual) */
} pcfd_mesh_data;

typedef struct pcfd_mesh pcfd_mesh;

void pcfd_mesh_default_spec(int kind, pcfd_mesh_spec* spec);
int  pcfd_mesh_build(const pcfd_mesh_spec* spec, pcfd_mesh** out);
/* Median-dual build from an external element list (what ucs.x's
 * BuildMaps + CalcMetrics do after a mesh reader, ucs/mesh.tcc:264-312,
 * 1576-1596).  Volume elements PCFD_TET/PYRAMID/PRISM/HEX in the local
 * numbering of the tables in mesh.cpp (tet any orientation, others with
 * their bottom face counter-clockwise seen from inside or outside: a
 * negatively oriented element is reflected); boundary faces PCFD_TRI/QUAD
 * are re-wound outward against their owning element (FixBoundaryWindings,
 * ucs/mesh.tcc:2336-2410).  Node ids are 0-based. */
int  pcfd_mesh_build_from(int nnode, const double* xyz, int nelem, const int* elem_type, const int* elem_ptr,
                          const int* elem_nodes, int nbface, const int* bface_type, const int* bface_ptr,
                          const int* bface_nodes, const int* bface_tag, int rcm, pcfd_mesh** out);
/* Mesh file readers (Mesh<Type>::Read*, ucs/mesh.tcc): read the file,
 * drop nodes no volume element uses, and build as pcfd_mesh_build_from.
 * PCFD_MESH_FMT_AUTO picks the format from the extension: .msh (GMSH 2.2 /
 * 4.0 / 4.1 ASCII, ReadGMSH_Master ucs/mesh.tcc:3553-3597), .ugrid (UGRID
 * ASCII, ReadUGRID_Ascii :6740-6932), .crunch/.fv/.uns (FieldView ASCII,
 * ReadCRUNCH_Ascii :6933-7300), .su2 (SU2 ASCII, ReadSU2_Ascii :3197-3548).
 * Face tags become bface_tag / bfactag; nfactag = max(PCFD_NFACTAG, largest
 * tag + 1). */
#define PCFD_MESH_FMT_AUTO      0
#define PCFD_MESH_FMT_GMSH      1
#define PCFD_MESH_FMT_UGRID     2
#define PCFD_MESH_FMT_FIELDVIEW 3
#define PCFD_MESH_FMT_SU2       4
int  pcfd_mesh_read(const char* path, int format, int rcm, pcfd_mesh** out);
const pcfd_mesh_data* pcfd_mesh_get(const pcfd_mesh* m);
void pcfd_mesh_free(pcfd_mesh* m);
const char* pcfd_mesh_error(void);

/* Distance-1 colouring of the node graph given by edges (used by the
 * multicolour SGS sweeps).  colour[nnode] receives colours 0..ncolor-1;
 * returns ncolor (<0 on error).  Deterministic: greedy (smallest free
 * colour) in smallest-last order, then eight rounds of iterated greedy over
 * whole colour classes, largest first.  Fewer colours mean fewer SGS passes
 * (2*nsgs*ncolor - (2*nsgs - 1)).  Colour 0 is the largest class and the
 * last colour the second largest: those are the passes the sweeps skip. */
int pcfd_color_nodes(int nnode, int nedge, const int* edges, int* colour);

#ifdef __cplusplus
}
#endif
#endif
