/* hexio.h -- mesh and feature-graph files: the C ABI of libhexopt.so for VTK, Medit .mesh,
 * OBJ and .fgraph files. Host code only: no call here touches a GPU.
 *
 * Each entry point replaces one member of the reference's h_io class (io.h:13-31, io.cpp)
 * and reads or writes the same bytes:
 *   - numbers are written as C's %g with six significant digits, which is what the
 *     reference's std::ostream with default flags prints;
 *   - numbers are read as sscanf's %lf / %d / %i read them (strtod, strtol base 10, strtol
 *     base 0).
 * Where the reference reads a malformed file silently (a short line keeps the previous
 * line's values, a missing section loops for ever, an out-of-range vertex id is undefined)
 * these calls fail with a message instead. Every file the reference writes reads back here
 * as the reference reads it.
 *
 * Errors: readers return NULL and writers return nonzero; hexio_error() then gives the message
 * of the calling thread's last failure. Nothing throws across the ABI.
 *
 * Mesh type codes are the reference's Mesh_type values (global_types.h:457-465). Cells and
 * faces are passed as CSR: off[n+1] (or NULL when every cell has k vertices) and idx, with
 * 0-based vertex ids. Vertex positions are row-major [nV][3] doubles.
 */
#ifndef HEXIO_H
#define HEXIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { HEXIO_TRI = 0, HEXIO_QUA = 1, HEXIO_HSUR = 2, HEXIO_TET = 3, HEXIO_HYB = 4, HEXIO_HEX = 5 };

/* Message of the calling thread's last failed hexio call ("" if none). */
const char* hexio_error(void);

/* ---- meshes read from files ------------------------------------------------------------ */
typedef struct hexio_mesh hexio_mesh;

/* Replaces h_io::read_hybrid_mesh_VTK (io.cpp:5-101): "POINTS n <type>" within the first 10
 * lines, one point per line, then the first "CELLS n m" line and one cell per line
 * ("k v0 .. v(k-1)", k in 4, 5, 6, 8). CELL_TYPES and POINT_DATA are not read. The reference
 * leaves the mesh type to its caller; here it is HEX when every cell has 8 vertices, TET when
 * every cell has 4, HYB otherwise. */
hexio_mesh* hexio_read_vtk(const char* path);
/* Replaces h_io::read_hybrid_mesh_MESH (io.cpp:181-294): "Vertices n" (or "Vertices" and n on
 * the next line), one vertex per line (x y z ref), then the first "Hexahedra" or "Tetrahedra"
 * section (1-based ids, a trailing ref ignored). Type HEX or TET. ("Tetrahedra n" on one line
 * is read as that count; the reference would take the next line for it.) */
hexio_mesh* hexio_read_mesh(const char* path);
/* Replaces h_io::read_hybrid_mesh_OBJ (io.cpp:326-356), which reads through igl::readOBJ
 * (extern/libigl/include/igl/readOBJ.cpp:12-225): "v" lines of 3 or 4 numbers (the first 3
 * kept), "vt"/"vn" lines checked and skipped, "f" lines of i, i/t, i//n or i/t/n words
 * (negative ids count back from the vertices read so far), comment, group, smoothing and
 * material lines ignored. Every face must have the same number of vertices. Type TRI for
 * triangles, QUA for quads, HYB otherwise. */
hexio_mesh* hexio_read_obj(const char* path);
/* out4 = { nV, number of cells, length of idx, type } */
int hexio_mesh_sizes(const hexio_mesh* m, int64_t* out4);
/* V[nV][3], off[nCells+1], idx[len]; any pointer may be NULL to skip that array */
int hexio_mesh_copy(const hexio_mesh* m, double* V, int32_t* off, int32_t* idx);
void hexio_mesh_free(hexio_mesh* m);

/* ---- mesh writers -------------------------------------------------------------------------- */
/* Replaces h_io::write_hybrid_mesh_VTK (io.cpp:102-180). type TRI/QUA/HYB writes the cells as
 * faces (cell type 5 for TRI, 9 for QUA, 7 for HYB); any other type writes them as volume
 * cells (10 for TET, 12 otherwise). Every cell line is written with the first cell's size, as
 * the reference does, except for HYB. boundary (NULL = all 0) becomes the "fixed" point data. */
int hexio_write_vtk(const char* path, int type, int nV, const double* V, int nC, int k, const int32_t* off,
                    const int32_t* idx, const uint8_t* boundary);
/* Replaces h_io::write_hybrid_mesh_MESH (io.cpp:295-325): vertices with ref 0; a Triangles
 * section (first 3 ids of each cell) for TRI and HSUR, a Hexahedra section for HEX, no cell
 * section for other types; 1-based ids; "End" with no newline. */
int hexio_write_mesh(const char* path, int type, int nV, const double* V, int nC, int k, const int32_t* off,
                     const int32_t* idx);
/* Replaces h_io::write_hybrid_mesh_OBJ (io.cpp:358-377): "v x y z" and "f a b c ..." (1-based). */
int hexio_write_obj(const char* path, int nV, const double* V, int nF, int k, const int32_t* off,
                    const int32_t* idx);
/* Replaces h_io::write_mesh_OBJ (io.cpp:378-411) for a vertex-per-row V[nV][dim] and
 * F[nF][k]: each "v" line ends with a space after its last coordinate, as there. */
int hexio_write_matrix_obj(const char* path, int nV, int dim, const double* V, int nF, int k, const int32_t* F);

/* ---- feature graphs (.fgraph) -------------------------------------------------------------- */
typedef struct hexio_fgraph hexio_fgraph;
typedef struct {
  double angle_threshold;
  int32_t orphan_curve;         /* 0 or 1 */
  int32_t orphan_curve_single;  /* 0 or 1 */
  int32_t n_corners;
  int32_t n_edges;
} hexio_fgraph_header;

/* Replaces h_io::read_feature_Graph_FGRAPH (io.cpp:412-435): "angle orphan orphan_single",
 * "n_corners n_edges", one corner vertex id per line, one "v0 v1" feature edge per line
 * (integers as %i reads them: 0x.. hexadecimal and 0.. octal too; ids are the reference's
 * uint32 values, so 4294967295 reads as -1). NULL when the file cannot be opened (the
 * reference returns false then) or is malformed. */
hexio_fgraph* hexio_read_fgraph(const char* path);
int hexio_fgraph_info(const hexio_fgraph* g, hexio_fgraph_header* out);
/* corners[n_corners], pairs[n_edges][2]; either may be NULL */
int hexio_fgraph_copy(const hexio_fgraph* g, int32_t* corners, int32_t* pairs);
void hexio_fgraph_free(hexio_fgraph* g);
/* Replaces h_io::write_feature_Graph_FGRAPH (io.cpp:436-447), given the corner ids and the
 * flagged feature edges' vertex pairs (pairs[n_edges][2]) in the order to write them. Ids are
 * written as the reference's uint32 values. */
int hexio_write_fgraph(const char* path, double angle_threshold, int orphan_curve, int orphan_curve_single,
                       int n_corners, const int32_t* corners, int n_edges, const int32_t* pairs);

#ifdef __cplusplus
}
#endif

#endif /* HEXIO_H */
