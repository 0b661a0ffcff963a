/* hexclean.h -- C ABI of the hex-mesh cleaning passes of clean_hex_mesh
 * (grid_meshing/grid_hex_meshing.cpp:1932-1964): the passes that turn the inside flags of an
 * octree hex mesh into the flags of a clean, manifold, connected sub-mesh.
 *
 * Host code, in libhexopt.so. Every pass is a serial sweep whose in-place updates depend on
 * the order of hexes, vertices, faces and edges, and the results equal the reference's only
 * when that order is kept. So the passes rebuild the reference's connectivity numbering
 * (build_connectivity, global_functions.cpp:16-262: faces and edges numbered by their sorted
 * vertex ids, neighbour lists in ascending id order) and run in the same order.
 *
 * Meshes are nV vertices (xyz rows, double) and nH hexes (8 vertex ids each, int32, the
 * reference's corner order). H_flag holds one byte per hex (1 = in the sub-mesh), read and
 * written in place. Every call returns the number of flagged hexes afterwards, or -1 with
 * hexclean_error() saying why: a vertex id out of range, a hex with a repeated vertex or a
 * face shared by more than two hexes. Thread-safe; the error message is per thread. */
#ifndef HEXCLEAN_H
#define HEXCLEAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* hexclean_error(void);

/* reorder_hex_mesh (global_functions.cpp:2199-2229): a hex whose 8 corner tetrahedra
 * (hex_tetra_table, global_types.h:164-174) have a negative summed determinant gets its
 * corners reordered to 3,2,1,0,7,6,5,4. H is rewritten in place; *flipped (nullable) is the
 * number of hexes reordered. Returns nH. */
int64_t hexclean_reorder_hex_mesh(int64_t nV, const double* V, int64_t nH, int32_t* H, int64_t* flipped);

/* tagging_uneven_element (grid_hex_meshing.cpp:1983-2004) over the whole mesh, hexes in
 * order and flags updated as it goes: a flagged hex with 5 faces on the flagged region's
 * boundary and 1 inside it is unflagged, an unflagged hex in the mirrored position is
 * flagged. */
int64_t hexclean_tagging_uneven_element(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag);

/* clean_non_manifold_ve (grid_hex_meshing.cpp:2006-2080) on the sub-mesh of flagged hexes
 * (re_indexing_connectivity, global_functions.cpp:664-698, keeps the hex and vertex order):
 * at a boundary vertex whose boundary faces fall into several edge-connected fans, the hexes
 * of the fans other than the first face's are unflagged; when no vertex is non-manifold, a
 * boundary edge with other than 2 boundary faces unflags the hex of its first boundary
 * face; the sub-mesh is rebuilt and the sweep repeats until neither pass changes a flag. */
int64_t hexclean_clean_non_manifold_ve(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag);

/* drop_small_pieces (grid_hex_meshing.cpp:2081-2124): the face-connected pieces of the
 * sub-mesh are walked from its lowest unvisited hex, each piece's size counted as the
 * reference counts it (its walk list, repeats included), and only the first largest piece
 * stays flagged. An empty sub-mesh is left as it is. */
int64_t hexclean_drop_small_pieces(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag);

/* The flag passes of clean_hex_mesh in the reference's order (grid_hex_meshing.cpp:1954-1964):
 * tagging_uneven_element, then, unless no hex is flagged, clean_non_manifold_ve and
 * drop_small_pieces. H_flag comes in as signed_dis < 0 of each hex's box centre
 * (hexopt_hex_inside). The scaffold layers and the medial-surface tags that follow in the
 * reference are not part of it (the tags: hexclean_medial_surface). */
int64_t hexclean_clean_flags(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag);

/* The medial-surface tags that end clean_hex_mesh (grid_hex_meshing.cpp:1966-1981): over the
 * faces of the whole mesh in build_connectivity's order, a face is on the medial surface when
 * it is a boundary face of a flagged hex or an inner face between a flagged and an unflagged
 * hex, and so are its 4 vertices. v_flag (nullable, nV bytes) gets 1 for the vertices on it;
 * quads (nullable, room for 6 nH x 4 ids) gets the faces on it in face order, each with the
 * corners of its first occurrence, as the reference's Hybrid::vs. Returns the number of faces
 * on the medial surface (the scaffold layers of scaffold_type 2 and 3, which come before these
 * tags in the reference, are not built). */
int64_t hexclean_medial_surface(int64_t nV, int64_t nH, const int32_t* H, const uint8_t* H_flag, uint8_t* v_flag,
                                int32_t* quads);

#ifdef __cplusplus
}
#endif

#endif /* HEXCLEAN_H */
