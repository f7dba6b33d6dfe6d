/*
 * fwn_io.h -- C ABI of the mesh reader / writer that feeds the winding-number engine.
 *
 * Counterpart of lagrange::io::load_mesh_obj / load_mesh_ply / save_mesh_obj / save_mesh_ply
 * (modules/io/include/lagrange/io/load_mesh_obj.h, load_mesh_ply.h, save_mesh_obj.h,
 * save_mesh_ply.h) restricted to what the engine consumes: vertex positions and polygonal facets
 * (other attributes -- normals, uvs, colours, materials -- are skipped when reading and not
 * written).  Host code only; no GPU is touched.
 *
 * A loaded mesh is an opaque fwn_mesh: positions in double (n_vertices x 3) and facets in CSR
 * form (facet f has corners [offsets[f], offsets[f+1]) of `corners`, zero-based vertex ids).
 * With triangulate != 0 every polygon of k > 3 corners becomes the fan (c0, c_i, c_i+1),
 * i = 1..k-2, as LoadOptions::triangulate does (modules/io/include/lagrange/io/types.h:99).
 *
 * Formats: OBJ (ASCII; "v x y z [w]" and "f i[/t][/n] ..." records, 1-based or negative
 * relative indices) and PLY (ascii, binary_little_endian, binary_big_endian; "vertex" element
 * with x, y, z of any scalar type, "face" element with a vertex_indices / vertex_index list;
 * other elements and properties skipped).
 */
#ifndef FWN_IO_H
#define FWN_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FWN_IO_OK = 0,
    FWN_IO_INVALID_ARGUMENT = 1,
    FWN_IO_FILE_ERROR = 2,
    FWN_IO_PARSE_ERROR = 3
};

enum {
    FWN_IO_FORMAT_AUTO = 0,       /* from the file extension (.obj / .ply, any case) */
    FWN_IO_FORMAT_OBJ = 1,
    FWN_IO_FORMAT_PLY = 2,        /* reading: any PLY encoding */
    FWN_IO_FORMAT_PLY_BINARY = 3  /* writing: binary_little_endian (FORMAT_PLY writes ascii) */
};

typedef struct fwn_mesh fwn_mesh;

/* Read a mesh file (format FWN_IO_FORMAT_AUTO, _OBJ or _PLY). */
int fwn_mesh_load(const char* path, int32_t format, int32_t triangulate, fwn_mesh** out);

/* Read a mesh from memory (format _OBJ or _PLY). */
int fwn_mesh_load_buffer(const char* data, int64_t size, int32_t format, int32_t triangulate, fwn_mesh** out);

/* Sizes of a loaded mesh: vertices, facets, corners (sum of facet sizes). */
int fwn_mesh_sizes(const fwn_mesh* m, int64_t* num_vertices, int64_t* num_facets, int64_t* num_corners);

/* Copy out: V (num_vertices x 3 double), offsets (num_facets + 1), corners (num_corners int32).
 * Any pointer may be NULL. */
int fwn_mesh_copy(const fwn_mesh* m, double* V, int64_t* offsets, int32_t* corners);

int fwn_mesh_free(fwn_mesh* m);

/* Write a mesh (format _OBJ, _PLY = ascii or _PLY_BINARY); positions are written with 17
 * significant digits (text) or as doubles (binary), so a save / load round trip is exact. */
int fwn_mesh_save(const char* path, int32_t format, const double* V, int64_t num_vertices, const int64_t* offsets,
                  const int32_t* corners, int64_t num_facets);

/* Message of the last failed call on this thread. */
const char* fwn_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FWN_IO_H */
