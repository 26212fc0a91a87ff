/* htgx_files.h -- readers of htrdr-atmosphere's ground input files.
 *
 * The ground of option -g is an htrdr-obj file (doc/htrdr-obj.5: an obj file
 * of triangles, each in a `usemtl front[:thin]:back' group naming the
 * materials on its two sides). The materials are listed by the
 * htrdr-materials file of option -M (doc/htrdr-materials.5: one material per
 * line, `name mrumtl-path|none [temperature]'), each opaque material's
 * spectral BRDF in an mrumtl file. Option -n names the sky material (the
 * reference's default is "air", config.mk:53).
 *
 * htgx_ground_files_load reads the three as htrdr_materials_create
 * (src/core/htrdr_materials.c:72-235, 327-361) and htrdr_geometry_create
 * (src/core/htrdr_geometry.c:177-566) do, and fills an htgx_ground_desc
 * (include/htgx.h) of type HTGX_GROUND_MESH with spectral BRDF lists and
 * interfaces, ready for htgx_ctx_create.
 *
 * mrumtl files (the mrumtl 0.2 library is not part of the reference; its
 * format is restated here):
 *
 *   wavelengths <count>                  bands <count>
 *   <nm> <brdf>                          <lo-nm> <hi-nm> <brdf>
 *   ...                                  ...
 *
 *   <brdf> ::= lambertian <reflectivity> | specular <reflectivity>
 *
 * with `#' comments and blank lines ignored, wavelengths strictly increasing,
 * bands increasing and not overlapping (they may touch), reflectivities in
 * [0, 1].
 *
 * Differences with the reference, all on invalid or unusual input:
 *  - the path and temperature of an htrdr-materials line are split at blanks
 *    (the reference runs them through wordexp(3): no `~', `$VAR' or quote
 *    expansion here);
 *  - a material listed twice keeps its last definition (as the reference's
 *    table does);
 *  - an interface none of whose materials has a BRDF is an error at load
 *    time (the reference asserts when a ray hits it).
 */
#ifndef HTGX_FILES_H
#define HTGX_FILES_H

#include "htgx.h"

#ifdef __cplusplus
extern "C" {
#endif

struct htgx_ground_files;

/* Read the ground of option -g (obj_path), the materials of option -M
 * (materials_path) and, through them, the mrumtl files; sky_mtl is the name
 * given by option -n (NULL: "air"; a name no material has is allowed, then
 * every rectangle-sensor sample whose up ray meets the ground is rejected).
 * Returns HTGX_IO_ERR if a file cannot be read, HTGX_BAD_ARG if one is
 * malformed (htgx_files_last_error() then says where and why). */
int htgx_ground_files_load
  (const char* obj_path,
   const char* materials_path,
   const char* sky_mtl,
   struct htgx_ground_files** out);

/* The ground description; its arrays belong to `files' and stay valid until
 * htgx_ground_files_destroy. */
const struct htgx_ground_desc* htgx_ground_files_get_desc(const struct htgx_ground_files* files);

/* Name of material m (index of the description's materials), NULL if out of
 * range; usemtl group name of interface i likewise. */
const char* htgx_ground_files_material_name(const struct htgx_ground_files* files, uint32_t m);
const char* htgx_ground_files_interface_name(const struct htgx_ground_files* files, uint32_t i);

int htgx_ground_files_destroy(struct htgx_ground_files* files);

/* Message of the last failed load of the calling thread ("" if none):
 * file, line and reason, as the reference logs them. */
const char* htgx_files_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HTGX_FILES_H */
