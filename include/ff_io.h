/*
 * ff_io -- the file formats around the form-factor path, as flat C over the
 * reference-layout structs of ff_plugin.h: Wavefront OBJ scenes with the
 * thermal MTL material blocks, .task frame files, the OBJ-like scene and
 * frame export, and the CSV export of F. Host code only (no device work).
 *
 * Return codes are the reference's: 0, or the negated OBJ_IMPORT_* /
 * CSV_EXPORT_* code (import_export/obj_import.h:31-36,
 * import_export/csv_export.h:26-27). Memory handed out is malloc'ed, as the
 * reference's (free the scene with ffio_scene_free).
 */
#ifndef FF_IO_H
#define FF_IO_H

#include <stdio.h>

#include "ff_plugin.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FFIO_OK 0
#define FFIO_OBJ_IMPORT_FILE_ERROR 21
#define FFIO_OBJ_IMPORT_FORMAT_ERROR 22
#define FFIO_OBJ_IMPORT_MATERIAL_NOT_DEFINED 23
#define FFIO_OBJ_IMPORT_MATERIAL_NOT_ENOUGH_PARAMETERS 24
#define FFIO_OBJ_IMPORT_MATERIAL_INVALID_PARAMETER_VALUE 25
#define FFIO_CSV_EXPORT_FILE_ERROR 31

/* heat_source_equation::params_t (thermal_equation/heat_source_cpu.h:34-39) */
typedef struct {
  int n_sources;
  ffp_heat_source* sources;
} ffp_heat_source_params;

/* obj_import::scene (import_export/obj_import.h:41, obj_import.cpp:196-349):
 * "v x y z" vertices ("vn"/"vt" skipped), triangles "f a b c", "f a//n ...",
 * "f a/t/n ..." (1-based; the first three indices of a longer face), a mesh
 * per "g"/"o" line, "usemtl" material of the current mesh, "mtllib" blocks
 * of 14 thermal properties (newmtl; shell.density ... rear.emissivity).
 * No mesh line: one mesh of every face; no material: the black body.
 * Faces before the first mesh line belong to no mesh. Vertex indices outside
 * the vertex list, "usemtl" before any mesh and a bare "newmtl"/"mtllib"
 * line are format errors (the reference's behaviour there is undefined). */
int ffio_obj_import(const char* filename, ffp_subject_scene** scene);

/* ffio_obj_import with options. FFIO_IMPORT_INLINE_MATERIALS also reads
 * "newmtl" blocks written inside the OBJ file itself, as
 * obj_export::scene writes them (obj_export.cpp:39-60) and as the reference's
 * models/*.obj hold them; the reference's importer treats such a line as an
 * unknown directive and then fails on the property lines. Without the flag
 * this is ffio_obj_import. */
#define FFIO_IMPORT_INLINE_MATERIALS 1
int ffio_obj_import_ex(const char* filename, int flags, ffp_subject_scene** scene);

/* subject::scene_free (subject/system.cpp): the scene and its arrays. */
void ffio_scene_free(ffp_subject_scene* scene);

/* obj_import::task (import_export/obj_import.h:54, obj_import.cpp:458-551):
 * reads lines from `in` until n_meshes "tmprt T" lines have been read
 * ("newfrm" starts a frame and rewinds the temperature index, "step dt",
 * "htsrc mesh power" appends a heat source, "dstsrc power x y z" sets the
 * distant source). t->temperatures and heat_source->sources are realloc'ed;
 * the distant source is written to *distant_source, malloc'ed when NULL
 * (the reference hands it out through its source callback). heat_source or
 * distant_source may be NULL to ignore those lines. Returns the number of
 * temperatures read into the current frame, or a negative code. */
int ffio_task_import(FILE* in, int n_meshes, ffp_ts_task* t, ffp_heat_source_params* heat_source,
                     ffp_source** distant_source);

/* obj_export::scene / task / task_binary (import_export/obj_export.h:29-33,
 * obj_export.cpp:20-110): vertices of every face, the material blocks, a
 * "g mesh_i" group per mesh; frames as "newfrm frame_<n_step>", "step",
 * one "tmprt" line per mesh (or "btmprt n" and n raw floats). */
int ffio_obj_export_scene(FILE* out, const ffp_subject_scene* scene);
int ffio_obj_export_task(FILE* out, int n_meshes, const ffp_ts_task* task);
int ffio_obj_export_task_binary(FILE* out, int n_meshes, const ffp_ts_task* task);

/* csv_export::export_csv (import_export/csv_export.h:30, csv_export.cpp:26-47):
 * one line per source mesh, ';'-separated F values in the default stream
 * format. */
int ffio_csv_export(const char* filename, const ffp_ff_scene* scene, const ffp_ff_task* task);

#ifdef __cplusplus
}
#endif

#endif /* FF_IO_H */
