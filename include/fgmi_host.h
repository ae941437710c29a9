/*
 * fgmi_host.h -- C-ABI of the host side of the LLG engine: mesh input, preprocessing, synthetic
 * mesh generators and .sol files. Pure C++17 on the CPU; no device code.
 *
 * Reference interfaces replaced:
 *   fgh_read_msh        Mesh::mesh::readMesh / readNodes / readTetraedrons / readTriangles
 *                       (src/read.cpp:12-197), gmsh-free, MSH 2.2 and 4.1, ASCII and binary
 *   fgh_prepare         Mesh::mesh ctor: scaling (src/read.cpp:75-90), Tet orientation
 *                       (src/tetra.cpp:410-424), controlTriangles (src/mesh.cpp:121-245),
 *                       sortNodes (src/mesh.cpp:334-367)
 *   fgh_read_sol        Mesh::mesh::readSol (src/read.cpp:199-249)
 *   fgh_write_sol       Mesh::mesh::savesol (src/save.cpp:164-186)
 *   fgh_gen_cuboid      python-modules/meshMaker.py Cuboid (:440-483), 6-tet Kuhn split
 */
#ifndef FGMI_HOST_H
#define FGMI_HOST_H

#include <stdint.h>
#include "fgmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fgh_mesh fgh_mesh;

/* Read a gmsh file (format 2.2 or 4.1, ASCII or little-endian binary). Returns NULL and fills
 * err on failure (including a truncated file). */
fgh_mesh *fgh_read_msh(const char *path, char *err, int errlen);
/* Structured cuboid [x0,x0+lx]x[y0,y0+ly]x[z0,z0+lz] of nx*ny*nz cells, each split into 6
 * tetrahedra; interior node coordinates jittered by U(-jitter, jitter) times the cell size,
 * tangentially on the faces (seeded std::mt19937_64). Volume physical group "volume" (tag 300),
 * boundary triangles physical group "surface" (tag 200). Coordinates in mesh units. */
fgh_mesh *fgh_gen_cuboid(double x0, double y0, double z0, double lx, double ly, double lz, int nx,
                         int ny, int nz, double jitter, uint64_t seed);
/* Ellipsoid of semi-axes (a,b,c) centred at (cx,cy,cz): the cells of a grid of spacing h whose
 * centres lie inside, split as in fgh_gen_cuboid, then boundary nodes projected radially onto
 * the ellipsoid surface. Volume physical group "volume". */
fgh_mesh *fgh_gen_ellipsoid(double cx, double cy, double cz, double a, double b, double c, double h);
/* Cylinder of radius r along z from z0 to z0 + length, axis (cx, cy): the cells of a grid of
 * spacing about h whose centres lie inside the disc, split as in fgh_gen_cuboid, with the
 * staircase rim nodes moved half-way towards the cylinder surface along the radius. Volume
 * physical group "volume". No counterpart in the reference (its meshes come from gmsh). */
fgh_mesh *fgh_gen_cylinder(double cx, double cy, double z0, double r, double length, double h);
/* Append another mesh's nodes and elements (physical tags kept, names merged by tag). */
int fgh_append(fgh_mesh *dst, const fgh_mesh *src);
void fgh_free(fgh_mesh *m);

/* Raw counts of the file/generator content. */
int fgh_raw_counts(const fgh_mesh *m, int64_t *n_nodes, int64_t *n_tets, int64_t *n_tris);
/* Evaluate an expression of the settings file at n points: vars is a comma-separated list of
 * variable names ("t", or "x,y,z"), vals[n * nvars] their values, out[n] the results. The language
 * is the arithmetic subset of the JavaScript FeeLLGood evaluates for initial_magnetization and Bext
 * (src/expression_parser.cpp:9-33: Math members without the prefix, cosh/sinh/tanh, ...), as
 * fgmi_run uses it. Returns FG_OK, or FG_ERR_ARG with the reason (position, unknown name, syntax)
 * in fgh_last_error(). No counterpart in the reference (it runs the expressions in Duktape). */
int fgh_eval_expression(const char *expr, const char *vars, int64_t n, const double *vals, double *out);

/* Write the raw mesh as ASCII MSH 2.2. */
int fgh_write_msh22(const fgh_mesh *m, const char *path);

/*
 * Preprocess like Mesh::mesh's constructor. Tetrahedra are kept when their physical group name
 * is one of vol_names, triangles when theirs is one of surf_names; region index k of the result
 * refers to vol[k] / surf[k]. Missing boundary and interface triangles are created in an extra
 * surface region (index n_surf, default parameters: Ks = 0, charges kept). Scale multiplies the
 * coordinates (mesh.length_unit). Returns FG_OK or FG_ERR_MESH with the reason in
 * fgh_last_error().
 */
int fgh_prepare(fgh_mesh *m, double scale, int n_vol, const char *const *vol_names,
                const fg_volume_region *vol, int n_surf, const char *const *surf_names,
                const fg_surface_region *surf);
/* Fill an fg_mesh view of the prepared mesh (pointers stay valid until fgh_free). */
int fgh_view(const fgh_mesh *m, fg_mesh *out);
/* node_index[file_position] = position after sorting (src/mesh.h:380-385). */
const int32_t *fgh_node_index(const fgh_mesh *m);
/* Bounding box of the prepared mesh: lengths l[3] and centre c[3] (src/mesh.h:51-61). */
int fgh_bbox(const fgh_mesh *m, double l[3], double c[3]);
const char *fgh_last_error(void);

/* .sol files: u and phi per node in file order, header "## time: t". */
int fgh_read_sol(const fgh_mesh *m, const char *path, double *u, double *phi, double *t);
/* fgh_write_sol with the spin accumulation s[3 * n_nodes] in three more columns (sx sy sz) when s
 * is not NULL (Mesh::mesh::savesol with spin_acc, src/save.cpp:164-186). */
int fgh_write_sol_s(const fgh_mesh *m, const char *path, const char *metadata, int precision,
                    const double *u, const double *phi, const double *s);
int fgh_write_sol(const fgh_mesh *m, const char *path, const char *metadata, int precision,
                  const double *u, const double *phi);

#ifdef __cplusplus
}
#endif

#endif
