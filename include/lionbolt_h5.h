/* liblbh5: results in the reference's HDF5 layout.
 *
 * Writes the groups, datasets and attributes LionboltHDF5.f90 writes, with the
 * names of src/IO/schema.yaml, so post-processing written for the reference
 * (Terpdose) reads this build's results unchanged:
 *
 *   /                       attribute problem_type ("general" | "slab")
 *   /mesh                   attributes problem_type, number_of_materials,
 *                           number_of_elements, number_of_spatial_dofs,
 *                           number_of_global_nodes; datasets connectivity
 *                           [NENK] (1-based global node of each spatial dof),
 *                           offset [NE+1], global_nodes [NK][3],
 *                           element_volumes [NE];
 *                           group material_to_spatial_dof with datasets
 *                           material_<m> (1-based spatial dofs of material m)
 *                           (WriteMeshH5, LionboltHDF5.f90:81-208)
 *   /<particle>             attribute object_type = "particle"
 *                           (OpenParticleH5, :15-46), attribute
 *                           number_of_angular_dofs and datasets angular_weights
 *                           [NI], abscissae [NI][3] (AppendOrdinatesH5,
 *                           :210-225), attribute number_of_energy_dofs and
 *                           dataset energy_grid [G+1] (AppendEnergyGridH5,
 *                           :227-238)
 *   /<particle>/angular_fluence/energy_<g>/angle_<i>   [NENK]
 *   /<particle>/uncollided_angular_fluence/energy_<g>/angle_<i>
 *                           (AppendAngularFluence, :282-344)
 *   /<particle>/fluence/energy_<g>, /<particle>/uncollided_fluence/energy_<g>
 *                           [NENK] (AppendSpatialArray, :240-280; Main.f90:448-470)
 *
 * Integers are 32-bit (the Fortran default integer), reals FP64, strings
 * fixed-length.  Two-dimensional Fortran arrays x(3, N) appear as [N][3], as
 * the Fortran HDF5 interface stores them.  Groups g and angles i are 1-based
 * in names, as in the reference.  Every function returns 0 or a negative code
 * with lbh5_last_error() set.  No reference counterpart as a C ABI: the
 * reference writes through NittanyPhysics' HDF5Tree. */
#ifndef LIONBOLT_H5_H
#define LIONBOLT_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbh5 lbh5;

/* Creates (truncates) the file.  OpenHDF5 (core, called at Main.f90:89). */
int lbh5_create(const char* path, lbh5** out);

/* Root attribute problem_type (WriteProblemTypeH5, LionboltHDF5.f90:72-79). */
int lbh5_write_problem_type(lbh5* h, const char* problem_type);

/* The mesh group.  node [NE][4] 0-based global nodes, rg [NK][3], mat [NE]
 * 1-based materials (1..Nm), vol [NE] element volumes. */
int lbh5_write_mesh(lbh5* h, int64_t NE, int64_t NK, const int64_t* node, const double* rg, const int64_t* mat,
                    const double* vol, int slab);

/* The particle group (name lower-cased as OpenParticleH5 does) with its
 * ordinates and energy grid (E may be NULL when G = 0). */
int lbh5_write_particle(lbh5* h, const char* particle, int64_t NI, const double* w, const double* k, int G,
                        const double* E);

/* Angular fluence of group g (1-based): v [NI][NENK] in the reference layout. */
int lbh5_append_angular_fluence(lbh5* h, const char* particle, int uncollided, int g, int64_t NI, int64_t NENK,
                                const double* v);

/* Scalar fluence of group g (1-based): phi [NENK]. */
int lbh5_append_fluence(lbh5* h, const char* particle, int uncollided, int g, int64_t NENK, const double* phi);

int lbh5_close(lbh5* h);
const char* lbh5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
