/*
 * lbgk_oracle.h — CPU oracle for the LBGK hot path (test infrastructure only).
 *
 * A plain C++ restatement of ASL's asl::LBGK and its LBGK boundary
 * conditions, used by tests/, by the smoke check and as bench.py's CPU
 * baseline.  It is never part of the shipped GPU path.
 *
 * What it restates (paths relative to the ASL source tree):
 *   - lattice templates d3q15/d3q19 (src/math/aslTemplates.cxx:30-83,146-185);
 *     d3q27 weights are an extension (ASL's d3q27 has none, :187-200)
 *   - pool/window streaming: createData, createCopyKernels, preProcessing
 *     (src/num/aslLBGK.cxx:129-185,237-250), including the linear-index wrap
 *     and the stale pool memory (filled with NaN here so that any dependence
 *     on it shows up)
 *   - the collision kernel LBGK::init0 (src/num/aslLBGK.cxx:187-235) over every
 *     cell of the ghost-inclusive block, in ASL's operation order
 *   - LBGKUtilities::initF (src/num/aslLBGK.cxx:262-280)
 *   - slice BCs with their point lists (src/num/aslBCond.cxx:373-510,
 *     src/num/aslLBGKBC.cxx:63-202) and map BCs with the out-of-block-safe
 *     views and the map guard (src/num/aslLBGKBC.cxx:219-432,
 *     src/aslGenerators.cxx:74-93, src/acl/aclMath/aclMathAlg.cxx:76-94)
 *   - periodic wrap (extension; ASL has none) as plain pull streaming.
 *
 * All host I/O is double; arithmetic is float or double as chosen at create.
 */
#ifndef LBGK_ORACLE_H
#define LBGK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

void *orc_create(int nx, int ny, int nz, int q, int fp64, unsigned periodic);
void orc_destroy(void *h);
int orc_set_threads(int n);
void orc_set_viscosity(void *h, double nu);
void orc_set_compressible(void *h, int c);
void orc_set_omega(void *h, const double *omega); /* NULL: off */
/* LBGKTurbulence's relaxation (fEq: the equilibrium of the cell's moments). */
void orc_set_turbulence(void *h, int on);
/* Learned closure (FP32 D3Q19 lattices): 1345 parameters (BF16-rounded
 * here), input shift/scale (7 each), output scale; params NULL: off.
 * Returns -1 for other lattices. */
int orc_set_closure(void *h, const float *params, const float *mu, const float *sc, float out);
/* ComputeSurfaceForceMap (flux 0: out 3 * len) / ComputeSurfaceFluxMap (flux 1: out len). */
void orc_surface(void *h, const float *map, int flux, double *out);
int orc_vis_fill(void *h, int kind, const float *map, const double u[3], const double rot[6]);
/* 0: the FP32 evaluation, 1: the BF16 evaluation of the GPU's matrix cores (default). */
void orc_set_closure_mode(void *h, int mode);
/* Closure features (7 per interior cell) of a periodic D3Q19 lattice. */
int orc_closure_features(void *h, float *out);
void orc_set_compute_macro(void *h, int comp_rho, int comp_vel);
void orc_set_geometry(void *h, double dx, const double pos[3]);
/* Extension: velocity-map density averages skip outer-ghost-layer neighbours. */
void orc_set_vmap_skip_outer(void *h, int on);
void orc_init_equilibrium(void *h, double rho, const double u[3]);
void orc_init_equilibrium_field(void *h, const double *rho, const double *u);
int orc_bc_add_slice(void *h, int kind, unsigned faces, double p, const double u[3]);
int orc_bc_add_map(void *h, int kind, const float *map, const float *cdomain,
                   double p, const double u[3], const double rot[6]);
int orc_bc_execute(void *h, int id);
void orc_lbgk_execute(void *h);
void orc_run(void *h, int nsteps, const int *order, int norder);
void orc_read_rho(void *h, double *out);
void orc_read_velocity(void *h, double *out);
void orc_read_f(void *h, double *out);
void orc_write_f(void *h, const double *in);
long orc_slice_entries(void *h, int id);
void orc_map_masks(void *h, int id, uint8_t *out);
void orc_conservation(void *h, const uint8_t *mask, double out[5]);
int orc_template(int q, int *c, double *w, int *inv);

/* Surface distance maps (surface_map.cpp): squared point-triangle distance,
 * and readSurface's map of ntri triangles (9 doubles each) on a grid of
 * size[3] nodes at pos + h * index: map = ((sqrt(min(cap^2, d^2)) - shift) / div1) / div2. */
double orc_triangle_d2(const double p[3], const double t[9]);
/* Distance-function map (df_map.cpp) of a postfix program (include/aslhip.h's
 * ASLHIP_DF_* op codes) on size[3] nodes at pos + dx * index, in float or double. */
void orc_df_map(const double *prog, int nwords, const int size[3], const double pos[3], double dx, int fp64,
                void *out);
void orc_surface_distance_map(const double *tri, long long ntri, const int size[3], const double pos[3],
                              const double h[3], double cap, double shift, double div1, double div2, float *map);

#ifdef __cplusplus
}
#endif
#endif
