/*
 * aslhip.h — C ABI of the MI355X (gfx950) LBGK library.
 *
 * This is the drop-in boundary for ASL's lattice Boltzmann hot path. ASL drives
 * that path through C++ objects (asl::LBGK, asl::LBGKUtilities and the LBGK
 * boundary-condition classes) whose device work is OpenCL generated at run time.
 * Here every such object maps to calls on one opaque lattice handle; the C++
 * compatibility layer (asl_amd/compat) and the Python host mirror (asl_amd/) are
 * both thin clients of these entry points.  Each declaration names the reference
 * interface it replaces (paths are relative to the ASL source tree).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative ASLHIP_E* code on
 *     failure; aslhip_last_error() then returns the message (thread local).
 *     ASL throws std::logic_error("ASL ERROR: ...") for the same conditions
 *     (src/aslUtilities.cxx:68-95); the C++ layer rethrows that way.
 *   - Host arrays use ASL's layout: the ghost-inclusive block of
 *     (nx+2)*(ny+2)*(nz+2) cells, linear index x*Sy*Sz + y*Sz + z (last index
 *     fastest, src/data/aslBlocks.h:118-129,189-196), one array per component
 *     (component-major: u is ux[len] then uy[len] then uz[len]; f is f0[len] ...
 *     f(Q-1)[len] in ASL's direction order, src/math/aslTemplates.cxx:146-200).
 *   - Populations are ASL's weight-normalised f_i = F_i / w_i
 *     (src/num/aslLBGK.cxx:46-72).
 *   - Element type of host buffers: float for ASLHIP_F32 / ASLHIP_BF16 /
 *     ASLHIP_F16 lattices, double for ASLHIP_F64 lattices.
 *   - All work of a lattice runs in order on its own HIP stream.  Host reads
 *     synchronise first, so what the host sees is synchronous like ASL's
 *     blocking queue (src/acl/Kernels/aclKernel.cxx:289-313).
 */
#ifndef ASLHIP_H
#define ASLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define ASLHIP_OK 0
#define ASLHIP_EINVAL (-1)    /* bad argument                              */
#define ASLHIP_ESTATE (-2)    /* call not allowed in the handle's state    */
#define ASLHIP_EHIP (-3)      /* HIP runtime error                         */
#define ASLHIP_ENOMEM (-4)    /* allocation failed                         */
#define ASLHIP_EUNDEF (-5)    /* boundary conditions leave a link undefined */

/* ---- storage / arithmetic type of the populations -------------------- */
#define ASLHIP_F32 0  /* FP32 storage, FP32 collision                      */
#define ASLHIP_F64 1  /* FP64 storage and collision (the FP64 GPU oracle)  */
#define ASLHIP_BF16 2 /* BF16 storage of f_i - 1, FP32 collision           */
#define ASLHIP_F16 3  /* FP16 storage of f_i - 1, FP32 collision           */

/* ---- lattice flags (aslhip_lattice_create) --------------------------- */
#define ASLHIP_PERIODIC_X 1u
#define ASLHIP_PERIODIC_Y 2u
#define ASLHIP_PERIODIC_Z 4u

/* ---- faces: bit (1 << asl::SlicesNames), src/num/aslBCond.h:309 ------ */
#define ASLHIP_X0 1u
#define ASLHIP_XE 2u
#define ASLHIP_Y0 4u
#define ASLHIP_YE 8u
#define ASLHIP_Z0 16u
#define ASLHIP_ZE 32u

/* ---- boundary-condition kinds ---------------------------------------- */
/* slice BCs: src/num/aslLBGKBC.cxx:91-202, factories :745-798            */
#define ASLHIP_BC_NOSLIP 1                   /* f_k(g) = f_kbar(g+c_k)           */
#define ASLHIP_BC_CONSTANT_PRESSURE 2        /* f_k(g) = p                       */
#define ASLHIP_BC_CONSTANT_VELOCITY 3        /* f_k(g) = f_kbar(g+c_k) + 6 u.c_k */
#define ASLHIP_BC_CONSTANT_PRESSURE_VELOCITY 4 /* f_k(g) = p + 3 u.c_k          */
/* map (level-set) BCs: src/num/aslLBGKBC.cxx:219-432, factories :801-875 */
#define ASLHIP_BC_NOSLIP_MAP 5               /* f_i(g) = f_ibar(g+c_i)           */
#define ASLHIP_BC_PRESSURE_VELOCITY_MAP 6    /* f_i(g) = feq(p, u)               */
#define ASLHIP_BC_VELOCITY_MAP 7             /* f_i(g) = rho_g + 3 rho_g u.c_i   */

/* ---- step flags ------------------------------------------------------ */
#define ASLHIP_STEP_WRITE_MACRO 1u /* last step of the call writes rho and u */
#define ASLHIP_STEP_DEFER 2u       /* keep the last step pending: it runs, writing
                                      rho and u, when the host next reads or
                                      synchronises, or before the next step */
#define ASLHIP_STEP_GRAPH 4u       /* run the plain steps (all but a first step
                                      that reads raw ghost values and a last
                                      one that writes rho and u) as replays of
                                      captured hipGraphs: chunks of
                                      ASLHIP_GRAPH_CHUNK steps, then one run of
                                      the remainder (captured on first use) */
#define ASLHIP_GRAPH_CHUNK 100

typedef struct aslhip_lattice aslhip_lattice;

/* Message of the last failed call on this thread (empty if none).  Returns
 * the message length. */
int aslhip_last_error(char *buf, size_t n);

/* Library version and the device it runs on (name written to buf). */
int aslhip_device_info(char *buf, size_t n, int *cu_count, double *hbm_bytes);

/* asl::d3q15() / d3q19() / d3q27() — src/math/aslTemplates.cxx:146-200:
 * the direction vectors c (q x 3, ASL's order), weights w (q) and inverse
 * directions inv (q) of the template this library uses for q (any output
 * may be NULL).  D3Q27 weights are the textbook ones (ASL defines none). */
int aslhip_template(int q, int *c, double *w, int *inv);

/* asl::LBGK::LBGK(Block, Param nu, const VectorTemplate*, bool compVel,
 * bool compRho, CommandQueue)  — src/num/aslLBGK.h:112-116,
 * src/num/aslLBGK.cxx:114-158.  nx, ny, nz: internal block size; q: 15, 19
 * or 27 (d3q15()/d3q19()/d3q27(), src/math/aslTemplates.cxx:146-200);
 * dtype: ASLHIP_F32/F64/BF16/F16; flags: ASLHIP_PERIODIC_* (periodic wrap is
 * an extension: ASL has none). */
int aslhip_lattice_create(aslhip_lattice **out, int nx, int ny, int nz, int q,
                          int dtype, unsigned flags);

/* Releases all device memory of the handle. */
int aslhip_lattice_destroy(aslhip_lattice *h);

/* Viscosity nu in lattice units; relaxation rate w = 1/(3 nu + 1/2) is
 * evaluated in the lattice's arithmetic type like ASL's generated kernel
 * (src/num/aslLBGK.cxx:203). */
int aslhip_set_viscosity(aslhip_lattice *h, double nu);

/* asl::LBGK::setCompressible — src/num/aslLBGK.h:129, equilibrium variants
 * at src/num/aslLBGK.cxx:220-231. */
int aslhip_set_compressible(aslhip_lattice *h, int compressible);

/* asl::LBGKTurbulence — src/num/aslLBGK.h:163-171,
 * src/num/aslLBGK.cxx:298-355: per cell, r = 4 |f - feq|^2 / |f|^2 and the
 * relaxation 1 / (3 min(nu r, 1/2) + 1/2) where r > 1 (else the plain one).
 * ASL reads feq there before assigning it (undefined); feq here is the
 * equilibrium of the cell's rho and momentum.  on = 0 returns to asl::LBGK. */
int aslhip_set_turbulence(aslhip_lattice *h, int on);

/* asl::LBGK::setOmega(Param w) — src/num/aslLBGK.h:120,
 * src/num/aslLBGK.cxx:207-208,252-255.  Rotating frame with angular velocity
 * omega[3] (lattice units): each collision first sets
 * momentum += (2/w) omega x momentum, component by component.  NULL turns
 * it off. */
int aslhip_set_omega(aslhip_lattice *h, const double omega[3]);

/* The lattice's asl::Block geometry: the internal block's dx and position
 * (asl::LBGK(Block b, ...), src/num/aslLBGK.cxx:114-127 and createData; the populations live
 * on offset(b, 1), whose node I sits at position - dx + dx * I).  Velocity-map
 * BCs with a rotation field evaluate omega x (r - centre) at these positions
 * (Index2PositionACL over the populations' block, src/num/aslLBGKBC.cxx:303-305,
 * src/math/aslIndex2Position.cxx:64-108).  Default: dx 1, position 0. */
int aslhip_set_geometry(aslhip_lattice *h, double dx, const double position[3]);

/* Extension (no counterpart in ASL), an opt-in defined value: velocity-map
 * BCs (src/num/aslLBGKBC.cxx:290-348) average the density over the
 * computation-node neighbours of a boundary node; ASL takes neighbours in
 * the block's outer ghost layer too and reads whatever its pool memory holds
 * there.  on = 1: such neighbours are left out of the average (and of its
 * weight sum); on = 0 (default): a lattice that would read them is refused
 * with ASLHIP_EUNDEF.  The asl:: layer sets 1 when the environment variable
 * ASLHIP_VMAP_SKIP_OUTER is 1. */
int aslhip_set_vmap_skip_outer(aslhip_lattice *h, int on);

/* compRho / compVel constructor flags (src/num/aslLBGK.cxx:210-218). */
int aslhip_set_compute_macro(aslhip_lattice *h, int comp_rho, int comp_vel);

/* asl::LBGKUtilities::initF(rho, vel) with constants —
 * src/num/aslLBGK.cxx:262-280.  Initialises every cell of the ghost-inclusive
 * block and writes rho and u. */
int aslhip_init_equilibrium(aslhip_lattice *h, double rho, const double u[3]);

/* asl::LBGKUtilities::initF(rho, vel) with per-cell fields (host arrays in
 * ASL layout; rho may be NULL meaning 1). */
int aslhip_init_equilibrium_field(aslhip_lattice *h, const void *rho,
                                  const void *u);

/* Slice BCs: asl::generateBCNoSlip / generateBCConstantPressure /
 * generateBCConstantVelocity / generateBCConstantPressureVelocity(lbgk, ...,
 * {slices}) — src/num/aslLBGKBC.h:281-300, src/num/aslLBGKBC.cxx:745-798,
 * point lists src/num/aslBCond.cxx:373-510.  faces: ASLHIP_X0 | ... ;
 * p and u as the factory's arguments (ignored where the kind has none). */
int aslhip_bc_add_slice(aslhip_lattice *h, int kind, unsigned faces, double p,
                        const double u[3], int *bc_id);

/* Map BCs: asl::generateBCNoSlip(lbgk, map), generateBCConstantPressure(lbgk,
 * p, map), generateBCConstantPressureVelocity(lbgk, p, u, map),
 * generateBCVelocity(lbgk, v, map[, computationalDomain]) —
 * src/num/aslLBGKBC.h:302-321, src/num/aslLBGKBC.cxx:801-875.  map and
 * cdomain: host float arrays of the ghost-inclusive block (cdomain may be
 * NULL).  For ASLHIP_BC_VELOCITY_MAP the wall velocity at position r is
 * u + omega x (r - centre) (asl::generatePFConstant / PFRotationField,
 * src/math/aslPositionFunction.cxx); rot = {omega[3], centre[3]} or NULL. */
int aslhip_bc_add_map(aslhip_lattice *h, int kind, const float *map,
                      const float *cdomain, double p, const double u[3],
                      const double rot[6], int *bc_id);

/* BC execute(): asl::BCLBGKCommon::execute / BCNoSlipMap::execute ...
 * (src/num/aslLBGKBC.cxx:80-83,257-260).  The fused step applies every BC
 * itself; this call records the driver's execute order, which decides who
 * wins where two BCs fill the same ghost link (the last one executed). */
int aslhip_bc_execute(aslhip_lattice *h, int bc_id);

/* Builds the packed cell-type map and the link-override tables from the BCs
 * and the current execute order.  Called implicitly by aslhip_step. */
int aslhip_finalize(aslhip_lattice *h);

/* asl::LBGK::execute() (preProcessing + collide, src/num/aslSingleKernelNM.cxx:
 * 48-53) followed by the BCs, nsteps times.  flags: ASLHIP_STEP_*.  Runs
 * asynchronously on the lattice's stream. */
int aslhip_step(aslhip_lattice *h, int nsteps, unsigned flags);

/* Blocks until the lattice's stream is idle. */
int aslhip_synchronize(aslhip_lattice *h);

/* asl::ComputeSurfaceForceMap / ComputeSurfaceFluxMap execute() —
 * src/num/aslLBGKBC.cxx:623-736, factories generateComputeSurfaceForce /
 * Flux (src/num/aslLBGKBC.h:329-336).  Momentum (force, 3 * len elements) or
 * mass (flux, len) exchanged across the links between the ghost nodes
 * (map <= 0) and computation nodes (map > 0) of map (len floats, ASL
 * layout), from the populations after the last step and its BCs; host
 * arrays of the host element type.  Cells with |map| >= 0.9999 get 0. */
int aslhip_surface_force(aslhip_lattice *h, const float *map, void *out);
int aslhip_surface_flux(aslhip_lattice *h, const float *map, void *out);

/* Visualisation fills of the lattice's rho / u at the boundary nodes of a map
 * (cells with map <= 0 and map > -0.9999), done on the device in place on the
 * values the last macro-writing step left (host reads return them):
 *   NOSLIP_VEL   generateBCNoSlipVel (src/num/aslLBGKBC.cxx:808-815) ->
 *                BCConstantValueMiddlePointMap(velocity, 0, map)
 *                (src/num/aslBasicBC.cxx:167-227): u(x) = -sum_i lc_i u(x+c_i)
 *                / max(sum_i lc_i, .01) over neighbours with map > 0;
 *   NOSLIP_RHO   generateBCNoSlipRho (:818-825) -> BCConstantGradientMap(rho,
 *                0, map) (src/num/aslBasicBC.cxx:283-351): rho(x) = weighted
 *                mean of rho over neighbours with map > 0, weights
 *                (n . c_i/|c_i| + 1)^2, n the normalised map gradient; cells
 *                whose weight sum is <= 0.1 keep their value;
 *   VELOCITY     generateBCVelocityVel (:845-850) -> BCValuePFMap(velocity,
 *                v, map) (src/num/aslBasicBC.cxx:124-145): u(x) = u (constant)
 *                or axis x (position(x) - centre) (rot = {axis, centre}), at
 *                the positions set by aslhip_set_geometry.
 * Map reads outside the block are -0.9; field reads wrap the linear index.
 * NOSLIP_VEL / NOSLIP_RHO need d3q15 or d3q19 (ASL's d3q27 has no stencil
 * coefficients).  aslhip_vis_add copies the map to the device (once); each
 * aslhip_vis_execute (an asl::NumMethod::execute of the fill) queues one
 * kernel on the lattice's stream. */
#define ASLHIP_VIS_NOSLIP_VEL 1
#define ASLHIP_VIS_NOSLIP_RHO 2
#define ASLHIP_VIS_VELOCITY 3
int aslhip_vis_add(aslhip_lattice *h, int kind, const float *map, const double u[3], const double rot[6],
                   int *vis_id);
int aslhip_vis_execute(aslhip_lattice *h, int vis_id);

/* asl::initData(map, DF) / generateDataContainer_SP(block, DF, gN, type) —
 * src/aslGeomInc.cxx:187-221 with the DistanceFunction tree of
 * src/math/aslDistanceFunction.cxx:55-164: the map of a distance-function
 * expression on a grid of size[3] nodes, node (i, j, k) at pos + dx * (i, j, k)
 * evaluated in the element type (fp64 ? double : float; constants rounded to
 * it where they are used, every operation in ASL's order, rounded once).
 * prog: nwords doubles in postfix order, each op code followed by its
 * parameters:
 *   SPHERE c0 c1 c2 r            sqrt(l2(p - c)) - r
 *   CYLINDER c0 c1 c2 o0 o1 o2 r  sqrt(l2(d - (d.o) o)) - r, d = p - c, |o| = 1
 *   CONE a0 a1 a2 o0 o1 o2 t      sqrt(l2(d - (d.o) o)) - t (d.o), d = p - a
 *   PLANE n0 n1 n2 b              p.n - b
 *   AND / OR                      max / min of the top two values
 *   NEG                           -x
 *   NORMALIZE f                   clamp(f x, -1, 1)
 * out: host array of size[0] * size[1] * size[2] elements, last index fastest. */
#define ASLHIP_DF_SPHERE 1
#define ASLHIP_DF_CYLINDER 2
#define ASLHIP_DF_CONE 3
#define ASLHIP_DF_PLANE 4
#define ASLHIP_DF_AND 5
#define ASLHIP_DF_OR 6
#define ASLHIP_DF_NEG 7
#define ASLHIP_DF_NORMALIZE 8
int aslhip_df_map(const double *prog, int nwords, const int size[3], const double pos[3], double dx, int fp64,
                  void *out);

/* asl::readSurface(file, ...) — src/readers/aslVTKFormatReaders.h:66-78,
 * src/readers/aslVTKFormatReaders.cxx:266-305: the triangles of a binary or
 * ASCII STL file, 9 doubles each (3 vertices, the file's x y z), rounded to
 * float as VTK's reader stores them.  count gives the number, read fills
 * ntri * 9 doubles. */
int aslhip_stl_count(const char *path, long long *ntri);
int aslhip_stl_read(const char *path, double *tri, long long ntri);

/* asl::readSurface(file, ...) for VTK XML PolyData files —
 * src/readers/aslVTKFormatReaders.cxx:219-265 (vtkXMLPolyDataReader, then
 * the same surfaceToData): the polygon and triangle-strip cells of every
 * piece as triangles (a polygon as the fan from its first point, a strip as
 * its consecutive triples), 9 doubles each (the file's x y z, points as
 * stored: Float32 or Float64).  DataArrays in ascii, binary or appended
 * (raw or base64) form, UInt32 or UInt64 headers, zlib-compressed or not,
 * either byte order.  Vertex and line cells are skipped. */
int aslhip_vtp_count(const char *path, long long *ntri);
int aslhip_vtp_read(const char *path, double *tri, long long ntri);

/* surfaceToData — src/readers/aslVTKFormatReaders.cxx:132-217 (VTK's
 * vtkImplicitModeller, capped, then rescaled): on a grid of size[3] nodes,
 * node (i, j, k) at pos + h * (i, j, k) (ASL axis order, last index fastest),
 *   d2  = min(cap^2, min over triangles of |node - T|^2 rounded to float),
 *   map = ((sqrt(d2) - shift) / div1) / div2  in float (constants rounded to float).
 * ASL uses cap = 3.6 dx, div1 = dx, div2 = 1.8 and shift = 1.2 dx (Block
 * variant) or 1.8 dx (dx variants).  tri: ntri * 9 doubles in ASL axis order
 * (readSurface's grid takes the file's x as ASL's last axis); map: host,
 * size[0] * size[1] * size[2] floats. */
int aslhip_surface_distance_map(const double *tri, long long ntri, const int size[3], const double pos[3],
                                const double h[3], double cap, double shift, double div1, double div2, float *map);

/* Writes the populations from a host view in the layout aslhip_read_f
 * returns (Q * len elements of the host element type, f_i of each cell,
 * weight-normalised as ASL stores them) and sets rho and u from them: the
 * write side of asl::LBGK::getF() (src/num/aslLBGK.h:125-127), e.g. to
 * restore a checkpoint.  Like initF, the lattice then starts afresh. */
int aslhip_write_f(aslhip_lattice *h, const void *f);

/* asl::LBGK::getRho() / getVelocity() / getF() host views —
 * src/num/aslLBGK.h:125-127.  out: len (rho), 3*len (u) or Q*len (f)
 * elements of the host element type.  Cells that are not computation nodes
 * hold unspecified values (ASL's hold collided ghost data). */
int aslhip_read_rho(aslhip_lattice *h, void *out);
int aslhip_read_velocity(aslhip_lattice *h, void *out);

/* Device addresses of rho (one array) and u (three arrays, component-major),
 * in the ghost-inclusive ASL layout and the element type aslhip_read_rho /
 * aslhip_read_velocity copy (double for ASLHIP_F64 lattices, float
 * otherwise), holding what those reads would return (a deferred step runs
 * first).  For a solver coupled on the GPU to lbgk->getVelocity()
 * (examples/flow/multiphase_flow.cc:182-184, which hands it to
 * asl::generateFDMultiPhase).  rho or u may be NULL.
 * Ordering: stream is the consumer's hipStream_t.  Work queued on it after
 * this call sees the values (it waits for the lattice's stream, the host does
 * not block).  The next call that overwrites them on this lattice (a step,
 * initF, write_f, a batch step) first waits for everything queued on that
 * stream up to the moment of that call; work the consumer queues later may
 * see the newer values.  stream NULL: the host synchronises with the
 * lattice's stream instead, and the caller must not read them after the next
 * step is requested. */
int aslhip_macro_device(aslhip_lattice *h, void *stream, const void **rho, const void **u);
int aslhip_read_f(aslhip_lattice *h, void *out);

/* No counterpart in ASL.  aslhip_read_f for callers that cannot allocate
 * device memory (tight host loops that checkpoint, harnesses that forbid
 * hipMalloc): the view is un-permuted through a fixed staging array of
 * ASLHIP_NOALLOC_SCRATCH_BYTES in the library's device code (one per GPU,
 * held until the process ends) instead of a buffer allocated per call, in
 * pieces of at most that size, so any lattice fits; the only limit, one
 * cell's Q values (216 bytes at most), is checked.  Same output as
 * aslhip_read_f.  The lattice's device must be the current one.  A pending
 * step runs first, as for aslhip_read_f; one whose boundary tables must be
 * rebuilt (BCs registered since the last step) is refused, since that
 * set-up allocates: call aslhip_finalize first.  The staging array keeps the
 * previous call's contents; calls that overlap on one GPU (from several host
 * threads) race on it.  It synchronises the host with the lattice's stream,
 * so it does not belong under graph capture. */
#define ASLHIP_NOALLOC_SCRATCH_BYTES (16u << 20)
int aslhip_read_f_noalloc(aslhip_lattice *h, void *out);

/* Packed cell-type map (one byte per ghost-inclusive cell):
 * 0 fluid, 1 fluid with BC links, 2 not computed (ghost/solid),
 * 3 computed ghost.  Bits 4-7 unused. */
int aslhip_read_cell_types(aslhip_lattice *h, uint8_t *out);

/* FP64 reductions over computation nodes: out = {mass, mx, my, mz, KE}
 * with mass = sum rho, m = sum rho u, KE = sum |m|^2 / (2 rho). */
int aslhip_reduce_conservation(aslhip_lattice *h, double out[5]);

/* Counts: computation nodes (counted in MLUPS), computed cells, boundary
 * cells, distinct link patterns. */
int aslhip_stats(aslhip_lattice *h, int64_t out[4]);

/* Launch tuning of the step kernel (no effect on results).
 *   ASLHIP_TUNE_XCD_REMAP: g > 1 = each of the 8 XCDs takes runs of g
 *   consecutive chunks (neighbouring chunks share one L2; default 4, 64 for
 *   FP16 / BF16 storage); 0 or 1 = the hardware round-robin order.
 *   ASLHIP_TUNE_BLOCK: workgroup size, a multiple of 64 up to 1024 (default
 *   512); a workgroup takes that many consecutive cells of a plane. */
#define ASLHIP_TUNE_XCD_REMAP 1
#define ASLHIP_TUNE_BLOCK 2
/* Batches only: graph chunks step groups of consecutive lattices of at most
 * this many MiB of populations one group at a time (0: all lattices each
 * step), so that a group can stay in the 256 MiB Infinity Cache. */
#define ASLHIP_TUNE_GROUP_MB 3
/* Lattices: 1 = address the populations with 64-bit offsets (the variant
 * that lattices over 4 GiB of populations use) whatever the size; 0 = by
 * size (default).  For testing that variant on small lattices. */
#define ASLHIP_TUNE_WIDE 4
/* Lattices: cells per lane along z of the step kernel: 1, 2 or 4, or 0 =
 * automatic (default: one cell per lane).  Runs of 2 or 4 cells apply to
 * non-periodic lattices under 4 GiB without a per-cell relaxation whose nz
 * is a multiple of the run (else the longest run that divides nz, down to
 * one cell) and take precedence over ASLHIP_TUNE_LDS; results are bit for
 * bit the same for every value. */
#define ASLHIP_TUNE_CPL 5
/* Lattices: 1 = the step kernel stages the populations of tiles of 256
 * consecutive cells through LDS (16-byte loads and stores of each
 * direction's slots for FP32, workgroup 256); 2 = tiles on even steps only,
 * odd steps direct; 0 = direct, one cell per lane (or the runs
 * ASLHIP_TUNE_CPL asks for).  Default: 1 for FP32 storage over 256 MiB of
 * populations (lattices the 256 MiB Infinity Cache cannot hold), else 0.  Tiles apply to the same lattices as runs of cells
 * when ASLHIP_TUNE_CPL is 0 or 1; results are bit for bit the same for
 * every value. */
#define ASLHIP_TUNE_LDS 6
int aslhip_set_tuning(aslhip_lattice *h, int knob, int value);
/* The value a tuning knob has now (ASLHIP_TUNE_CPL: the cells per lane the
 * step kernel runs with; ASLHIP_TUNE_LDS: 0, 1 or 2 as it runs LDS-staged
 * tiles on no step, every step or even steps). */
int aslhip_tuning_value(aslhip_lattice *h, int knob, int *value);
/* The workgroup size the step kernel runs with (runs of 2 or 4 cells per lane:
 * at most 512; LDS-staged tiles on every step: 256; tiles on even steps only:
 * the odd steps' workgroup) and the XCD run length (either may be NULL). */
int aslhip_get_tuning(aslhip_lattice *h, int *block, int *remap);

/* The autotuner: times nsteps (rounded up to even) real steps of the lattice
 * for every candidate setting (where the lattice can run them: LDS-staged
 * tiles on every step x XCD runs 0/2/4/8/16/64, tiles on even steps x XCD
 * runs x odd-step workgroup 256/512/1024; direct with cells per lane 1/2/4
 * where it can run them x XCD runs x workgroup 256/512/1024) with HIP
 * events, keeps the fastest and returns its workgroup size and XCD run
 * length (either pointer may be NULL; the variant: aslhip_tuning_value).
 * The trial steps advance the simulation like any other steps. */
int aslhip_autotune(aslhip_lattice *h, int nsteps, int *best_block, int *best_remap);

/* Times nsteps steps with HIP events recorded on the lattice's stream.
 * ms_total: wall time of the whole call between the events; ms_kernel:
 * summed time of the step kernels (events around each launch; 0 with
 * ASLHIP_STEP_GRAPH, whose replays carry no events between launches);
 * ms_each: nsteps doubles, the duration of each launch (0 with
 * ASLHIP_STEP_GRAPH).  Any output may be NULL. */
int aslhip_time_steps(aslhip_lattice *h, int nsteps, unsigned flags,
                      double *ms_total, double *ms_kernel, double *ms_each);

/* ---- learned closure ------------------------------------------------
 * Extension (no counterpart in ASL): a per-cell relaxation-time correction
 * from a 7-32-32-1 network of the pre-collision non-equilibrium moments
 *   x = (Pi_xx, Pi_yy, Pi_zz, Pi_xy, Pi_xz, Pi_yz, rho - 1),
 *   Pi_ab = sum_i w_i c_ia c_ib f_i - rho/3 delta_ab - m_a m_b / rho,
 *   xn = (x - mu) * sc, h1 = relu(W1 xn + b1), h2 = relu(W2 h1 + b2),
 *   y = w3.h2 + b3, tau = max(tau0 + y * out_scale, 1/2 + (tau0 - 1/2) / 2)
 *   with tau0 = 3 nu + 1/2 (the correction lowers the viscosity by at most half).
 * params (1345 floats): W1 (32x7, row per unit), b1, W2 (32x32), b2, w3, b3;
 * rounded to BF16 and evaluated in FP32 inside the step kernel.  D3Q19
 * lattices with FP32 collision (FP32, BF16 or FP16 storage); params NULL
 * switches back to plain LBGK. */
#define ASLHIP_CLOSURE_NPARAM 1345
int aslhip_closure_set(aslhip_lattice *h, const float *params, const float mu[7], const float sc[7],
                       float out_scale);
/* How the step kernel evaluates the network: ASLHIP_CLOSURE_BF16_MFMA
 * (default) — BF16 inputs and hidden activations on the matrix cores
 * (v_mfma_f32_32x32x8bf16_1k, FP32 accumulation), the output layer in FP32;
 * ASLHIP_CLOSURE_FP32 — every product and sum in FP32 on the vector ALUs, in
 * a fixed order (reproducible bit for bit on the CPU). */
#define ASLHIP_CLOSURE_FP32 0
#define ASLHIP_CLOSURE_BF16_MFMA 1
int aslhip_closure_set_mode(aslhip_lattice *h, int mode);

/* The network inputs x of every interior cell for the next step (periodic
 * D3Q19 lattices): out holds 7 floats per cell, cells in ASL order. */
int aslhip_closure_features(aslhip_lattice *h, float *out);

/* Trainer of the closure network on one GPU (FP32 master weights, Adam:
 * Kingma & Ba 2015, beta1 0.9, beta2 0.999, eps 1e-8).  Loss: half the mean
 * squared error of y against the targets. */
typedef struct aslhip_mlp aslhip_mlp;
int aslhip_mlp_create(aslhip_mlp **out, const float *params);
int aslhip_mlp_destroy(aslhip_mlp *m);
/* Training set: x (n x 7, already normalised: xn above) and targets t (n);
 * copied to the device. */
int aslhip_mlp_set_data(aslhip_mlp *m, const float *x, const float *t, long long n);
/* Gradient (1345 floats) and loss over samples first .. first+count-1. */
int aslhip_mlp_gradient(aslhip_mlp *m, long long first, int count, float *grad, double *loss);
/* steps Adam steps on batches of `batch` samples drawn (with replacement) by
 * a counter-based hash of seed; losses (may be NULL): loss of every batch. */
int aslhip_mlp_train(aslhip_mlp *m, int steps, int batch, double lr, unsigned long long seed, double *losses);
int aslhip_mlp_get_params(aslhip_mlp *m, float *out);

/* ---- batches of independent lattices ---------------------------------
 * Extension (ASL advances each lattice with its own blocking launches):
 * many lattices of uneven sizes, each with its own viscosity and BCs, advance
 * together, one launch of the step kernel per time step for all of them.
 * All lattices of a batch share the template, population type and
 * periodicity, and have no velocity-map BCs.  ASLHIP_STEP_GRAPH replays
 * hipGraph-captured chunks of ASLHIP_GRAPH_CHUNK steps. */
typedef struct aslhip_batch aslhip_batch;
int aslhip_batch_create(aslhip_batch **out, aslhip_lattice *const *lattices, int n);
int aslhip_batch_destroy(aslhip_batch *b);
/* nsteps steps of every lattice (flags: ASLHIP_STEP_WRITE_MACRO,
 * ASLHIP_STEP_GRAPH).  Lattices' own reads see the result in order. */
int aslhip_batch_step(aslhip_batch *b, int nsteps, unsigned flags);
/* As aslhip_batch_step, timed with HIP events on the batch's stream. */
int aslhip_batch_time_steps(aslhip_batch *b, int nsteps, unsigned flags, double *ms_total);
/* Workgroup size (ASLHIP_TUNE_BLOCK), LDS-staged tiles (ASLHIP_TUNE_LDS 1
 * or 0; default: tiles for non-periodic FP32 batches over 256 MiB of
 * populations) or group
 * size (ASLHIP_TUNE_GROUP_MB) of the batch's launches; the autotuners choose
 * between tiles and the direct kernel's workgroup size (256/512/1024,
 * timing nsteps real eager steps each; returns 256 for tiles) and the group
 * size (0, 64 .. 384 MiB, timing `chunks` real graph chunks each). */
int aslhip_batch_set_tuning(aslhip_batch *b, int knob, int value);
int aslhip_batch_autotune(aslhip_batch *b, int nsteps, int *best_block);
int aslhip_batch_autotune_groups(aslhip_batch *b, int chunks, int *best_mb);

#ifdef __cplusplus
}
#endif

#endif /* ASLHIP_H */
