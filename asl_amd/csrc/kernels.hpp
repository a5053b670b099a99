// kernels.hpp — argument blocks and launchers of the gfx950 LBGK kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "closure.hpp"
#include "lattice_consts.hpp"

namespace aslhip {

// Threads per workgroup of the step kernel (each thread one cell), at most;
// of its variant with runs of 2 or 4 cells per lane, at most; of its
// LDS-staged variant (one tile of STEP_TILE cells per workgroup), exactly.
constexpr int STEP_BLOCK = 1024;
constexpr int STEP_RUN_BLOCK = 512;
constexpr int STEP_TILE = 256;

// Step-kernel flags.
enum : unsigned {
  STEP_F_RHO = 1u,           // write rho of computed cells
  STEP_F_VEL = 2u,           // write u of computed cells
  STEP_F_RAW = 4u,           // read ghost links raw (no BC resolution)
  STEP_F_COMPRESSIBLE = 8u,  // compressible equilibrium (ASL default)
  STEP_F_OMEGA = 16u,        // rotating frame: momentum += (2/w) omega x momentum
  STEP_F_CLOSURE = 32u,      // per-cell relaxation time from the closure network
  STEP_F_CLOSURE_MFMA = 64u, // ... evaluated with BF16 matrix instructions
  STEP_F_TURB = 128u,        // LBGKTurbulence's relaxation limiter
  STEP_F_COUNT = 256u,       // number of flag combinations
};

// One fused collide+stream pass (AA pattern) over the ghost-inclusive block.
// The launch is a 1-D grid of chunks of `block` lanes of the interior planes
// x = 1..nx: with cpl = 1 a lane is one cell of rows y = 1..ny, every z
// including the two ghost columns; with cpl = 2 or 4 a lane is a run of cpl
// interior cells z0 .. z0 + cpl - 1 (nz a multiple of cpl).
//
// Links into cells of the six face layers are resolved from the cell's
// coordinates: for each face, fface[f] = slot * 4 + mode of the last
// executed slice BC on that face (0: none) and fval*[f][k] its value for
// direction k.  A link k into a cell is claimed by the faces the cell lies on
// whose inward direction e_f has c_k . e_f > 0; the highest slot wins (the
// BC executed last).  Cells whose links this rule does not describe (next to
// map obstacles, computed ghost cells) carry CODE_FLUID_BC / CODE_GHOST and a
// link pattern instead.
struct StepArgs {
  char *fd[27];               // base of each direction array
  unsigned qs32;              // bytes between direction arrays (if all fit in 4 GiB)
  int sx, sy;                 // strides of the layout: Sy*Sz and Sz
  int nx, ny, nz;             // interior size
  int len;                    // ghost-inclusive cell count
  int block;                  // workgroup size (a multiple of 64, at most STEP_BLOCK)
  int bpp;                    // block-cell chunks per plane
  int nblocks;                // bpp * nx
  unsigned long long zmagic;  // ceil(2^40 / sy): row = (off * zmagic) >> 40
  int remap;                  // > 1: runs of that many chunks per XCD (see xcd_remap)
  int cpl;                    // cells per lane along z (1: one cell per lane, ghost columns included)
  int nzv;                    // runs per row, nz / cpl (cpl > 1)
  unsigned long long vmagic;  // ceil(2^40 / nzv): row = (lane * vmagic) >> 40
  int tile;                   // 1: LDS-staged tiles (block = STEP_TILE, cpl = 1)
  int fface[6];               // per face: slot * 4 + mode of its winning slice BC
  float fval32[6][27];        // per face and direction: addend / constant (FP32 lattices)
  double fval64[6][27];       // the same for FP64 lattices
  float om32[3];              // angular velocity of the frame (STEP_F_OMEGA)
  double om64[3];
  const ClosureParams *cl;    // closure network (STEP_F_CLOSURE)
  float cl_tau0;              // 3 nu + 1/2 in float
  float nu32;                 // viscosity (STEP_F_TURB)
  double nu64;
  const unsigned char *code;  // packed cell types
  const unsigned short *pat;  // link pattern of CODE_FLUID_BC / CODE_GHOST cells
  const LinkRec *rec;         // link patterns
  const unsigned char *mode;  // resolution mode of each BC slot
  const void *val;            // [slot][Q] addend or constant, compute type
  void *rho, *u;              // macroscopic outputs (ASL layout)
  double w;                   // relaxation rate (exact in the compute type)
  unsigned flags;
};

// Several lattices advanced by one launch of the step kernel: desc[l] is
// lattice l's StepArgs (device memory), workgroup b works on lattice lat[b],
// chunk b - first[lat[b]].
struct BatchLaunch {
  const StepArgs *desc;
  const uint2 *ent;  // per workgroup: lattice, chunk
  int nblocks;
  int block;
  int tile;  // 1: LDS-staged tiles (block = STEP_TILE; non-periodic, under 4 GiB)
};

// Boundary-link kernel entries (velocity map BCs).
struct VLink {
  int g;              // ghost cell (linear index)
  unsigned comp;      // directions i whose neighbour g+c_i is a computation node
  unsigned own;       // links g -> g+c_i this BC resolves
  int slot;           // BC slot (parameters in VLinkArgs::bc)
};

struct VLinkBC {
  double u[3];
  double omega[3], centre[3];
  int has_rot;
};

struct VLinkArgs {
  char *f;
  long long qs_bytes;
  int sx, sy;
  const VLink *links;
  int nlinks;
  const VLinkBC *bc;  // indexed by slot
  double dx, pos0[3]; // position of ghost-inclusive index 0 is pos0 - dx
  int after_odd;      // state after an odd step (else after an even one / init)
};

struct InitArgs {
  char *f;
  long long qs_bytes;
  int len;
  double rho0, u0[3];       // constants (used when rho_in/u_in are null)
  const void *rho_in, *u_in;  // fields in the macroscopic type
  void *rho, *u;
};

// per: periodic lattice of interior size nx x ny x nz (after an odd step the
// populations of a face cell sit in the wrapped neighbour's slots).
// Surface force (3 x len) or flux (len) of a level-set map, in the
// macroscopic type, from post-collision and incoming population views.
struct SurfaceArgs {
  const float *map;
  const void *post, *in;  // Q * len each
  void *out;
  int sx, sy, len;
  int flux;               // 1: ComputeSurfaceFluxMap, 0: ComputeSurfaceForceMap
};

// Visualisation fill of a lattice's rho or u in place at the guarded boundary
// nodes of a map (map <= 0 and > -0.9999): ASLHIP_VIS_NOSLIP_VEL,
// ASLHIP_VIS_NOSLIP_RHO or ASLHIP_VIS_VELOCITY (include/aslhip.h).
struct VisArgs {
  const float *map;    // len values, ASL layout
  void *rho, *u;       // macroscopic type
  int kind;
  int len, sx, sy;
  double u0[3];        // ASLHIP_VIS_VELOCITY: the velocity (without rotation)
  double rot[6];       // ASLHIP_VIS_VELOCITY with has_rot: axis, centre
  int has_rot;
  double dx, pos0[3];  // position of ghost-inclusive index 0 is pos0 - dx
};

// Populations from a host view (Q x len, f_i of each cell) into the layout a
// step of odd parity reads (slot (x, inv(i))), with rho and u of each cell.
struct WriteFArgs {
  char *f;
  long long qs_bytes;
  int len;
  int compressible;
  const void *in;  // Q * len macroscopic type
  void *rho, *u;
};

// The cells first .. first + count - 1 (count = len: all of them) go to
// out[i * count + cell - first].
struct UnpermuteArgs {
  const char *f;
  long long qs_bytes;
  int sx, sy, len;
  int per, nx, ny, nz;
  int after_odd;
  int first, count;
  void *out;  // Q * count macroscopic type
};

struct ReduceArgs {
  const char *f;
  long long qs_bytes;
  int sx, sy, len;
  int per, nx, ny, nz;
  int after_odd;
  const unsigned char *code;
  double *partial;  // 5 per block
};

// One step of one lattice: args points to a StepArgs in device memory.
struct StepLaunch {
  const StepArgs *args;
  int nblocks;
  int block;
  int variant;   // per-cell relaxation: 0 none, 1 LBGKTurbulence or the FP32 closure, 2 the matrix-core closure
  int cpl;       // cells per lane (2 or 4: the run variant; non-periodic, under 4 GiB, no closure)
  int tile;      // 1: LDS-staged tiles of STEP_TILE cells (same restrictions, cpl 1, block STEP_TILE)
};

// wide: the populations span more than 4 GiB (64-bit per-lane addresses).
hipError_t launch_step(int q, int dtype, bool odd, bool periodic, bool wide, const StepLaunch &a, hipStream_t s);
hipError_t launch_step_batch(int q, int dtype, bool odd, bool periodic, bool wide, const BatchLaunch &b,
                             hipStream_t s);
hipError_t launch_vlinks(int q, int dtype, const VLinkArgs &a, hipStream_t s);
// Closure features (7 floats per interior cell) of a periodic D3Q19 lattice
// with FP32 collision; args: the lattice's StepArgs in device memory.
hipError_t launch_closure_features(int dtype, bool odd, const StepArgs *args, long long ncell, float *out,
                                  hipStream_t s);
hipError_t launch_init(int q, int dtype, const InitArgs &a, hipStream_t s);
hipError_t launch_write_f(int q, int dtype, const WriteFArgs &a, hipStream_t s);
// The populations every computed cell receives in the next step (Q * len,
// macroscopic type, other cells untouched); lattices under 4 GiB.
hipError_t launch_peek(int q, int dtype, bool odd, bool periodic, const StepLaunch &a, void *out, hipStream_t s);
hipError_t launch_surface(int q, int dtype, const SurfaceArgs &a, hipStream_t s);
hipError_t launch_vis(int q, int dtype, const VisArgs &a, hipStream_t s);
hipError_t launch_unpermute(int q, int dtype, const UnpermuteArgs &a, hipStream_t s);
// The allocation-free staging array (ASLHIP_NOALLOC_SCRATCH_BYTES bytes of
// device memory in the library's code object, one per device): its address
// on the current device, looked up once per device.
hipError_t noalloc_scratch(void **ptr);
// partial must hold 5 * reduce_blocks(len) doubles; out: 5 doubles (device).
int reduce_blocks(int len);
hipError_t launch_reduce(int q, int dtype, const ReduceArgs &a, double *out, hipStream_t s);

}  // namespace aslhip
