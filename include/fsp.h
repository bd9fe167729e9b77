/*
 * fsp.h — C ABI of the MI355X pressure projection for the δ-3D volumetric
 * solver (Fluidisim water-core). Plain pointers and sizes only.
 *
 * Layout (the reference's, code/water-core/src/delta3d.rs:26-28, 187-206):
 *   cells  c  = (k*ny + j)*nx + i          x fastest, k (z) up
 *   u faces   = (k*ny + j)*(nx+1) + i      (nx+1)*ny*nz
 *   v faces   = (k*(ny+1) + j)*nx + i      nx*(ny+1)*nz
 *   w faces   = (k*ny + j)*nx + i          nx*ny*(nz+1)   (k = nz is the lid)
 *   columns   = j*nx + i                   nx*ny
 *
 * Threading: one caller per solver handle; separate handles are independent.
 * All device memory is reserved by fsp_create (the reference's invariant
 * I-06, delta3d.rs:208-241): fsp_project* allocates nothing.
 */
#ifndef FSP_H
#define FSP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSP_ABI_VERSION 1

/* Replaces the reference's Error (code/water-core/src/delta_projection.rs:133-150);
 * FSP_DEVICE and FSP_INVALID have no counterpart (device failure, bad handle/options). */
typedef enum {
  FSP_OK = 0,
  FSP_DOMAIN = 1,       /* Error::Domain: empty domain, unsupported geometry */
  FSP_NOT_FINITE = 2,   /* Error::NotFinite: non-finite input or intermediate */
  FSP_SHAPE = 3,        /* Error::Shape: wrong array length */
  FSP_CONVERGENCE = 4,  /* Error::Convergence: degraded result refused (fsp_options.refuse_degraded) */
  FSP_DEVICE = 5,
  FSP_INVALID = 6
} fsp_status;

/* Replaces Domain3 (delta3d.rs:44-51). */
typedef struct {
  int32_t nx, ny, nz;
  float dx;
} fsp_domain;

/* FSP_PRECOND_LEARNED: the network of fsp_learned.h (fsp_attach_net) inside
 * flexible CG, falling back to MGPCG (V-cycle as configured) when an FCG
 * pass of learned_budget iterations ends unconverged. */
enum { FSP_PRECOND_JACOBI = 0, FSP_PRECOND_MULTIGRID = 1, FSP_PRECOND_LEARNED = 2 };
enum { FSP_SMOOTH_JACOBI = 0, FSP_SMOOTH_RBGS = 1 };
enum { FSP_STORE_F32 = 0, FSP_STORE_F16 = 1, FSP_STORE_BF16 = 2 };
/* Coarse operators: the reference's re-discretisation (averaged apertures,
 * Dirichlet at a coarse half cell; delta3d_multigrid.rs:141-255), or the
 * Galerkin product P^T A P / 16 with piecewise-constant P (equal to the
 * re-discretisation in the interior, exact at the surface and at solids). */
enum { FSP_COARSE_REDISCRETISED = 0, FSP_COARSE_GALERKIN = 1 };
/* Fine-level red-black smoothing: both colours of a sweep in one launch, a
 * brick of z with its halo in LDS per workgroup (the default), or one launch
 * per colour. The V-cycles are identical bit for bit. */
enum { FSP_SWEEP_FUSED = 0, FSP_SWEEP_SINGLE = 1 };

/* Solver settings. fsp_default_options gives the reference's multigrid
 * preconditioner (delta3d_multigrid.rs:28-34: damped Jacobi ω=6/7, 2 pre,
 * 2 post, 8 coarsest sweeps, every coarse level the reference rule allows). */
typedef struct {
  int32_t preconditioner;  /* FSP_PRECOND_* */
  int32_t smoother;        /* FSP_SMOOTH_* */
  int32_t max_levels;      /* coarse levels cap; -1 = reference rule (delta3d_multigrid.rs:73-87) */
  int32_t pre_sweeps, post_sweeps, coarse_sweeps;
  int32_t gauge;           /* 1: remove the mean of b and p when no Dirichlet face exists */
  int32_t vcycle_storage;  /* FSP_STORE_*: storage type of the V-cycle vectors */
  int32_t replace_every;   /* true-residual replacement period in iterations (0 = never) */
  int32_t refuse_degraded; /* 1: a degraded projection returns FSP_CONVERGENCE, outputs untouched */
  int32_t device;          /* HIP device ordinal */
  int32_t coarse_operator; /* FSP_COARSE_* */
  int32_t graph;           /* 1: replay each CG iteration from a captured HIP graph */
  int32_t learned_budget;  /* FSP_PRECOND_LEARNED: FCG iterations before the MGPCG fallback (0: 50) */
  int32_t learned_smooth;  /* FSP_PRECOND_LEARNED: red-black sweeps on the network's output (default 1) */
  int32_t learned_vcycle;  /* FSP_PRECOND_LEARNED: 1 = the network's output is the initial guess of one
                              V-cycle (needs f32 storage and red-black smoothing) in place of the
                              learned_smooth sweeps; default 0 */
  int32_t fine_sweeps;     /* FSP_SWEEP_*: how the fine level's red-black sweeps are launched */
} fsp_options;

/* Surface state (Volume3 fields rho, g_eff, rest, homogeneous_ghost;
 * delta3d.rs:102-110, delta3d_mobile.rs:126-147). */
typedef struct {
  float rho, g_eff, rest;
  int32_t homogeneous_ghost; /* 1: ghost values 0 (p = 0 at the surface) */
  int32_t closed_lid;        /* 1: every top face closed (aperture 0): Neumann lid */
} fsp_surface;

/* Replaces Report (delta_projection.rs:164-188). */
typedef struct {
  uint32_t refinements;
  uint32_t iterations;
  int32_t degraded;
  int32_t floor;
  double residual;
  double divergence;
  double backward_error;
  double divergence_plain;
  /* timing of the call, milliseconds (no reference counterpart) */
  double ms_device;          /* device time of the projection (events around it) */
  double ms_transfer;        /* host<->device copies of fsp_project (0 for the device variant) */
} fsp_report;

typedef struct fsp_solver fsp_solver;

int32_t fsp_abi_version(void);
const char* fsp_status_string(fsp_status s);
void fsp_default_options(fsp_options* out);

/* Replaces Volume3::configure + enable_multigrid (delta3d.rs:210-294,
 * delta3d_multigrid.rs:339-349): reserves every device buffer. */
fsp_status fsp_create(const fsp_domain* domain, const fsp_options* options, fsp_solver** out);
void fsp_destroy(fsp_solver* s);
/* Coarse levels the solver holds (Volume3::multigrid_levels, delta3d_multigrid.rs:352-354). */
int32_t fsp_levels(const fsp_solver* s);

/* Replaces set_free_surface and the cut geometry (delta3d_mobile.rs:9-25,
 * delta3d.rs:300-319, delta3d_cut.rs:172-258). Host pointers. eta: nx*ny
 * column heights (required); eta_roundoff: nx*ny or NULL; solid: cells,
 * nonzero = solid (the cut's zero volume fraction), or NULL; open_u/v/w: face
 * apertures in [0, 1] (the cut's, weighting the operator and the divergence
 * as apply_cut/divergence_cut and mobile_row do, delta3d.rs:769-855,
 * delta3d_mobile.rs:152-188), or NULL (derived: a face touching a solid cell
 * or the box walls is closed, the lid open). Fractional apertures need the
 * re-discretised coarse operator (FSP_COARSE_GALERKIN returns FSP_DOMAIN) and
 * are not taken by the surface-step entry points (FSP_DOMAIN); this is a
 * configuration call and may reserve their device copy. */
fsp_status fsp_set_geometry(fsp_solver* s, const fsp_surface* surface,
                            const float* eta, size_t n_eta,
                            const float* eta_roundoff,
                            const uint8_t* solid, size_t n_solid,
                            const float* open_u, size_t n_open_u,
                            const float* open_v, size_t n_open_v,
                            const float* open_w, size_t n_open_w);

/* Cell labels of a level-set geometry (apic3d.rs:33-36). */
enum { FSP_LABEL_AIR = 0, FSP_LABEL_WATER = 1, FSP_LABEL_SOLID = 2 };

/* The geometry of the APIC-3D projection (apic3d.rs:940-1051), replacing its
 * labels and particle level set: labels (cells, FSP_LABEL_*) and phi (cells,
 * the level set at the cell centres, negative in water). Water rows couple
 * to water neighbours; an air neighbour is a ghost face with the pressure 0
 * on the zero level, theta = max(phi_c / (phi_c - phi_a), theta_min) (0.01
 * in the reference, :724); solid cells are walls that still carry their
 * faces' velocities into the divergence; the domain's boundary is a wall.
 * fsp_project then solves the reference's system (its operator times 1/dx^2,
 * its right-hand side times 1/dx^2: the same pressure) with the configured
 * preconditioner, scale = -rho/dt and k1 = dt/rho. A configuration call: the
 * first one reserves the labels' and phi's device copies; fsp_set_geometry
 * returns to the column geometry. The surface-step entry points refuse a
 * level-set geometry (FSP_DOMAIN). */
fsp_status fsp_set_level_set(fsp_solver* s, const uint8_t* labels, size_t n_labels, const float* phi, size_t n_phi,
                             float theta_min);

/* The coupled surface of the perturbation step over a background wave
 * (step_perturbation_mobile, delta3d_coupling.rs:720-864), as its projection
 * reads it (delta3d_mobile.rs:92-147 with surface_coupled; the terms
 * prepare_background3 computes, delta3d_coupling.rs:97-189). Host pointers.
 * total: nx*ny heights of the total surface (eta + the background's
 * elevation), which set wetness and the ghost fractions in place of eta;
 * bg_up: nx*ny terms added to the up ghost's value; bg_x: (nx+1)*ny*nz terms
 * added to the value of the side ghost across that u face; bg_y: nx*(ny+1)*nz
 * for v faces; a NULL term array is zeros. The up ghost's value keeps
 * rho*g*((eta - rest) - eta_roundoff) from fsp_set_geometry's eta. total ==
 * NULL switches the coupling off. Needs a column geometry without
 * fractional apertures (FSP_DOMAIN otherwise, as the coupled step refuses
 * a cut, :729); stays on over later fsp_set_geometry calls of that kind;
 * fsp_set_level_set and fractional apertures switch it off. The surface-step
 * entry points refuse a coupled surface (FSP_DOMAIN). A configuration call:
 * the first one reserves the device copies. */
fsp_status fsp_set_coupling(fsp_solver* s, const float* total, size_t n_total, const float* bg_up, size_t n_up,
                            const float* bg_x, size_t n_x, const float* bg_y, size_t n_y);

/* The same from device pointers (copied on the solver's stream; sizes as
 * above, finiteness not checked). */
fsp_status fsp_set_coupling_device(fsp_solver* s, const float* d_total, const float* d_bg_up, const float* d_bg_x,
                                   const float* d_bg_y);

/* Replaces Volume3::enable_graded (delta3d_graded.rs:64-100): the graded
 * column. Below the cubic top layers the pressure lives only at the node
 * layers `nodes` (fine layers from 0 to nz - 1, strictly increasing:
 * FSP_SHAPE otherwise) and is linear between two nodes; fsp_project and
 * fsp_project_device then solve P^T A P x = P^T b with the condensed
 * diagonal as preconditioner and correct u, v, w from P x, as
 * project_mobile3_graded does (:310-447), reporting the restricted divergence
 * metric (:451-486) and backward_error 0. Refused on a cut or level-set
 * geometry (FSP_DOMAIN, here and at the projection); the projection refuses
 * a surface that leaves the lowest cubic node's layer or the one above it
 * dry in some column (FSP_DOMAIN, nothing written). nodes == NULL or
 * count == 0 returns to the fine scheme. A configuration call: it reserves
 * seven vectors of nx*ny*count floats the first time (or when count grows). */
fsp_status fsp_enable_graded(fsp_solver* s, const int32_t* nodes, size_t count);

/* Volume3::pressure_unknowns_per_column (delta3d_graded.rs:102-105): the
 * graded column's nodes, or nz. */
int32_t fsp_pressure_unknowns_per_column(const fsp_solver* s);

/* Replaces Volume3::refine_mobile3 (delta3d_coupling.rs:226-249), the
 * coupled step's second solve after a degraded floor stop (:836-840): the
 * corrected faces d_u/d_v/d_w become the predicted faces, the pressure is
 * solved from 0 with homogeneous ghosts (value 0), the faces are corrected
 * again and the result is added to the base pressure d_p in f32. Device
 * pointers, all in/out. report->refinements = 1 and report->iterations the
 * second solve's (the step adds the first's). On a non-OK status d_p holds
 * the base pressure. */
fsp_status fsp_refine_device(fsp_solver* s, float* d_p, float* d_u, float* d_v, float* d_w, float scale, float k1,
                             double rel_tol, uint32_t max_iters, fsp_report* report);

/* Replaces Volume3::project_mobile3 (delta3d_mobile.rs:465-609). Host
 * pointers. us/vs/ws: predicted faces; p: in = warm start (p_warm != 0) or
 * ignored, out = pressure; u/v/w out: corrected faces. scale = -rho/dt,
 * k1 = dt/rho (delta3d_mobile.rs:836-837). rel_tol: relative true residual
 * (1e-6 in the reference, :507). On any non-OK status no output is written. */
fsp_status fsp_project(fsp_solver* s,
                       const float* us, size_t n_us, const float* vs, size_t n_vs, const float* ws, size_t n_ws,
                       float* p, size_t n_p, int32_t p_warm,
                       float* u, size_t n_u, float* v, size_t n_v, float* w, size_t n_w,
                       float scale, float k1, double rel_tol, uint32_t max_iters, fsp_report* report);

/* The same on device pointers (resident use; no host copies). On a non-OK
 * status d_p is left as it was; d_u, d_v, d_w may hold an intermediate
 * correction. Face arrays that are 16-byte aligned (as every hipMalloc or
 * torch allocation is) are read and written four cells at a time; others
 * take the per-cell kernels (same results, slower). */
fsp_status fsp_project_device(fsp_solver* s, const float* d_us, const float* d_vs, const float* d_ws,
                              float* d_p, int32_t p_warm, float* d_u, float* d_v, float* d_w,
                              float scale, float k1, double rel_tol, uint32_t max_iters, fsp_report* report);

/* Batched projection of independent patches (no counterpart in the
 * reference, which projects one volume per call): each patch is its own
 * solver with its geometry set; `concurrency` patches run at a time, each on
 * its solver's HIP stream from its own host thread (<= 0: 4, the hardware
 * queues of a process). Device pointers. Every patch is projected; the
 * first non-OK status is returned and reports[i] holds each patch's report. */
typedef struct fsp_patch_io {
  const float* us;
  const float* vs;
  const float* ws;
  float* p;
  int32_t p_warm;
  float* u;
  float* v;
  float* w;
  float scale;
  float k1;
} fsp_patch_io;

fsp_status fsp_project_batch(fsp_solver* const* solvers, const fsp_patch_io* io, int32_t n, double rel_tol,
                             uint32_t max_iters, int32_t concurrency, fsp_report* reports);

/* Tile autotuner (no counterpart in the reference): times the fine-level
 * red-black pass of `s` (geometry set) for each workgroup tile width in x
 * (quads) that fits the rows and keeps the fastest. Writes the chosen width
 * to *tile_x and its mean pass time to *us (either may be null). */
fsp_status fsp_autotune(fsp_solver* s, int32_t* tile_x, double* us);

/* Replaces apply_pressure_operator_for_trials (delta3d_mobile.rs:936-949):
 * out = A p with the current geometry. Device pointers. */
fsp_status fsp_apply_operator_device(fsp_solver* s, const float* d_p, float* d_out);

/* Replaces assemble_pressure_problem_for_trials (delta3d_mobile.rs:960-979):
 * rhs and Jacobi prec from a divergence field. Device pointers. */
fsp_status fsp_assemble_device(fsp_solver* s, const float* d_div, float scale, float* d_rhs, float* d_prec);

/* One preconditioner application z = M^-1 r on the current geometry
 * (v_cycle3, delta3d_multigrid.rs:393-429), with the fine kernels a
 * projection on this geometry runs. Device pointers. */
fsp_status fsp_precondition_device(fsp_solver* s, const float* d_r, float* d_z);

/* Divergence of faces (delta3d.rs:747-763, :806-823). Device pointers. */
fsp_status fsp_divergence_device(fsp_solver* s, const float* d_u, const float* d_v, const float* d_w, float* d_out);

/* Balance3 of the uncoupled surface step (delta3d_balance.rs Balance3;
 * delta3d_mobile.rs:866-877): perturbation volume after the step, its change,
 * the wall discharge (zero: the walls are closed) and residual = delta -
 * perturbation_in; band_in, sponge_out and outgoing are 0 on this path. m^3. */
typedef struct {
  double volume, delta, band_in, perturbation_in, sponge_out, residual, outgoing;
} fsp_balance;

/* Replaces Volume3::step_surface_mobile (delta3d_mobile.rs:819-927): advects
 * (u, v, w) over duration_us microseconds into the predicted faces
 * (advect_mobile3, :615-732), projects them with a warm start from p
 * (project_mobile3 with scale = -rho/dt, k1 = dt/rho, relative tolerance
 * 1e-6), extrapolates the solved faces up the columns (:345-394) and moves
 * the column heights by the wetted discharges (transport_mobile3,
 * :734-782). The solver's eta / eta_roundoff (set by fsp_set_geometry) are
 * the surface state; fsp_get_surface reads them. Device pointers, updated in
 * place. Refusals, as the reference's: duration 0 or above 2^53 us ->
 * FSP_NOT_FINITE; g_eff <= 0, dt^2 g / dx > 1 or a surface out of bounds
 * (below the column floor + 2 dx or above (nz-1) dx, before or after the
 * step) -> FSP_DOMAIN; a degraded projection -> FSP_CONVERGENCE; a
 * non-finite field -> FSP_NOT_FINITE. A refusal after the checks restores u,
 * v, w, p, eta and eta_roundoff bit for bit. Scope: no sparse column set,
 * the second-order advection term off (the reference's default). */
fsp_status fsp_step_surface_mobile_device(fsp_solver* s, float* d_u, float* d_v, float* d_w, float* d_p,
                                          uint64_t duration_us, uint32_t max_iters, fsp_report* report,
                                          fsp_balance* balance);

/* The solver's column heights and their roundoff (Volume3::surface,
 * delta3d.rs; host arrays of nx*ny; eta_roundoff may be NULL). */
fsp_status fsp_get_surface(fsp_solver* s, float* eta, float* eta_roundoff);

/* Stages of the surface step on their own, for tests. advect_mobile3
 * (delta3d_mobile.rs:615-732): predicted faces from (u, v, w), dt seconds. */
fsp_status fsp_advect_device(fsp_solver* s, const float* d_u, const float* d_v, const float* d_w, float dt,
                             float* d_us, float* d_vs, float* d_ws);
/* extrapolate_mobile3 (delta3d_mobile.rs:345-394), in place, current geometry. */
fsp_status fsp_extrapolate_device(fsp_solver* s, float* d_u, float* d_v, float* d_w);
/* transport_mobile3 (delta3d_mobile.rs:734-782): updates the solver's eta and
 * eta_roundoff; copies the discharges ((nx+1)*ny and nx*(ny+1)) out when the
 * flux pointers are not NULL. factor = dt/dx. */
fsp_status fsp_transport_device(fsp_solver* s, const float* d_u, const float* d_v, float factor, float* d_flux_x,
                                float* d_flux_y);
/* perturbation_volume (delta3d_balance.rs:105-113) and mobile_in_bounds
 * (delta3d_mobile.rs:801-813) of the current surface. */
fsp_status fsp_surface_scalars(fsp_solver* s, double* volume, int32_t* in_bounds);

/* Live timing of one kernel (no reference counterpart; a measurement hook).
 * fsp_profile_begin reserves max_launches event pairs and brackets every
 * launch of the kernel named `kernel` (its symbol name, e.g. "k_fgs") with
 * them; fsp_profile_end synchronises and returns the launch count, the
 * summed device time and, when the arrays are not NULL, up to max_durations
 * per-launch times and loop-active flags in launch order. active[i] is the
 * CG loop's active flag launch i started with: the iteration kernels return
 * at once when it is 0 (launches queued past convergence). */
fsp_status fsp_profile_begin(fsp_solver* s, const char* kernel, uint32_t max_launches);
fsp_status fsp_profile_end(fsp_solver* s, uint32_t* launches, double* total_ms, float* durations_ms,
                           int32_t* active, uint32_t max_durations);

/* Seeded synthetic inputs (host memory). cfg: 0 = closed box, i.i.d.
 * U[-1,1] faces; 1 = free-surface tank with 8 spheres and a 16-mode smooth
 * velocity. Sizes in reference orientation (nz is vertical). Arrays may be
 * NULL when not wanted. scale_ref scales heights/radii (edge/128 for patches). */
fsp_status fsp_generate(int32_t cfg, int32_t nx, int32_t ny, int32_t nz, uint64_t seed,
                        float* eta, uint8_t* solid, float* us, float* vs, float* ws);

#ifdef __cplusplus
}
#endif
#endif /* FSP_H */
