/*
 * lbm_capi.h — C ABI of the MI355X lattice Boltzmann time-step library (liblbm_amd.so).
 *
 * The reference (michaltakac/lbm-arrayfire) has no library boundary: each example program runs
 * its time loop inline against ArrayFire's array API (e.g. cpp/d2q9_bgk_lid.cpp:121-180,
 * rust/examples/d3q27_mrt_lid.rs:302-444).  Each entry point below replaces one group of those
 * ArrayFire calls; the replaced call sites are cited per declaration.  The calling convention
 * mirrors ArrayFire's own C API (af_err status codes + a last-error string, asynchronous work on
 * one device queue fenced by an explicit sync).
 *
 * Conventions
 *  - Plain C types only.  The caller owns every host buffer; the library owns device memory.
 *  - Layout is the reference's column-major, x fastest: node n = x + nx*(y + ny*z); population i
 *    of node n lives at n + N*i (N = nx*ny*nz), cpp/d2q9_bgk_lid.cpp:86.
 *  - Every call returns lbm_status; on failure lbm_last_error() holds a thread-local message.
 *    No C++ exception crosses the ABI.
 *  - One handle is bound to one device and one stream.  Calls on a handle must be serialised
 *    by the caller; distinct handles may be driven from distinct threads.
 */
#ifndef LBM_CAPI_H
#define LBM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 3

typedef struct lbm_ctx *lbm_handle;
typedef struct lbm_batch_ctx *lbm_batch;
typedef struct lbm_server_ctx *lbm_server;

typedef enum {
  LBM_OK = 0,
  LBM_EINVAL = 1,     /* bad argument or descriptor */
  LBM_ENOMEM = 2,     /* device allocation failed / over budget */
  LBM_EDEVICE = 3,    /* HIP runtime error or no device */
  LBM_EUNSTABLE = 4,  /* non-finite populations detected by the monitors */
  LBM_EINTERNAL = 5,
  LBM_EAGAIN = 6      /* server: job does not fit now, retry later */
} lbm_status;

/* Velocity sets (rows of the reference tables cpp/d2q9_bgk_lid.cpp:72-79 and
 * rust/examples/d3q27_mrt_lid.rs:98-104).  D3Q19 (the next lattice the reference's work in
 * progress cpp/d3q19_mrt_channel.cpp aims at) is D3Q27's rest, face and edge directions, in
 * D3Q27's order, with weights 1/3, 1/18, 1/36. */
typedef enum { LBM_D2Q9 = 9, LBM_D3Q19 = 19, LBM_D3Q27 = 27 } lbm_lattice;

/* Collision operators. BGK = cpp/d2q9_bgk_lid.cpp:149-153.  MRT_RAW = the reference's raw
 * monomial moment matrix (rust/examples/d3q27_mrt_lid.rs:167-199).  MRT_GS = the same rows
 * Gram-Schmidt orthogonalised (not in the reference); for D3Q19, the orthogonal basis of
 * d'Humieres et al. (2002), rows rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pi_xx, pww, pi_ww,
 * pxy, pyz, pxz, mx, my, mz, with projected equilibrium moments (s[0..18]).  D2Q9 is BGK only. */
typedef enum { LBM_BGK = 0, LBM_MRT_RAW = 1, LBM_MRT_GS = 2 } lbm_collision;

/* Population storage in HBM.  F16S stores the shifted value (f - w_i) * 2^15 as IEEE half,
 * BF16S stores f - w_i as bfloat16, with FP32 collision arithmetic on the deviations; the steps
 * round them stochastically (random bits from the node index and the lattice's step count), the
 * initial states to nearest even.  F64 computes in FP64, all others in FP32.  Lattices hold fewer than 2^32 nodes, and F16S /
 * BF16S lattices fewer than 2^31 (lbm_create returns LBM_EINVAL otherwise). */
typedef enum { LBM_F64 = 0, LBM_F32 = 1, LBM_F16S = 2, LBM_BF16S = 3 } lbm_storage;

/* Boundary semantics.  REFERENCE reproduces the reference: every axis periodic, full-way
 * bounce-back at solid nodes, Dirichlet overrides of rho/u before the equilibrium
 * (cpp/d2q9_bgk_lid.cpp:94-98,141-155; rust/examples/d2q9_bgk_channel.rs:190-201).
 * PHYSICAL: half-way (and moving-wall) bounce-back at links into solid nodes, non-equilibrium
 * bounce-back (Zou-He) velocity inlet, zero-gradient outlet at a reference density (no reference
 * counterpart).  Solid nodes' populations play no part in a PHYSICAL step.  Each step of a 3D
 * lattice sets a solid node to the rest equilibrium f_i = w_i when its aligned block of 64 node
 * indices holds a non-solid node (the kernels write whole blocks); other solid nodes, and those
 * of 2D lattices, keep their values. */
typedef enum { LBM_BC_REFERENCE = 0, LBM_BC_PHYSICAL = 1 } lbm_bc_mode;

/* How the MRT moment transforms are evaluated.  The MFMA engines evaluate the collision in
 * correction form, f' = f - M^-1 S M g with g = f - f_eq(rho, u) (equal to the reference's
 * f' = M^-1 (m - S (m - meq)), rust/examples/d3q27_mrt_lid.rs:336,385-387, because meq = M f_eq;
 * the reference table's row-26 difference is carried along), so the matrix units see only the
 * small non-equilibrium part. */
typedef enum {
  LBM_MRT_AUTO = 0,         /* SCALAR: the fastest measured engine on gfx950 for every storage */
  LBM_MRT_SCALAR = 1,       /* per-node unrolled FMA code with compile-time matrices */
  LBM_MRT_MFMA_F32 = 2,     /* d = M g and c = M^-1 S d on v_mfma_f32_16x16x4_f32, 64-node tiles */
  LBM_MRT_MFMA_BF16X3 = 3,  /* the same transforms on v_mfma_f32_16x16x32_bf16: g and S d split into
                               three bfloat16 terms each, M and 24 M^-1 (or M^T) exact in bfloat16,
                               FP32 accumulation */
  LBM_MRT_MFMA_FUSED = 4    /* c = M^-1 S M g as one FP32 MFMA contraction with the operator
                               precomputed on the host */
} lbm_mrt_engine;

/* Equilibrium moments for MRT_RAW: REFERENCE = the reference's hand-written table
 * (rust/examples/d3q27_mrt_lid.rs:357-383, row 26 = |u|^2 + rho/27); PROJECTED = M * f_eq.
 * MRT_GS always uses PROJECTED. */
typedef enum { LBM_MEQ_REFERENCE = 0, LBM_MEQ_PROJECTED = 1 } lbm_meq_mode;

typedef enum { LBM_INLET_UNIFORM = 0, LBM_INLET_PARABOLIC = 1 } lbm_inlet_profile;

/* How PHYSICAL mode closes a velocity-inlet node (no reference counterpart: the reference's
 * inlets are the REFERENCE-mode velocity overrides).  ZOU_HE: the c_x = +1 unknowns by
 * non-equilibrium bounce-back with the transverse correction (Zou & He 1997; Hecht & Harting 2010
 * in 3D).  EXTRAPOLATED: every population f_eq(rho_e, u_in) + (g - f_eq(rho_e, u_e)), g the
 * downstream (x + 1) neighbour's populations with density rho_e and velocity u_e (Guo, Zheng &
 * Shi 2002), the outlet's closure with the roles of rho and u swapped.  On D3Q27 ducts at
 * tau near 0.5 the Zou-He inlet feeds a growing mode at the inlet plane (the 128x64x64 sphere duct
 * at tau = 0.536 diverges by step 1,500) where the extrapolated inlet stays stable. */
typedef enum { LBM_INLET_ZOU_HE = 0, LBM_INLET_EXTRAPOLATED = 1 } lbm_inlet_closure;

/* Node-type byte, one per node (the reference's BOUND mask, cpp/d2q9_bgk_lid.cpp:90-94,
 * generalised).  Bits combine. */
#define LBM_NODE_FLUID 0x00u
#define LBM_NODE_SOLID 0x01u   /* REFERENCE: full-way bounce-back here, rho=u=0 override.
                                  PHYSICAL: wall; links into it bounce back half-way. */
#define LBM_NODE_MOVING 0x02u  /* REFERENCE: u := u_wall override (lid, :141-142).
                                  PHYSICAL (with SOLID): the wall moves with u_wall. */
#define LBM_NODE_INLET 0x04u   /* REFERENCE: ux := inlet velocity override (channel :190).
                                  PHYSICAL: velocity inlet on the x = 0 face, closed as
                                  lbm_desc.inlet_closure selects. */
#define LBM_NODE_OUTLET 0x08u  /* PHYSICAL: zero-gradient outlet on the x = nx-1 face: every
                                  population = f_eq(rho_bc, u) + f_neq of the upstream neighbour
                                  (u = the neighbour's velocity). */
#define LBM_NODE_RHO_FIX 0x10u /* REFERENCE: rho := rho_bc after the solid override
                                  (channel set_row, rust/examples/d2q9_bgk_channel.rs:197-201). */

typedef struct {
  int32_t lattice;        /* lbm_lattice: 9, 19 or 27 */
  int32_t nx, ny, nz;     /* nz must be 1 for D2Q9 */
  int32_t collision;      /* lbm_collision */
  int32_t storage;        /* lbm_storage */
  int32_t bc_mode;        /* lbm_bc_mode */
  int32_t mrt_engine;     /* lbm_mrt_engine */
  int32_t meq_mode;       /* lbm_meq_mode */
  int32_t inlet_profile;  /* lbm_inlet_profile */
  int32_t inlet_closure;  /* lbm_inlet_closure (PHYSICAL mode) */
  int32_t reserved;       /* 0 */
  double omega;           /* BGK relaxation rate 1/tau */
  double s[27];           /* MRT rates, one per moment row */
  double u_wall[3];       /* lid / moving-wall velocity */
  double u_in;            /* inlet velocity (peak value for PARABOLIC) */
  double rho_bc;          /* density imposed by RHO_FIX and used for moving walls (1.0) */
} lbm_desc;

/* ---- device / library ------------------------------------------------------------------- */
int lbm_abi_version(void);
/* Selects the device later lbm_create calls bind to.  Replaces af::setDevice(argv[1])
 * (cpp/d2q9_bgk_lid.cpp:211-214). */
lbm_status lbm_set_device(int device);
/* Thread-local message of the last failed call ("" if none). */
const char *lbm_last_error(void);
/* Device bytes a lattice with this descriptor holds (both population buffers, node types, in
 * PHYSICAL mode the 4-byte wall link masks, monitor scratch).  Used for admission by HBM
 * footprint. */
size_t lbm_footprint_bytes(const lbm_desc *d);

/* ---- one lattice ------------------------------------------------------------------------ */
/* Allocates the lattice and uploads the node-type mask.  Replaces the setup of
 * cpp/d2q9_bgk_lid.cpp:81-98 (index maps, BOUND, ON, TO_REFLECT/REFLECTED), which the fused
 * kernel does not need.  node_type: N bytes, x fastest. */
lbm_status lbm_create(const lbm_desc *d, const uint8_t *node_type, lbm_handle *out);
lbm_status lbm_destroy(lbm_handle h);

/* F = f_eq(rho, u) at every node.  Any field pointer may be NULL (rho -> 1, u -> 0).  Replaces
 * cpp/d2q9_bgk_lid.cpp:100-107 and the channel variant cpp/d2q9_bgk_channel.cpp:107-118. */
lbm_status lbm_init_fields(lbm_handle h, const double *rho, const double *ux, const double *uy,
                           const double *uz);
/* F = f_eq(rho0, u0 + noise).  SplitMix64(seed) draws D numbers per node for every node in node
 * order (x component, then y, then z); draw r maps to noise_amp * (2 * (r >> 11) * 2^-53 - 1),
 * uniform in [-noise_amp, noise_amp), and is added only at non-solid nodes.  No reference
 * counterpart (the reference starts noise-free). */
lbm_status lbm_init_equilibrium(lbm_handle h, double rho0, const double u0[3], uint64_t seed,
                                double noise_amp);
/* MRT only: f = rho0/27 everywhere followed by one moment-space collision with rho = rho0,
 * u = 0 and no bounce-back (rust/examples/d3q27_mrt_lid.rs:228-287).  BGK: f = f_eq(rho0, 0). */
lbm_status lbm_init_reference(lbm_handle h, double rho0);

/* Enqueues nsteps fused stream+collide steps on the handle's stream and returns.  Replaces the
 * loop body cpp/d2q9_bgk_lid.cpp:124-155 / rust/examples/d3q27_mrt_lid.rs:304-389.  2D D2Q9 BGK
 * lattices with FP32-arithmetic storage advance several steps per launch from an on-chip tile
 * (bit-identical to single steps; LBM_MULTISTEP=off in the environment at lbm_create turns it
 * off, LBM_MULTISTEP=<shape index> picks another launch shape). */
lbm_status lbm_step(lbm_handle h, int64_t nsteps);
/* Steps one launch of lbm_step advances (1 = one launch per step).  No reference counterpart. */
int32_t lbm_steps_per_launch(lbm_handle h);
/* As lbm_step, then waits; *ms = device time between HIP events recorded on the handle's
 * stream around the launches. */
lbm_status lbm_step_timed(lbm_handle h, int64_t nsteps, double *ms);
/* Waits for the handle's stream.  Replaces af::sync() (cpp/d2q9_bgk_lid.cpp:118,182). */
lbm_status lbm_sync(lbm_handle h);
/* Device self-check, no reference counterpart: compares the FP32 kernels' reciprocal (hardware
 * estimate + one Newton step, used for 1/rho) with IEEE division for every float bit pattern in
 * [lo, hi) on the current device and stores the number of differing results in *mismatches.
 * Zero over the normal floats with a normal reciprocal is the identity FP32 bit-exactness rests
 * on. */
lbm_status lbm_check_reciprocal(uint32_t lo, uint32_t hi, uint64_t *mismatches);
int64_t lbm_steps_done(lbm_handle h);

/* Macroscopic fields the reference holds after a step (DENSITY/UX/UY[/UZ] of
 * cpp/d2q9_bgk_lid.cpp:131-146): moments of the pulled populations with the boundary overrides
 * applied (REFERENCE) or of the post-boundary populations (PHYSICAL).  N doubles each; any
 * pointer may be NULL.  Blocks. */
lbm_status lbm_get_fields(lbm_handle h, double *rho, double *ux, double *uy, double *uz);
/* The same fields at `count` selected nodes (sensors); each output holds count doubles, NULL ok.
 * Blocks.  No reference counterpart (the reference reads whole fields for display). */
lbm_status lbm_get_fields_at(lbm_handle h, const int64_t *nodes, int32_t count, double *rho, double *ux,
                             double *uy, double *uz);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid hipMalloc): lbm_get_fields_at with the same
 * arguments and the same values, bit for bit, computed at the listed nodes only and calling no
 * device allocation.  At most LBM_NOALLOC_MAX_NODES nodes per call (LBM_EINVAL beyond).  The node
 * list and the results are staged in one fixed device array per GPU (640 KiB, held until the
 * process ends) that every call on that GPU shares: calls that overlap on one GPU, from several
 * threads, race.  Blocks. */
#define LBM_NOALLOC_MAX_NODES 16384
lbm_status lbm_get_fields_at_noalloc(lbm_handle h, const int64_t *nodes, int32_t count, double *rho,
                                     double *ux, double *uy, double *uz);
/* Current post-collision populations as doubles (decoded for shifted storage), N*Q.  Blocks. */
lbm_status lbm_get_populations(lbm_handle h, double *f);
/* Restarts the step count of the stochastic rounding (shifted storages), as the lbm_init_* calls do. */
lbm_status lbm_set_populations(lbm_handle h, const double *f);
/* Bitwise checkpoint / restore: the raw storage-format bytes of the current populations (Q arrays
 * of N elements), then 8 bytes holding the lattice's step count (uint32, then 4 zero bytes), which
 * the shifted storages' stochastic rounding draws on, so a restored lattice continues exactly as
 * the original would have. */
size_t lbm_state_bytes(lbm_handle h);
lbm_status lbm_get_state(lbm_handle h, void *dst);
lbm_status lbm_set_state(lbm_handle h, const void *src);
/* Wavefront reductions over the current populations with FP64 accumulation:
 * mass = sum_n sum_i f, kinetic_energy = sum_n |sum_i c_i f|^2 / (2 sum_i f) over non-solid
 * nodes, nonfinite = count of non-finite populations.  Any pointer may be NULL.  Blocks. */
lbm_status lbm_monitors(lbm_handle h, double *mass, double *kinetic_energy, int64_t *nonfinite);
/* Force on the solid nodes inside box = {x0, x1, y0, y1, z0, z1} (half-open node ranges; NULL =
 * every solid node) by momentum exchange over the links from non-solid nodes into them:
 * F = sum c_i (2 f_i - 6 w_i rho_bc c_i.u_wall), the second term only for moving walls, f = the
 * post-collision populations of the current state, FP64 accumulation.  Drag coefficient of a
 * body in a uniform stream: C_d = 2 F_x / (rho u^2 A).  No reference counterpart (the reference
 * reports no forces).  Blocks. */
lbm_status lbm_force(lbm_handle h, const int32_t box[6], double force[3]);

/* ---- batching: many lattices, one launch per group (kernel configuration and size class, node
 * counts within a factor of 4) per step, the groups on concurrent streams ------------------- */
/* No reference counterpart (each reference program steps one lattice).  Members keep their own
 * state and stay readable through their handles (lbm_get_*, lbm_monitors, lbm_set_*); while a
 * lattice is a member its work runs on the batch's stream and lbm_step on it is refused.  Stepping
 * a batch gives every member exactly the populations it would have reached stepped alone. */
lbm_status lbm_batch_create(lbm_batch *out);
lbm_status lbm_batch_add(lbm_batch b, lbm_handle h);
lbm_status lbm_batch_remove(lbm_batch b, lbm_handle h);
/* Enqueues nsteps steps of every member (replays of a captured two-step hipGraph). */
lbm_status lbm_batch_step(lbm_batch b, int64_t nsteps);
lbm_status lbm_batch_step_timed(lbm_batch b, int64_t nsteps, double *ms);
lbm_status lbm_batch_sync(lbm_batch b);
/* Autotuner: times each configuration group's persistent grid size over `trials` launches per
 * candidate and keeps the fastest.  Members' states are not changed. */
lbm_status lbm_batch_tune(lbm_batch b, int32_t trials);
int32_t lbm_batch_size(lbm_batch b);
lbm_status lbm_batch_destroy(lbm_batch b);

/* ---- serving: continuous batching of simulation jobs ------------------------------------ */
/* No reference counterpart.  Jobs are admitted while their footprint fits the HBM budget;
 * admitted jobs stay resident and are stepped together, and jobs that arrive while others run
 * join at the next tick.  A job's populations are exactly those it would reach run alone.  The
 * server keeps the device buffers of finished jobs for later ones, in size classes (2^k times 1,
 * 1.25, 1.5 or 1.75): a job's footprint is lbm_footprint_bytes with each buffer rounded up to its
 * class (at most 25 % more), plus, for a job submitted with want_fields, the 4 N doubles its field
 * readout takes on the device; idle kept buffers count against the budget until a job needs the
 * room.  A job whose field readout fails is reported with that status and keeps no fields. */
typedef struct {
  lbm_desc desc;
  const uint8_t *node_type; /* N bytes, copied at submit */
  double rho0;              /* initial state: lbm_init_equilibrium(rho0, u0, seed, noise_amp) */
  double u0[3];
  uint64_t seed;
  double noise_amp;
  int64_t nsteps;
  int32_t want_fields;      /* keep rho/u at completion for lbm_server_take */
  int32_t reserved;
} lbm_job_spec;

typedef struct {
  int64_t job_id;
  int32_t status;           /* LBM_OK, LBM_EUNSTABLE (non-finite populations) or a create error */
  int32_t reserved;
  double submit_s, admit_s, finish_s; /* server clock (lbm_server_clock), seconds */
  double mass, kinetic_energy;        /* monitors at completion */
  int64_t nodes, steps;
} lbm_job_result;

/* budget_bytes: device memory the server may hold for resident jobs; steps_per_tick: most steps
 * a tick advances the resident jobs by.  The server runs on the device current at its creation
 * (lbm_set_device), and its jobs' lattices share its one stream: make that device current again
 * before each lbm_server_* call (a job admitted with another device current fails with
 * LBM_EINVAL in its result). */
lbm_status lbm_server_create(size_t budget_bytes, int32_t steps_per_tick, lbm_server *out);
/* Queues a job (LBM_ENOMEM if it could never fit the budget). */
lbm_status lbm_server_submit(lbm_server s, const lbm_job_spec *spec, int64_t *job_id);
/* Admits queued jobs that fit, advances all resident jobs by min(steps_per_tick, fewest steps
 * left) in one batched launch sequence, retires finished (or non-finite) jobs and admits again.
 * Blocks until the tick's steps are done. */
lbm_status lbm_server_tick(lbm_server s, int32_t *active, int32_t *queued);
/* Moves up to cap finished results into out; *n = number moved. */
lbm_status lbm_server_poll(lbm_server s, lbm_job_result *out, int32_t cap, int32_t *n);
/* Fields of a finished job submitted with want_fields (N doubles each, NULL ok); frees them. */
lbm_status lbm_server_take(lbm_server s, int64_t job_id, double *rho, double *ux, double *uy, double *uz);
/* Drops a queued or resident job without a result. */
lbm_status lbm_server_cancel(lbm_server s, int64_t job_id);
/* Monitors of every resident job are read every `ticks` ticks (0 = only at completion). */
lbm_status lbm_server_set_check_every(lbm_server s, int32_t ticks);
/* Footprint of the resident jobs (idle kept buffers not included). */
size_t lbm_server_used_bytes(lbm_server s);
double lbm_server_clock(lbm_server s);

/* Where a server's time went (host wall-clock seconds, cumulative since creation). */
typedef struct {
  int64_t ticks;            /* ticks that launched steps */
  int64_t steps;            /* batch steps launched (sum of tick lengths) */
  int64_t admitted, retired, unstable;
  int64_t batch_rebuilds;   /* membership changes that rewrote the batch tables */
  int64_t graph_captures;   /* step-graph captures */
  double admit_s;           /* creating and initialising admitted jobs */
  double retire_s;          /* final monitors, kept fields, releasing finished jobs */
  double launch_s;          /* enqueueing tick steps (table rewrites and captures included) */
  double wait_s;            /* waiting for the device at the end of ticks */
  double check_s;           /* periodic non-finite checks of resident jobs */
} lbm_server_stats;
lbm_status lbm_server_get_stats(lbm_server s, lbm_server_stats *out);
lbm_status lbm_server_destroy(lbm_server s);

/* ---- two-component Shan-Chen D2Q9 (cpp/d2q9_shanchen.cpp:39-181) --------------------------
 * Two BGK populations f (fluid 1) and g (fluid 2) on a fully periodic nx x ny D2Q9 lattice,
 * coupled through the common velocity u' = (omega1 j1 + omega2 j2) / (omega1 rho1 + omega2 rho2)
 * and the pseudopotential interaction psi = rho: fluid s relaxes towards f_eq(rho_s, u_s) with
 * u_s = u' - (G / omega_s) sum_{i>=1} w_i rho_t(x - c_i) c_i (t = the other fluid; densities after
 * streaming).  This is the reference program with its work-in-progress defects corrected: the
 * populations are streamed before the collision, each fluid's interaction uses the other fluid's
 * density, the interaction sums start from zero every step, and u' weights the momenta j_s (the
 * reference divides them by rho_s first).  Storage F64 or F32.  No reference counterpart for the
 * handle itself: the reference runs the loop inline. */
typedef struct lbm_sc_ctx *lbm_sc_handle;
typedef struct {
  int32_t nx, ny;    /* periodic D2Q9 lattice */
  int32_t storage;   /* LBM_F64 or LBM_F32 */
  int32_t pad_;
  double omega1;     /* relaxation rate of fluid 1 */
  double omega2;     /* relaxation rate of fluid 2 */
  double g;          /* interaction amplitude G (the reference uses -1.2) */
} lbm_sc_desc;
lbm_status lbm_sc_create(const lbm_sc_desc *d, lbm_sc_handle *out);
/* f = w_i rho1, g = w_i rho2 (both at rest), the reference's initial state (:76-79);
 * rho1, rho2: N doubles each, x fastest. */
lbm_status lbm_sc_init(lbm_sc_handle h, const double *rho1, const double *rho2);
/* Enqueues nsteps steps (stream both fluids, densities, interaction, collision); asynchronous. */
lbm_status lbm_sc_step(lbm_sc_handle h, int64_t nsteps);
/* As lbm_sc_step, then waits; *ms = device time of the launches (HIP events). */
lbm_status lbm_sc_step_timed(lbm_sc_handle h, int64_t nsteps, double *ms);
/* Fields after the last step, N doubles each (any may be NULL), moments of the current
 * (post-collision) populations: the fluids' densities (the collision conserves them, so they are
 * the step's densities that the reference displays, :113-114,183, up to rounding) and the mixture
 * velocity (j1 + j2) / (rho1 + rho2).  Blocks. */
lbm_status lbm_sc_get_fields(lbm_sc_handle h, double *rho1, double *rho2, double *ux, double *uy);
/* Populations as doubles: 9 N of fluid 1 then 9 N of fluid 2, population i of node n at
 * n + N i.  Blocks. */
lbm_status lbm_sc_get_populations(lbm_sc_handle h, double *f);
lbm_status lbm_sc_set_populations(lbm_sc_handle h, const double *f);
lbm_status lbm_sc_destroy(lbm_sc_handle h);

#ifdef __cplusplus
}
#endif
#endif /* LBM_CAPI_H */
