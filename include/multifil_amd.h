/*
 * multifil_amd.h -- C ABI of the MI355X (gfx950) half-sarcomere step engine.
 *
 * The reference has no FFI: its hot path is pure-Python JAX.  The entry
 * points below replace these reference functions, and are what a ctypes
 * binding on the reference side would call:
 *
 *   mf_tables  <- transitions.py:99-190  expm_pade6_batch, with the generator
 *                 builders transitions.py:434-513 (TM, 27 x 4x4) and
 *                 transitions.py:632-922 (XB, 2*n_bins x 6x6; absorbing
 *                 variant transitions.py:1029-1078)
 *   mf_run     <- simulation.py:584-747  _run_sim_kernel: per instance the
 *                 initial create_and_equilibrate (:633-642) and the lax.scan
 *                 over timesteps of scan_fn (:679-721) = timestep.py:136-199
 *                 (kinetics_step + solve_equilibrium) + axial_force_at_mline
 *                 (forces.py:588-627) + compute_all_metrics (metrics_fn.py:
 *                 154-417)
 *   mf_run with n_steps = 1 and host-given uniforms is the teacher-forced
 *                 single step, timestep.py:136-199 with random_values=
 *                 (transitions.py:521,1086).
 *
 * Conventions: every pointer in mf_launch is DEVICE memory owned by the
 * caller; calls are stream-ordered on the given hipStream_t (0 = null
 * stream); a return value != 0 is an error whose text mf_last_error()
 * returns.  No C++ exception crosses this boundary.  Calls that share
 * buffers must not overlap.
 */
#ifndef MULTIFIL_AMD_H
#define MULTIFIL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_ABI_VERSION 5
#define MF_N_PARAMS 44      /* DynamicParams fields, params.py:73-553 order */
#define MF_N_METRICS 52     /* metrics_fn.py:343-417 order                */
#define MF_MAX_BINS 256     /* n_xb_bins upper bound                       */
#define MF_MAX_POPS 4       /* subpopulations per instance                 */

/* Solver modes. */
#define MF_SOLVER_REFERENCE 0  /* FP32 Newton-CG, caps and tolerance floor of
                                  solver.py:369-427,714-718                  */
#define MF_SOLVER_REFINED   1  /* FP64 Newton residual + FP32 CG correction,
                                  iterated to refine_tol                     */

/* Instance flags. */
#define MF_INST_INIT    1   /* run create_and_equilibrate before step 0     */
#define MF_INST_DYNAMIC 2   /* dynamic lattice spacing (K_lat, nu)           */

/* Subpopulation modes (core/subpopulation.py; transitions.py:580-602,
 * 925-975): populations share the mechanics and differ in their rates. */
#define MF_POP_NONE       0
#define MF_POP_MEAN_FIELD 1  /* one generator, sum_k f_k Q_k               */
#define MF_POP_EXPLICIT   2  /* per-head / per-site labels pick table sets  */
/* pop_flags: which kinetics the populations change (xb_* / tm_* scales). */
#define MF_POP_XB 1
#define MF_POP_TM 2

/* Launch flags. */
#define MF_LAUNCH_DYNAMIC 1 /* some instance has MF_INST_DYNAMIC: run the
                               kernels that carry the bordered (N+1) solve
                               of solver.py:599-656                        */
#define MF_LAUNCH_KINETICS_ONLY 2 /* each step is kinetics_step alone
                               (timestep.py:65-129): nearest sites, TM and
                               XB transitions; no solve, no metrics row    */
#define MF_LAUNCH_REBUILD 8  /* an instance whose pCa or spacing driver
                               changes from one step to the next rebuilds
                               its transition tables inside the kernel
                               (the rows that step reads), so a trace runs
                               in one launch; without it the caller splits
                               launches at driver changes and calls
                               mf_tables between them                    */
#define MF_LAUNCH_SOLVE_ONLY 4 /* each step is solve_equilibrium alone
                               (solver.py:663-771) at the step's Z-line and
                               spacing (dynamic: d_ref = spacing, the state's
                               spacing the first guess); the metrics row
                               carries force, residual, spacing and
                               iterations, its kinetic columns are void    */

/* One lattice (SarcTopology.create output, packed).  All offsets are in
 * elements into the float / int32 lattice blobs. */
typedef struct mf_lattice {
  int32_t n_thick, n_crowns, n_thin, n_sites;
  int32_t n_xb_per_crown, n_faces, width, n_bins;
  int64_t crown_offsets;    /* f32 [n_thick*n_crowns]                      */
  int64_t crown_rests;      /* f32 [n_thick*n_crowns]                      */
  int64_t binding_offsets;  /* f32 [n_thin*n_sites]                        */
  int64_t binding_rests;    /* f32 [n_thin*n_sites]                        */
  int64_t thin_rest_cum;    /* f32 [n_thin*n_sites] cumsum of binding_rests */
  int64_t bin_edges;        /* f32 [n_bins+1]                              */
  int64_t bin_centers;      /* f32 [n_bins]                                */
  int64_t xb_thin;          /* i32 [X] target thin filament                */
  int64_t xb_slot;          /* i32 [X] thin*n_faces + face                 */
  int64_t xb_valid;         /* i32 [X] 0/1                                 */
  int64_t face_sites;       /* i32 [n_thin*n_faces*width], padded like
                               xb_to_site_indices (sarc_geometry.py:1480) */
  int64_t face_count;       /* i32 [n_thin*n_faces]                        */
  int64_t tm_prev, tm_next; /* i32 [n_thin*n_sites]                        */
  float overlap_lo;         /* min(crown_offsets) - 13 nm (metrics_fn.py:108) */
  float overlap_hi;         /* max(crown_offsets) + 13 nm (metrics_fn.py:109) */
} mf_lattice;

/* One simulated half-sarcomere in a launch. */
typedef struct mf_instance {
  int32_t lattice;     /* index into mf_launch.lattices                    */
  int32_t n_steps;     /* steps to advance in this launch (may be 0)       */
  int32_t t0;          /* global index of the first step: trace row, RNG
                          counter, metrics row                             */
  int32_t flags;       /* MF_INST_*                                        */
  uint64_t seed;       /* Philox4x32-10 key                                */
  int64_t state_off;   /* byte offset of the instance state in `state`     */
  int64_t table_off;   /* float offset of the instance tables in `tables`  */
  float K_lat;         /* dynamic spacing: lattice stiffness x n_thick     */
  float nu;            /* dynamic spacing: Poisson exponent                */
  int64_t label_off;   /* MF_POP_EXPLICIT: byte offset of this instance's
                          labels in pop_labels (u8 [X] XB then u8 [S] TM) */
} mf_instance;

typedef struct mf_solver_cfg {
  int32_t mode;        /* MF_SOLVER_*                                      */
  int32_t n_newton;    /* Newton cap (params.py:655: 4)                    */
  int32_t n_cg;        /* CG iterations per Newton step (params.py:656: 6) */
  int32_t n_refine;    /* Newton cap in MF_SOLVER_REFINED; up to 4x while  *
                        * each iteration still halves max|F|             */
  float tol;           /* NaN: max(solver_tol, thick_k*1e-4, 0.25)         */
  float refine_tol;    /* pN, MF_SOLVER_REFINED stopping residual          */
  float dt;            /* ms                                               */
  float cg_forcing;    /* MF_SOLVER_REFINED: CG stops at a relative r.u of
                          max(cg_forcing refine_tol / max|F|,
                          1e-5 max|F|)^2; 0 runs n_cg iterations (or to
                          FP32 rounding)                                  */
} mf_solver_cfg;

typedef struct mf_launch {
  const mf_lattice* lattices;   /* [n_lattices]                            */
  const float* lat_f32;         /* lattice float blob                      */
  const int32_t* lat_i32;       /* lattice int blob                        */
  const mf_instance* instances; /* [B]                                     */
  const float* params;          /* [B][MF_N_PARAMS]                        */
  const float* drivers;         /* [B][t_stride][3] = z_line, pCa, spacing */
  const float* uniforms;        /* [B][t_stride][u_stride] (TM draws then
                                   XB draws) or NULL for device Philox    */
  uint8_t* state;               /* instance states, see mf_state_layout    */
  float* tables;                /* per-instance transition tables          */
  float* metrics;               /* [B][t_stride][MF_N_METRICS] or NULL     */
  double* force;                /* [B][t_stride] FP64 M-line force or NULL */
  int32_t n_lattices, B, t_stride, u_stride;
  mf_solver_cfg cfg;
  int32_t block;                /* threads per workgroup: 256, 512, 768 or 1024 */
  int32_t t_base;               /* added to every instance's t0            */
  int32_t launch_flags;         /* MF_LAUNCH_*                             */
  int32_t pad;
  const mf_lattice* lattices_host; /* optional HOST copy of `lattices`; when
                                   NULL the library reads the device copy  */
  uint8_t* workspace;           /* [B][workspace_stride] device scratch for
                                   lattices whose per-instance arrays exceed
                                   LDS; NULL when mf_workspace_bytes gives 0 */
  int64_t workspace_stride;     /* bytes per instance, >= mf_workspace_bytes */
  /* Subpopulations (all instances of a launch share the mode, K and flags,
   * as run() requires, simulation.py:533-540). */
  int32_t n_pops;               /* K (1 = none; <= MF_MAX_POPS)             */
  int32_t pop_mode;             /* MF_POP_*                                 */
  int32_t pop_flags;            /* MF_POP_XB | MF_POP_TM                    */
  int32_t pad2;
  const float* pop_params;      /* [B][K][MF_N_PARAMS] population constants,
                                   row 0 = the wild type (= params)        */
  const float* pop_fractions;   /* [B][K] mean-field weights               */
  const uint8_t* pop_labels;    /* explicit: labels, see label_off          */
} mf_launch;

/* Byte layout of one instance state (all sections 8-byte aligned):
 *   f64 pos[N]  (N = n_thick*n_crowns + n_thin*n_sites; thick then thin)
 *   f64 spacing   (carried lattice spacing)
 *   i32 xb_bound[X]   (-1 unbound)          X = total_xbs
 *   i32 xb_key[X]     (2*n_bins table row of the last kinetics step)
 *   i32 xb_near[X]    (nearest site along the target thin filament and
 *   f32 xb_xdist[X]    its axial offset from the crown, from the launch's
 *                      last nearest-site update: the reference's
 *                      xb_nearest_bs and xb_distances[..., 0]; written by
 *                      the kernel, never read)
 *   i32 thin_bound[S] (-1 unbound)          S = n_thin*n_sites
 *   u8  xb_state[X], u8 tm_state[S]
 * mf_state_bytes returns the size for a lattice. */
int64_t mf_state_bytes(int32_t n_thick, int32_t n_crowns, int32_t n_thin,
                       int32_t n_sites, int32_t n_xb_per_crown);
/* Floats of per-instance table storage for n_bins. */
int64_t mf_table_floats(int32_t n_bins);
/* ... with n_pop_tm TM and n_pop_xb XB table sets (MF_POP_EXPLICIT: K for
 * the kinetics the populations change, else 1).  Layout: header, TM sets,
 * XB sets, absorbing sets; labels select the set. */
int64_t mf_table_floats_pop(int32_t n_bins, int32_t n_pop_tm, int32_t n_pop_xb);

int mf_abi_version(void);
const char* mf_last_error(void);

/* Transition tables for every instance from its drivers at row t0 (pCa,
 * spacing) or, for MF_INST_DYNAMIC, its carried spacing. */
int mf_tables(const mf_launch* L, void* stream);

/* Advance every instance n_steps (one persistent workgroup per instance). */
int mf_run(const mf_launch* L, void* stream);

/* Device workspace bytes per instance that mf_run needs for this launch's
 * lattices, block and mode (0: everything fits in LDS).  No counterpart in
 * the reference, which leaves all buffers to XLA. */
int mf_workspace_bytes(const mf_launch* launch, int64_t* per_instance);

/* Dynamic LDS bytes and the workgroup size mf_run would use (diagnostics). */
int mf_run_lds_bytes(const mf_launch* L, int64_t* lds_bytes);

/* Size of the library's static device scratch array (per GPU, held until the
 * process ends). */
#define MF_NOALLOC_SCRATCH_BYTES (64LL << 20)

/* mf_run without a caller workspace, for callers that cannot allocate device
 * memory (code under stream / graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference, which leaves all
 * buffers to XLA.  L->workspace and L->workspace_stride are ignored: the
 * kernels that keep per-instance arrays in device memory (lattices too large
 * for LDS) use a file-scope __device__ array of MF_NOALLOC_SCRATCH_BYTES
 * instead, B * mf_workspace_bytes (rounded up to 256) of it; a larger launch
 * is rejected with an error.  The call itself allocates nothing (no
 * hipMalloc, no host allocation); its device address is looked up once per
 * device on the first call.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (two streams, two threads) race.  Nothing in
 * it has to start at zero.  Results are identical to mf_run's. */
int mf_run_noalloc(const mf_launch* L, void* stream);

#ifdef __cplusplus
}
#endif
#endif
