/* dimcv.h -- C ABI of libdimcv.so, the gfx950 implementation of the DIMCV-SPH
 * particle step.
 *
 * The reference has no FFI: its hot path is a set of Taichi @ti.kernel methods
 * on Python classes.  Each entry point below replaces one of those kernels (or
 * a fixed sequence of them); the comment names the reference lines.  The
 * Python drop-in classes (particle_system.py, sph_base.py, dimcv_sph.py,
 * karman_vortex.py under dimcv-sph_amd/) call these entries through ctypes.
 *
 * Conventions
 *  - All pointers in dimcv_state are device pointers owned by the caller
 *    (PyTorch tensors).  The library never allocates device memory; scratch
 *    space is the caller-provided workspace of dimcv_workspace_bytes() bytes.
 *  - `stream` is a hipStream_t passed as void*; every launch is asynchronous on
 *    it.  Entries that return a host value (documented) synchronise the stream.
 *  - Return value: 0 on success, a negative dimcv_status on failure; the
 *    message is available from dimcv_last_error().  No C++ exception crosses
 *    the ABI.
 *  - Per-particle vectors are packed float[3] (x, y, z), grad_v is a row-major
 *    float[9], colour is int32[3]; bool fields are uint8.
 *  - Device-resident scalars: counters[] (int32, dimcv_counter) and dscal[]
 *    (double, dimcv_dscalar).
 */
#ifndef DIMCV_H
#define DIMCV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIMCV_ABI_VERSION 1

/* Largest particle capacity n_max: the neighbour passes address a lane's
 * list entries and its neighbours' fields by 32-bit byte offsets
 * (4 * entry and 12 * slot).  Entry points that take a workspace or build
 * neighbour lists return DIMCV_ERR_ARG above it. */
#define DIMCV_MAX_SLOTS (1 << 28)

typedef enum dimcv_status {
  DIMCV_OK = 0,
  DIMCV_ERR_ARG = -1,        /* invalid argument / shape */
  DIMCV_ERR_HIP = -2,        /* HIP runtime error */
  DIMCV_ERR_WORKSPACE = -3,  /* workspace too small */
  DIMCV_ERR_KEY = -4,        /* active solid particle whose cell key lies past the cell table */
  DIMCV_ERR_NBR = -5,        /* neighbour list width exceeded */
  DIMCV_ERR_RANGE = -6       /* size beyond a documented maximum */
} dimcv_status;

/* Indices into state->counters (int32). */
typedef enum dimcv_counter {
  DIMCV_C_PARTICLE_NUM = 0,   /* ps.particle_num            particle_system.py:48 */
  DIMCV_C_FLUID_NUM = 1,      /* ps.fluid_particle_num      :113 */
  DIMCV_C_SOLID_NUM = 2,      /* ps.solid_particle_num      :115 */
  DIMCV_C_NUM_SAMPLES = 3,    /* solver.num_samples         dimcv_sph.py:46 */
  DIMCV_C_FLUID_LIST = 4,     /* entries in state->fluid_list (= fluid slots) */
  DIMCV_C_NBR_MAX = 5,        /* largest neighbour count seen by the last build */
  DIMCV_C_KEY_ERROR = 6,      /* active solids whose cell key lies past the cell table */
  DIMCV_C_PREV_NUM = 7,       /* particle count before the last activity update */
  DIMCV_C_SAMPLE_SLOTS = 8,   /* is_sample count (compaction scratch) */
  DIMCV_C_JACOBI_DONE = 9,    /* native substep: current solve has converged */
  DIMCV_C_JACOBI_ITERS = 10,  /* native substep: iterations of the current solve */
  DIMCV_C_REDUCE_TICKET = 11, /* native substep: groups done in the running FP64 reduction (0 between calls) */
  DIMCV_C_STEP_ERROR = 12,    /* native substep: nonzero when the grid or list build failed (1 solid key past
                                 the table, 2 list width over nbr_kmax, 4 no fluid); the substep's writers of
                                 x, v, is_sample and life_time then return, so a failing step leaves them as
                                 the build left them.  Cleared by every grid build. */
  DIMCV_C_BAR_ARRIVE = 13,    /* native step, device-side Jacobi loop: grid-barrier arrivals */
  DIMCV_C_BAR_GEN = 14,       /* ... barrier generation */
  DIMCV_C_BAR_ABORT = 15,     /* ... nonzero when a barrier wait timed out */
  DIMCV_C_COUNT = 16
} dimcv_counter;

/* Indices into state->dscal (double). */
typedef enum dimcv_dscalar {
  DIMCV_D_KDEN_SUM = 0,   /* solver.kden_sum (FP64)              dimcv_sph.py:52 */
  DIMCV_D_RESIDUAL = 1,   /* last compute_density_error sum      :407-419 */
  DIMCV_D_CDF_TOTAL = 2,  /* total weight of the importance CDF */
  DIMCV_D_COUNT = 8
} dimcv_dscalar;

/* Velocity-recovery far-field evaluation (vort2vel, dimcv_sph.py:693-714). */
typedef enum dimcv_far_mode {
  DIMCV_FAR_EXACT = 0,  /* reference expression order, IEEE div/sqrt: bit-exact */
  DIMCV_FAR_FAST = 1,   /* per-sample weights folded, reciprocal-based, FP32 */
  DIMCV_FAR_KAHAN = 2   /* FAST with Kahan-compensated FP32 accumulation */
} dimcv_far_mode;

/* Scalar configuration.  Every float is the f32 value the reference kernels
 * see: Python f64 arithmetic first, then one rounding to f32 (the host
 * wrapper computes these exactly as particle_system.py:41-58,
 * sph_base.py:9-48 and dimcv_sph.py:13-64 do). */
typedef struct dimcv_params {
  int32_t grid_num[3];          /* ceil(domain_size / support)  particle_system.py:54 */
  int32_t n_cells;              /* prod(grid_num)               :55 */
  int64_t n_max;                /* particle_max_num             :117-123 */
  int32_t nbr_kmax;             /* neighbour-list width (build-defined) */
  int32_t periodic;             /* 0: reference clipping; 1: minimum-image wrap */
  float grid_size;              /* f32(support_radius), pos_to_index :382-384 */
  float support_radius;         /* f32(4 r) */
  float inv_support_radius;     /* f32(1 / (4 r)) */
  float kernel_k;               /* 2.5465f * f32(inv^3)         sph_base.py:33-35 */
  float kernel_dk;              /* (6 * 2.5465f) * f32(inv^3)   sph_base.py:56-57 */
  float density0;               /* f32(density0) */
  float dt;                     /* solver.dt[None] */
  float gravity[3];             /* f32(gravitation) */
  float visc_coef;              /* f32(2 (dim+2) nu)            dimcv_sph.py:170,177 */
  float visc_h2;                /* f32(0.01 support^2)          :178-179 */
  float m_eps;                  /* f32(1e-5)                    :24 */
  float domain_start[3];        /* f32, update_activity         particle_system.py:420 */
  float domain_end[3];
  float box_size[3];            /* periodic wrap lengths (periodic = 1 only) */
  int32_t has_vortex_domain;    /* dimcv_sph.py:76-84 */
  float vortex_start[3];
  float vortex_end[3];
  int32_t has_transit;          /* dimcv_sph.py:401-405 */
  float transit_a;              /* f32(area[0]) */
  float transit_ba;             /* f32(area[1] - area[0]) */
  float vp_factor;              /* f32(0.1592 * volumeScale * m_V0)  :691,714 */
  float eps;                    /* f32(1e-6)                    :15 */
  float inv_eps;                /* f32(1 / 1e-6)                :16 */
  float neg_gen_a;              /* f32(-GenProbPara)            :298 */
  float neg_del_b;              /* f32(-DelProbPara)            :303 */
  float far_r2_max_rejected;    /* largest f32 t with sqrt_rn(t) <= support */
  float bounce_padding;         /* f32(padding) for enforce_boundary_3D (sph_base.py:149-179) */
  float bounce_hi[3];           /* f32(domain_size - padding), same kernel */
  int32_t far_mode;             /* dimcv_far_mode */
  float m_V0;                   /* f32(0.8 (2 r)^3)             particle_system.py:46 */
  float nbr_r2_limit;           /* smallest f32 t with sqrt_rn(t) >= support: |r| < h <=> r.r < t */
  double vp_factor_f64;         /* 0.1592 * volumeScale * m_V0 in FP64 (direct FP64 sum) */
  int32_t dyn_rigid;            /* 1: the scene holds dynamic rigid particles (reaction writes,
                                   moving boundary volumes each step) */
  float inv_dt;                 /* f32(1 / dt), solver.inv_dt        dimcv_sph.py:28 */
} dimcv_params;

/* One scene of a batch (build-defined batching, dimcv_batch_*).  All scenes
 * of a batch share the particle radius, rest density, time step, DFSPH
 * tolerances and far-field mode (the dimcv_params of the batch); each has its
 * own box [0, domain_end] (pos_to_index ignores domain_start, so boxes start
 * at the origin), cell grid, viscosity and sample count.  The reference has
 * no batching; viscosity is a module constant there (sph_base.py:14). */
typedef struct dimcv_scene {
  int32_t grid_num[3];      /* ceil(domain size / grid_size) (particle_system.py:52-56) */
  int32_t cell_base;        /* first global cell key of the scene */
  int32_t n_cells;          /* grid_num[0] * grid_num[1] * grid_num[2] */
  int32_t m;                /* importance samples per step (> 0) */
  int32_t sample_base;      /* offset of the scene's samples in sample_idx and u_cdf */
  int32_t pad0;
  float domain_start[3];    /* activity bounds (particle_system.py:413-427) */
  float domain_end[3];
  float visc_coef;          /* f32(2 (dim + 2) nu_scene) (sph_base.py:14, dimcv_sph.py:186) */
  float pad1;
} dimcv_scene;

/* Per-scene device counters of a batch: scene_counters[s * DIMCV_SC_COUNT + k]. */
typedef enum dimcv_scene_counter {
  DIMCV_SC_DONE = 0,          /* Jacobi: converged */
  DIMCV_SC_ITERS = 1,         /* Jacobi: iterations of the current solve */
  DIMCV_SC_NUM_SAMPLES = 2,   /* samples drawn this step (m, or 0 when kden_sum is 0) */
  DIMCV_SC_FLUID = 3,         /* fluid-list entries of the scene (fluid_particle_num) */
  DIMCV_SC_FLUID_START = 4,   /* first fluid-list entry (multiple of 256) */
  DIMCV_SC_SLOT_START = 5,    /* first slot */
  DIMCV_SC_SLOTS = 6,         /* active slots (particle_num of the scene) */
  DIMCV_SC_TILE_START = 7,    /* first CDF scan tile */
  DIMCV_SC_SPLITS = 8,        /* far-field source splits */
  DIMCV_SC_PART_BASE = 9,     /* first far-field partial */
  DIMCV_SC_ITEM_START = 10,   /* first far-field work item */
  DIMCV_SC_ITERS_V = 11,      /* divergence iterations of the last substep */
  DIMCV_SC_ITERS_P = 12,      /* pressure iterations of the last substep */
  DIMCV_SC_FLUID_SRC = 13,    /* first entry in the build's unpadded fluid list */
  DIMCV_SC_COUNT = 16
} dimcv_scene_counter;

/* Per-scene FP64 scalars: scene_dscal[s * DIMCV_SD_COUNT + k]. */
typedef enum dimcv_scene_dscalar {
  DIMCV_SD_KDEN_SUM = 0,
  DIMCV_SD_RESIDUAL = 1,
  DIMCV_SD_CDF_TOTAL = 2,
  DIMCV_SD_COUNT = 4
} dimcv_scene_dscalar;

/* Device buffers.  Field names follow the reference (particle_system.py:143-220,
 * dimcv_sph.py:30-53).  Fields marked [sorted] are permuted by the grid build:
 * the build writes the permuted copy into the matching *_buffer pointer and the
 * caller swaps the pair afterwards (dimcv_grid_build). */
typedef struct dimcv_state {
  /* ParticleSystem */
  int32_t* object_id;        /* [sorted] */
  float* x;                  /* [sorted] */
  float* x_0;                /* [sorted] */
  float* v;                  /* [sorted] */
  float* acceleration;
  float* m_V;                /* [sorted] */
  float* m;                  /* [sorted] */
  float* density;            /* [sorted] */
  float* pressure;           /* [sorted] */
  int32_t* material;         /* [sorted] */
  int32_t* color;            /* [sorted] */
  uint8_t* is_dynamic;       /* [sorted] */
  float* vorticity;
  float* vorticity_vis;
  float* dfsph_factor;       /* [sorted] */
  float* density_adv;        /* [sorted] */
  uint8_t* is_active;        /* [sorted] */
  uint8_t* is_sample;        /* [sorted] */
  float* life_time;          /* [sorted] */
  int32_t* grid_ids;         /* [sorted] cell key per slot */
  /* permutation targets of the [sorted] fields */
  int32_t* object_id_buffer;
  float* x_buffer;
  float* x_0_buffer;
  float* v_buffer;
  float* m_V_buffer;
  float* m_buffer;
  float* density_buffer;
  float* pressure_buffer;
  int32_t* material_buffer;
  int32_t* color_buffer;
  uint8_t* is_dynamic_buffer;
  float* dfsph_factor_buffer;
  float* density_adv_buffer;
  uint8_t* is_active_buffer;
  uint8_t* is_sample_buffer;
  float* life_time_buffer;
  int32_t* grid_ids_buffer;
  /* cell table: inclusive prefix sum of per-cell counts (the reference's
   * grid_particles_num after PrefixSumExecutor.run), and the per-cell
   * remainder the counting sort leaves (all zero after a build) */
  int32_t* grid_particles_num;
  int32_t* grid_particles_num_temp;
  /* DIMCVSPHSolver */
  float* d_vort;
  float* d_vort_smoothed;
  float* v_star;
  float* grad_v;
  float* x_temp;
  int32_t* sample_idx;
  float* kinematic_vort_num;
  float* kernel_density;
  int32_t* num_particles_each_emitter;
  int32_t n_emitters;
  int32_t pad0;
  /* build-owned */
  int32_t* fluid_list;       /* ascending fluid slots, counters[DIMCV_C_FLUID_LIST] entries */
  int32_t* nbr_count;        /* per fluid-list entry: neighbours found */
  uint32_t* nbr;             /* [nbr_kmax][n_max]: slot | (is_fluid << 31), ascending slot order */
  float* v_vort;             /* optional [n_max][3]: recovered velocity per processed particle */
  int32_t* counters;         /* [DIMCV_C_COUNT] */
  double* dscal;             /* [DIMCV_D_COUNT] */
  void* workspace;
  int64_t workspace_bytes;
  /* batching (all NULL / 0 outside dimcv_batch_*) */
  int32_t* scene_id;           /* [sorted] scene of each slot */
  int32_t* scene_id_buffer;
  const dimcv_scene* scenes;   /* [n_scenes] (device) */
  int32_t n_scenes;
  int32_t pad1;
  int32_t* scene_counters;     /* [n_scenes][DIMCV_SC_COUNT] */
  double* scene_dscal;         /* [n_scenes][DIMCV_SD_COUNT] */
  int32_t* block_scene;        /* [n_max / 256 + 1]: scene of each 256-entry fluid-list block, -1 none */
  void* batch_workspace;       /* dimcv_batch_workspace_bytes */
  int64_t batch_workspace_bytes;
  /* optional [sorted] particle identity (insertion index), carried by the
   * grid build's gather like every sorted field; NULL = not tracked */
  int32_t* particle_id;
  int32_t* particle_id_buffer;
} dimcv_state;

/* ---- library ---------------------------------------------------------------- */
int dimcv_abi_version(void);
const char* dimcv_last_error(void);

/* Vector-ALU issue-rate calibration; no counterpart in the reference (it
 * serves the roofline of the far field, a kernel bound by vector
 * instruction issue).  Runs two dense streams of independent FP32 fused
 * multiply-adds, 4 waves per SIMD, on the current device and synchronises:
 * operands all in vector registers, then one operand a scalar register.
 * out[0] / out[2]: the shader clock during each stream (GHz, from s_memtime
 * against the 100 MHz s_memrealtime); out[1] / out[3]: wave-instructions
 * issued per SIMD and quad-cycle at that clock.  scratch: device memory of
 * at least dimcv_valu_issue_probe_bytes() (-1 without a device). */
int64_t dimcv_valu_issue_probe_bytes(void);
int dimcv_valu_issue_probe(void* scratch, int64_t scratch_bytes, double* out, void* stream);
/* Workspace bytes needed for n_max particles, n_cells cells and up to
 * max_samples importance samples. */
int64_t dimcv_workspace_bytes(int64_t n_max, int64_t n_cells, int64_t max_samples);

/* ---- particle insertion and scene set-up ------------------------------------ */
/* ParticleSystem.add_particles / _add_particles / add_particle
 * (particle_system.py:316-380).  Inputs are device arrays of n entries; slots
 * [particle_num, particle_num + n) are written, entries beyond n_max are
 * dropped (reference semantics, :361).  The dropped count is returned through
 * *dropped (host) after a stream synchronisation. */
int dimcv_add_particles(const dimcv_state* st, const dimcv_params* p, int32_t object_id, int32_t n,
                        const float* pos, const float* vel, const float* density,
                        const float* pressure, const int32_t* material, const int32_t* is_dynamic,
                        const int32_t* color, int32_t* dropped, void* stream);
/* One fluid emitter (FluidEmitters entry of the scene JSON).  `pos` is the
 * device copy of the layer add_cube builds for it (dimcv_sph.py:616-637:
 * lower corner squareCenter - squareSize/2, size squareSize + r|v|/|v|, the
 * numpy lattice of particle_system.py:647-696), the same at every emission. */
typedef struct dimcv_emitter {
  const float* pos;      /* device float[3 n] */
  int32_t n;             /* compute_cube_particle_num of the layer */
  int32_t object_id;
  float velocity[3];     /* f32(velocity) */
  float density;         /* f32(density) */
  int32_t color[3];
} dimcv_emitter;
/* DIMCVSPHSolver.emit_particle (dimcv_sph.py:616-637) with the emitter counts
 * of dump_num_particles_each_emitters_* (:639-647) kept on the device: each
 * emitter's layer is appended at particle_num (dropped beyond n_max, as
 * particle_system.py:361), particle_num and fluid_particle_num grow by the
 * slots written, num_particles_each_emitter[object_id] by the whole layer.
 * Asynchronous: no host copy, no synchronisation. */
int dimcv_emit(const dimcv_state* st, const dimcv_params* p, const dimcv_emitter* emitters, int32_t n_emitters,
               void* stream);
/* KarmanVortexSolver.export_ply (karman_vortex.py:122-214): the ASCII data
 * rows of the PLY file, formatted on the device from x_temp, vorticity,
 * color and object_id of the active slots below particle_num (slot order),
 * each float as Python's "{:.7f}" of its float32 value, the particle type
 * from obj_type[object_id] (device int32[n_obj_type]; 4 outside it).
 * _size returns the row count and byte size (synchronises); _rows writes the
 * rows into the device buffer `out` of out_bytes (at least that size).  The
 * header is the caller's (the reference's fixed text with the row count). */
int dimcv_export_ply_size(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound,
                          const int32_t* obj_type, int32_t n_obj_type, int64_t* rows, int64_t* bytes,
                          void* stream);
int dimcv_export_ply_rows(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound,
                          const int32_t* obj_type, int32_t n_obj_type, char* out, int64_t out_bytes,
                          void* stream);
/* The row formatter's float conversion, on the host (no device work): v as
 * Python's "{:.7f}" writes float(v), NUL-terminated into out (at least 49
 * bytes); returns the length.  For checking the device rows' digits. */
int32_t dimcv_format_f7(float v, char* out);
/* KarmanVortexSolver.init_cylinder (karman_vortex.py:27-36). */
int dimcv_init_cylinder(const dimcv_state* st, const dimcv_params* p, int32_t object_id, float cx,
                        float cy, float radius, void* stream);

/* ---- uniform grid ------------------------------------------------------------ */
/* ParticleSystem.initialize_particle_system (particle_system.py:555-560):
 * update_activity (:413-427), particle_partition (:429-488), update_grid_id
 * (:490-498), PrefixSumExecutor.run (:559) and counting_sort (:500-553),
 * restated as: activity -> cell keys (inactive slots keyed past the last cell)
 * -> stable LSD radix sort of (key, slot) -> cell table -> one gather of the
 * [sorted] fields into the *_buffer pointers -> ascending fluid list.
 * Slot order equals the serial reference order (ascending compaction, stable
 * counting sort).  The caller swaps every [sorted] field with its buffer
 * afterwards.  Synchronises the stream and copies counters[] to host_counters
 * (DIMCV_C_COUNT int32).  Cell keys are the reference's flat index; a fluid
 * particle whose key lies past the cell table (exactly on the upper x face)
 * is deactivated as one outside the domain, and DIMCV_ERR_KEY is returned
 * when an active solid's key does (both undefined behaviour in the
 * reference). */
int dimcv_grid_build(const dimcv_state* st, const dimcv_params* p, int32_t prev_particle_num,
                     int32_t* host_counters, void* stream);
/* Fixed-radius neighbour list of every fluid slot, the set and order that
 * ParticleSystem.for_all_neighbors (particle_system.py:562-586) visits: 27
 * cells in ndrange order, cell range [gpn[max(0,c-1)], gpn[c]) (cell 0 is never
 * visited), strict |x_i - x_j| < support test, j != i.  Returns DIMCV_ERR_NBR
 * (after a stream synchronisation) when a list exceeds nbr_kmax. */
int dimcv_neighbor_build(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                         void* stream);

/* ---- boundary volumes (sph_base.py:91-112) ---------------------------------- */
int dimcv_compute_static_boundary_volume(const dimcv_state* st, const dimcv_params* p,
                                         int32_t particle_bound, void* stream);
int dimcv_compute_moving_boundary_volume(const dimcv_state* st, const dimcv_params* p,
                                         int32_t particle_bound, void* stream);

/* Shape matching of a dynamic rigid object (sph_base.py:183-222,
 * compute_com and solve_constraints; the reference's step never calls
 * them).  dimcv_rigid_moments: FP64 moments over the object's dynamic rigid
 * particles, summed per block and then over blocks in block order (one
 * fixed order, where the reference's Taichi reduction order is unspecified);
 * mode 0: out[0] = sum m, out[1..3] = sum m x, m = f32(m_V0 * density);
 * mode 1: out[0..8] = sum m (x - cm)(x_0 - rest_cm)^T, row-major, cm and
 * rest_cm host float[3].  out: host double[9]; uses the workspace's far-field
 * partial region; synchronises `stream`.
 * dimcv_rigid_apply: x += (goal - x) * 1, goal = cm + R (x_0 - rest_cm), for
 * the same particles; T: host float[15] = cm[3], R[9] row-major, rest_cm[3]. */
int dimcv_rigid_moments(const dimcv_state* st, const dimcv_params* p, int32_t object_id, int32_t mode,
                        const float* cm, const float* rest_cm, int32_t particle_bound, double* out,
                        void* stream);
int dimcv_rigid_apply(const dimcv_state* st, const dimcv_params* p, int32_t object_id, const float* T,
                      int32_t particle_bound, void* stream);

/* ---- SPH / DFSPH passes (dimcv_sph.py).  `fluid_bound` / `particle_bound` are
 * host-side upper bounds of the fluid-list length / particle count used to
 * size the launch; the kernels read the exact counts from counters[]. ------- */
int dimcv_compute_densities(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);        /* :126-148 */
int dimcv_compute_DFSPH_factor(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);     /* :86-124 */
int dimcv_compute_density_change(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);   /* :337-374 */
/* field *= scale on fluid slots (:395-399). */
int dimcv_multiply_time_step(const dimcv_state* st, const dimcv_params* p, float* field, float scale,
                             int32_t fluid_bound, void* stream);
/* Canonical FP64 sum of w(x)(density_adv - offset) over fluid slots
 * (:407-419) into dscal[DIMCV_D_RESIDUAL]; if host_sum is non-NULL the stream
 * is synchronised and the sum copied there. */
int dimcv_compute_density_error(const dimcv_state* st, const dimcv_params* p, float offset,
                                int32_t fluid_bound, double* host_sum, void* stream);
int dimcv_divergence_solver_iteration(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :421-468 */
int dimcv_compute_non_pressure_forces(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, int32_t fluid_bound, void* stream); /* :150-208 */
int dimcv_predict_velocity(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream); /* :220-226 */
int dimcv_advect(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream);           /* :210-218 */
int dimcv_compute_density_adv(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :470-493 */
int dimcv_pressure_solve_iteration(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :495-549 */
int dimcv_copy_x_temp(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream);      /* :797-801 */
int dimcv_compute_vorticity_vis(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :778-795 */
/* sph_base.py:149-179 (fluid only, as called from karman_vortex.py:226-230). */
int dimcv_enforce_boundary_3d(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream);

/* ---- DIMCV passes ------------------------------------------------------------ */
int dimcv_compute_v_star(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);      /* :264-271 */
int dimcv_compute_vorticity(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);   /* :228-248,284-293 */
int dimcv_compute_grad_v(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);      /* :250-262,273-282 */
/* init() (:803-806) as one neighbour pass: v*, vorticity and grad_v. */
int dimcv_dimcv_init(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);
int dimcv_compute_d_vorticity(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :551-583 */
int dimcv_compute_all_d_vort_smoothed(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :660-677 */
/* compute_kinematic_vorticity_num_sample (:295-324).  u1/u2: device f32 arrays
 * indexed by slot replacing the two ti.random() draws; with do_bernoulli = 0
 * only the KVN is written (fixed-M importance sampling). */
int dimcv_compute_kvn_sample(const dimcv_state* st, const dimcv_params* p, const float* u1,
                             const float* u2, int32_t do_bernoulli, int32_t fluid_bound, void* stream);
/* record_vort_part_sample_idx (:326-335): ascending compaction. */
int dimcv_record_sample_idx(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream);
/* compute_kernel_density (:721-751) + canonical FP64 kden_sum. */
int dimcv_compute_kernel_density(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);
/* Build-defined inverse-CDF importance sampler: weights kden over in-domain
 * fluid slots, canonical FP64 inclusive scan, u (device f64[m]) mapped to
 * indices by upper_bound.  Writes sample_idx[0..m) and num_samples. */
int dimcv_cdf_sample(const dimcv_state* st, const dimcv_params* p, const double* u, int32_t m,
                     int32_t fluid_bound, void* stream);
/* No counterpart in the reference: dimcv_cdf_sample for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same arguments, kernels and results; st->workspace is
 * not used (it may be NULL).  Its scratch (FP64 weights, CDF and scan block
 * totals) is a fixed device array inside the library, one per GPU, held until
 * the process ends; fluid_bound above DIMCV_NOALLOC_MAX_FLUID returns
 * DIMCV_ERR_RANGE.  Allocates nothing.  Call with the stream's device
 * current; calls that overlap on one GPU race (they share its scratch). */
#define DIMCV_NOALLOC_MAX_FLUID (1 << 20)
int dimcv_cdf_sample_noalloc(const dimcv_state* st, const dimcv_params* p, const double* u, int32_t m,
                             int32_t fluid_bound, void* stream);
/* vort2vel (:649-719): near field over the neighbour list, far field over the
 * sample set (p->far_mode), velocity gate |v_vort| < |v|.  max_samples is the
 * capacity of the packed sample table (pass n_max; num_samples never exceeds
 * it) and must fit the workspace. */
int dimcv_vort2vel(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                   int32_t max_samples, void* stream);
/* The two halves of dimcv_vort2vel, in this order: near field + S_i
 * (dimcv_sph.py:686-691, into the workspace), then far field, combination and
 * velocity gate (:693-719). */
int dimcv_vort2vel_near(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream);
int dimcv_vort2vel_far(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                       int32_t max_samples, void* stream);
/* Number of source splits the fast/Kahan far field uses for `targets` fluid
 * particles, in the single-scene and the batched steps alike (host query, no
 * device work; -1 for a negative count).  Build-defined: a launch-shape
 * policy, no reference counterpart; splits * targets <= 2^21 always. */
int32_t dimcv_far_split_count(int64_t targets);
int dimcv_update_particle_life_time(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream); /* :753-760 */

/* ---- native substep ---------------------------------------------------------- */
/* Options of dimcv_substep.  The host-side scalars are the values the
 * reference's Python computes (f64 arithmetic, f32 kernel arguments). */
typedef struct dimcv_step_opts {
  const float* u1;          /* Bernoulli uniforms per slot (device) or NULL in fixed-M mode */
  const float* u2;
  const double* u_cdf;      /* fixed-M mode: m uniforms (device) */
  int32_t m;                /* fixed-M sample count; 0 selects the Bernoulli sample set */
  int32_t max_iter_v;       /* m_max_iterations_v (dimcv_sph.py:22) */
  int32_t max_iter_p;       /* m_max_iterations (:23) */
  int32_t chunk;            /* Jacobi iterations issued per host convergence check (>= 1); 0: the
                               first check after out_kv / out_kp iterations as passed in (the
                               previous step's counts, at least 1), then after every iteration.
                               Negative: dimcv_step converges on the device (after out_kv / out_kp
                               launched iterations, one persistent launch per solve runs any
                               further ones; the host waits once, at the end of the step);
                               dimcv_substep treats it as 0 and scenes with dynamic rigid
                               particles keep the host checks.  Unused by the batched step. */
  double eta_v;             /* inv_dt * 0.1 * 0.01 * rho0 (:389) */
  double eta_p;             /* 0.05 * 0.01 * rho0 (:611) */
  float inv_dt;             /* f32(1 / dt) (:27, :381) */
  float dt;                 /* f32(dt) (:393) */
  float inv_dt2;            /* f32(1 / dt^2) (:28, :602) */
  float density0;           /* offset of the pressure residual (:588) */
  int32_t fluid_count;      /* fluid_particle_num (residual divisor, :589,595) */
  int32_t bounce_back;      /* enforce_boundary_3D after the substep (karman_vortex.py:226-230) */
  int32_t out_kv;           /* output: divergence iterations */
  int32_t out_kp;           /* output: pressure iterations */
  int32_t time_far;         /* 1: time the far-field kernel with HIP events on `stream` */
  float out_far_ms;         /* output: far-field kernel duration (time_far = 1) */
} dimcv_step_opts;

/* KarmanVortexSolver.substep (karman_vortex.py:232-242) with the two Jacobi
 * solves (dimcv_sph.py:376-393, 597-614) run inside the library: each
 * iteration's kernels are skipped on the device once the solve has
 * converged, and the host checks convergence every `chunk` iterations
 * (chunk = 1 is the reference's cadence).  The kernels and their order are
 * those of the per-pass entries, so results are identical to calling them
 * one by one.  Requires the neighbour lists of the current positions. */
int dimcv_substep(const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* opts,
                  int32_t particle_bound, int32_t fluid_bound, void* stream);

/* One step after emission: ParticleSystem.initialize_particle_system
 * (particle_system.py:555-560) and the step's neighbour lists, then the
 * substep, in one call (karman_vortex.py:219-224 without the host round
 * trips between them).  `pre` is the state before the grid build; `st` is
 * the same state with every sorted field exchanged with its _buffer twin
 * (the state the caller uses after the build).  particle_bound is the
 * particle count before the build (or any larger value); all launches are
 * bounded by it and the kernels read the exact counts from the device.
 * opts->fluid_count is ignored (the residual divisor is the device's fluid
 * count).  The counter array (DIMCV_C_COUNT) is copied to host_counters at
 * the end of the step.  A step whose build fails (an active solid keyed past
 * the cell table, DIMCV_ERR_KEY; a list wider than nbr_kmax, DIMCV_ERR_NBR)
 * returns the error with x, v, life_time and the sample set unchanged: a
 * guard after the build makes every state writer of the substep skip. */
int dimcv_step(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* opts,
               int32_t particle_bound, int32_t* host_counters, void* stream);
/* dimcv_step in two halves, called in this order with the same arguments:
 * _begin enqueues the grid build, the lists and the substep up to the
 * divergence solve's first convergence check and returns without waiting;
 * _finish does the rest (dimcv_step = _begin + _finish).  The uniforms
 * (opts->u_cdf, or u1/u2) are first read by work _finish enqueues, so the
 * caller may fill them between the two calls with work enqueued on the same
 * stream (karman_vortex.py:219-224 draws them inside the step, after the
 * build). */
int dimcv_step_begin(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                     dimcv_step_opts* opts, int32_t particle_bound, void* stream);
int dimcv_step_finish(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                      dimcv_step_opts* opts, int32_t particle_bound, int32_t* host_counters, void* stream);
/* No counterpart in the reference (which steps synchronously): the second
 * half of a device-side step (opts->chunk < 0, no dynamic rigid particles)
 * without the wait at its end.  dimcv_step_finish_async enqueues the rest of
 * the step and a copy of the counters and returns; dimcv_step_wait waits for
 * that copy, fills host_counters, opts->out_kv / out_kp / out_far_ms and
 * returns the step's status as dimcv_step_finish would.  The caller may
 * enqueue the next step's dimcv_step_begin (whose particle_bound must still
 * bound the particle count, e.g. the count before the pending step when no
 * particles are inserted) before dimcv_step_wait, so the device runs the next
 * build while the host prepares it; a failed step is reported by the wait.
 * One asynchronous step may be outstanding: dimcv_step_finish,
 * dimcv_step_finish_async and dimcv_substep return DIMCV_ERR_ARG while one
 * is.  For drivers that pipeline steps (the bench's timed loop); the same
 * kernels, so results are those of dimcv_step. */
int dimcv_step_finish_async(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                            dimcv_step_opts* opts, int32_t particle_bound, void* stream);
int dimcv_step_wait(const dimcv_params* p, dimcv_step_opts* opts, int32_t* host_counters);

/* ---- comparators and initial conditions (build-defined) --------------------- */
/* Direct O(N^2) Biot-Savart over every in-domain fluid particle: the sum the
 * far-field Monte Carlo term estimates.  precision: 0 = FP32, 1 = FP32 Kahan,
 * 2 = FP64.  Writes v_vort (or out_f64 [n_max][3] for precision 2); does not
 * touch v. */
int dimcv_direct_bs(const dimcv_state* st, const dimcv_params* p, int32_t precision,
                    double* out_f64, int32_t fluid_bound, void* stream);
/* Vortex-ring initial velocity at n f32 positions: pts is a device table of
 * npts ring points, 7 doubles each (position, line element dl, G/(4 pi));
 * u(x) = sum g dl x (x - p) / |x - p|^3 (1 - exp(-|x - p|^2 inv_sigma2)),
 * FP64 accumulation, written as f32. */
int dimcv_ring_velocity(int64_t npts, const double* pts, double inv_sigma2, int64_t n,
                        const float* x, float* vel, void* stream);

/* ---- segment solver: O(N^2) regularised Biot-Savart --------------------------- */
/* The segment pool of SegmentSystem (segment_system.py:16-29) and the
 * constants of SegmentSolver (segment_solver.py:9-20).  Device arrays hold
 * segment_num = n entries (vectors packed float[3]); `sources` is scratch of at
 * least 32 n bytes. */
typedef struct dimcv_segments {
  float* x_minus;
  float* x_plus;
  float* gamma;
  const int32_t* active;   /* 1 = active */
  float* age;
  void* sources;           /* scratch: the packed source table */
  int32_t n;               /* segment_num */
  float reg_radius;        /* f32(regularizationRadiusR) */
  float dt;                /* f32(timeStepSize) */
  float dt6;               /* f32(dt / 6) */
  float inv4pi;            /* f32(1 / (4 pi)) */
  float gamma_decay;       /* f32(gammaDecay) */
  float u_inf[3];          /* backgroundVelocity */
} dimcv_segments;
/* _compute_endpoint_velocity_bs (form 0, segment_solver.py:378-418: at the
 * endpoints, vortex vector gamma L t) or _compute_velocity_at_factor (form 1,
 * :468-521: at x + factor dt k, vortex vector gamma d).  Writes the two
 * endpoint velocities of every active segment; inactive ones are left as
 * they are.  Sums run over active j != i in ascending j, bitwise the serial
 * reference arithmetic. */
int dimcv_segment_velocity(const dimcv_segments* sg, int32_t form, float factor, const float* k_minus,
                           const float* k_plus, float* out_minus, float* out_plus, void* stream);
/* advect_segments_rk4 (segment_solver.py:420-447, 458-542): k1 = (v_minus,
 * v_plus), k2..k4 by form 1 at factors 0.5, 0.5, 1, then the RK4 endpoint
 * update, gamma *= gamma_decay, age += dt.  k is scratch of 24 n floats (the
 * eight stage arrays). */
int dimcv_segment_rk4(const dimcv_segments* sg, const float* v_minus, const float* v_plus, float* k,
                      void* stream);

/* ---- batched scenes (build-defined) ----------------------------------------- */
/* Many independent scenes advance together in one set of launches.  Slots of
 * scene s are contiguous and ordered by scene (inactive slots of every scene
 * follow the last scene); cell keys are cell_base + the scene's own flattened
 * cell index, so the stable sort keeps each scene in its solo order.  The
 * fluid list gives each scene a segment starting at a multiple of 256 (unused
 * entries hold -1), so every per-block FP64 partial, every canonical
 * reduction and scan, every Jacobi convergence test and the far-field source
 * splitting of a scene are those of the same scene run alone: each scene's
 * results are bitwise equal to its solo run.  Batches are open boxes with no
 * solids, emitters, vortex domain or transit area, in fixed-M (inverse-CDF)
 * sampling mode.  The Jacobi solves converge on the device: after each
 * solve's first iteration (two ordinary launches), one persistent launch runs
 * the per-scene convergence tests and any further iterations of every scene,
 * its phases separated by grid barriers (its grid is at most the resident
 * block count; a barrier wait that exceeds about a second aborts the launch
 * and the step returns DIMCV_ERR_HIP).  Nothing here synchronises the host
 * except the final status readback of dimcv_batch_substep and
 * dimcv_batch_graph_step, and dimcv_batch_status.
 *
 * Capacities: n_max >= total particles + 256 * n_scenes; block_scene holds
 * n_max / 256 + 1 entries; n_cells (params) = sum of scene cells.
 * scene_particles: host array of the scenes' particle counts (must be given;
 * the layout depends only on n_scenes and n_max). */
int64_t dimcv_batch_workspace_bytes(int32_t n_scenes, const int32_t* scene_particles, int64_t n_max);
/* Grid build of every scene (activity, keys, sort, cell table, gather into
 * the *_buffer fields including scene_id, padded fluid list, per-scene
 * counters).  Every slot below n_max is scanned (slots at or past
 * particle_num get the sentinel key).  The caller swaps the sorted buffers
 * afterwards, then builds the neighbour lists. */
int dimcv_batch_grid_build(const dimcv_state* st, const dimcv_params* p, void* stream);
/* Neighbour lists of the current (swapped) buffers. */
int dimcv_batch_neighbor_build(const dimcv_state* st, const dimcv_params* p, void* stream);
/* Reads the error counters of the last build (key outside the grid,
 * neighbour-list overflow) and the counters; synchronises `stream`. */
int dimcv_batch_status(const dimcv_state* st, const dimcv_params* p, int32_t* host_counters, void* stream);
/* The substep of every scene (dimcv_substep per scene).  opts->u_cdf holds the
 * uniforms of all scenes (scene s at scenes[s].sample_base); opts->m is the
 * total sample count; fluid_count and the u1/u2 inputs are unused.
 * out_iters (host, optional): [n_scenes][2] divergence / pressure iterations. */
int dimcv_batch_substep(const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* opts,
                        int32_t* out_iters, void* stream);

/* Launch shapes of the batched far field and neighbour build
 * (result-invariant; chosen by the autotuner of the Python layer):
 * far_workgroups in [0, 65535] (0 = default 2048), nbr_block in {0 (= 256),
 * 64, 128, 256}. */
int dimcv_batch_set_launch(int32_t far_workgroups, int32_t nbr_block);

/* Step graphs: the whole batched step (grid build, neighbour lists, a guard
 * that stops the state writers after a failed build, substep with both
 * solves converging on the device) captured as one HIP graph, for each
 * parity of the sorted-buffer swap: a step is one graph launch and one
 * status readback.  st_even / st_odd: the state before an even / odd step
 * (buffers swapped).  `stream` must not be the null stream.  The uniforms are
 * read from opts->u_cdf at replay time, so the caller refreshes that buffer
 * before each step. */
typedef struct dimcv_batch_graph dimcv_batch_graph;
int dimcv_batch_graph_create(const dimcv_state* st_even, const dimcv_state* st_odd, const dimcv_params* p,
                             const dimcv_step_opts* opts, void* stream, dimcv_batch_graph** out);
/* One step from the captured graph; parity selects st_even / st_odd.  Returns
 * the build's errors (as dimcv_batch_status) after the step, whose state
 * writers were skipped on the device. */
int dimcv_batch_graph_step(dimcv_batch_graph* g, int32_t parity, int32_t* out_iters, void* stream);
void dimcv_batch_graph_destroy(dimcv_batch_graph* g);

/* ---- adjoint passes (build-defined differentiation) ------------------------ */
/* Vector-Jacobian products of the step's passes with the step's discrete
 * structure frozen (neighbour lists, samples, iteration counts and every gate
 * keep their forward values).  One entry, dimcv_adj_pass, runs the VJP of one
 * pass over the fluid list of `st` (the forward's neighbour lists and, for the
 * far field, its samples), reading forward values and the upstream adjoint
 * and ADDING into the adjoint outputs.  x_bar is the adjoint of the positions
 * x of `st`.  Open single-scene systems only.  The neighbour passes also read
 * grid_particles_num[0] of the grid build that made the lists (cell 0's
 * particles, which no list contains); FACTOR uses the workspace's near-field
 * region and the far-field passes its far-field and sample regions as scratch.
 *
 *  pass                 reference          f0 / f1 / f2                  g_in        outputs
 *  DENSITY              dimcv_sph.py:126   -                             rho_bar     x_bar
 *  FACTOR               :86-124            alpha                         alpha_bar   x_bar
 *  DIV       mode 0     :337-374           v / b (gate b > 0)            b_bar       x_bar, b1 = v_bar
 *            mode 1     :470-493           v / da (gate da > rho0)       da_bar      x_bar, b1 = v_bar, b0 = rho_bar
 *  JACOBI    mode 0     :421-468           b / alpha / rho (k = b a rho) P_bar       x_bar, b0 = kappa_bar
 *            mode 1     :495-549           da / alpha (k = (da-rho0) a)  P_bar       x_bar, b0 = kappa_bar
 *  NONPRESSURE          :150-208           v                             a_bar       x_bar, b1 = v_bar
 *  CURL                 :228-248, :551-583 u (v or v*)                   C_bar       x_bar, b1 = u_bar
 *  GRAD_V               :250-262           v                             G_bar (9)   x_bar, b1 = v_bar
 *  SMOOTH               :660-677           d_vort                        w_bar       x_bar, b1 = d_vort_bar
 *  KVN                  :295-312           grad_v (9)                    kvn_bar     b1 = grad_v_bar (9)
 *  KDEN                 :721-751           kvn                           kden_bar    x_bar, b1 = kvn_bar
 *  GATE                 :717-719           v_vort / v                    -           b2 = gate (1 / 0)
 *  FAR_TARGETS          :693-716           w / kden; kden_sum; mode = M  u_bar       x_bar, b0 = S_bar, b2 = weights
 *  FAR_SOURCES          :693-716           w / kden / weights; mode = M  -           x_bar, b0 = w_bar, b1 = kden_bar
 *  NEAR                 :649-658, 687-692  w / S_bar                     u_bar       x_bar, b0 = w_bar, b1 = kden_bar
 *
 * P_bar is the adjoint of the sum a Jacobi iteration subtracts from v (minus
 * the adjoint of the iteration's output velocity); u_bar is the gated
 * adjoint of the recovered velocity.  The far-field passes take the sample
 * bound M in `mode` (kden_sum in `dscalar`) and pack samples and targets into
 * the workspace's far-field regions (DIMCV_ERR_WORKSPACE when they do not
 * fit). */
typedef enum dimcv_adj_pass_id {
  DIMCV_ADJ_DENSITY = 0,
  DIMCV_ADJ_FACTOR = 1,
  DIMCV_ADJ_DIV = 2,
  DIMCV_ADJ_JACOBI = 3,
  DIMCV_ADJ_NONPRESSURE = 4,
  DIMCV_ADJ_CURL = 5,
  DIMCV_ADJ_GRAD_V = 6,
  DIMCV_ADJ_SMOOTH = 7,
  DIMCV_ADJ_KVN = 8,
  DIMCV_ADJ_KDEN = 9,
  DIMCV_ADJ_GATE = 10,
  DIMCV_ADJ_NEAR = 11,
  DIMCV_ADJ_FAR_TARGETS = 12,
  DIMCV_ADJ_FAR_SOURCES = 13
} dimcv_adj_pass_id;

typedef struct dimcv_adj_args {
  const float* f0;
  const float* f1;
  const float* f2;
  const float* g_in;
  float* x_bar;
  float* b0;
  float* b1;
  float* b2;
  double dscalar;
  float scalar;
  int32_t mode;
} dimcv_adj_args;

int dimcv_adj_pass(const dimcv_state* st, const dimcv_params* p, int32_t pass, const dimcv_adj_args* args,
                   int32_t fluid_bound, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DIMCV_H */
