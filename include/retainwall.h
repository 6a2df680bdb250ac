/*
 * retainwall.h -- C ABI of the MI355X (gfx950) SPH retaining-wall step.
 *
 * The library (libretainwall.so) owns every device array of one simulation
 * context; the caller owns host buffers.  No device pointer crosses this
 * boundary, no HIP/torch type appears in a signature, and every entry point
 * returns an int status (0 = LQ_OK) with the message kept in
 * lq_last_error(ctx).  One context = one GPU = one HIP stream; a context is
 * not thread-safe.
 *
 * Each entry point names the reference routine it replaces (file:line in the
 * reference's src/ directory).  The reference reaches all of them as private
 * methods of LoquatSph / LoquatGrid (loquat.h:97-120, loquat_grid.h:46-66); a
 * host that keeps those class names calls these functions underneath.
 */
#ifndef RETAINWALL_H
#define RETAINWALL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LQ_MODEL_NUMBER_MAX 10  /* types.h:38 */
#define LQ_MATERIAL_SIZE_MAX 15 /* types.h:39 */
#define LQ_RIGID_MOTION_MAX 10  /* types.h:40 */
#define LQ_TABLE_MAX 10         /* types.h:80-81 (table = scalar[10]) */

enum lq_status {
  LQ_OK = 0,
  LQ_ERR_ARG = 1,      /* bad argument / size / unsupported option          */
  LQ_ERR_HIP = 2,      /* a HIP runtime call failed                         */
  LQ_ERR_DELETE = 3,   /* npb + npm + np_delete_m != np (loquat_grid.cu:182) */
  LQ_ERR_DOMAIN = 4,   /* a boundary particle left the cell grid            */
  LQ_ERR_STATE = 5     /* call out of order (e.g. step before upload)       */
};

/* types.h:47-50 */
enum lq_model_mat { LQ_RIGID = 0, LQ_DP_EPP = 1, LQ_MUI_DP_EPP = 2, LQ_HYPO_BASIC = 3, LQ_HYPO_CS = 4 };
enum lq_fix { LQ_FIX = 0, LQ_NONFIX = 1 };           /* types.h:46 */
enum lq_shape { LQ_CUBOID = 0, LQ_CIRCLE = 1 };      /* types.h:49 */

/* Cell-key function.  LINEAR reproduces the reference hash
 * key = cx*ny*nz + cy*nz + cz (+ncell_material for soil), loquat_grid.cu:88-90,
 * bit for bit.  MORTON interleaves the bits of the cell coordinates (x, y, z
 * from the high end, z lowest) on the cube of 2^b cells per axis that holds
 * the grid, so consecutive slots are compact blocks of space; lq_create then
 * sets ncell_material = 2^(3b), ncell_delete_m = 2^(3b+1), ncell = that + 1
 * (b <= 9).  MORTON needs a single-case context and the REFERENCE_ORDER,
 * NEIGHBOR_LIST or PAIR_LIST variant; neighbour sets, the per-particle cell
 * order of every sum and hence the physics are those of LINEAR keys (the
 * order of particles inside a cell can differ once particles change cells,
 * which changes sums by rounding only). */
enum lq_key_mode { LQ_KEY_LINEAR = 0, LQ_KEY_MORTON = 1 };

/* Interaction kernel variant.  The first four produce bit-identical results
 * (the reference's per-particle summation order).  REFERENCE_ORDER walks the
 * 27 cells per thread exactly as particle_interaction.cu:117-140; TILE stages
 * each cell-column tile's neighbourhood in LDS and scans it from there;
 * NEIGHBOR_LIST (the default of the Python and C++ hosts) scans global memory
 * per lane and compacts hits into LDS lists before evaluating them;
 * PAIR_LIST scans once per step into per-slot pair lists in HBM (cut-off
 * plus a 1% skin, traversal order kept) that both interaction calls of the
 * step evaluate, falling back to the cell scan on the device when the
 * predictor moved a particle further than the skin allows or a list
 * overflowed.  COLUMNS searches h-sized sub-cells: per call it indexes every
 * particle by 2h column, sub-column and sub-cell height, and each soil
 * particle scans, in the 5 x 5 sub-columns around it, only the sub-cell
 * window that can hold a neighbour (about 180 candidates where the 27-cell
 * scan tests 450).  Its neighbour sets are the 27-cell scan's, bit for bit
 * (same FP32 predicate), but it sums the pairs in column order with
 * hardware reciprocal / square root / exp2 / log2, so its results agree with
 * the others to FP32 rounding, not bit for bit.
 * Size limits: REFERENCE_ORDER, TILE, NEIGHBOR_LIST and COLUMNS handle every
 * context (fewer than 2^26 particles, the lq_create limit); PAIR_LIST needs
 * fewer than 2^25 particles; COLUMNS needs LQ_KEY_LINEAR, no periodic planes
 * and at most 256 cells along z.  lq_autotune picks the fastest for the
 * current state within the bit-identical four, or between the two block
 * orders and register builds of COLUMNS for a context set to COLUMNS.  In a
 * batched context (lq_create_batch) every variant gives each case the
 * results, bit for bit, of a one-case context running the same variant. */
enum lq_interaction_variant {
  LQ_INTERACTION_REFERENCE_ORDER = 0,
  LQ_INTERACTION_TILE = 1,
  LQ_INTERACTION_NEIGHBOR_LIST = 2,
  LQ_INTERACTION_PAIR_LIST = 3,
  LQ_INTERACTION_COLUMNS = 4
};

/* RigidMotion, types.h:66-76 (the per-step imposed motion of one body). */
typedef struct lq_rigid_motion {
  int32_t is_trans;
  int32_t is_rotat;
  float trans_vel[3];
  float trans_acc[3];
  int32_t pattern; /* lq_shape */
  float pos_rp[3];
  float rotat_vel[3];
  float rotat_acc[3];
} lq_rigid_motion;

/* MotionTable, types.h:81-97: one body's piecewise-constant acceleration
 * tables, read from <case>.motion (loquat_io.cu:276-288). */
typedef struct lq_motion_table {
  int32_t is_translation;
  int32_t table_size_trans;
  float trans_vel[LQ_TABLE_MAX][3];
  float trans_time[LQ_TABLE_MAX];
  float trans_acc[LQ_TABLE_MAX][3];
  int32_t is_rotation;
  int32_t pattern;
  int32_t table_size_rotat;
  float pos_rp[3];
  float rotat_vel[3];
  float rotat_time[LQ_TABLE_MAX];
  float rotat_acc[LQ_TABLE_MAX][3];
} lq_motion_table;

/* POD mirror of Parameters (types.h:100-165): same names, units and meaning.
 * model_par[k][0..14] follows GenCase.cu:52-68 for HypoCS.  Derived fields
 * (h, i_h, eta, bound_thickness, vol_def, cell_num, ncell*, kernel_*) are
 * filled by lq_initialize_parameters (loquat.cu:130-190). */
typedef struct lq_params {
  int32_t dim;
  int32_t periodbound;
  float planeleft[4];
  float planeright[4];
  int32_t n_motion;
  lq_rigid_motion mv[LQ_RIGID_MOTION_MAX];
  int32_t np, npb, npm, np_delete_m;
  float domain_size[3];
  float domain_min[3];
  int32_t cell_num[3];
  int32_t ncell, ncell_material, ncell_delete_m;
  float h, i_h, dr, hdr, vol_def, bound_thickness;
  float gravity[3];
  float rho0_m, damping;
  int32_t n_model;
  float model_par[LQ_MODEL_NUMBER_MAX][LQ_MATERIAL_SIZE_MAX];
  float kernel_normalization_par, kernel_gradient_normalization_par;
  int32_t stress_shepard_filter, dynamic_time_step, integration_method, viscosity_type;
  float artificial_viscosity_alpha_m, cs_m, eta, cfl;
  float artificial_pressure_coefficient_material, kernel_zero, kernel_dr;
  float time_max, time, dt, output_frequency;
  int32_t output_number, output_format;
  /* --- extensions with no reference field --- */
  int32_t key_mode;            /* lq_key_mode */
  int32_t interaction_variant; /* lq_interaction_variant */
} lq_params;

/* Host AoS views in the .dat column semantics (loquat_io.cu:229-254).
 * Positions are absolute on upload.  str/stn are xx yy zz xy xz yz. */
typedef struct lq_host_state {
  const int32_t* codes;   /* [np][4] block id, Fix/NonFix, model, motion id */
  const float* pospres;   /* [np][4] x y z p                                */
  const float* velrhop;   /* [np][4] vx vy vz rho                           */
  const float* idmassidx; /* [np][4] id mass idx M_crt (idx is overwritten) */
  const float* str;       /* [np][6]                                        */
  const float* stn;       /* [np][6]                                        */
  const float* normal;    /* [npb][4] nx ny nz slip(-1 no-slip)             */
} lq_host_state;

/* Field ids for lq_get_field / lq_set_field.  Every field is per sorted slot
 * with np rows in the host AoS layout noted; positions are cell-relative. */
enum lq_field {
  LQ_F_CODES = 0,     /* int32[np][4]                                      */
  LQ_F_CELLPOS = 1,   /* int32[np][4]  cx cy cz key                        */
  LQ_F_POS = 2,       /* f32[np][4]    relative x y z, p                   */
  LQ_F_POS_ORI = 3,   /* f32[np][4]    accumulated displacement            */
  LQ_F_VEL = 4,       /* f32[np][4]    vx vy vz rho                        */
  LQ_F_IDM = 5,       /* f32[np][4]    id mass idx M_crt                   */
  LQ_F_POS_INI = 6,   /* f32[np][4]                                        */
  LQ_F_VEL_INI = 7,   /* f32[np][4]                                        */
  LQ_F_ACCAD = 8,     /* f32[np][4]    ax ay az drho/dt                    */
  LQ_F_SPIN = 9,      /* f32[np][4]    w_xy w_xz w_yz 0 (SkewTensor order) */
  LQ_F_CVAR = 10,     /* f32[np][4]    K d slip wall-id  (contact_var)     */
  LQ_F_NORMAL = 11,   /* f32[np][4]    boundary rows meaningful            */
  LQ_F_STR = 12,      /* f32[np][6]                                        */
  LQ_F_STN = 13,      /* f32[np][6]                                        */
  LQ_F_STR_INI = 14,  /* f32[np][6]                                        */
  LQ_F_STN_INI = 15,  /* f32[np][6]                                        */
  LQ_F_ACESTN = 16,   /* f32[np][6]    strain rate D                       */
  LQ_F_CMAG = 17,     /* f32[np][9]    force(xx xy xz) normal(yx yy yz) wall-vel(zx zy zz) */
  LQ_F_PART_CEL = 18, /* int32[np]     sorted keys                         */
  LQ_F_PART_IDX = 19, /* int32[np]     gather map new slot -> old slot     */
  LQ_F_CELL_BEG = 20, /* int32[ncell]                                      */
  LQ_F_CELL_END = 21, /* int32[ncell]                                      */
  LQ_F_ORIG_ID = 22,  /* int32[np]     original .dat line of each slot     */
  LQ_F_COUNT = 23
};

typedef struct lq_ctx lq_ctx;

/* ---- context ------------------------------------------------------------ */

/* sizeof(lq_params), sizeof(lq_motion_table), sizeof(lq_rigid_motion),
 * sizeof(lq_host_state) into sizes[0..n) so bindings can check their layout.
 * No reference counterpart (binding aid). */
int lq_abi_sizes(size_t* sizes, int n);

/* InitializeParameters, loquat.cu:130-190: fill the derived fields of *par
 * in place from dim, dr, hdr, domain_size, kernel choice. */
int lq_initialize_parameters(lq_params* par);

/* LoquatSph ctor + AllocateGpuMemory + GridConfig (loquat.cu:29-69, 193-220;
 * loquat_grid.cu:59-68).  par must already be initialised.  With
 * periodbound set the interaction variant must be REFERENCE_ORDER or
 * NEIGHBOR_LIST (the variants that run the periodic image pass); batch
 * contexts reject periodbound. */
int lq_create(const lq_params* par, int device, lq_ctx** out);
void lq_destroy(lq_ctx* ctx);
const char* lq_last_error(const lq_ctx* ctx);

/* CopyDataFromHostToDevice (loquat.cu:223-234), then
 * PositionFromAbsolute2Relative (loquat_grid.cu:253-271). */
int lq_upload(lq_ctx* ctx, const lq_host_state* host);

/* LoquatIo::MotionInfoALL (loquat_io.cu:266-303): motion tables per body.
 * The tables live on the device and advance in lq_rigid_body_update;
 * lq_get_motion_tables reads their current state back. */
int lq_set_motion_tables(lq_ctx* ctx, int n, const lq_motion_table* tables);
int lq_get_motion_tables(lq_ctx* ctx, int n, lq_motion_table* tables);

/* Raw field access for checkpoints, IO and tests (host AoS layouts above). */
int lq_get_field(lq_ctx* ctx, int field, void* dst, size_t bytes);
int lq_set_field(lq_ctx* ctx, int field, const void* src, size_t bytes);

/* Current parameters as seen by the device (time, dt, mv, npm, np_delete_m). */
int lq_get_params(lq_ctx* ctx, lq_params* out);
/* Host-side time (FP64, loquat.cu:558) and step counter. */
int lq_get_time(lq_ctx* ctx, double* time, int64_t* step);
int lq_set_time(lq_ctx* ctx, double time, int64_t step);

/* Checkpoint / restart.  lq_checkpoint_write stores the whole device state
 * a step reads (every per-slot array, the predictor snapshots and rates,
 * cell ranges, imposed-motion tables, time, dt, step and particle counts)
 * as raw bits; lq_checkpoint_read restores it into a context created with
 * the same parameters (same shape, else LQ_ERR_ARG), after which steps
 * continue bit for bit as the run that wrote it.  Replaces the reference's
 * restart from N.dat + output_N.xml (loquat_io.cu:315-468, read back by
 * LoadParameters/LoadParticles), which loses the contact state, the
 * displacement record and the motion tables' progress.  The surrogate's
 * networks are not part of a checkpoint. */
int lq_checkpoint_write(lq_ctx* ctx, const char* path);
int lq_checkpoint_read(lq_ctx* ctx, const char* path);

/* ---- step pieces (each enqueues work on the context stream) -------------- */

/* StepCalculation, loquat.cu:366-374: dt = cfl*h/cs, par.time = time. */
int lq_step_calculation(lq_ctx* ctx);
/* LoquatGrid::NeighborSearch, loquat_grid.cu:147-191: hash, stable sort,
 * begin/end, deleted-particle bookkeeping (device-side, no host sync). */
int lq_neighbor_search(lq_ctx* ctx);
/* LoquatSph::SortArraysPrimary / SortArraysAll, loquat.cu:254-287 (contact
 * state is carried with its particle as well). */
int lq_sort_arrays_primary(lq_ctx* ctx);
int lq_sort_arrays_all(lq_ctx* ctx);
/* LoquatSph::RigidBodyUpdate, adami_boundary.cu:175-203 (runs on device). */
int lq_rigid_body_update(lq_ctx* ctx);
/* LoquatSph::AdamiBoundary, adami_boundary.cu:89-169. */
int lq_adami_boundary(lq_ctx* ctx);
/* LoquatSph::FindCornerParticle, adami_boundary.cu:209-306. */
int lq_find_corner_particle(lq_ctx* ctx);
/* LoquatSph::StressRegularization, field_smoothing.cu:74-101. */
int lq_stress_regularization(lq_ctx* ctx);
/* LoquatSph::StressSavingRegularization, field_smoothing.cu:103-112: the
 * Shepard-smoothed stress written only to dst f32[np][6] for output (soil
 * rows smoothed, other rows copied); the state is not changed.  Synchronises. */
int lq_stress_saving_regularization(lq_ctx* ctx, float* dst, size_t bytes);
/* ParticleInteractionMaterial, particle_interaction.cu:260-305, with the
 * wall-contact force resolved in two deterministic passes. */
int lq_particle_interaction(lq_ctx* ctx);
/* LoquatSph::Predictor, predictor_corrector.cu:52-103 (snapshot + update). */
int lq_predictor(lq_ctx* ctx);
/* LoquatSph::Corrector, predictor_corrector.cu:109-214. */
int lq_corrector(lq_ctx* ctx);

/* One or more whole time steps (loquat.cu:548-561 without Saving): the
 * predictor and corrector sections in order.  use_graph != 0 captures two
 * steps in a hipGraph once (again after a layout, variant or timing change)
 * and replays it; an odd step is enqueued directly.  No host
 * synchronisation inside. */
int lq_step(lq_ctx* ctx, int nsteps, int use_graph);

/* Wall earth-pressure probe, utilities.cu:578-670: F = 0.001 * sum fx [kN]
 * over slots npb..np (soil, deleted particles' last contact force
 * included, as the reference's loop) with x_abs <= box[0],
 * box[1] <= z_abs <= box[2].  The
 * reference hard-codes box = {0.04+bt+3dr, bt+0.04, bt+1.04} and sums
 * serially in FP32; here the fx values are summed exactly (128-bit fixed
 * point, 2^-64 resolution, independent of order) and scaled once.
 * Synchronises. */
int lq_wall_force(lq_ctx* ctx, const float box[3], double* force_kn);
/* Wall pressure profile, GetValuesAlongLine (utilities.cu:224-367, written
 * to stress_zz<N>.txt at each save).  The reference hard-codes the wall face
 * x = 0.04 + bt - 0.005 t, the depth range [bt+0.04, bt+1.04] and
 * int(1/dr) bins; here geom = {x_face, z_lo, z_hi, z_top} and nbins are
 * arguments.  Soil slots (deleted ones included, as the reference loops to
 * np) in the bands [x_face + (k-1) dr, x_face + k dr], k = 1..3, with
 * z_lo <= z <= z_hi fall in bin int((z_top - z) / dr); per bin and band the
 * last slot in slot order gives contact-force x / dr.  out f32[nbins][4] =
 * (depth = bin dr where band 1 has a slot, p1, p2, p3); bins without a slot
 * read 0 (the reference leaves them unwritten).  Single-case contexts.
 * Synchronises. */
int lq_wall_pressure_profile(lq_ctx* ctx, const float geom[4], int nbins, float* out, size_t bytes);
/* lq_wall_pressure_profile without allocation: the same inputs, outputs
 * and results, with the per-(band, bin) markers and the staged rows held in
 * a fixed device array of the library (one per GPU, kept until the process
 * ends) instead of memory allocated per call, so no hipMalloc/hipFree or
 * allocating routine runs.  nbins is limited to LQ_PROFILE_NOALLOC_MAX_BINS
 * (else LQ_ERR_ARG).  Calls that overlap on one GPU (several host threads,
 * or contexts on separate streams) share that array and race; serialise
 * them.  No counterpart in the reference: for callers that cannot allocate
 * device memory (code under graph capture, serving loops, harnesses that
 * forbid allocation).  Synchronises. */
#define LQ_PROFILE_NOALLOC_MAX_BINS 4096
int lq_wall_pressure_profile_noalloc(lq_ctx* ctx, const float geom[4], int nbins, float* out, size_t bytes);
/* Stress sensors at caller-given points, GetSoilStress/cuGetStressZz
 * (utilities.cu:470-528; the reference's InterpolateResults call at
 * loquat.cu:463 is commented out and its point is hard-coded).  points
 * f32[n][3] absolute; per point the Shepard-weighted stress
 * sum_j sigma_j W V_j / sum_j W V_j over soil slots with |x - x_j| < 2h,
 * x_j the absolute position (relative + cell 2h in FP32) and V_j = m_j/rho_j
 * (summed serially in cell order).  reference_rows != 0: the reference's
 * cells (the point's own y row of cells, x +-1, every z level); 0: the 27
 * cells around the point.  out f32[n][8] = (xx yy zz xy xz yz, sum W V,
 * number of slots); the reference computes xx alone.  Single-case
 * contexts.  Synchronises. */
int lq_sensor_stress(lq_ctx* ctx, const float* points, int n, int reference_rows, float* out, size_t bytes);
/* Free-surface heights at caller-given points (x, y), GetSurface
 * (utilities.cu:26-91) and GetSurfaceAtPoint (:371-427), which the
 * reference's driver never calls and which compare their absolute points
 * with the solver's cell-relative positions; here positions are absolute.
 * points f32[n][2]; out f32[n] = the largest absolute z of a soil slot near
 * the vertical line through the point, -1e10 where there is none.
 * radius > 0: horizontal distance < radius, cells x +-1 and y +-1 around
 * the point, every z (GetSurface, radius 0.5 dr); radius < 0: |x - x_j| <
 * -radius over the point's own y row of cells, x +-1, every z
 * (GetSurfaceAtPoint, -2h).  |radius| <= 2h.  Single-case contexts.
 * Synchronises. */
int lq_surface_heights(lq_ctx* ctx, const float* points, int n, float radius, float* out, size_t bytes);
/* Strip surcharge on the free surface: every soil slot whose block id
 * (codes.x) equals block carries a downward force pressure * dr^2 (the load
 * on one lattice column), added with gravity in the interaction.  block -1
 * removes it; contact-wall ids (>= 10) are rejected.  No reference
 * counterpart: the 10 kPa strip load of the large sloping-backfill workload,
 * whose generator marks the top particle of each column under the strip. */
int lq_set_surcharge(lq_ctx* ctx, int block, float pressure);

/* Absolute positions (PositionFromRelative2Absolute, loquat_grid.cu:273-289)
 * into dst f32[np][4].  Synchronises. */
int lq_positions_absolute(lq_ctx* ctx, float* dst, size_t bytes);

/* Counts after deletions: npb, npm, np_delete_m.  Synchronises. */
int lq_get_counts(lq_ctx* ctx, int32_t* npb, int32_t* npm, int32_t* np_delete_m);

/* Block until the context stream is idle; reports device-side errors
 * (deletion bookkeeping, boundary out of grid). */
int lq_synchronize(lq_ctx* ctx);

/* ---- instrumentation (no reference counterpart) --------------------------
 * When enabled, lq_step and the piecewise calls record HIP events around
 * each kernel class on the stream it runs on (Adami and the corner search of
 * COLUMNS steps run on a second stream beside the column index and the
 * interaction; their times are spans while they share the GPU with those
 * kernels, not their time alone); lq_kernel_times returns the
 * accumulated milliseconds and launch counts per class.  Graph replays are
 * timed too: the captured graph holds the events as event-record nodes,
 * pointed at fresh events before every replay.  Once 8192 event pairs are
 * waiting, lq_step synchronises between steps to fold them into the totals. */
enum lq_kernel_class {
  LQ_K_NEIGHBOR_SEARCH = 0, LQ_K_SORT_ARRAYS = 1, LQ_K_ADAMI = 2, LQ_K_CORNER = 3,
  LQ_K_INTERACTION = 4, LQ_K_CONTACT = 5, LQ_K_PREDICTOR = 6, LQ_K_CORRECTOR = 7,
  LQ_K_SHEPARD = 8, LQ_K_PAIR_LIST_BUILD = 9,
  /* COLUMNS: the sub-cell index and the pair records it writes (two timed
   * sections per call when Adami runs beside it: soil records, then the
   * wall records once Adami is done); LQ_K_INTERACTION then times the
   * interaction kernel alone */
  LQ_K_COLUMN_INDEX = 10, LQ_K_CLASS_COUNT = 11
};
int lq_enable_timing(lq_ctx* ctx, int enable);
/* Times only the classes whose bit (1 << lq_kernel_class) is set in mask
 * (0 turns timing off) and resets the counters; fewer timed sections keep
 * more of the graph replay's advantage over per-kernel launches. */
int lq_enable_timing_classes(lq_ctx* ctx, unsigned mask);
/* STREAM-style copy of `bytes` between two device scratch buffers, `reps`
 * launches, average milliseconds per launch (HBM bandwidth probe and the
 * known-size copy used to calibrate hardware byte counters). */
int lq_stream_copy(lq_ctx* ctx, size_t bytes, int reps, double* ms_per_launch);
int lq_kernel_times(lq_ctx* ctx, double* ms, int64_t* launches, int n);

/* Tile shape of the TILE interaction variant: shape index (0: 1 wave,
 * 1: 2 waves, 2: 4 waves per workgroup) and tz = cells per tile along z.
 * Contact arrays are restored after tuning.  No reference counterpart (the
 * reference has one fixed 128-thread launch). */
int lq_set_tile(lq_ctx* ctx, int shape, int tz);
/* Switch the interaction variant of a live context; xcd_remap != 0 makes
 * the particle-parallel variants hand each of the 8 XCDs one contiguous run
 * of particle blocks (hardware places block b on XCD b % 8).  A periodic
 * context accepts REFERENCE_ORDER and NEIGHBOR_LIST only.  No reference
 * counterpart. */
int lq_set_interaction(lq_ctx* ctx, int variant, int xcd_remap);
/* Register budget of the COLUMNS interaction kernel: 0 at most 96 VGPRs
 * (5 waves per SIMD), 1 at most 128 (the compiler may use fewer; the step
 * form compiles to 96 either way).  Both give
 * bit-identical results; lq_autotune picks one for a context set to
 * COLUMNS; the default is 0.  No reference counterpart. */
int lq_set_columns_build(lq_ctx* ctx, int build);
/* COLUMNS steps: run Adami (cuAdamiBoundaryMaterial, adami_boundary.cu:
 * 26-158) and the corner search (cuFindCornerParticleMaterial,
 * adami_boundary.cu:209-292) on a second stream beside the column build and
 * the interaction.  The walls' pair records are written once Adami is done;
 * the interaction leaves the wall contact of particles whose cell
 * neighbourhood holds a wall particle to a pass after the search (on by
 * default; fused steps do not use it).  Results are bit-identical either
 * way.  Their timed sections then sit on the second stream.  No reference
 * counterpart. */
int lq_set_overlap(lq_ctx* ctx, int on);
/* Fused steps of the COLUMNS variant (off by default): lq_step then advances
 * every soil particle whose contact force needs no redistribution (pass B)
 * through cuPredictorMaterial / cuCorrectorMaterial (predictor_corrector.cu
 * :52-85, 139-199, with the HypoCS update) in the interaction kernel itself,
 * and only the rest after pass B; the results are bit-identical to the
 * unfused order.  It is slower on MI355X (the interaction's lanes follow the
 * sub-cell index, so the per-particle state it then updates is gathered
 * rather than streamed).  The piecewise calls (lq_particle_interaction,
 * lq_predictor, lq_corrector) never fuse.  No reference counterpart. */
int lq_set_fusion(lq_ctx* ctx, int on);
/* PAIR_LIST capacity (entries per soil slot, 1..160; default 160) and skin
 * (list radius (1 + skin) * 2h, 0..0.1; default 0.01).  The lists are
 * reused in the corrector stage while no particle moved more than
 * 0.4 * skin * 2h; otherwise, or when a slot overflows, the step falls back
 * to the cell scan on the device.  No reference counterpart. */
int lq_set_pair_list(lq_ctx* ctx, int cap, float skin);
/* After a step: whether the last pair-list build overflowed, the largest
 * particle displacement since that build and the displacement the skin
 * allows (max_disp > disp_limit means the corrector stage used the cell
 * scan).  Synchronises. */
int lq_pair_list_status(lq_ctx* ctx, int* overflow, float* max_disp, float* disp_limit);
/* lq_autotune times every interaction variant (reference order, neighbour
 * lists, pair lists, each tile shape over a few depths) on the current state
 * as the two interaction calls of one step, makes the fastest the active one
 * and reports it (shape/tz meaningful for TILE, xcd_remap for the others). */
int lq_autotune(lq_ctx* ctx, int reps, int* variant, int* shape, int* tz, int* xcd_remap, double* ms_per_step);

/* Squared pair cut-off used by every neighbour kernel: the largest float
 * r2_cut with fl(fl(sqrt(r2_cut)) * i_h) < 2, so r^2 <= r2_cut is exactly
 * the reference's q < 2 test (particle_interaction.cu:33-35,
 * adami_boundary.cu:60-62).  Host-only, no device needed. */
int lq_interaction_cutoff_sq(float i_h, float* r2_cut);

/* Neighbour list of every soil slot (debug/parity): for soil slot s the
 * original ids (.dat line) of all particles j with q_ij < 2, in traversal
 * order, written to ids[s*max_nb ...], counts[s] = number found (may exceed
 * max_nb; only the first max_nb are stored).  Synchronises. */
int lq_neighbor_lists(lq_ctx* ctx, int max_nb, int32_t* counts, int32_t* ids, size_t ids_bytes);

/* ---- batched contexts (no reference counterpart; the reference runs one
 * case per process, types.h:100-165) ---------------------------------------
 * lq_create_batch builds one context for n_cases independent cases, each
 * described by its own lq_params (after lq_initialize_parameters).  The
 * cases must agree on every parameter the step reads globally (dr, hdr,
 * gravity, damping, kernel constants, viscosity, cs, cfl, eta, artificial
 * pressure, Shepard period, time and dt); geometry, particle counts,
 * materials (model_par) and motions (mv, motion tables) are per case.  The
 * cases are laid out as cell-aligned blocks of one global grid separated by
 * empty cells, so every kernel of a step runs once over all cases and no
 * neighbour range crosses a case; each case's particles occupy the slots, in
 * the order, that a context of that case alone would give them, so every
 * per-case result equals that context's bit for bit.  lq_step,
 * lq_synchronize, lq_autotune and the instrumentation work on the whole
 * batch; lq_upload / lq_set_motion_tables / lq_wall_force / lq_set_field
 * are for one-case contexts only. */
int lq_create_batch(const lq_params* cases, int n_cases, int device, lq_ctx** out);
int lq_case_count(const lq_ctx* ctx);
/* Upload case k (all cases before the first step). */
int lq_upload_case(lq_ctx* ctx, int k, const lq_host_state* host);
int lq_set_case_motion_tables(lq_ctx* ctx, int k, int n, const lq_motion_table* tables);
/* Case k's rows of a per-slot field (lq_field; not the grid arrays), in its
 * own slot order, codes and cellpos in its own frame: bytes = np_k x row. */
int lq_get_case_field(lq_ctx* ctx, int k, int field, void* dst, size_t bytes);
int lq_case_positions_absolute(lq_ctx* ctx, int k, float* dst, size_t bytes);
int lq_case_counts(lq_ctx* ctx, int k, int32_t* npb, int32_t* npm, int32_t* np_delete_m);
/* Wall force of every case: boxes[3k..3k+2] as lq_wall_force's box in case
 * k's frame, force_kn[k] out. */
int lq_wall_force_cases(lq_ctx* ctx, const float* boxes, double* force_kn);

/* ---- HypoCS surrogate (no reference counterpart; replaces the cuMatModHypo
 * call of predictor_corrector.cu:75 and :172 when enabled) -----------------
 * MLP 14 -> 128 -> 128 -> 128 -> 7 with GELU (tanh form) on the hidden
 * layers.  Inputs x: s = sigma - cohesion (xx yy zz xy xz yz), D*dt (6), the
 * clamped void ratio e, dt.  Outputs y: (rate - Jaumann term)*dt (6) and
 * M_crt = e - e_c.  The Jaumann rotation, the I1 > -4 and tensile branches
 * and the shear cap stay exact.  Parameters (LQ_MLP_NPARAM floats): W1[14][128]
 * b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][7] b4[7];
 * norm (LQ_MLP_NNORM floats): x_mean[14] x_std[14] y_mean[7] y_std[7]. */
#define LQ_MLP_IN 14
#define LQ_MLP_OUT 7
#define LQ_MLP_HIDDEN 128
#define LQ_MLP_NPARAM 35847
#define LQ_MLP_NNORM 42

/* Standalone (no context, current device).  lq_mlp_train runs Adam over
 * shuffled mini-batches of normalised tuples (FP32 forward/backward
 * kernels); lq_mlp_loss_grad returns mean((y_n(x) - y_n)^2) and its
 * gradient; lq_mlp_forward evaluates in FP32 or with the BF16-MFMA kernel
 * the step uses. */
int lq_mlp_normalisation(const float* x, const float* y, size_t n, float* norm);
int lq_mlp_train(const float* x, const float* y, size_t n, const float* norm, int epochs, int batch, float lr,
                 uint64_t seed, int init, float* params, float* loss_per_epoch);
int lq_mlp_loss_grad(const float* params, const float* norm, const float* x, const float* y, size_t n, float* grad,
                     double* loss);
int lq_mlp_forward(const float* params, const float* norm, const float* x, size_t n, float* y, int bf16);
const char* lq_mlp_last_error(void);

/* Context side.  lq_surrogate_set installs 1 or 2 ensemble members, the
 * normalisation, the guard box [lo, hi] of the raw inputs, the ensemble
 * tolerance (largest normalised disagreement accepted) and the audited
 * fraction.  With the surrogate enabled every soil HypoCS call evaluates the
 * members (BF16 MFMA); a slot outside the box or with members disagreeing
 * takes the exact update; audit_frac of the accepted slots (hash of id, step
 * and stage) also run the exact update and record the normalised error.
 * lq_surrogate_stats: counts[0] surrogate, [1] range fallback, [2] ensemble
 * fallback, [3] audited, [4] I1 > -4 branch; audit_max = largest audited
 * error; reset != 0 clears them. */
int lq_surrogate_set(lq_ctx* ctx, int members, const float* params, const float* norm, const float* lo,
                     const float* hi, float ens_tol, float audit_frac);
int lq_surrogate_enable(lq_ctx* ctx, int on);
int lq_surrogate_stats(lq_ctx* ctx, int64_t counts[5], float* audit_max, int reset);
/* Logging of exact (x, y) tuples while the exact update runs: every
 * `every`-th soil slot per call (rotating), up to capacity tuples.
 * lq_surrogate_log_read copies min(logged, max_rows) tuples out. */
int lq_surrogate_log(lq_ctx* ctx, int every, int capacity);
int lq_surrogate_log_read(lq_ctx* ctx, float* x, float* y, int max_rows, int* rows);

#ifdef __cplusplus
}
#endif
#endif /* RETAINWALL_H */
