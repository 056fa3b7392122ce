/* C-ABI of the MI355X level-set hot path (seeds, FIM redistancing,
 * narrowband construction, upwind advection) on one GPU.
 *
 * Plain pointers and sizes only. Host buffers are caller-owned and copied in
 * and out; the context owns all device memory. Every call returns LS_OK (0)
 * or a negative LS_ERR_* code, with a message in ls_last_error(); nothing
 * exits the process or throws across this boundary. One context per host
 * thread; calls on one context are serialized by the caller.
 */
#ifndef LEVELSET_HIP_H
#define LEVELSET_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ls_ctx ls_ctx;

enum {
  LS_OK = 0,
  LS_ERR_ARG = -1,      /* invalid argument */
  LS_ERR_STATE = -2,    /* call out of order (e.g. step before mesh) */
  LS_ERR_HIP = -3,      /* HIP runtime error */
  LS_ERR_LIMIT = -4,    /* a patch exceeds device limits; use smaller patches */
  LS_ERR_INTERNAL = -5
};

enum { LS_F32 = 0, LS_F64 = 1 };
enum { LS_MODE_ADVECT = 0, LS_MODE_MCF = 1 };
/* Execution of a timestep: one persistent cooperative kernel (default), the
 * same kernel replayed from a captured hipGraph, or one launch per phase with
 * the FIM loop driven from the host. Results are identical. */
enum { LS_EXEC_PERSISTENT = 0, LS_EXEC_KERNELS = 1, LS_EXEC_GRAPH = 2 };
/* FIM sweep organisation (results are identical): one thread per element
 * copy, or the (element, corner) solves left by the exact lower-bound skip
 * spread one per thread (needs more LDS). AUTO (default) is per element,
 * the faster of the two on MI355X in measurements so far. By default a
 * visit stages each element copy's corner ids and heights in LDS when that
 * costs the persistent kernel no workgroups per CU; OR-ing
 * LS_SWEEP_UNSTAGED into the mode makes every sweep read them from memory
 * (same results). Tet corners whose solve cannot lower their value are
 * skipped by an exact lower bound: the height bound, and by default also a
 * supporting-plane bound that removes about a quarter of the remaining
 * solves on the torus; LS_SWEEP_NO_PLANE_BOUND keeps the height bound only (same
 * results, more solves). */
enum { LS_SWEEP_AUTO = 0, LS_SWEEP_ELEMENTS = 1, LS_SWEEP_TASKS = 2, LS_SWEEP_UNSTAGED = 4,
       LS_SWEEP_NO_PLANE_BOUND = 8 };
/* Where the patch layout is built: on the context's GPU (default; radix
 * sorts, identical arrays) or on the host. Host-only contexts always build
 * on the host. */
enum { LS_LAYOUT_DEVICE = 0, LS_LAYOUT_HOST = 1 };
/* Warm start of the FIM (no reference counterpart): none, |phi| after the
 * previous step (identity predictor), or the predictor distances held by the
 * context (ls_mlp_predict, or caller values from ls_set_warm). Non-seed
 * vertices of the previous narrowband start the FIM at
 * prediction * (1 + warm_margin) + warm_abs_margin instead of LARGENUM. The
 * FIM only lowers values and every value it lowers comes from an element
 * solve, so the only unsupported values at its end are warm values still
 * held: a step in which such a vertex inside the band lies more than
 * warm_res_tol below what its elements give it from the final distances
 * (an undercut) re-runs the plain method from the same phi in the same
 * launch, and returns the plain method's bits; in a batch this holds per
 * mesh (a mesh's undercut re-runs that mesh only). */
enum { LS_WARM_NONE = 0, LS_WARM_IDENTITY = 1, LS_WARM_LEARNED = 2 };

/* Per-step parameters. Replaces the argument list of
 * meshFIM3d::GenerateData(char*, int nsteps, float timestep, int inside_niter,
 * int nside, int block_size, float bandwidth, int part_type, int metis_size,
 * bool verbose) (src/core3d/include/meshFIM3d.h:32-36; 2D takes double
 * timestep/bandwidth, src/core2d/include/meshFIM2d.h:36-39). */
typedef struct ls_step_params {
  double dt;           /* timestep (timeStep_) */
  double bandwidth;    /* narrowband width (bandwidth_) */
  double tol;          /* FIM convergence tolerance (SMALLNUM) */
  double lambda_eps;   /* interior tet-solve regulariser (SMALLNUM in localSolverTet1) */
  int32_t niter;       /* sweeps per patch visit (NUM_ITER: 5 in 3D, 10 in 2D) */
  int32_t inside_niter;/* advection passes per step (insideIterations_) */
  int32_t mode;        /* LS_MODE_ADVECT, or LS_MODE_MCF: mean-curvature flow phi_t = kappa |grad phi| */
  int32_t max_fim_iters;/* 0: run the FIM loop to convergence; > 0: stop after that many iterations */
  int32_t warm;         /* LS_WARM_*; needs LS_EXEC_PERSISTENT or LS_EXEC_GRAPH */
  int32_t cfl;          /* 1: advect with min(dt, CFL bound of the step's new narrowband) (ls_compute_dt) */
  double warm_margin;   /* warm value = predicted distance * (1 + warm_margin) + warm_abs_margin */
  double warm_res_tol;  /* undercut above which a warm step re-runs plain (0: tol) */
  double warm_abs_margin;
} ls_step_params;

typedef struct ls_stats {
  int32_t fim_iters;       /* active-list iterations of the FIM loop */
  int32_t nb_count;        /* narrowband patches after the step */
  int64_t patch_visits;    /* sum of active-list sizes */
  int64_t probes;          /* neighbour probes */
  int64_t nb_verts;        /* vertex-updates of the step (owned vertices of narrowband patches) */
  int32_t n_seed_patches;
  int32_t sweep_launches;  /* launches of the FIM patch-sweep kernel (visits + probes) */
  double ms_seed, ms_fim, ms_nb, ms_advect, ms_total; /* HIP-event stage timings */
  double ms_sweep_kernels;  /* summed durations of the FIM patch-sweep launches */
  double ms_advect_kernels; /* summed durations of the advection launches */
  /* Sums over patch visits of element copies, owned and halo vertices, for
   * algorithmic-byte accounting: FIM visits, neighbour probes, advection. */
  int64_t visit_copies, visit_verts, visit_halo;
  int64_t probe_copies, probe_verts, probe_halo;
  int64_t adv_copies, adv_verts, adv_halo;
  int64_t seed_copies, seed_verts, seed_halo;
  double ms_step_kernel;    /* duration of the persistent step kernel (LS_EXEC_PERSISTENT/GRAPH) */
  int32_t exec_mode;        /* execution mode used for the step */
  int32_t sweep_staged;     /* 1: the FIM sweeps staged corner ids and heights in LDS (LS_SWEEP_UNSTAGED) */
  /* Persistent kernel phase times (device wall clock, barrier to barrier):
   * reset+seeds, FIM visits with neighbour activation, (unused: 0),
   * copyback with the probes and the next list, (unused: 0), narrowband,
   * (unused: 0; advection runs as its own launches after the kernel and is
   * part of ms_total). */
  double ms_ph_seed, ms_ph_visit, ms_ph_copyback, ms_ph_probe, ms_ph_next, ms_ph_nb, ms_ph_advect;
  int64_t elem_updates;     /* element updates executed by FIM visits and probes */
  int64_t local_solves;     /* local solves executed (the rest were skipped exactly) */
  double warm_residual;     /* warm start: largest undercut, max(F_v - DT_v) over vertices inside the band
                               still holding their warm value (0: none) */
  int32_t warm_patches;     /* warm start: patches that received warm values */
  int32_t warm_fallback;    /* warm start: meshes whose residual check failed and which re-ran plain
                               (0 or 1; for a batch, each mesh is accepted or rejected on its own) */
  int32_t sweep_tasks;      /* 1: the FIM sweeps ran task-compacted */
  int32_t step_wg_per_cu;   /* persistent step: workgroups per CU of its grid (the kernel build
                               compiled for it); 0 when the step ran one launch per phase */
  double dt_used;           /* the advection's timestep: dt, or min(dt, cfl_dt) with cfl = 1 */
  double cfl_dt;            /* cfl = 1: the new narrowband's CFL bound (0 otherwise) */
  double cfl_max_rin;       /* cfl = 1: the largest inscribed radius over it (0 otherwise) */
  int32_t fast_len_roots;   /* tets: 1 when every element's least height is at least 2^-350 and
                               2^-20 of its size, so the local solves' length roots skip sqrt's
                               small-argument rescaling (same bits); 0 on meshes with flatter elements */
  int32_t plane_bound;      /* tets: 1 when the sweeps skipped solves by the supporting-plane bound
                               (default; LS_SWEEP_NO_PLANE_BOUND, or element sizes outside
                               [2^-48, 2^48] where its FP32 evaluation is not used, give 0) */
} ls_stats;

typedef struct ls_layout_info {
  int64_t n_verts, n_cells, n_copies;
  int32_t n_patches, max_patch_copies, max_patch_verts, max_patch_halo;
  int32_t max_vertex_slots, dim;
  int64_t n_halo, n_slots, n_adjacent;  /* lengths of halo_ids, vslots, padj */
} ls_layout_info;

/* Context lifetime. Replaces LevelSet's device selection and reset
 * (src/LevelSet.cu:146-166); no device reset is performed. */
int ls_create(ls_ctx** out, int device);
void ls_destroy(ls_ctx* ctx);
const char* ls_last_error(const ls_ctx* ctx);

/* Mesh: dim 2 (triangles, FP64 path) or 3 (tetrahedra, FP32 path).
 * xyz is n_verts x 3, cells is n_cells x (dim+1). Replaces
 * LevelSet::initializeMesh (src/LevelSet.cu:111-144). */
int ls_set_mesh(ls_ctx* ctx, int dim, int64_t n_verts, const double* xyz, int64_t n_cells,
                const int32_t* cells);

/* Storage precision of the tet engine: LS_F32 (default; element arithmetic is
 * FP64, distances and geometry are stored in FP32) or LS_F64 (FP64 storage,
 * identical to an FP64 host build). Triangle meshes always use LS_F64.
 * Call before partitioning. */
int ls_set_precision(ls_ctx* ctx, int dtype);

/* Selects the execution mode (LS_EXEC_*); may be called at any time. */
int ls_set_execution(ls_ctx* ctx, int mode);
/* Selects the FIM sweep organisation (LS_SWEEP_*); may be called at any time. */
int ls_set_sweep(ls_ctx* ctx, int mode);
/* Selects where the next partition call builds the layout (LS_LAYOUT_*). */
int ls_set_layout_builder(ls_ctx* ctx, int builder);

/* Partition + patch layout (meshFIM3d::InitPatches/InitPatches2/
 * GenerateBlockNeighbors, src/core3d/cuda/meshFIM3d.cu:423-570).
 * ls_partition_grid restates GraphPartition_Square (:269-352) for an
 * L x W x H structured vertex numbering; ls_partition_rcb is the METIS
 * stand-in (Partition_METIS, :354-421) with exactly n_parts parts;
 * ls_partition_labels takes caller labels in [0, n_parts). */
int ls_partition_grid(ls_ctx* ctx, int L, int W, int H, int bl, int bw, int bh);
int ls_partition_rcb(ls_ctx* ctx, int n_parts);
int ls_partition_labels(ls_ctx* ctx, const int32_t* labels, int n_parts);
int ls_get_layout_info(ls_ctx* ctx, ls_layout_info* info);
/* Any pointer may be NULL. vert_perm: N (new -> original vertex);
 * copy_src: F (original element of each copy); vert_off/ele_off: P+1;
 * copy_verts: F x (dim+1) renumbered vertex ids. */
int ls_get_layout(ls_ctx* ctx, int32_t* vert_perm, int32_t* copy_src, int32_t* vert_off,
                  int32_t* ele_off, int32_t* copy_verts);
/* Per-patch lists of the layout (no reference counterpart; for inspection
 * and tests). Any pointer may be NULL. halo_off/padj_off: P+1; halo_ids:
 * n_halo renumbered ids, sorted per patch; copy_local: F x 4 patch-local
 * corner ids (owned vertices first, then the halo in order); vslot_off: N+1;
 * vslots: n_slots patch-local slots (copy * (dim+1) + corner) per renumbered
 * owned vertex, ordered by (corner, copy); padj: n_adjacent neighbour
 * patches per patch, sorted. */
int ls_get_patch_lists(ls_ctx* ctx, int32_t* halo_off, int32_t* halo_ids, uint16_t* copy_local,
                       int32_t* vslot_off, uint16_t* vslots, int32_t* padj_off, int32_t* padj);

/* Initial level set, one value per vertex in the caller's order
 * (LevelSet::initializeVertices, src/LevelSet.cu:52-80). dtype LS_F32/LS_F64. */
int ls_set_phi(ls_ctx* ctx, const void* phi, int dtype);
/* Velocity modes of ls_set_velocity: each element's own velocity
 * (LS_VEL_ELEMENT); the first element copy's velocity for a whole patch, as
 * the reference's 3D advection reads it (LS_VEL_PATCH_FIRST,
 * src/core3d/include/meshFIM_kernels.h:51-53); or, triangles only, the
 * reference's 2D advection, which ignores the given velocity and moves each
 * element with (phi_A, phi_B) / |grad phi| in its local frame
 * (LS_VEL_PHI_2D, src/core2d/include/meshFIM2d_kernels.h:133-134,178-181;
 * an element with |grad phi| = 0 does not move). The CFL bound uses the
 * given velocity in every mode, as the reference's does. */
enum { LS_VEL_ELEMENT = 0, LS_VEL_PATCH_FIRST = 1, LS_VEL_PHI_2D = 2 };
/* One velocity (x,y,z) per element (LevelSet::initializeAdvection,
 * src/LevelSet.cu:82-108) and the velocity mode (LS_VEL_*). */
int ls_set_velocity(ls_ctx* ctx, const double* vel, int mode);

/* One timestep: seeds, sign save, FIM, narrowband + sign restore, advection
 * (loop body of meshFIM3d::GenerateData, src/core3d/cuda/meshFIM3d.cu:639-678;
 * redistance3d::FindSeedPoint/ReInitTsign/GenerateData,
 * src/core3d/include/redistance3d.h:27-39). stats may be NULL. When the new
 * narrowband is empty, no advection runs and stats->nb_count is 0 (the
 * reference ends its time loop there, meshFIM3d.cu:655-659). */
int ls_step(ls_ctx* ctx, const ls_step_params* params, ls_stats* stats);

/* Current level set in the caller's vertex order (kernel_compute_vertT_before_permute3d,
 * src/core3d/include/meshFIM_kernels.h:661-670). */
int ls_get_phi(ls_ctx* ctx, void* out, int dtype);
/* Narrowband patch ids, sorted ascending; ids may be NULL to query *n. */
int ls_get_narrowband(ls_ctx* ctx, int32_t* ids, int32_t* n);

/* Batches of independent meshes, advanced together by one launch per step.
 * No reference counterpart: LevelSet solves one mesh per object and resets
 * the device for every solve (src/LevelSet.cu:146-166). Each mesh keeps the
 * results it would have alone; per-mesh timestep and narrowband width
 * replace the uniform ls_step_params values.
 *
 * Mesh m owns vertices [vert_off[m], vert_off[m+1]) and cells
 * [cell_off[m], cell_off[m+1]) of the concatenated arrays; cell entries are
 * global vertex ids inside the mesh's own range. Offsets have n_meshes+1
 * entries starting at 0. Partition with ls_partition_batch, or with
 * ls_partition_labels when no patch spans two meshes. */
int ls_set_mesh_batch(ls_ctx* ctx, int dim, int n_meshes, const int64_t* vert_off, const double* xyz,
                      const int64_t* cell_off, const int32_t* cells);
/* Recursive bisection of every mesh into ceil(n_m / patch_verts) patches. */
int ls_partition_batch(ls_ctx* ctx, int patch_verts);
/* Per-mesh timestep and narrowband width (n_meshes values each); both NULL
 * restores the uniform ls_step_params values. */
int ls_set_batch_params(ls_ctx* ctx, const double* dt, const double* bandwidth);
/* Narrowband patch count of every mesh after the last step (n_meshes values;
 * 0 marks a mesh whose narrowband emptied). */
int ls_get_batch_nb_counts(ls_ctx* ctx, int32_t* counts);

/* Unsigned distances after the last step's redistancing, before advection
 * (the DT array of redistance3d::GenerateData, src/core3d/cuda/redistance3d.cu:225-273),
 * in the caller's vertex order; LARGENUM where the FIM did not reach. */
int ls_get_distances(ls_ctx* ctx, double* out);

/* CFL timestep bound over the current narrowband, or over every patch
 * before the first step (meshFIM2d::compute_deltaT, src/core2d/cuda/meshFIM2d.cu:609-674,
 * with kernel_compute_timestep2d, src/core2d/include/meshFIM2d_kernels.h:405-475,
 * whose result the reference discards). Per element copy, with R its inscribed
 * radius (3V / face area sum for tets, 2A / perimeter for triangles), d the
 * dimension, Ceik and Ccurv the element's coefficients and |Cadv| the length of
 * its velocity: dt1 = 4R^2 / |Ccurv| / 2 / d when |Ccurv| > 1e-12,
 * dt2 = 2R / (|Ceik| + |Cadv|) / d when |Ceik| + |Cadv| > 1e-12; *dt is the
 * least min(dt1, dt2) (LARGENUM when no element bounds it) and *max_rin the
 * largest R. Computed in FP64 from the stored local coordinates. */
int ls_compute_dt(ls_ctx* ctx, double* dt, double* max_rin);
/* Eikonal and curvature coefficients of the CFL bound, one per element in
 * the caller's order (the per-element arrays m_ceik_global_d and
 * m_ccurv_global_d, src/core2d/cuda/meshFIM2d.cu:724-731); NULL restores the
 * reference's values for that array, Ceik = 10 and Ccurv = 0. */
int ls_set_cfl_coefficients(ls_ctx* ctx, const double* ceik, const double* ccurv);

/* Predictor distances for the next ls_step with warm = LS_WARM_LEARNED, one
 * per vertex in the caller's order (no reference counterpart): a warm start
 * from any external predictor, checked and finished by the FIM exactly as the
 * network's values are. Values at or above LARGENUM/2 leave a vertex cold. */
int ls_set_warm(ls_ctx* ctx, const double* warm);
/* The predictor distances the context holds (from ls_mlp_predict or
 * ls_set_warm, rounded to the working type), in the caller's order. */
int ls_get_warm(ls_ctx* ctx, double* out);

/* Learned warm start (no reference counterpart; tet and triangle meshes). A
 * 22-64-64-4 ReLU MLP per element copy: inputs are the four corner phi values
 * and the six edge vectors in the copy's local frame, divided by |AB|;
 * outputs y_i = (D_i - |phi_i|) / |AB| with D_i the redistanced distance.
 * A triangle copy fills the same inputs with its three corners and edges
 * AB, AC, BC (the rest zero) and uses three outputs.
 * Inference and training run BF16 MFMA with FP32 accumulation. Parameters
 * are a flat FP32 vector: W1 (22x64, by input), b1 (64), W2 (64x64), b2,
 * W3 (64x4), b3 (4). */
#define LS_MLP_PARAMS 5892
int ls_mlp_set_params(ls_ctx* ctx, const float* params, int n);
int ls_mlp_get_params(ls_ctx* ctx, float* params, int n);
/* Warm values for the next ls_step with warm = LS_WARM_LEARNED: per vertex
 * of the current narrowband, the mean predicted distance over its copies. */
int ls_mlp_predict(ls_ctx* ctx);
/* Training pairs the method generates: phase 0 before ls_step records the
 * inputs of every copy of the current narrowband, phase 1 after it records
 * their targets from the step's distances. The first call sizes the device
 * dataset to `capacity` pairs; pairs beyond it are dropped. */
int ls_mlp_collect(ls_ctx* ctx, int phase, int64_t capacity);
/* Collected pairs: *n always; X (n x 22) and Y (n x 4) when not NULL. */
int ls_mlp_dataset(ls_ctx* ctx, float* X, float* Y, int64_t* n);
/* Sum of squared errors over n caller samples (X n x 22, Y n x 4) and the
 * gradient of their mean (sum over the 4 outputs, mean over samples). */
int ls_mlp_gradient(ls_ctx* ctx, const float* X, const float* Y, int64_t n, float* grad, double* loss);
/* ls_mlp_gradient without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * stream capture, tight loops, hosts that forbid hipMalloc). The samples are
 * staged and the sums accumulated in the library's static device scratch,
 * one copy per GPU held until the process ends (about 2.3 MB), so n may not
 * exceed 16384 (LS_ERR_LIMIT beyond). Calls that overlap on one GPU share
 * that scratch and race. Same results as ls_mlp_gradient up to the order of
 * floating-point sums. */
int ls_mlp_gradient_noalloc(ls_ctx* ctx, const float* X, const float* Y, int64_t n, float* grad, double* loss);
/* Adam on the GPU over the collected pairs (shuffled each epoch with seed);
 * *loss is the last epoch's mean squared error per output. */
int ls_mlp_train(ls_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* LEVELSET_HIP_H */
