/*
 * psierra_amd -- MI355X (gfx950) element-evaluation and global-assembly engine
 * for pSierra's Total Lagrangian Tet4/Tet10 compressible Neo-Hookean path.
 *
 * C ABI: plain pointers and sizes only.  Device pointers are HIP device
 * memory (e.g. torch tensors' data_ptr()); host pointers are ordinary memory.
 * Every device entry point is stream-ordered on the hipStream_t passed as
 * `stream` (NULL = the default stream), performs no allocation, no
 * host<->device copy and no synchronisation, so it can be captured in a
 * hipGraph.
 *
 * Return codes: 0 success; > 0 invalid argument or unsupported input
 * (psx_last_error() has the message); < 0 a negated hipError_t.
 *
 * Element inversion (det F <= 0, pSierra raises
 * "Element inversion detected: det(F) = ... <= 0.") is reported through a
 * device int32 status[2] = {flag, lowest offending element index}; the host
 * wrapper turns it into the same ValueError.
 *
 * Numbering conventions are pSierra's: DOF = 3 * node + component
 * (src/psierra/utility/indexing.py:35-59); the global matrix is the CSR that
 * scipy's coo_matrix(...).tocsr() builds from the element triplets
 * (src/psierra/utility/assembly.py:37-76): rows sorted, columns sorted,
 * duplicates summed, explicit zeros kept.
 */
#ifndef PSIERRA_AMD_H
#define PSIERRA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psx_plan psx_plan;

enum {
  PSX_OK = 0,
  PSX_EINVAL = 1,       /* bad argument (null pointer, size, range) */
  PSX_EUNSUPPORTED = 2, /* input outside what the kernels support */
  PSX_EVOLUME = 3,      /* element with det([1 X]) <= 0 (or det J <= 0) */
  PSX_ENOTREADY = 4     /* e.g. no GPU visible */
};

/* Element arithmetic precision.  PSX_FP64: everything in double.
 * PSX_FP32: coordinate and displacement differences in double, then element
 * arithmetic in float; every accumulation (global scatter, f) in double. */
enum { PSX_FP64 = 0, PSX_FP32 = 1 };

/* Info slots filled by psx_plan_info. */
enum {
  PSX_INFO_ELEMENTS = 0,
  PSX_INFO_NODES = 1,
  PSX_INFO_NPE = 2,
  PSX_INFO_NNZ = 3,
  PSX_INFO_TILES = 4,
  PSX_INFO_TILE = 5,
  PSX_INFO_MAX_DEGREE = 6,
  PSX_INFO_DEVICE_BYTES = 7,
  PSX_INFO_INCIDENCES = 8,
  PSX_INFO_MAX_INCIDENCES = 9, /* most incident elements of any node */
  PSX_INFO_WIDE = 10,          /* 1: a node has more than 64 incident elements or more than 255
                                  neighbours; such a Tet4 plan runs its mask-free row kernel
                                  (runs of up to `tile` incidences), a Tet10 plan its general
                                  path (element rows summed per CSR value in element order);
                                  the mass and F-bar kernels refuse it */
  PSX_INFO_COUNT = 11
};

/* Version of this interface: 7 (6 had no psx_vjp_batch; 5 had no psx_batch_* /
 * psx_assemble_batch;
 * 4: psx_ecsw_training took an element-major
 * work block and gave no sums of squares; 3 had no psx_ecsw_training_noalloc;
 * 2: psx_ecsw_* took rank 64 only; 1 had psx_ecsw_training without
 * work_elements and no psx_plan_bind_material). */
int32_t psx_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* psx_last_error(void);

/* ---- Structural pattern (host only, no GPU needed) -----------------------
 * Replaces compute_element_dofs (src/psierra/utility/indexing.py:35-59),
 * create_coo_sparsity_indices (src/psierra/utility/assembly.py:6-34) and the
 * index half of coo_matrix(...).tocsr() (src/psierra/utility/assembly.py:75-76).
 * conn: n_elements x npe node ids (row-major int64).  nnz < 2^31 required. */
int32_t psx_pattern_nnz(const int64_t* conn, int64_t n_elements, int32_t npe,
                        int64_t n_nodes, int64_t* nnz);
int32_t psx_pattern_build(const int64_t* conn, int64_t n_elements, int32_t npe,
                          int64_t n_nodes, int32_t* indptr /* 3N+1 */,
                          int32_t* indices /* nnz */);

/* ---- Assembly plan --------------------------------------------------------
 * Built once per mesh, like SolidDomain.__init__'s precomputation
 * (src/psierra/domain/solid.py:29-69).  Host inputs; the plan uploads the
 * pattern, a node->element incidence map with slot ranks, a locality order
 * of the nodes and the reference coordinates to the current HIP device.
 * Rejects elements with det([1 X]) <= 0 (Tet4, src/psierra/element/
 * tet4_base.py:56-61) or det J <= 0 at a quadrature point (Tet10,
 * src/psierra/element/tet10_total_lagrangian.py:104-108) with PSX_EVOLUME.
 * tile: incidences per work-group (0 = default; 64, 128 or 256).  A mesh
 * with a node of more than 64 incident elements or more than 255 neighbours
 * gives a "wide" plan (PSX_INFO_WIDE): Tet4 then needs every node's incident
 * elements to fit one tile (at most `tile`) and assembles through a
 * mask-free row kernel; Tet10 assembles through a general path (per
 * element row into scratch, then each CSR value and f entry summed over its
 * contributions in element order, up to 2,386,092 elements); the mass and
 * F-bar kernels refuse a wide plan with PSX_EUNSUPPORTED. */
int32_t psx_plan_create(const int64_t* conn, int64_t n_elements, int32_t npe,
                        const double* X /* N x 3 */, int64_t n_nodes,
                        int32_t tile, psx_plan** out);
int32_t psx_plan_destroy(psx_plan* plan);
int32_t psx_plan_info(const psx_plan* plan, int64_t* info, int32_t n_info);
/* Host copies of the CSR pattern (indptr 3N+1, indices nnz). */
int32_t psx_plan_pattern(const psx_plan* plan, int32_t* indptr, int32_t* indices);
/* Change the work-group tile of an existing plan (re-tiles on the host). */
int32_t psx_plan_set_tile(psx_plan* plan, int32_t tile);

/* ---- Global assembly (device) ----------------------------------------------
 * Replaces SolidDomain.assemble_internal_forces
 * (src/psierra/domain/solid.py:230-274, with assemble_vector,
 * src/psierra/utility/assembly.py:79-108) and
 * SolidDomain.assemble_tangent_stiffness_matrix
 * (src/psierra/domain/solid.py:280-325, with assemble_sparse_matrix,
 * src/psierra/utility/assembly.py:37-76), evaluated in one pass.
 * u: N x 3 double (device), mu/lam: per-element Lame parameters (device,
 * n_elements each; the material is NeoHookean with bulk modulus
 * K = lam + 2 mu / 3, src/psierra/material/parameters.py:135).
 * mu = lam = NULL: the material bound by psx_plan_bind_material (Tet4).
 * f: 3N doubles (device, overwritten); kval: nnz doubles (device,
 * overwritten) in the CSR order of psx_plan_pattern, or NULL for f only.
 * status: device int32[2], reset by the call. */
int32_t psx_assemble(psx_plan* plan, const double* u, const double* mu,
                     const double* lam, double* f, double* kval,
                     int32_t precision, void* stream, int32_t* status);

/* Binds per-element Lame parameters (device, n_elements each) to a Tet4
 * plan, as pSierra's element holds its material (the assemble_* hooks of
 * src/psierra/domain/solid.py:238-241, 288-291 take no parameters): the plan
 * copies mu and lam, and its FP32 rows then read (V mu, V K) from their
 * geometry records instead of gathering two parameters per element (one
 * pass at bind time, 72 B per element of device memory).  Call it again when
 * the parameters change; psx_assemble with mu = lam = NULL uses the last
 * binding.  Stream-ordered on `stream`. */
int32_t psx_plan_bind_material(psx_plan* plan, const double* mu, const double* lam, void* stream);

/* ---- Batched assembly (device) -------------------------------------------
 * Many Tet4 (or many Tet10) plans -- meshes of uneven size, or one plan
 * repeated for parameter samples -- assembled by one launch over the union of their
 * tiles, each instance's f and K bit-identical to its own psx_assemble.  No
 * single reference function: it replaces a Python loop of
 * SolidDomain.assemble_internal_forces / assemble_tangent_stiffness_matrix
 * (src/psierra/domain/solid.py:230-325) over many domains, e.g. the forward
 * evaluations least_squares makes for one 2-point Jacobian
 * (src/psierra/inverse/parameter_identification.py:370-379).
 *
 * psx_batch_create takes n_plans plans of one element type (host array of
 * handles; one plan may appear several times) that share a device and a
 * tile size and are not wide (PSX_EUNSUPPORTED for a wide plan: assemble it
 * with psx_assemble; PSX_EINVAL for mixed types, devices or tiles).  The
 * plans must outlive the batch.  The batch holds its own displacement
 * buffer per instance (32 B per node) and a table of the plans' views,
 * refreshed by the next psx_assemble_batch after psx_plan_set_tile or
 * psx_plan_bind_material on one of its plans (that call synchronises the
 * device once; do not make it under graph capture).
 *
 * psx_assemble_batch: u, mu, lam, f, kval are DEVICE arrays of n_plans
 * device pointers, instance i's arrays as psx_assemble takes them (u 3N_i,
 * mu/lam n_elements_i, f 3N_i, kval nnz_i).  mu = lam = NULL: every plan's
 * bound material (psx_plan_bind_material on each; Tet4 only).  kval = NULL:
 * f only.
 * status: device int32[2 n_plans], instance i's inversion flag and lowest
 * inverted element at 2i, 2i+1, reset by the call.  Two launches on
 * `stream`: the displacement permutation of all instances, then the tile
 * kernel.
 *
 * psx_batch_set_persistent: 0 (default) launches one work-group per tile of
 * the global list (XCD-contiguous); 1 launches a persistent grid of as many
 * work-groups as stay resident, each walking the list (XCD x takes the x-th
 * eighth).  Same results either way; at 128 meshes of 3k-663k elements the
 * first is faster (2.53 against 2.98 ms in FP32: the hardware dispatcher
 * already balances the list, and the loop keeps more registers live).
 *
 * psx_batch_info fills up to n_info of {instances, tiles, nodes, tile,
 * persistent, total nnz}. */
typedef struct psx_batch psx_batch;
int32_t psx_batch_create(psx_plan* const* plans, int32_t n_plans, psx_batch** out);
int32_t psx_batch_destroy(psx_batch* batch);
int32_t psx_batch_set_persistent(psx_batch* batch, int32_t persistent);
int32_t psx_batch_info(const psx_batch* batch, int64_t* info, int32_t n_info);
int32_t psx_assemble_batch(psx_batch* batch, const double* const* u, const double* const* mu,
                           const double* const* lam, double* const* f, double* const* kval,
                           int32_t precision, void* stream, int32_t* status);
/* The adjoint psx_vjp of every instance of a Tet4 batch in one launch: u, mu,
 * lam, w, g_u (and g_mu, g_lam, both or neither) are DEVICE arrays of
 * n_plans device pointers laid out as psx_vjp takes them; each instance's
 * outputs are the bits of its own psx_vjp.  It stands for a loop of the
 * forward evaluations that pSierra's parameter identification runs per
 * parameter sample or load case
 * (src/psierra/inverse/parameter_identification.py:300-379).  The first call
 * allocates the batch's cotangent buffer (32 B per node). */
int32_t psx_vjp_batch(psx_batch* batch, const double* const* u, const double* const* mu,
                      const double* const* lam, const double* const* w, double* const* g_u,
                      double* const* g_mu, double* const* g_lam, int32_t precision, void* stream,
                      int32_t* status);

/* ---- Other hyperelastic models (device) -----------------------------------
 * The psx_assemble path for the compressible models of
 * src/psierra/material/mooney_rivlin.py:69-152 (Psi = c10 (I1bar - 3) +
 * c01 (I2bar - 3) + K/2 (ln J)^2) and src/psierra/material/yeoh.py:69-148
 * (Psi = c10 (I1bar - 3) + c20 (I1bar - 3)^2 + c30 (I1bar - 3)^3 +
 * K/2 (ln J)^2) under Tet4TotalLagrangian and Tet10TotalLagrangian (the
 * same call sites as psx_assemble).  params: host array of device pointers,
 * one value per element each -- PSX_NEO_HOOKEAN: mu, lambda (as
 * psx_assemble); PSX_MOONEY_RIVLIN: c10, c01, K; PSX_YEOH: c10, c20, c30, K. */
enum { PSX_NEO_HOOKEAN = 0, PSX_MOONEY_RIVLIN = 1, PSX_YEOH = 2 };
int32_t psx_assemble_hyperelastic(psx_plan* plan, int32_t model, const double* const* params,
                                  const double* u, double* f, double* kval, int32_t precision,
                                  void* stream, int32_t* status);

/* ---- F-bar Tet4 assembly (device) ------------------------------------------
 * Replaces Tet4Fbar's domain hooks (src/psierra/element/tet4_fbar.py:
 * prepare_for_assembly :226-255, assemble_internal_forces :264-330,
 * assemble_tangent_stiffness_matrix :332-459) for a Tet4 plan and the
 * compressible Neo-Hookean material with per-element (mu, lambda).
 *
 * psx_fbar_create builds the pattern of K_dev + K_vol: node pairs that share
 * an element or a neighbouring node (the nodal patches of :86-109), laid out
 * like psx_plan_pattern (node rows contiguous, columns ascending).  The plan
 * must outlive the F-bar object.  psx_fbar_pattern fills indptr (3N+1) and
 * indices (nnz, psx_fbar_nnz) on the host.
 *
 * psx_fbar_assemble: u (3N), mu, lam (n_elements) on the device; bulk is the
 * single K of Tet4Fbar._get_bulk_modulus (:112-118) used by K_vol.  Writes
 * f (3N) and, when kval != NULL, every value of kval (psx_fbar_nnz) -- the
 * reference's K_dev + K_vol with explicit zeros kept (its CSR sum drops
 * exact-zero sums).  Element inversion as psx_assemble (element index in
 * status[1]).  Working memory is part of the F-bar object: calls on one
 * object must not overlap. */
typedef struct psx_fbar psx_fbar;
int32_t psx_fbar_create(psx_plan* plan, psx_fbar** out);
/* The same pattern from the Tet4 connectivity alone (host only, no GPU):
 * conn n_elements x 4 node ids (row-major int64). */
int32_t psx_fbar_pattern_nnz(const int64_t* conn, int64_t n_elements, int64_t n_nodes, int64_t* nnz);
int32_t psx_fbar_pattern_build(const int64_t* conn, int64_t n_elements, int64_t n_nodes, int32_t* indptr,
                               int32_t* indices);
int32_t psx_fbar_destroy(psx_fbar* fbar);
int32_t psx_fbar_nnz(const psx_fbar* fbar, int64_t* nnz);
int32_t psx_fbar_pattern(const psx_fbar* fbar, int32_t* indptr, int32_t* indices);
int32_t psx_fbar_assemble(psx_fbar* fbar, const double* u, const double* mu, const double* lam, double bulk,
                          double* f, double* kval, int32_t precision, void* stream, int32_t* status);

/* Free-dof CSC over the F-bar pattern: as psx_csc_create (below), for the
 * kval of psx_fbar_assemble (the reference's K_tan[free][:, free].tocsc(),
 * src/psierra/analysis/nonlinear_static.py:354-355). */
struct psx_csc;
int32_t psx_fbar_csc_create(const psx_fbar* fbar, const int64_t* free_dofs, int64_t n_free, struct psx_csc** out);

/* ---- Consistent mass matrix (device, Tet4 and Tet10, FP64) ---------------
 * Replaces SolidDomain.assemble_mass_matrix (src/psierra/domain/solid.py:
 * 208-228) with Tet4Base.compute_mass_matrix
 * (src/psierra/element/tet4_base.py:138-151) and Tet10Base.compute_mass_matrix
 * (src/psierra/element/tet10_base.py:344-372): rho (n_elements, device) ->
 * kval (nnz, device, overwritten) in the CSR order of psx_plan_pattern; the
 * off-diagonal entries of every 3x3 block are explicit zeros.  V_e is the
 * volume of the element's corner tetrahedron (straight-edged Tet10).  The
 * first call on a plan also allocates and fills the plan's mass tensor and
 * element volumes (n_elements doubles), on `stream`. */
int32_t psx_assemble_mass(psx_plan* plan, const double* rho, double* kval, void* stream);

/* ---- Adjoint of the internal force (device, Tet4) ------------------------
 * No counterpart in the reference, which differentiates only by finite
 * differences: MaterialParameterIdentifier.identify calls
 * scipy.optimize.least_squares with its default 2-point Jacobian
 * (src/psierra/inverse/parameter_identification.py:370-379).
 * For a cotangent w (3N, device) computes g_u = K_T w (3N), and, if g_mu and
 * g_lam are non-NULL, d(w^T f_int)/d mu_e and d/d lambda_e (n_elements each).
 * Matrix-free; the same node-row scatter as psx_assemble (deterministic). */
int32_t psx_vjp(psx_plan* plan, const double* u, const double* mu, const double* lam,
                const double* w, double* g_u, double* g_mu, double* g_lam,
                int32_t precision, void* stream, int32_t* status);

/* ---- Free-DOF solver matrix and residual (device) -------------------------
 * Replaces K_total = K_tan + K_contact; K_total[free][:, free].tocsc()
 * (src/psierra/analysis/nonlinear_static.py:354-355) and
 * r_free = f_step_free - f_int[free], ||r|| (nonlinear_static.py:322-323).
 * psx_csc_create (host inputs): free dofs strictly increasing; records, per
 * column node, where its 3x3 blocks sit in the plan's CSR values, and the
 * structural entry count of K[free][:, free] (the object keeps its own copies:
 * the plan may be destroyed first).  psx_csc_extract: from the
 * plan's CSR values kval, writes colptr (n_free + 1), rowind and vals
 * (capacity nnz_struct each) with row indices ascending; with drop_zeros,
 * entries whose value is exactly 0.0 are dropped, as scipy's CSR '+' does.
 * The number of stored entries is colptr[n_free].  psx_free_residual: r_free
 * (n_free) = f_ext_free - f_int[free] (f_ext_free may be NULL = 0) and
 * sumsq[0] = sum r^2, accumulated in FP64 in a fixed order. */
typedef struct psx_csc psx_csc;
int32_t psx_csc_create(const psx_plan* plan, const int64_t* free_dofs, int64_t n_free,
                       psx_csc** out);
int32_t psx_csc_destroy(psx_csc* csc);
int32_t psx_csc_info(const psx_csc* csc, int64_t* n_free, int64_t* nnz_struct);
int32_t psx_csc_extract(psx_csc* csc, const double* kval, int32_t drop_zeros,
                        int32_t* colptr, int32_t* rowind, double* vals, void* stream);
int32_t psx_free_residual(psx_csc* csc, const double* f_int, const double* f_ext_free,
                          double* r_free, double* sumsq, void* stream);

/* ---- Batched ECSW hyper-reduced evaluation (device, Tet4, rank 1..64) ----
 * Replaces HyperReducedECSWModel.evaluate_forces_and_stiffness
 * (src/psierra/hyperreduction/model.py:254-318) for many independent
 * instances in one persistent launch; Phi_e is cached at creation as the
 * model does (model.py:185-213).  Host inputs to psx_ecsw_create, concatenated
 * over instances with rid_offsets (n_instances + 1): X_e (n x 4 x 3), phi_e
 * (n x 12 x rank, rows of the basis for the element's 12 DOFs; one rank for
 * every instance, 1 <= rank <= 64, stored padded with zero columns to the
 * 64-wide matrix-core tiles), weights (> 0), mu, lam (n each).
 * psx_ecsw_evaluate: u_r (n_instances x rank, device) -> f_r
 * (n_instances x rank) and K_r (n_instances x rank x rank, symmetrised
 * 0.5 (K + K^T) as model.py:317), FP64 outputs of FP32 element arithmetic
 * and FP32 matrix-core accumulation.
 * Info slots: n_instances, n_rid_elements, n_work_items, grid, device_bytes. */
typedef struct psx_ecsw psx_ecsw;
int32_t psx_ecsw_create(int32_t n_instances, const int64_t* rid_offsets, const double* X_e,
                        const double* phi_e, const double* weights, const double* mu,
                        const double* lam, int32_t rank, psx_ecsw** out);
int32_t psx_ecsw_destroy(psx_ecsw* ecsw);
int32_t psx_ecsw_info(const psx_ecsw* ecsw, int64_t* info, int32_t n_info);
int32_t psx_ecsw_evaluate(psx_ecsw* ecsw, const double* u_r, double* f_r, double* K_r,
                          void* stream, int32_t* status);

/* ---- ECSW training matrix (device, Tet4, FP64) ----------------------------
 * Replaces ECSWTrainingMatrix.build, single-domain path
 * (src/psierra/hyperreduction/ecsw.py:171-365, loops at :319-365) for a
 * Tet4 plan.  phi: basis (3N x rank, row-major, user DOF order, 1 <= rank
 * <= 64); snapshots: (3N x n_snapshots, row-major); mu, lam: per element.
 * Writes A (rows x n_elements, row-major) with rows = n_snapshots * rank
 * (+ rank (rank + 1) / 2 when include_stiffness):
 *   A[k rank + i][e] = (Phi_e^T f_e(u_k))_i,
 *   A[n_snapshots rank + t][e] = (Phi_e^T K0_e Phi_e)[triu_t] (unscaled;
 *   t in numpy.triu_indices order, K0_e the element tangent at u = 0),
 * written directly in that layout (no element-major copy, no transpose).
 * sumsq (device, 2 doubles, or NULL): the sums of squares of the force rows
 * and of the stiffness rows, in a fixed order, from which the caller forms
 * the reference's balancing scale (ecsw.py:350-357) without another pass
 * over A; it needs work (device) of at least
 * psx_ecsw_training_work_doubles(plan) doubles for the per-group partial
 * sums (work may be NULL when sumsq is).  b = A 1 is left to the caller.
 * status as psx_assemble (element inversion). */
int32_t psx_ecsw_training(psx_plan* plan, const double* phi, int32_t rank, const double* snapshots,
                          int32_t n_snapshots, const double* mu, const double* lam,
                          int32_t include_stiffness, double* A, double* work, int64_t work_doubles,
                          double* sumsq, void* stream, int32_t* status);
/* Doubles of psx_ecsw_training's work array: 2 per group of 16 elements. */
int64_t psx_ecsw_training_work_doubles(const psx_plan* plan);

/* psx_ecsw_training without the work argument, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  The per-group partial sums of squares live in a 32 MiB device
 * array of the library (one per GPU, held until the process ends), so A and
 * sumsq are the same as psx_ecsw_training's; plans of more than 33,554,432
 * elements (2 doubles per group of 16 beyond 4 Mi doubles) are refused with
 * PSX_EUNSUPPORTED.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (different streams or threads) race.  The plan's
 * device must be the current device. */
int32_t psx_ecsw_training_noalloc(psx_plan* plan, const double* phi, int32_t rank, const double* snapshots,
                                  int32_t n_snapshots, const double* mu, const double* lam,
                                  int32_t include_stiffness, double* A, double* sumsq, void* stream,
                                  int32_t* status);

/* ---- Element batch (device) -------------------------------------------------
 * Replaces Tet4TotalLagrangian.compute_nonlinear_force_and_tangent_stiffness
 * (src/psierra/element/tet4_total_lagrangian.py:77-139) and the Tet10 one
 * (src/psierra/element/tet10_total_lagrangian.py:63-136) for n independent
 * elements.  X, u: n x npe x 3; mu, lam: n; k: n x 3npe x 3npe (or NULL);
 * f: n x 3npe.  status as above (volume failures set flag 2). */
int32_t psx_element_batch(int32_t npe, int64_t n, const double* X, const double* u,
                          const double* mu, const double* lam, double* k, double* f,
                          int32_t precision, void* stream, int32_t* status);

/* ---- Material batch (device) ------------------------------------------------
 * Replaces NeoHookean.compute_stress_and_tangent
 * (src/psierra/material/neohookean.py:55-122) for n deformation gradients.
 * F: n x 3 x 3; mu, bulk: n; S: n x 3 x 3; C: n x 6 x 6 (Voigt order
 * xx, yy, zz, xy, yz, zx, src/psierra/utility/tensor.py:5-12). */
int32_t psx_nh_material_batch(int64_t n, const double* F, const double* mu,
                              const double* bulk, double* S, double* C,
                              int32_t precision, void* stream, int32_t* status);

#ifdef __cplusplus
}
#endif

#endif /* PSIERRA_AMD_H */
