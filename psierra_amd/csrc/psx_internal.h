// Internal declarations shared by the host plan builder and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/psierra_amd.h"

namespace psx {

// Thread-local error message behind psx_last_error().
void set_error(const std::string& msg);
void clear_error();
int32_t fail(int32_t code, const std::string& msg);

#define PSX_HIP_TRY(call)                                                          \
  do {                                                                             \
    hipError_t psx_e_ = (call);                                                    \
    if (psx_e_ != hipSuccess)                                                      \
      return ::psx::fail(-static_cast<int32_t>(psx_e_),                            \
                         std::string(#call) + ": " + hipGetErrorString(psx_e_));   \
  } while (0)

// Incidence packing: element id in the low 28 bits, local node in the top 4.
constexpr int kIncElemBits = 28;
constexpr uint32_t kIncElemMask = (1u << kIncElemBits) - 1u;
// A tile of TILE lanes holds at most TILE incidences and at most TILE
// reduction lanes (Tet4; Tet10: kTet10LanePasses * TILE): kTet4DiagLanes
// (Tet10: kTet10DiagLanes) per node for the diagonal block and f, plus one
// per (node, slot) unit.  The kernels with feed bit masks or slot-rank bytes
// (Tet10, mass, F-bar) need every node to have at most kMaxNodeIncidences
// incidences and at most kMaxRankDegree neighbours; a plan outside those
// bounds ("wide") is tiled for the Tet4 kernels only (a node's run at most
// TILE incidences, any degree) and the other kernels refuse it.
constexpr int kTet4DiagLanes = 8;
constexpr int kTet10DiagLanes = 4;
constexpr int kMaxNodeIncidences = 64;
constexpr int kMaxRankDegree = 255;
// Tet4 row kernel (tet4_rows): per incidence the 10-bit staging positions of
// its three off-diagonal blocks (one in area A, two in area B, each area
// unit-major); per unit (node, slot) its first A and first B positions
// (10 bits each) and its tile-local node (bits 20..28).
constexpr int kPosBits = 10;
constexpr uint32_t kPosMask = (1u << kPosBits) - 1u;
constexpr int kTet10LanePasses = 4;

// Host-side structural data (all int32 unless noted), built once per mesh.
struct HostPattern {
  int64_t E = 0, N = 0;
  int npe = 0;
  std::vector<int32_t> adj_ptr;  // N+1: node adjacency (user ids, sorted)
  std::vector<int32_t> adj;      // node neighbours
  std::vector<int32_t> n2e_ptr;  // N+1: incident elements per node
  std::vector<int32_t> n2e;      // e * npe + p, sorted by e
  int64_t nnz() const { return 9 * static_cast<int64_t>(adj_ptr.empty() ? 0 : adj_ptr.back()); }
};

int32_t build_host_pattern(const int64_t* conn, int64_t E, int npe, int64_t N, HostPattern& out);
void fill_csr(const HostPattern& hp, int32_t* indptr, int32_t* indices);
}  // namespace psx

struct psx_csc;

namespace psx {
// Free-dof CSC of the node-block CSR whose node rows are given by an
// adjacency (user ids ascending per node, node a's row contiguous).
int32_t create_csc_adj(const std::vector<int32_t>& adj_ptr, const std::vector<int32_t>& adj,
                       const std::vector<int32_t>& order, int64_t N,
                       const int64_t* free_dofs, int64_t n_free, psx_csc** out);

// Device view of a plan, passed by value to kernels.
struct DevPlan {
  int32_t E, N, npe, n_tiles, tile;
  int32_t max_inc;                // most incidences of any node
  int32_t conn_stride;            // ints per element in conn (4 or 12)
  const int32_t* conn;            // internal node ids, AoS [E][conn_stride]
  const double4* Xi;              // reference coordinates, internal order
  const float4* geo;              // Tet4: per element and local node (h_b, V), FP32, 4 float4
  const float4* geo_bound;        // Tet4 with bound material: (h_b, V mu) for b < 2, (h_b, V K) for b >= 2
  const int32_t* perm;            // internal -> user node id
  const int32_t* inc_ptr;         // N+1, internal order
  const uint32_t* inc_elem;       // packed (element | local << 28)
  const uint32_t* inc_rank;       // slot ranks, rank_words per incidence
  int32_t rank_words;             // 1 (Tet4) or 3 (Tet10)
  const int32_t* kofs;            // N: first CSR value of the node's 3 rows
  const int32_t* deg;             // N: neighbours of the node
  const int32_t* uoff;            // N+1: prefix of deg in internal order
  const int32_t* tile_node;       // n_tiles+1
  const int4* tile_info;          // n_tiles: first node, nodes, first incidence, incidences
  const int4* node_rec;           // N+1: first incidence, first unit, kofs, self slot rank
  const uint32_t* inc_pos;        // Tet4, per incidence: staging positions of its 3 off-diagonal blocks
  const uint32_t* unit_rec;       // Tet4, per unit: first A | first B << 10 | tile-local node << 20
  int32_t wide;                   // 1: a node exceeds the mask/rank-byte bounds (Tet4 row kernels only)
  // General path of a wide Tet10 plan: per element and row node the 90 K
  // values and 3 f values into ke / fe, then every CSR value (and f entry)
  // sums its contributions in increasing element order.
  int64_t gen_nnz;                // CSR values of the plan
  const int32_t* gen_kptr;        // gen_nnz + 1: contributions of each CSR value
  const int32_t* gen_kidx;        // index into ke (e 900 + 30 row + column)
  const int32_t* gen_fptr;        // 3N + 1: contributions of each f entry
  const int32_t* gen_fidx;        // index into fe (e 30 + 3 p + c)
  double* gen_ke;                 // E x 900 element tangents (scratch)
  double* gen_fe;                 // E x 30 element forces (scratch)
  double4* ui;                    // scratch: displacement in internal order
  double4* wi;                    // scratch: adjoint cotangent in internal order
  // Tet4 tiles of a plan that is not wide: each tile's halo (the nodes of its
  // elements outside the tile) so the FP32 row kernels stage the tile's
  // displacements in LDS once per node instead of gathering four per
  // incidence.
  const int2* tile_halo;          // n_tiles: first halo node, halo nodes (-1: too many nodes to stage)
  const int32_t* halo;            // the tiles' halo nodes (internal ids, ascending per tile)
  const uint32_t* inc_loc;        // per incidence: tile-local index (own nodes first, then the halo)
                                  //   of the element's nodes in the row's order, byte q
};

// Own + halo nodes whose displacements a Tet4 row tile stages in LDS.
constexpr int kTileLdsNodes = 128;
// F-bar K_vol by slot: one wave per node stages its terms in LDS (24 B per
// term, 24 B per unit, a byte per term and two per slot), each lane sums at
// most two slots, and the row goes back through the term area.
constexpr int kKvolSlotMaxUnits = 64;
constexpr int kKvolSlotMaxTerms = 512;
constexpr int kKvolSlotMaxSlots = 128;
// doubles of the term area: the terms' g_a[q], then the row's values
__host__ __device__ inline int kvol_slot_area(int max_terms, int max_deg2) {
  return 3 * max_terms > 9 * max_deg2 ? 3 * max_terms : 9 * max_deg2;
}

// Device view of a batch of Tet4 plans assembled by one persistent launch:
// the tiles of all plans form one global list (plan i owns tiles
// tile_start[i] .. tile_start[i+1]), as do their nodes for the displacement
// permutation.  Each plan's view is copied with ui pointing at the batch's
// own buffer, so one plan may appear several times (parameter samples).
struct BatchDev {
  const DevPlan* plans = nullptr;         // n
  const int32_t* tile_start = nullptr;    // n + 1
  const int32_t* node_start = nullptr;    // n + 1
  const double* const* bound_mu = nullptr;   // n: the plans' bound (mu, lambda) (FP64 rows)
  const double* const* bound_lam = nullptr;
  int32_t n = 0, n_tiles = 0, n_nodes = 0, tile = 0, max_inc = 0;
  int32_t npe = 4;                        // nodes per element of every plan (4 or 10)
};

// F-bar store layout in the distance-2 pattern of K_dev + K_vol (device view):
// per internal node {first value, row length in blocks, diagonal slot, 0},
// per (node, slot) unit of the element pattern its slot in the wider row.
struct FbarView {
  const int4* node_rec2 = nullptr;
  const int32_t* slot2 = nullptr;
};

// F-bar working set (device): the wider pattern's layout and the per-call
// element and node quantities of the nodal volume averaging.
struct FbarDev {
  const int4* node_rec2 = nullptr;     // N: {first value, row length, diagonal slot, 0}
  const int32_t* slot2 = nullptr;      // per element-pattern unit: slot in the distance-2 row
  const int4* kv_unit = nullptr;       // per element-pattern unit (p, a): {first unit of a, deg a,
                                       //   unit (a, p), first entry of the slot map}
  const uint16_t* kv_slot = nullptr;   // per (p, a, q in N1(a)): slot of q in p's distance-2 row
  int32_t max_deg2 = 0;                // longest distance-2 row (blocks)
  // K_vol by slot (when every node has at most kKvolSlotMaxUnits units and
  // kKvolSlotMaxTerms terms; else kv_unit / kv_slot): node p's terms (a, q)
  // sorted by q's slot in p's row, a ascending within a slot
  int32_t kv_slots = 0;                // 1: the slot layout below is built
  int32_t kv_max_terms = 0;            // most terms of any node
  int32_t kv_max_units = 0;            // most units of any node
  const int4* kv_node = nullptr;       // N: {first term, first slot offset, terms, units}
  const uint16_t* kv_term = nullptr;   // per term: index of (p, a) among p's units | rank of q in N1(a) << 8
  const uint16_t* kv_soff = nullptr;   // per node deg2 + 1: first term of each slot (node-relative)
  double* qv = nullptr;                // per element: V / 4
  double* vjm1 = nullptr;              // per element: V (J - 1) / 4
  double* wq = nullptr;                // per element: 12 values V J / 4 F h_b
  double* jm1 = nullptr;               // per internal node: J_node - 1
  double4* g = nullptr;                // per element-pattern unit (a, b): {g_a[b], V0_a}
};

// Structural CSC of K[free][:, free] (device view).
// A node's free mask has bit r set when dof 3n + r is free; its free base is
// the number of free dofs below 3n, so dof 3n + r (free) has free index
// base + popcount(mask & ((1 << r) - 1)).
struct DevCsc {
  int32_t n_free;
  int32_t n_nodes;          // column nodes with a free dof, in locality order (waves close in time share cache lines)
  const int4* node;         // 2 n_nodes: {first unit, units, free base, free mask}, {structural start of column 0, 1, 2, 0}
  const int4* unit;         // per unit (b, a), a ascending: {K position of K[3a][3b], row stride 3 deg(a), a's free base, a's free mask}
  int32_t nnz_struct;       // structural entries of K[free][:, free]
  const int32_t* free;      // free dof ids (ascending)
  int32_t* scratch;         // per-block counts / offsets
  double* partial;          // per-block residual partial sums
};

}  // namespace psx

struct psx_csc {
  int device = 0;
  int64_t n_free = 0, nnz_struct = 0;
  psx::DevCsc d{};
  std::vector<void*> allocations;
};

struct psx_fbar {
  psx_plan* plan = nullptr;            // borrowed: must outlive this object
  std::vector<int32_t> adj2_ptr, adj2; // distance-2 adjacency, user order
  int64_t nnz2 = 0;
  int64_t device_bytes = 0;
  psx::FbarDev d{};
  std::vector<void*> allocations;
};

struct psx_plan {
  int device = 0;
  psx::HostPattern hp;
  std::vector<int32_t> perm_h;     // internal -> user
  std::vector<int32_t> inc_ptr_h;  // internal order
  std::vector<int32_t> deg_h;      // internal order
  std::vector<uint16_t> inc_slot_h;  // Tet4: per incidence, the slot of each other node (local order) in the row
  int32_t max_deg = 0;
  int32_t max_run = 0;             // most incidences of any node
  double* mass_coef = nullptr;     // device npe x npe reference mass tensor (lazy)
  double* elem_vol = nullptr;      // device per-element corner volume (lazy, first mass assembly)
  double* mu_bound = nullptr;      // device copies of the bound material (psx_plan_bind_material)
  double* lam_bound = nullptr;
  bool bound = false;
  int64_t version = 0;             // bumped by psx_plan_set_tile and psx_plan_bind_material
  int64_t n_inc = 0;
  int64_t device_bytes = 0;
  psx::DevPlan d{};
  std::vector<void*> allocations;
};

namespace psx {
// Kernel launchers (defined in the .hip files).
int32_t launch_tet4_hyper(const DevPlan& p, int model, const double* const* params, const double* u,
                          double* f, double* kval, int precision, hipStream_t s, int32_t* status);
int32_t launch_tet4_assemble(const DevPlan& p, const double* u, const double* mu, const double* lam,
                             double* f, double* kval, int precision, hipStream_t s, int32_t* status);
int32_t launch_t4_fbar_rows(const DevPlan& p, const double* mu, const double* lam, const double* jm1_node,
                            const FbarView& fv, double* f, double* kval, int precision, hipStream_t s,
                            int32_t* status);
int32_t launch_permute_u(const DevPlan& p, const double* u, hipStream_t s, int32_t* status);
int32_t launch_batch_permute(const BatchDev& b, const double* const* u, hipStream_t s, int32_t* status);
int32_t launch_tet4_batch_vjp(const BatchDev& b, const double* const* u, const double* const* mu,
                              const double* const* lam, const double* const* w, double* const* gu,
                              double* const* gmu, double* const* glam, int precision, hipStream_t s,
                              int32_t* status);
int32_t launch_tet10_batch(const BatchDev& b, const double* const* u, const double* const* mu,
                           const double* const* lam, double* const* f, double* const* kval, int precision,
                           int cus, bool persistent, hipStream_t s, int32_t* status);
int32_t launch_tet4_batch(const BatchDev& b, const double* const* u, const double* const* mu,
                          const double* const* lam, double* const* f, double* const* kval, int precision,
                          int cus, bool persistent, hipStream_t s, int32_t* status);
int32_t launch_tet4_assemble_bound(const DevPlan& p, const double* u, const double* mu, const double* lam,
                                   double* f, double* kval, int precision, hipStream_t s, int32_t* status);
int32_t launch_bind_material(const DevPlan& p, const double* mu, const double* lam, float4* out, hipStream_t s);
int32_t launch_tet4_fbar(const DevPlan& p, const FbarDev& fb, const double* u, const double* mu,
                         const double* lam, double bulk, double* f, double* kval, int precision, hipStream_t s,
                         int32_t* status);
int32_t launch_tet10_hyper(const DevPlan& p, int model, const double* const* params, const double* u, double* f,
                           double* kval, int precision, hipStream_t s, int32_t* status);
int32_t launch_tet10_assemble(const DevPlan& p, const double* u, const double* mu, const double* lam,
                              double* f, double* kval, int precision, hipStream_t s, int32_t* status);
int32_t launch_tet4_vjp(const DevPlan& p, const double* u, const double* mu, const double* lam,
                        const double* w, double* gu, double* gmu, double* glam, int precision,
                        hipStream_t s, int32_t* status);
int32_t launch_ecsw_training(const DevPlan& p, const double* phi, int r, const double* snaps, int m,
                             const double* mu, const double* lam, int stiff, double* A, double* partial,
                             double* sumsq, hipStream_t s, int32_t* status);
int32_t launch_element_volumes(const DevPlan& p, double* vol, hipStream_t s);
int32_t launch_mass(const DevPlan& p, const double* rho, const double* vol, const double* coef, double* kval,
                    hipStream_t s, int64_t n_units);
int32_t launch_csc_extract(const DevCsc& c, const double* kval, int drop, int* colptr, int* rowind,
                           double* vals, hipStream_t s);
int32_t launch_free_residual(const DevCsc& c, const double* f_int, const double* f_ext_free, double* r,
                             double* sumsq, hipStream_t s);
int32_t launch_element_batch(int npe, int64_t n, const double* X, const double* u, const double* mu,
                             const double* lam, double* k, double* f, int precision, hipStream_t s,
                             int32_t* status);
int32_t launch_material_batch(int64_t n, const double* F, const double* mu, const double* bulk,
                              double* S, double* C, int precision, hipStream_t s, int32_t* status);
}  // namespace psx
