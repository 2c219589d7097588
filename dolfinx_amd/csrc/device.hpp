// Device-side internals of libdfx: HIP context, device handle layouts and
// the reduction helpers shared by the kernels. gfx950 (MI355X) only.
#pragma once

#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>
#include <memory>
#include <string>

namespace dfx
{

#define DFX_HIP(call)                                                          \
  do                                                                           \
  {                                                                            \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::dfx::Error(DFX_ERR_HIP, std::string(#call) + ": "                \
                                          + hipGetErrorString(e_));            \
  } while (0)

/// Library-wide device state: the library stream, a second stream for
/// work that overlaps it inside one call (both at default priority), and
/// the events that fork and join them.
struct Context
{
  bool ready = false;
  int device = 0;
  int num_cu = 256;
  bool cooperative = false; // cooperative launches (co-resident grids) supported
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t events[64] = {};
};

Context& ctx();
void ensure_init();
hipStream_t stream();

/// Routes the library's launches and copies to stream s while it lives.
struct StreamScope
{
  hipStream_t saved;
  explicit StreamScope(hipStream_t s) : saved(ctx().stream) { ctx().stream = s; }
  ~StreamScope() { ctx().stream = saved; }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;
};
void check_launch(const char* what);
/// Workgroups of `kernel` (threads each, lds bytes of dynamic LDS) resident
/// on the device at once: the grid of a persistent kernel, so that no
/// workgroup waits for another to finish before it starts its share.
int resident_blocks(const void* kernel, int threads, std::size_t lds);

inline std::size_t dtype_size(int dtype) { return dtype == DFX_F64 ? 8 : 4; }

/// Whether the calling thread is recording a graph (dfx_graph_begin).
bool capturing();
/// Raises DFX_ERR_CAPTURE when the calling thread is recording a graph:
/// for steps that need the host part-way (allocation, uploads, waits).
void no_capture(const char* what);

/// Device allocation helpers (hipMalloc on the library's device).
void* dmalloc(std::size_t bytes);
void dfree(void* p);
void h2d(void* dst, const void* src, std::size_t bytes);
void d2h(void* dst, const void* src, std::size_t bytes);

/// Blocks dispatch round-robin over the 8 XCDs of an MI355X
/// (block b runs on XCD b % 8). Map b to a logical index so that each XCD
/// receives one contiguous range of logical blocks, keeping gathers of
/// neighbouring rows inside one XCD's L2.
__device__ inline int xcd_contiguous(int b, int nb)
{
  const int xcd = b & 7, idx = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

/// Items of a persistent workgroup: XCD x (= blockIdx % 8) owns the item
/// range [x n / 8, (x + 1) n / 8) and its workgroups stride through it, so
/// each XCD sweeps one contiguous part (one L2 sees each neighbourhood
/// once). Grids of fewer than 8 workgroups stride over all items.
struct XcdRange
{
  int first, end, step;
};
__device__ inline XcdRange xcd_range(int n)
{
  const int nb = gridDim.x, b = blockIdx.x;
  if (nb < 8)
    return {b, n, nb};
  const int x = b & 7;
  const int lo = static_cast<int>(static_cast<long long>(n) * x / 8);
  const int hi = static_cast<int>(static_cast<long long>(n) * (x + 1) / 8);
  return {lo + (b >> 3), hi, (nb - x + 7) >> 3};
}

/// Sum over a 256-thread block (4 wavefronts of 64); result valid in
/// thread 0. Fixed order, so deterministic.
__device__ inline double block_sum_256(double v, double* lds /*[4]*/)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    lds[wave] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    s = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return s;
}

/// Reads a device flag once per block (thread 0) and broadcasts it, so
/// that every wavefront of the block takes the same branch.
__device__ inline int block_flag(const int* flag)
{
  __shared__ int f;
  if (threadIdx.x == 0)
    f = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int v = f;
  __syncthreads();
  return v;
}

/// Per-block partial sums of NV values: block b stores its sums at
/// partials[k * gridDim.x + b]. No atomics: a later kernel (or the one-block
/// finish_sums_k) adds the partials in block order, so every reduction is
/// deterministic.
template <int NV>
__device__ inline void block_partials(const double (&v)[NV], double* partials)
{
  __shared__ double lds[4];
#pragma unroll
  for (int k = 0; k < NV; ++k)
  {
    const double s = block_sum_256(v[k], lds);
    if (threadIdx.x == 0)
      partials[k * gridDim.x + blockIdx.x] = s;
  }
}

/// Sum of n partials, computed by the whole block in a fixed order; the
/// result is returned to every thread.
__device__ inline double sum_partials(const double* partials, int n)
{
  __shared__ double lds[4];
  __shared__ double total;
  double a = 0;
  for (int i = threadIdx.x; i < n; i += 256)
    a += partials[i];
  const double s = block_sum_256(a, lds);
  if (threadIdx.x == 0)
    total = s;
  __syncthreads();
  const double t = total;
  __syncthreads();
  return t;
}

} // namespace dfx

/// la::MatrixCSR on the device (compact blocks, row-major inside blocks).
struct dfx_csr
{
  std::int32_t nrows = 0, ncols = 0, bs = 1, dtype = DFX_F64;
  std::int64_t nnz = 0;          // block entries
  std::int32_t max_row = 0;      // longest row (blocks)
  std::int32_t* row_ptr = nullptr; // device, nrows+1 (int32: nnz < 2^31)
  std::int32_t* cols = nullptr;    // device, nnz
  std::int32_t* diag = nullptr;    // device, position of (i,i) or -1
  void* values = nullptr;          // device, nnz*bs*bs of dtype
  std::uint64_t pattern_id = 0;    // identifies the pattern for slot caches
  bool zeroed = true;              // values known to be all zero
  // zeroed but not yet written: the zeros are stored just before the
  // values are first read (a store-mode assembly overwrites every entry
  // and needs no zero fill)
  mutable bool zero_pending = false;
  int lanes = 0;                   // SpMV variant: 0 default, 1 stream, G>1 vector
  // Row blocks of the stream SpMV: consecutive rows whose entries fit one
  // LDS chunk (or a single longer row); device, nblk + 1 boundaries as
  // (row, entry) pairs.
  std::int32_t* rblk = nullptr;
  std::int32_t nblk = 0;
  // pattern arrays borrowed from another matrix (not freed by this one)
  const dfx_csr* pattern_owner = nullptr;
  // Sliced-ELL copy used inside the solvers (bs = 1), refreshed from the CSR
  // values at the start of every solve: slices of 64 consecutive rows,
  // a slice's entries stored column-major (entry k of its rows together),
  // padded to the slice's longest row. sell_off: nslices + 1 entry offsets,
  // then nslices + 1 byte offsets into sell_cols; sell_base: per slice the
  // base column of its 16-bit column offsets, or -1 - (first column) for a
  // slice with 32-bit columns. Layout and columns are kept while the
  // pattern (sell_pattern) is unchanged; sell_use says whether the solvers
  // take this copy. Owned by this matrix.
  mutable void* sell_vals = nullptr;
  mutable void* sell_cols = nullptr;
  mutable std::int64_t* sell_off = nullptr;
  mutable std::int32_t* sell_base = nullptr;
  mutable std::int32_t nslices = 0;
  mutable std::uint64_t sell_pattern = 0;
  mutable bool sell_use = false;
};

struct dfx_vec
{
  std::int32_t n = 0, dtype = DFX_F64;
  void* data = nullptr;
};

struct dfx_bc
{
  std::int32_t n_total = 0, ndofs = 0;
  std::int32_t* dofs = nullptr;   // device, unrolled dof indices
  std::int8_t* marker = nullptr;  // device, n_total
  double* g = nullptr;            // device, n_total (FP64), nullptr if g = 0
  // pattern whose diagonal slots were checked to hold every BC dof
  mutable std::uint64_t diag_checked = 0;
};

struct dfx_form
{
  dfx_form_desc d{};          // sizes and kernel id (host pointers cleared)
  std::uint64_t id = 0;       // unique per form created in this process
  // id of a form found to hold the same geometry (x and x_dofmap) on the
  // device, and whether the cell pass of the next assembly was already
  // done by a system cell pass (cell records / element vectors current)
  std::uint64_t same_geometry_as = 0;
  mutable bool cell_data_ready = false;
  mutable bool cell_vec_ready = false;
  // Facet integrals whose kernel has the reference's ABI (run on the
  // host): 1 exterior, given as (cell, local facet) pairs; 2 interior,
  // given as (cell0, local facet0, cell1, local facet1) quads; 0 none. The
  // host copy of the dofmap gives the rows of each facet's tensor.
  int host_facets = 0;
  std::vector<std::int32_t> h_entities;
  std::vector<std::int32_t> h_dofmap;
  double constants[8] = {0};
  bool affine = true;         // every cell an affine image of the reference
  // Q1 hexahedra whose Jacobians are all exactly diagonal (edges along the
  // axes) and whose dofmap is the vertex map: variant 15 then keeps 4
  // values per cell record (the diagonal of G and c |det J|), the BC
  // column masks in cell_mask
  bool compact_records = false;
  std::int32_t max_dof = -1;  // largest node index in the dofmap
  void* x = nullptr;          // device, num_nodes x 3 of dtype
  std::int32_t* x_dofmap = nullptr;
  std::int32_t* dofmap = nullptr; // may alias x_dofmap
  void* coeffs = nullptr;     // device, num_cells x cstride of dtype
  // cell -> CSR slot offsets (position inside the row), built on first
  // assembly into a pattern and reused while the pattern is unchanged
  std::uint64_t slot_pattern = 0;
  void* slots = nullptr;
  int slot_bytes = 0;         // 1 (uint8) or 2 (uint16)
  // Row-owner plan: incidences of each row (cell, local index), in cell
  // order, and for each incidence the offsets inside the row of the
  // cell's k column dofs. Built once per pattern.
  std::int32_t* inc_off = nullptr; // device, (max_dof + 2)
  std::int32_t* inc = nullptr;     // device, num_cells * k: c * k + i
  void* islots = nullptr;          // device, num_cells * k * k
  std::uint64_t plan_pattern = 0;
  int islot_bytes = 0;
  std::int32_t max_inc = 0;        // most incidences of one row
  // Per-cell factors for the affine row kernel (8 values of dtype) and
  // per-cell BC column masks; reference tensors (7 x k x k of dtype).
  void* cell_data = nullptr;
  std::uint32_t* cell_mask = nullptr;
  void* ref_tensors = nullptr;
  void* cell_vec = nullptr;        // per-cell element vectors (linear forms)
  // Incidence-parallel assembly: row of each incidence, row blocks of at
  // most one incidence per thread (gblk, ngblk + 1 row boundaries), and the
  // element matrices of all cells (num_cells x k x k of dtype) when they are
  // formed as one batched product.
  std::int32_t* inc_row = nullptr;
  std::int32_t* gblk = nullptr;
  std::int32_t ngblk = 0;
  std::int32_t gblk_slice = 0;     // longest slice of a row block (entries)
  std::uint64_t gather_pattern = 0;
  // Structured Q1 rows: the cell at each of the 8 corners of a row vertex
  // (-1 absent) and, per row, the CSR position of each of the 27 neighbour
  // offsets (255 absent; 28 bytes per row). struct_ok: -1 not checked for
  // struct_pattern, 0 the mesh is not a structured box, 1 usable.
  std::int32_t* corner_cell = nullptr;
  std::uint8_t* slotmap = nullptr;
  std::uint64_t struct_pattern = 0;
  int struct_ok = -1;
  // Fixed-width rows (hexahedra whose vertices lie in at most 8 cells):
  // per row its incidences c * 8 + i in cell order, padded with -1 to 8
  // (inc8, rows x 8), and their slot offsets (islot8, rows x 8 x 8 bytes),
  // so a row's plan is read without a first load of its offsets.
  std::int32_t* inc8 = nullptr;
  std::uint8_t* islot8 = nullptr;
  std::uint64_t fixed_pattern = 0;
  // Forms whose cell tensors come from a host element kernel with the
  // reference's kernel ABI (kernelptr_t): the kernel and its custom data,
  // host copies of what it reads per cell (geometry, packed coefficients,
  // constants) and the device array of the tensors of its last run.
  void* host_kernel = nullptr;
  void* host_kernel_data = nullptr;
  std::vector<double> h_x;
  std::vector<std::int32_t> h_x_dofmap;
  std::vector<double> h_coeffs;
  std::vector<double> h_constants;
  void* elem_tensors = nullptr;
  // Device packing of coefficients: for each packed position the index of
  // its value in the caller's array (per input cell or input vertex), the
  // mesh and layout it was built for, and a staging buffer for the values.
  std::int32_t* pack_map = nullptr;
  const dfx_mesh* pack_mesh = nullptr;
  int pack_layout = -1;
  double* pack_stage = nullptr;
  std::int64_t pack_stage_n = 0;
  // Row teams of variant 16: per row, in team order, (row, first
  // incidence, incidences, 0); rows ordered by incidence count inside
  // chunks of 4096 consecutive rows; built for trow_pattern.
  std::int32_t* trow = nullptr;
  std::int32_t ntrow = 0;
  std::uint64_t trow_pattern = 0;
  // Elasticity row groups: nteam + 1 (first row position, slice length)
  // pairs, then from team_rows_at the (row, LDS slice offset, first
  // incidence, incidences) of all rows in group order; built for
  // team_pattern with team_cap values of slice per group.
  std::int32_t* team_grp = nullptr;
  std::int32_t nteam = 0;
  std::int32_t team_rows_at = 0;
  std::int64_t team_cap = 0;
  std::uint64_t team_pattern = 0;
  // Facet integrals. Exterior: (cell, local facet) pairs, per row the
  // facet incidences (4 e + facet vertex, entity order) and each facet's
  // vertex contributions (nfacet x 4) or matrix (nfacet x 16) of dtype.
  // Interior: (cell0, lf0, cell1, lf1) quads, incidences 8 e + macro
  // index, macro-element matrices (nfacet x 64).
  std::int32_t nfacet = 0;
  std::int32_t* facets = nullptr;
  std::int32_t* finc_off = nullptr;
  std::int32_t* finc = nullptr;
  void* fvals = nullptr;
  // pattern whose slots were found for every facet entry (later
  // assemblies into it skip the check and its host synchronisation), the
  // slots found (offset in the row of entry b of incidence p at
  // fslot[p W + b], W = 4 exterior / 8 interior), and the rows that have
  // facet incidences
  std::uint64_t facet_pattern = 0;
  std::uint8_t* fslot = nullptr;
  std::int32_t* frows = nullptr;
  std::int32_t nfrows = 0;
  void* cell_elem = nullptr;
  int row_variant = 1;             // row-kernel variant (dfx_form_autotune)
};

namespace dfx
{
/// Store the pending zeros of A's values (no-op unless zero_pending).
/// Every reader of the values calls this first.
void materialize_zero(const dfx_csr& A);
/// MatrixCSR::add of nblocks host blocks (m x n block entries each, values
/// (m bs) x (n bs) row-major in A's dtype), every entry summed in staging
/// order on the device; "Entry not in sparsity" for an absent entry.
void csr_add_blocks(dfx_csr& A, std::int64_t nblocks, int m, int n, const std::int32_t* rows,
                    const std::int32_t* cols, const void* values);
/// b[bs rows[q] + comp] += values[q bs + comp] for nblocks blocks of m rows
/// (host values in b's dtype), summed in staging order on the device.
void vec_add_blocks(dfx_vec& b, std::int64_t nblocks, int m, int bs, const std::int32_t* rows,
                    const void* values);
} // namespace dfx
