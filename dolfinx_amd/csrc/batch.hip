// Batched (segmented) Jacobi-preconditioned CG for many independent
// systems of uneven size: one launch per CG phase serves every system in
// the batch, and a fixed number of iterations is captured once into a
// hipGraph that is replayed while systems join and leave (the graph's
// kernels read the current systems from device tables, so membership
// changes need no re-capture).
//
// Each system's rows are cut into chunks of a fixed number of rows that
// depends only on the system, reductions are summed per system in chunk
// order, and nothing is shared between systems: a system's iterates are
// bitwise the same alone and in any mix. gfx950 (MI355X) only.
#include "device.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace dfx
{
namespace
{

constexpr int kBlock = 256;
constexpr int kLanes = 4;                      // lanes per row (P1 rows ~15)
constexpr int kRowsPerChunk = kBlock / kLanes; // rows per chunk / workgroup

/// Device view of one system (FP64).
struct SysDesc
{
  const std::int32_t* rp;
  const std::int32_t* ci;
  const double* v;
  const double* b;
  double* x;
  double* r;
  double* p;
  double* q;
  double* dinv;
  const std::int32_t* diag; // CSR position of each row's diagonal (-1: none)
  std::int32_t nrows;
  std::int32_t chunk0; // first chunk of the system
  std::int32_t nchunks;
  std::int32_t pad;
};

/// Per-system CG scalars. iter is read by the phase kernels and advanced
/// from iter_next by the first chunk of the next SpMV (never in a kernel
/// that reads it).
struct SysState
{
  double rz[2];
  double rr0, rr, tol2;
  std::int32_t iter, iter_next, done, max_it;
  std::int32_t active, converged, pad0, pad1;
};

struct ChunkDesc
{
  std::int32_t sys, row0, row1, first; // first: chunk is the system's first
};

__device__ inline double sum_range(const double* partials, int c0, int n)
{
  __shared__ double lds[4];
  __shared__ double total;
  double a = 0;
  for (int i = threadIdx.x; i < n; i += kBlock)
    a += partials[c0 + i];
  const double s = block_sum_256(a, lds);
  if (threadIdx.x == 0)
    total = s;
  __syncthreads();
  const double t = total;
  __syncthreads();
  return t;
}

/// Whether a system takes no part in this phase, read once per block so that
/// every wavefront of the block takes the same branch.
__device__ inline bool block_skip(const SysState& S)
{
  __shared__ int skip;
  if (threadIdx.x == 0)
    skip = (__hip_atomic_load(&S.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0
            || __hip_atomic_load(&S.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
               ? 1
               : 0;
  __syncthreads();
  const bool r = skip != 0;
  __syncthreads();
  return r;
}

/// Block sum of one value, stored by thread 0 at partials[slot].
__device__ inline void store_block_sum(double v, double* partials, int slot)
{
  __shared__ double lds[4];
  const double s = block_sum_256(v, lds);
  if (threadIdx.x == 0)
    partials[slot] = s;
}

__global__ void __launch_bounds__(kBlock)
    bcg_spmv_k(const int* nchunk, const ChunkDesc* chunks, const SysDesc* sys, SysState* st,
               double* p_pq)
{
  const int nch = *nchunk;
  for (int t = blockIdx.x; t < nch; t += gridDim.x)
  {
    const ChunkDesc ch = chunks[t];
    SysState& S = st[ch.sys];
    if (ch.first && threadIdx.x == 0)
      S.iter = S.iter_next;
    if (block_skip(S))
      continue;
    const SysDesc& D = sys[ch.sys];
    const std::int32_t row = ch.row0 + threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    double s = 0;
    if (row < ch.row1)
    {
      const std::int32_t e = D.rp[row + 1];
      for (std::int32_t j = D.rp[row] + lane; j < e; j += kLanes)
        s += D.v[j] * D.p[D.ci[j]];
    }
#pragma unroll
    for (int off = kLanes / 2; off > 0; off >>= 1)
      s += __shfl_down(s, off, kLanes);
    double d = 0;
    if (lane == 0 && row < ch.row1)
    {
      D.q[row] = s;
      d = s * D.p[row];
    }
    store_block_sum(d, p_pq, t);
  }
}

__global__ void __launch_bounds__(kBlock)
    bcg_update_k(const int* nchunk, const ChunkDesc* chunks, const SysDesc* sys,
                 const SysState* st, const double* p_pq, double* p_rz, double* p_rr)
{
  const int nch = *nchunk;
  for (int t = blockIdx.x; t < nch; t += gridDim.x)
  {
    const ChunkDesc ch = chunks[t];
    const SysState& S = st[ch.sys];
    if (block_skip(S))
      continue;
    const SysDesc& D = sys[ch.sys];
    const double pq = sum_range(p_pq, D.chunk0, D.nchunks);
    const double alpha = S.rz[S.iter & 1] / pq;
    double srz = 0, srr = 0;
    for (std::int32_t i = ch.row0 + threadIdx.x; i < ch.row1; i += kBlock)
    {
      D.x[i] += alpha * D.p[i];
      const double ri = D.r[i] - alpha * D.q[i];
      D.r[i] = ri;
      srz += ri * D.dinv[i] * ri;
      srr += ri * ri;
    }
    store_block_sum(srz, p_rz, t);
    store_block_sum(srr, p_rr, t);
  }
}

__global__ void __launch_bounds__(kBlock)
    bcg_direction_k(const int* nchunk, const ChunkDesc* chunks, const SysDesc* sys,
                    SysState* st, const double* p_rz, const double* p_rr)
{
  const int nch = *nchunk;
  for (int t = blockIdx.x; t < nch; t += gridDim.x)
  {
    const ChunkDesc ch = chunks[t];
    SysState& S = st[ch.sys];
    if (block_skip(S))
      continue;
    const SysDesc& D = sys[ch.sys];
    const double rz = sum_range(p_rz, D.chunk0, D.nchunks);
    const double rr = sum_range(p_rr, D.chunk0, D.nchunks);
    const int k = S.iter;
    const bool conv = rr / S.rr0 < S.tol2;
    const bool stop = conv || k + 1 >= S.max_it;
    const double beta = rz / S.rz[k & 1];
    __syncthreads(); // every thread has read the scalars before the writes
    if (ch.first && threadIdx.x == 0)
    {
      S.rz[(k + 1) & 1] = rz;
      S.rr = rr;
      S.iter_next = k + 1;
      if (stop)
      {
        S.converged = conv ? 1 : 0;
        S.done = 1;
      }
    }
    if (stop)
      continue;
    for (std::int32_t i = ch.row0 + threadIdx.x; i < ch.row1; i += kBlock)
      D.p[i] = D.dinv[i] * D.r[i] + beta * D.p[i];
  }
}

/// Start of a system's solve from x = 0: r = b, p = z = dinv b, partial
/// sums of (r.z, r.r); dinv from the CSR diagonal.
__global__ void __launch_bounds__(kBlock)
    bcg_start_k(SysDesc D, const std::int32_t* diag, int c0, double* p_rz, double* p_rr)
{
  const std::int32_t row0 = blockIdx.x * kRowsPerChunk;
  const std::int32_t row1 = min(row0 + kRowsPerChunk, D.nrows);
  double srz = 0, srr = 0;
  for (std::int32_t i = row0 + threadIdx.x; i < row1; i += kBlock)
  {
    const std::int32_t pd = diag[i];
    const double a = pd >= 0 ? D.v[pd] : 0.0;
    const double di = a != 0.0 ? 1.0 / a : 1.0;
    D.dinv[i] = di;
    const double bi = D.b[i];
    D.x[i] = 0.0;
    D.r[i] = bi;
    D.p[i] = di * bi;
    srz += bi * di * bi;
    srr += bi * bi;
  }
  store_block_sum(srz, p_rz, c0 + blockIdx.x);
  store_block_sum(srr, p_rr, c0 + blockIdx.x);
}

__global__ void __launch_bounds__(kBlock)
    bcg_start_finish_k(const double* p_rz, const double* p_rr, int c0, int n, SysState* S,
                       double tol2, int max_it)
{
  const double rz = sum_range(p_rz, c0, n);
  const double rr = sum_range(p_rr, c0, n);
  if (threadIdx.x == 0)
  {
    S->rz[0] = rz;
    S->rz[1] = 0;
    S->rr0 = rr;
    S->rr = rr;
    S->tol2 = tol2;
    S->iter = 0;
    S->iter_next = 0;
    S->max_it = max_it;
    S->active = 1;
    S->converged = rr == 0.0 ? 1 : 0;
    S->done = (rr == 0.0 || max_it <= 0) ? 1 : 0;
  }
}

/// Restart of every active system from x = 0, chunk by chunk over the
/// batch's tables (the values of bcg_start_k for each chunk).
__global__ void __launch_bounds__(kBlock)
    bcg_restart_k(const int* nchunk, const ChunkDesc* chunks, const SysDesc* sys,
                  const SysState* st, double* p_rz, double* p_rr)
{
  const int nch = *nchunk;
  for (int t = blockIdx.x; t < nch; t += gridDim.x)
  {
    const ChunkDesc ch = chunks[t];
    if (st[ch.sys].active == 0)
      continue;
    const SysDesc& D = sys[ch.sys];
    double srz = 0, srr = 0;
    for (std::int32_t i = ch.row0 + threadIdx.x; i < ch.row1; i += kBlock)
    {
      const std::int32_t pd = D.diag[i];
      const double a = pd >= 0 ? D.v[pd] : 0.0;
      const double di = a != 0.0 ? 1.0 / a : 1.0;
      D.dinv[i] = di;
      const double bi = D.b[i];
      D.x[i] = 0.0;
      D.r[i] = bi;
      D.p[i] = di * bi;
      srz += bi * di * bi;
      srr += bi * bi;
    }
    store_block_sum(srz, p_rz, t);
    store_block_sum(srr, p_rr, t);
  }
}

/// Per system (one workgroup per slot): the restart's reductions and
/// counters, as bcg_start_finish_k with the system's rtol and max_it kept.
__global__ void __launch_bounds__(kBlock)
    bcg_restart_finish_k(int nsys, const SysDesc* sys, SysState* st, const double* p_rz,
                         const double* p_rr)
{
  for (int s = blockIdx.x; s < nsys; s += gridDim.x)
  {
    SysState& S = st[s];
    if (S.active == 0)
      continue;
    const SysDesc& D = sys[s];
    const double rz = sum_range(p_rz, D.chunk0, D.nchunks);
    const double rr = sum_range(p_rr, D.chunk0, D.nchunks);
    if (threadIdx.x == 0)
    {
      S.rz[0] = rz;
      S.rz[1] = 0;
      S.rr0 = rr;
      S.rr = rr;
      S.iter = 0;
      S.iter_next = 0;
      S.converged = rr == 0.0 ? 1 : 0;
      S.done = (rr == 0.0 || S.max_it <= 0) ? 1 : 0;
    }
    __syncthreads();
  }
}

} // namespace
} // namespace dfx

using namespace dfx;

/// A batch of independent FP64 CG solves.
struct dfx_batch
{
  std::int32_t max_sys = 0, max_chunks = 0;
  std::vector<SysDesc> hsys;
  std::vector<int> slot_used;
  std::vector<std::int32_t> chunks_of; // per slot
  SysDesc* sys = nullptr;
  SysState* st = nullptr;
  ChunkDesc* chunks = nullptr;
  int* nchunk = nullptr;
  double* p_pq = nullptr;
  double* p_rz = nullptr;
  double* p_rr = nullptr;
  // graph of `graph_iters` iterations
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_iters = 0;
  // work vectors (r, p, q, diag^-1) of systems that left, kept by row count
  // for the next system of that size (no device allocation or free, and so
  // no device-wide synchronisation, while serving)
  struct Work
  {
    std::int64_t n;
    double* v[4];
  };
  std::vector<Work> pool;
};

namespace
{

void batch_rebuild_chunks(dfx_batch& B)
{
  no_capture("a change of batch membership");
  std::vector<ChunkDesc> ch;
  for (int s = 0; s < B.max_sys; ++s)
  {
    if (!B.slot_used[s])
      continue;
    SysDesc& d = B.hsys[s];
    d.chunk0 = static_cast<std::int32_t>(ch.size());
    d.nchunks = (d.nrows + kRowsPerChunk - 1) / kRowsPerChunk;
    for (int c = 0; c < d.nchunks; ++c)
      ch.push_back({s, c * kRowsPerChunk, std::min((c + 1) * kRowsPerChunk, d.nrows), c == 0});
  }
  require(static_cast<int>(ch.size()) <= B.max_chunks, "batch is full (rows)");
  const int n = static_cast<int>(ch.size());
  if (n)
    DFX_HIP(hipMemcpyAsync(B.chunks, ch.data(), sizeof(ChunkDesc) * n, hipMemcpyHostToDevice, stream()));
  DFX_HIP(hipMemcpyAsync(B.nchunk, &n, sizeof(int), hipMemcpyHostToDevice, stream()));
  DFX_HIP(hipMemcpyAsync(B.sys, B.hsys.data(), sizeof(SysDesc) * B.max_sys, hipMemcpyHostToDevice,
                         stream()));
  DFX_HIP(hipStreamSynchronize(stream()));
}

/// Workgroups of the phase kernels: a fixed grid (so that a captured graph
/// stays valid) striding over the live chunks.
int phase_grid(const dfx_batch& B)
{
  return std::max(1, std::min(B.max_chunks, ctx().num_cu * 4));
}

void launch_iterations(dfx_batch& B, int iters)
{
  hipStream_t s = stream();
  const int g = phase_grid(B);
  for (int i = 0; i < iters; ++i)
  {
    bcg_spmv_k<<<g, kBlock, 0, s>>>(B.nchunk, B.chunks, B.sys, B.st, B.p_pq);
    bcg_update_k<<<g, kBlock, 0, s>>>(B.nchunk, B.chunks, B.sys, B.st, B.p_pq, B.p_rz, B.p_rr);
    bcg_direction_k<<<g, kBlock, 0, s>>>(B.nchunk, B.chunks, B.sys, B.st, B.p_rz, B.p_rr);
  }
}

} // namespace

extern "C" {

dfx_status dfx_batch_create(int32_t max_systems, int64_t max_rows, dfx_batch** out)
{
  return guard(
      [&]
      {
        require(out && max_systems > 0 && max_rows > 0, "bad argument");
        auto B = std::make_unique<dfx_batch>();
        B->max_sys = max_systems;
        B->max_chunks = static_cast<std::int32_t>((max_rows + kRowsPerChunk - 1) / kRowsPerChunk
                                                  + max_systems);
        B->hsys.assign(max_systems, SysDesc{});
        B->slot_used.assign(max_systems, 0);
        B->sys = static_cast<SysDesc*>(dmalloc(sizeof(SysDesc) * max_systems));
        B->st = static_cast<SysState*>(dmalloc(sizeof(SysState) * max_systems));
        B->chunks = static_cast<ChunkDesc*>(dmalloc(sizeof(ChunkDesc) * B->max_chunks));
        B->nchunk = static_cast<int*>(dmalloc(sizeof(int)));
        B->p_pq = static_cast<double*>(dmalloc(sizeof(double) * B->max_chunks));
        B->p_rz = static_cast<double*>(dmalloc(sizeof(double) * B->max_chunks));
        B->p_rr = static_cast<double*>(dmalloc(sizeof(double) * B->max_chunks));
        DFX_HIP(hipMemsetAsync(B->st, 0, sizeof(SysState) * max_systems, stream()));
        batch_rebuild_chunks(*B);
        *out = B.release();
      });
}

dfx_status dfx_batch_destroy(dfx_batch* B)
{
  if (B)
  {
    if (B->exec)
      (void)hipGraphExecDestroy(B->exec);
    if (B->graph)
      (void)hipGraphDestroy(B->graph);
    for (int s = 0; s < B->max_sys; ++s)
      if (B->slot_used[s])
      {
        dfree(B->hsys[s].r);
        dfree(B->hsys[s].p);
        dfree(B->hsys[s].q);
        dfree(B->hsys[s].dinv);
      }
    for (auto& w : B->pool)
      for (double* v : w.v)
        dfree(v);
    dfree(B->sys);
    dfree(B->st);
    dfree(B->chunks);
    dfree(B->nchunk);
    dfree(B->p_pq);
    dfree(B->p_rz);
    dfree(B->p_rr);
    delete B;
  }
  return DFX_OK;
}

dfx_status dfx_batch_add(dfx_batch* B, const dfx_csr* A, const dfx_vec* b, dfx_vec* x,
                         double rtol, int32_t max_it, int32_t* slot)
{
  return guard(
      [&]
      {
        require(B && A && b && x && slot, "null argument");
        require(A->dtype == DFX_F64 && b->dtype == DFX_F64 && x->dtype == DFX_F64,
                "batched CG runs in FP64");
        require(A->bs == 1 && A->nrows == A->ncols, "scalar square matrix required");
        require(b->n == A->nrows && x->n == A->nrows, "size mismatch");
        int s = -1;
        for (int i = 0; i < B->max_sys; ++i)
          if (!B->slot_used[i])
          {
            s = i;
            break;
          }
        require(s >= 0, "batch is full (systems)");
        materialize_zero(*A);
        SysDesc d{};
        d.rp = A->row_ptr;
        d.ci = A->cols;
        d.v = static_cast<const double*>(A->values);
        d.b = static_cast<const double*>(b->data);
        d.x = static_cast<double*>(x->data);
        d.diag = A->diag;
        d.nrows = A->nrows;
        const std::int64_t n = std::max(A->nrows, 1);
        auto it = std::find_if(B->pool.begin(), B->pool.end(),
                               [&](const dfx_batch::Work& w) { return w.n == n; });
        if (it != B->pool.end())
        {
          d.r = it->v[0];
          d.p = it->v[1];
          d.q = it->v[2];
          d.dinv = it->v[3];
          B->pool.erase(it);
        }
        else
        {
          d.r = static_cast<double*>(dmalloc(sizeof(double) * n));
          d.p = static_cast<double*>(dmalloc(sizeof(double) * n));
          d.q = static_cast<double*>(dmalloc(sizeof(double) * n));
          d.dinv = static_cast<double*>(dmalloc(sizeof(double) * n));
        }
        B->hsys[s] = d;
        B->slot_used[s] = 1;
        try
        {
          batch_rebuild_chunks(*B);
        }
        catch (...)
        {
          B->slot_used[s] = 0;
          dfree(d.r);
          dfree(d.p);
          dfree(d.q);
          dfree(d.dinv);
          throw;
        }
        const SysDesc& dd = B->hsys[s];
        if (dd.nrows > 0)
          bcg_start_k<<<dd.nchunks, kBlock, 0, stream()>>>(dd, A->diag, dd.chunk0, B->p_rz,
                                                          B->p_rr);
        bcg_start_finish_k<<<1, kBlock, 0, stream()>>>(B->p_rz, B->p_rr, dd.chunk0, dd.nchunks,
                                                       B->st + s, rtol * rtol, max_it);
        check_launch("batch_add");
        *slot = s;
      });
}

dfx_status dfx_batch_remove(dfx_batch* B, int32_t slot)
{
  return guard(
      [&]
      {
        require(B && slot >= 0 && slot < B->max_sys && B->slot_used[slot], "bad slot");
        SysDesc& d = B->hsys[slot];
        // queued work of this system precedes any later use of its vectors
        // in stream order, so they go back to the pool at once
        B->pool.push_back({std::max<std::int64_t>(d.nrows, 1), {d.r, d.p, d.q, d.dinv}});
        d = SysDesc{};
        B->slot_used[slot] = 0;
        SysState zero{};
        DFX_HIP(hipMemcpyAsync(B->st + slot, &zero, sizeof(SysState), hipMemcpyHostToDevice, stream()));
        batch_rebuild_chunks(*B);
      });
}

dfx_status dfx_batch_step(dfx_batch* B, int32_t iterations, int32_t use_graph)
{
  return guard(
      [&]
      {
        require(B && iterations > 0, "bad argument");
        // inside an open capture the iterations are recorded into it
        if (!use_graph || capturing())
        {
          launch_iterations(*B, iterations);
          check_launch("batch_step");
          return;
        }
        if (!B->exec || B->graph_iters != iterations)
        {
          if (B->exec)
            (void)hipGraphExecDestroy(B->exec);
          if (B->graph)
            (void)hipGraphDestroy(B->graph);
          B->exec = nullptr;
          B->graph = nullptr;
          DFX_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
          launch_iterations(*B, iterations);
          DFX_HIP(hipStreamEndCapture(stream(), &B->graph));
          DFX_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
          B->graph_iters = iterations;
        }
        DFX_HIP(hipGraphLaunch(B->exec, stream()));
      });
}

dfx_status dfx_batch_restart(dfx_batch* B)
{
  return guard(
      [&]
      {
        require(B, "null batch");
        hipStream_t s = stream();
        bcg_restart_k<<<phase_grid(*B), kBlock, 0, s>>>(B->nchunk, B->chunks, B->sys, B->st,
                                                        B->p_rz, B->p_rr);
        bcg_restart_finish_k<<<std::min(B->max_sys, 1024), kBlock, 0, s>>>(
            B->max_sys, B->sys, B->st, B->p_rz, B->p_rr);
        check_launch("batch_restart");
      });
}

dfx_status dfx_batch_status(dfx_batch* B, int32_t n, const int32_t* slots, int32_t* done,
                            int32_t* converged, int32_t* iterations, double* rel_residual)
{
  return guard(
      [&]
      {
        require(B && (n == 0 || slots), "bad argument");
        std::vector<SysState> h(B->max_sys);
        d2h(h.data(), B->st, sizeof(SysState) * B->max_sys);
        for (int i = 0; i < n; ++i)
        {
          require(slots[i] >= 0 && slots[i] < B->max_sys, "bad slot");
          const SysState& S = h[slots[i]];
          if (done)
            done[i] = S.done;
          if (converged)
            converged[i] = S.converged;
          if (iterations)
            iterations[i] = S.iter_next;
          if (rel_residual)
            rel_residual[i] = S.rr0 > 0 ? std::sqrt(S.rr / S.rr0) : 0.0;
        }
      });
}

} // extern "C"
