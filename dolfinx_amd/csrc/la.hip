// Device linear algebra for the assemble-and-solve path: la::MatrixCSR and
// la::Vector on HBM, CSR SpMV, FP64-accumulated inner products, and the
// device-resident (Jacobi-preconditioned) CG and mixed-precision
// refinement solvers. gfx950 (MI355X) only.
#include "device.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace dfx
{

namespace
{
thread_local std::string g_last_error;
Context g_ctx;
std::atomic<std::uint64_t> g_pattern_counter{1};
} // namespace

namespace
{
// graph capture on this thread (dfx_graph_begin .. dfx_graph_end) and the
// first error raised inside it
thread_local bool g_capturing = false;
thread_local std::string g_capture_error;
} // namespace

void set_last_error(const std::string& msg)
{
  g_last_error = msg;
  if (g_capturing && g_capture_error.empty())
    g_capture_error = msg;
}

bool capturing() { return g_capturing; }

void no_capture(const char* what)
{
  if (g_capturing)
    throw Error(DFX_ERR_CAPTURE, std::string(what)
                                     + " cannot be recorded into a graph capture (run the sequence "
                                       "once before capturing it, so that its set-up is done)");
}

void init_noalloc_scratch(int device);

Context& ctx() { return g_ctx; }

void ensure_init()
{
  if (g_ctx.ready)
    return;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    throw Error(DFX_ERR_NODEV, "no HIP device available");
  DFX_HIP(hipSetDevice(0));
  hipDeviceProp_t prop;
  DFX_HIP(hipGetDeviceProperties(&prop, 0));
  g_ctx.num_cu = prop.multiProcessorCount;
  g_ctx.cooperative = prop.cooperativeLaunch != 0;
  DFX_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  DFX_HIP(hipStreamCreateWithFlags(&g_ctx.aux, hipStreamNonBlocking));
  DFX_HIP(hipEventCreateWithFlags(&g_ctx.fork, hipEventDisableTiming));
  DFX_HIP(hipEventCreateWithFlags(&g_ctx.join, hipEventDisableTiming));
  for (auto& e : g_ctx.events)
    DFX_HIP(hipEventCreate(&e));
  init_noalloc_scratch(0);
  g_ctx.ready = true;
}

hipStream_t stream()
{
  ensure_init();
  return g_ctx.stream;
}

int resident_blocks(const void* kernel, int threads, std::size_t lds)
{
  ensure_init();
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, std::size_t>, int> known;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_tuple(kernel, threads, lds);
  auto it = known.find(key);
  if (it != known.end())
    return it->second;
  int per_cu = 0;
  DFX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
  const int n = std::max(per_cu, 1) * g_ctx.num_cu;
  known.emplace(key, n);
  return n;
}

void check_launch(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(DFX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void* dmalloc(std::size_t bytes)
{
  ensure_init();
  no_capture("device allocation");
  void* p = nullptr;
  if (bytes == 0)
    bytes = 256;
  DFX_HIP(hipMalloc(&p, bytes));
  return p;
}

void dfree(void* p)
{
  if (p)
    (void)hipFree(p);
}

void h2d(void* dst, const void* src, std::size_t bytes)
{
  no_capture("host-to-device copy");
  if (bytes)
    DFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  DFX_HIP(hipStreamSynchronize(stream()));
}

void d2h(void* dst, const void* src, std::size_t bytes)
{
  no_capture("device-to-host copy");
  if (bytes)
    DFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  DFX_HIP(hipStreamSynchronize(stream()));
}

std::uint64_t next_pattern_id() { return g_pattern_counter++; }

// ---------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------

constexpr int kBlock = 256;

/// Blocks for a grid-stride elementwise/reduction kernel over n items.
int reduce_blocks(std::int64_t n)
{
  std::int64_t b = (n + kBlock * 4 - 1) / (kBlock * 4);
  return static_cast<int>(std::clamp<std::int64_t>(b, 1, 2048));
}

template <typename T>
__global__ void __launch_bounds__(kBlock) fill_k(T* v, std::int64_t n, T a)
{
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    v[i] = a;
}

template <typename T, typename S>
__global__ void __launch_bounds__(kBlock) convert_k(T* dst, const S* src, std::int64_t n)
{
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    dst[i] = static_cast<T>(src[i]);
}

/// r = alpha x + y
template <typename T>
__global__ void __launch_bounds__(kBlock)
    axpy_k(T* r, double alpha, const T* x, const T* y, std::int64_t n)
{
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    r[i] = static_cast<T>(alpha) * x[i] + y[i];
}

/// Partial sums of a_i b_i (FP64) per block.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    dot_k(const T* a, const T* b, std::int64_t n, double* partials)
{
  double s[1] = {0.0};
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    s[0] += double(a[i]) * double(b[i]);
  block_partials<1>(s, partials);
}

/// One block: out[k] = sum of the nb partials of value k.
__global__ void __launch_bounds__(kBlock)
    finish_sums_k(const double* partials, int nb, int nv, double* out)
{
  for (int k = 0; k < nv; ++k)
  {
    const double t = sum_partials(partials + std::int64_t(k) * nb, nb);
    if (threadIdx.x == 0)
      out[k] = t;
  }
}

/// Position of the diagonal entry of each block row.
__global__ void __launch_bounds__(kBlock)
    diag_pos_k(std::int32_t nrows, const std::int32_t* rp, const std::int32_t* ci,
               std::int32_t* diag)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrows)
    return;
  std::int32_t lo = rp[r], hi = rp[r + 1];
  while (lo < hi)
  {
    const std::int32_t mid = (lo + hi) >> 1;
    if (ci[mid] < r)
      lo = mid + 1;
    else
      hi = mid;
  }
  diag[r] = (lo < rp[r + 1] && ci[lo] == r) ? lo : -1;
}

/// Workgroup barrier ordering LDS accesses only: waits for this wave's LDS
/// operations, not for its outstanding global loads (which a full
/// __syncthreads fence would drain).
__device__ inline void lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
}

/// Entries per LDS chunk of the stream SpMV (16 KB of products).
template <typename T>
constexpr int stream_chunk()
{
  return 16384 / static_cast<int>(sizeof(T));
}

/// Stream CSR SpMV, y (+)= A x. Each row block (consecutive rows whose
/// entries fit one chunk) is processed by one workgroup: the products
/// v_j x[c_j] of the block's contiguous entry range are formed with fully
/// coalesced loads into LDS, then one thread per row sums its products left
/// to right, the order of the reference's row-serial loop. Row blocks are
/// given as (first row, first entry) pairs (rbj, nblk + 1 of them), so a
/// block's bounds are one load. While the rows of the current block are
/// summed, the next block's entries and its rows' positions (and y, dotv)
/// are in flight, and the bounds of the block after it: each block waits
/// for one memory round trip (its x gather) instead of a chain of them. A
/// block holding a single longer row is reduced with a strided sum. With
/// dotv, partial sums of y.dotv go to partials.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    spmv_stream_k(std::int32_t nblk, const int2* __restrict__ rbj,
                  const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                  const T* __restrict__ v, const T* __restrict__ x, T* __restrict__ y,
                  int accumulate, const T* __restrict__ dotv, double* partials,
                  const int* stop)
{
  if (stop && block_flag(stop))
    return;
  constexpr int NC = stream_chunk<T>();
  constexpr int U = NC / kBlock;
  __shared__ T prod[NC];
  __shared__ double red[4];
  double d[1] = {0.0};
  struct Bounds
  {
    std::int32_t r0, r1, j0, j1;
  };
  // this thread's row of a block: its products [kb, ke) in the chunk, and
  // y and dotv at the row
  struct Row
  {
    std::int32_t kb, ke;
    T yold, dv;
  };
  auto bounds = [&](int b) -> Bounds
  {
    const int2 lo = rbj[b], hi = rbj[b + 1];
    return {lo.x, hi.x, lo.y, hi.y};
  };
  auto row_of = [&](const Bounds& B) -> Row
  {
    Row w{0, 0, T(0), T(0)};
    const std::int32_t r = B.r0 + static_cast<std::int32_t>(threadIdx.x);
    if (r < B.r1)
    {
      w.kb = rp[r] - B.j0;
      w.ke = rp[r + 1] - B.j0;
      if (accumulate)
        w.yold = y[r];
      if (dotv)
        w.dv = dotv[r];
    }
    return w;
  };
  T vv[U];
  std::int32_t cc[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
  {
    vv[u] = T(0);
    cc[u] = 0;
  }
  // unconditional loads at clamped positions (no branch around a load, so
  // all of them are in flight together); entries past j1 are not used
  auto stage = [&](const Bounds& B)
  {
    const std::int32_t last = max(B.j1 - 1, B.j0);
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const std::int32_t j = min(B.j0 + static_cast<std::int32_t>(threadIdx.x) + u * kBlock, last);
      vv[u] = v[j];
      cc[u] = ci[j];
    }
  };
  const XcdRange xr = xcd_range(nblk);
  int bi = xr.first;
  if (bi < xr.end)
  {
    Bounds cur = bounds(bi);
    const bool cur_short = cur.j1 - cur.j0 <= NC;
    if (cur_short && cur.j1 > cur.j0)
      stage(cur);
    Row row = row_of(cur);
    int nbi = bi + xr.step;
    Bounds nxt = nbi < xr.end ? bounds(nbi) : Bounds{0, 0, 0, 0};
    for (;;)
    {
      const bool more = nbi < xr.end;
      const int nnbi = nbi + xr.step;
      Row nrow{0, 0, T(0), T(0)};
      Bounds nnxt{0, 0, 0, 0};
      if (cur.j1 - cur.j0 <= NC)
      {
        if (cur.j1 > cur.j0)
        {
          T xv[U];
#pragma unroll
          for (int u = 0; u < U; ++u)
            xv[u] = x[cc[u]];
#pragma unroll
          for (int u = 0; u < U; ++u)
          {
            const std::int32_t j = cur.j0 + threadIdx.x + u * kBlock;
            if (j < cur.j1)
              prod[j - cur.j0] = vv[u] * xv[u];
          }
        }
        lds_barrier();
        // the next block's loads go out before the sums (a row block has
        // at most kBlock rows: one row per thread)
        if (more)
        {
          if (nxt.j1 - nxt.j0 <= NC && nxt.j1 > nxt.j0)
            stage(nxt);
          nrow = row_of(nxt);
          if (nnbi < xr.end)
            nnxt = bounds(nnbi);
        }
        const std::int32_t r = cur.r0 + threadIdx.x;
        if (r < cur.r1)
        {
          T sum = 0;
          for (std::int32_t k = row.kb; k < row.ke; ++k)
            sum += prod[k];
          const T yr = accumulate ? row.yold + sum : sum;
          y[r] = yr;
          if (dotv)
            d[0] += double(yr) * double(row.dv);
        }
        lds_barrier();
      }
      else
      {
        double sum = 0;
        for (std::int32_t j = cur.j0 + threadIdx.x; j < cur.j1; j += kBlock)
          sum += double(v[j] * x[ci[j]]);
        const double t = block_sum_256(sum, red);
        if (threadIdx.x == 0)
        {
          const T yr = accumulate ? row.yold + static_cast<T>(t) : static_cast<T>(t);
          y[cur.r0] = yr;
          if (dotv)
            d[0] += double(yr) * double(row.dv);
        }
        if (more)
        {
          if (nxt.j1 - nxt.j0 <= NC && nxt.j1 > nxt.j0)
            stage(nxt);
          nrow = row_of(nxt);
          if (nnbi < xr.end)
            nnxt = bounds(nnbi);
        }
      }
      if (!more)
        break;
      cur = nxt;
      row = nrow;
      nxt = nnxt;
      nbi = nnbi;
    }
  }
  if (partials)
    block_partials<1>(d, partials);
}

/// Rows per slice of the sliced-ELL solver copy (one wavefront).
constexpr int kSlice = 64;

/// Slice shapes of the sliced-ELL copy: for slice s, out[3 s] = its
/// longest row and out[3 s + 1], out[3 s + 2] = its smallest and largest
/// column (one wavefront per slice, lane = row; columns are sorted in a row).
__global__ void __launch_bounds__(kBlock)
    sell_shape_k(std::int32_t nrows, std::int32_t nslices, const std::int32_t* __restrict__ rp,
                 const std::int32_t* __restrict__ ci, std::int32_t* __restrict__ out)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  const std::int32_t sl = static_cast<std::int32_t>(t / kSlice);
  if (sl >= nslices)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t);
  std::int32_t len = 0, lo = 0x7fffffff, hi = -1;
  if (r < nrows)
  {
    const std::int32_t j0 = rp[r], j1 = rp[r + 1];
    len = j1 - j0;
    if (len > 0)
    {
      lo = ci[j0];
      hi = ci[j1 - 1];
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
  {
    len = max(len, __shfl_xor(len, o, 64));
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  if (t % kSlice == 0)
  {
    out[3 * std::int64_t(sl)] = len;
    out[3 * std::int64_t(sl) + 1] = lo;
    out[3 * std::int64_t(sl) + 2] = hi;
  }
}

/// Sliced-ELL copy of a CSR matrix: thread (slice, lane) writes entry k of
/// row slice*64 + lane at voff[slice] + 64 k + lane, padded with value 0.
/// Columns (written when cols is given) are 16-bit offsets from the slice's
/// base column where the slice's columns span less than 2^16 (base >= 0),
/// else 32-bit indices (base = -1); padding columns are the slice's first
/// column. off holds voff (nslices + 1 entries), then the column byte
/// offsets (nslices + 1).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    sell_build_k(std::int32_t nrows, std::int32_t nslices, const std::int64_t* __restrict__ off,
                 const std::int32_t* __restrict__ cbase, const std::int32_t* __restrict__ rp,
                 const std::int32_t* __restrict__ ci, const T* __restrict__ v,
                 unsigned char* __restrict__ cols, T* __restrict__ sval)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  const std::int32_t sl = static_cast<std::int32_t>(t / kSlice);
  if (sl >= nslices)
    return;
  const int lane = static_cast<int>(t % kSlice);
  const std::int32_t r = sl * kSlice + lane;
  const std::int64_t base = off[sl];
  const std::int32_t w = static_cast<std::int32_t>((off[sl + 1] - base) / kSlice);
  const std::int32_t j0 = r < nrows ? rp[r] : 0;
  const std::int32_t len = r < nrows ? rp[r + 1] - j0 : 0;
  if (sval)
    for (std::int32_t k = 0; k < w; ++k)
      sval[base + std::int64_t(k) * kSlice + lane] = k < len ? v[j0 + k] : T(0);
  if (!cols)
    return;
  const std::int32_t b = cbase[sl];
  unsigned char* c = cols + off[nslices + 1 + sl];
  if (b >= 0)
  {
    std::uint16_t* c16 = reinterpret_cast<std::uint16_t*>(c);
    for (std::int32_t k = 0; k < w; ++k)
      c16[std::int64_t(k) * kSlice + lane] = static_cast<std::uint16_t>(k < len ? ci[j0 + k] - b : 0);
  }
  else
  {
    std::int32_t* c32 = reinterpret_cast<std::int32_t*>(c);
    const std::int32_t pad = -b - 1; // the slice's first column, encoded
    for (std::int32_t k = 0; k < w; ++k)
      c32[std::int64_t(k) * kSlice + lane] = k < len ? ci[j0 + k] : pad;
  }
}

/// Sliced-ELL SpMV for the solvers, y (+)= A x: one wavefront per slice of
/// 64 rows, lane = row, each load instruction reading entry k of 64
/// consecutive rows (coalesced, no LDS), 16-bit column offsets where the
/// slice allows. Entries are taken sixteen at a time (values and columns,
/// then the x gathers) and each row is summed left to right with separate
/// multiply and add, the order and rounding of the stream kernel and of the
/// reference's row loop; padding is skipped. Workgroups take groups of 4
/// slices, each XCD a contiguous range. With dotv, partial sums of y.dotv
/// go to partials.
template <typename T, int G>
__global__ void __launch_bounds__(kBlock)
    sell_spmv_k(std::int32_t nrows, std::int32_t nslices, const std::int64_t* __restrict__ off,
                const std::int32_t* __restrict__ cbase, const std::int32_t* __restrict__ rp,
                const unsigned char* __restrict__ cols, const T* __restrict__ sval,
                const T* __restrict__ x, T* __restrict__ y, int accumulate,
                const T* __restrict__ dotv, double* partials, const int* stop)
{
  if (stop && block_flag(stop))
    return;
  double d[1] = {0.0};
  const int lane = threadIdx.x % kSlice, wave = threadIdx.x / kSlice;
  const std::int32_t ngroups = (nslices + 3) / 4;
  const XcdRange xr = xcd_range(ngroups);
  for (std::int32_t g = xr.first; g < xr.end; g += xr.step)
  {
    const std::int32_t sl = g * 4 + wave;
    if (sl >= nslices)
      continue;
    const std::int64_t vb = off[sl] + lane;
    const std::int32_t w = static_cast<std::int32_t>((off[sl + 1] - off[sl]) / kSlice);
    const std::int32_t b = cbase[sl];
    const unsigned char* c = cols + off[nslices + 1 + sl];
    const std::int32_t r = sl * kSlice + lane;
    const std::int32_t len = r < nrows ? rp[r + 1] - rp[r] : 0;
    T sum = T(0);
    for (std::int32_t k = 0; k < w; k += G)
    {
      T vv[G], xv[G];
      std::int32_t cc[G];
      if (b >= 0)
      {
        const std::uint16_t* c16 = reinterpret_cast<const std::uint16_t*>(c) + lane;
#pragma unroll
        for (int u = 0; u < G; ++u)
        {
          const std::int64_t e = std::int64_t(min(k + u, w - 1)) * kSlice;
          vv[u] = sval[vb + e];
          cc[u] = b + c16[e];
        }
      }
      else
      {
        const std::int32_t* c32 = reinterpret_cast<const std::int32_t*>(c) + lane;
#pragma unroll
        for (int u = 0; u < G; ++u)
        {
          const std::int64_t e = std::int64_t(min(k + u, w - 1)) * kSlice;
          vv[u] = sval[vb + e];
          cc[u] = c32[e];
        }
      }
#pragma unroll
      for (int u = 0; u < G; ++u)
        xv[u] = x[cc[u]];
      {
        // product then sum, not a fused multiply-add
#pragma clang fp contract(off)
#pragma unroll
        for (int u = 0; u < G; ++u)
          if (k + u < len)
            sum = sum + vv[u] * xv[u];
      }
    }
    if (r < nrows)
    {
      const T yr = accumulate ? y[r] + sum : sum;
      y[r] = yr;
      if (dotv)
        d[0] += double(yr) * double(dotv[r]);
    }
  }
  if (partials)
    block_partials<1>(d, partials);
}

/// Stream SpMV of a block-CSR matrix with BS x BS blocks (row-major inside
/// a block), y (+)= A x. Row blocks hold at most kBlock / BS block rows and
/// stream_chunk / BS blocks. Phase 1: task t = (block b, block-row component
/// a) loads the BS contiguous values of that block row (consecutive tasks
/// read consecutive addresses) and the BS entries of x, and stores their
/// product sum in LDS; phase 2: one thread per (block row, component) sums
/// its row's block products left to right. The next row block's entries
/// are loaded while the current rows are summed.
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock)
    bsr_stream_k(std::int32_t nblk, const int2* __restrict__ rbj,
                 const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                 const T* __restrict__ v, const T* __restrict__ x, T* __restrict__ y,
                 int accumulate, const T* __restrict__ dotv, double* partials,
                 const int* stop)
{
  if (stop && block_flag(stop))
    return;
  constexpr int NC = stream_chunk<T>();        // products per chunk
  constexpr int U = NC / kBlock;               // tasks per thread
  __shared__ T prod[NC];
  double d[1] = {0.0};
  T vv[U][BS];
  std::int32_t cc[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
  {
#pragma unroll
    for (int k = 0; k < BS; ++k)
      vv[u][k] = T(0);
    cc[u] = 0;
  }
  std::int32_t r0 = 0, r1 = 0, j0 = 0, j1 = 0;
  auto bounds = [&](int b)
  {
    const int2 lo = rbj[b], hi = rbj[b + 1];
    r0 = lo.x;
    r1 = hi.x;
    j0 = lo.y;
    j1 = hi.y;
  };
  // task t of the row block: block j0 + t / BS, component t % BS; loads at
  // clamped positions so that no load sits behind a branch
  auto stage = [&]()
  {
    const std::int32_t last = (j1 - j0) * BS - 1;
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const std::int32_t t = min(static_cast<std::int32_t>(threadIdx.x) + u * kBlock, last);
      const std::int64_t j = j0 + t / BS;
      const T* a = v + (j * BS + t % BS) * BS;
#pragma unroll
      for (int k = 0; k < BS; ++k)
        vv[u][k] = a[k];
      cc[u] = ci[j];
    }
  };
  const XcdRange xr = xcd_range(nblk);
  int bi = xr.first;
  if (bi < xr.end)
  {
    bounds(bi);
    if (j1 > j0)
      stage();
  }
  while (bi < xr.end)
  {
    const int nbi = bi + xr.step;
    const std::int32_t ntask = (j1 - j0) * BS;
    T xv[U][BS];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const T* xx = x + std::int64_t(cc[u]) * BS;
#pragma unroll
      for (int k = 0; k < BS; ++k)
        xv[u][k] = xx[k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const std::int32_t t = threadIdx.x + u * kBlock;
      if (t < ntask)
      {
        T s = vv[u][0] * xv[u][0];
#pragma unroll
        for (int k = 1; k < BS; ++k)
          s += vv[u][k] * xv[u][k];
        prod[t] = s;
      }
    }
    lds_barrier();
    const std::int32_t br = r0 + static_cast<std::int32_t>(threadIdx.x) / BS;
    const int comp = static_cast<int>(threadIdx.x) % BS;
    const bool has = br < r1;
    std::int32_t kb = 0, ke = 0;
    const std::int64_t yi = std::int64_t(br) * BS + comp;
    T yold = T(0), dv = T(0);
    if (has)
    {
      kb = rp[br] - j0;
      ke = rp[br + 1] - j0;
      if (accumulate)
        yold = y[yi];
      if (dotv)
        dv = dotv[yi];
    }
    if (nbi < xr.end)
    {
      bounds(nbi);
      if (j1 > j0)
        stage();
    }
    if (has)
    {
      T s = 0;
      for (std::int32_t k = kb; k < ke; ++k)
        s += prod[k * BS + comp];
      const T yr = accumulate ? yold + s : s;
      y[yi] = yr;
      if (dotv)
        d[0] += double(yr) * double(dv);
    }
    lds_barrier();
    bi = nbi;
  }
  if (partials)
    block_partials<1>(d, partials);
}

/// y (+)= A x for scalar CSR: G lanes per row, each lane strides the row,
/// partial sums combined with wavefront shuffles. Rows are handed out in
/// XCD-contiguous chunks of kBlock/G rows. Accumulation in T.
template <typename T, int G>
__global__ void __launch_bounds__(kBlock)
    spmv_k(std::int32_t nrows, const std::int32_t* __restrict__ rp,
           const std::int32_t* __restrict__ ci, const T* __restrict__ v,
           const T* __restrict__ x, T* __restrict__ y, int accumulate)
{
  constexpr int RPB = kBlock / G;
  const int nchunk = (nrows + RPB - 1) / RPB;
  const int lane = threadIdx.x % G;
  for (int chunk = xcd_contiguous(blockIdx.x, gridDim.x); chunk < nchunk;
       chunk += gridDim.x)
  {
    const std::int32_t row = chunk * RPB + threadIdx.x / G;
    T s = 0;
    if (row < nrows)
    {
      const std::int32_t e = rp[row + 1];
      for (std::int32_t j = rp[row] + lane; j < e; j += G)
        s += v[j] * x[ci[j]];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
      s += __shfl_down(s, off, G);
    if (lane == 0 && row < nrows)
      y[row] = accumulate ? y[row] + s : s;
  }
}

/// Block (bs x bs) CSR: one lane per (block row, component k0), looping
/// over the row's blocks; y[bs*i + k0] (+)= sum_j sum_k1 A_j[k0,k1] x[...].
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock)
    bsr_spmv_k(std::int32_t nrows, const std::int32_t* __restrict__ rp,
               const std::int32_t* __restrict__ ci, const T* __restrict__ v,
               const T* __restrict__ x, T* __restrict__ y, int accumulate)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nrows) * BS)
    return;
  const std::int32_t row = static_cast<std::int32_t>(t / BS);
  const int k0 = static_cast<int>(t % BS);
  T s = 0;
  for (std::int32_t j = rp[row]; j < rp[row + 1]; ++j)
  {
    const T* a = v + (std::int64_t(j) * BS + k0) * BS;
    const T* xx = x + std::int64_t(ci[j]) * BS;
#pragma unroll
    for (int k1 = 0; k1 < BS; ++k1)
      s += a[k1] * xx[k1];
  }
  y[t] = accumulate ? y[t] + s : s;
}

/// Lanes per row of the scalar CSR kernels for a mean row length.
int lanes_per_row(const dfx_csr& A)
{
  if (A.lanes == -1)
    return 1; // sliced-ELL snapshot: one lane per row
  if (A.lanes)
    return A.lanes;
  if (A.nblk > 0)
    return 1; // stream kernel
  const double mean = A.nrows ? double(A.nnz) / A.nrows : 1.0;
  int G = 2;
  while (G < 64 && G * 6 < mean)
    G *= 2;
  return G;
}

void materialize_zero(const dfx_csr& A)
{
  if (!A.zero_pending)
    return;
  const std::int64_t n = A.nnz * A.bs * A.bs;
  const int nb = reduce_blocks(n);
  if (A.dtype == DFX_F64)
    fill_k<double><<<nb, kBlock, 0, stream()>>>(static_cast<double*>(A.values), n, 0.0);
  else
    fill_k<float><<<nb, kBlock, 0, stream()>>>(static_cast<float*>(A.values), n, 0.0f);
  check_launch("csr_zero");
  A.zero_pending = false;
}

/// The stream SpMV row blocks of A as (first row, first entry) pairs.
inline const int2* row_blocks(const dfx_csr& A) { return reinterpret_cast<const int2*>(A.rblk); }

/// Queries the resident grids of the stream SpMV kernels once, at matrix
/// creation, so that no solve pays for the first occupancy query.
void prime_stream_grids()
{
  static std::once_flag once;
  std::call_once(once,
                 []
                 {
                   const void* k[] = {reinterpret_cast<const void*>(spmv_stream_k<double>),
                                      reinterpret_cast<const void*>(spmv_stream_k<float>),
                                      reinterpret_cast<const void*>(bsr_stream_k<double, 3>),
                                      reinterpret_cast<const void*>(bsr_stream_k<float, 3>),
                                      reinterpret_cast<const void*>(bsr_stream_k<double, 2>),
                                      reinterpret_cast<const void*>(bsr_stream_k<float, 2>)};
                   for (const void* f : k)
                     resident_blocks(f, kBlock, 0);
                 });
}

constexpr int kMaxPartials = 2048;

/// y (+)= A x through the sliced-ELL copy of A (its values as last
/// copied), with the partial sums of y . dotv into partials when given.
/// Returns the number of workgroups.
template <typename T>
int launch_sell(const dfx_csr& A, const T* x, T* y, int accumulate, const T* dotv,
                double* partials, const int* stop)
{
  // 1024 workgroups (16 wavefronts per CU) keep the stream fed and halve
  // the partials a reduction sums
  const int ngroups = (A.nslices + 3) / 4;
  const int nb = std::max(1, std::min(ngroups, kMaxPartials / 2));
  // entries per step: the width among 16, 12, 9, 8 that pads the longest
  // row least (27-entry hexahedral rows: 3 steps of 9, no clamped loads)
  int G = 16;
  for (int g : {12, 9, 8})
    if ((A.max_row + g - 1) / g * g < (A.max_row + G - 1) / G * G)
      G = g;
  auto go = [&](auto kern)
  {
    kern<<<nb, kBlock, 0, stream()>>>(A.nrows, A.nslices, A.sell_off, A.sell_base, A.row_ptr,
                                      static_cast<const unsigned char*>(A.sell_cols),
                                      static_cast<const T*>(A.sell_vals), x, y, accumulate, dotv,
                                      partials, stop);
  };
  if (G == 9)
    go(sell_spmv_k<T, 9>);
  else if (G == 12)
    go(sell_spmv_k<T, 12>);
  else if (G == 8)
    go(sell_spmv_k<T, 8>);
  else
    go(sell_spmv_k<T, 16>);
  return nb;
}

template <typename T>
void launch_spmv(const dfx_csr& A, const T* x, T* y, bool accumulate)
{
  materialize_zero(A);
  hipStream_t s = stream();
  if (A.lanes == -1)
  {
    // the sliced-ELL snapshot taken by dfx_csr_set_lanes(A, -1)
    launch_sell<T>(A, x, y, accumulate ? 1 : 0, nullptr, nullptr, nullptr);
    return;
  }
  if (A.bs == 1 && lanes_per_row(A) == 1)
  {
    const int nb = std::max(
        1, std::min(A.nblk, resident_blocks(reinterpret_cast<const void*>(spmv_stream_k<T>),
                                            kBlock, 0)));
    spmv_stream_k<T><<<nb, kBlock, 0, s>>>(A.nblk, row_blocks(A), A.row_ptr, A.cols,
                                           static_cast<const T*>(A.values), x, y,
                                           accumulate, nullptr, nullptr, nullptr);
  }
  else if (A.bs == 1)
  {
    const int G = lanes_per_row(A);
    const int rpb = kBlock / G;
    const int nchunk = (A.nrows + rpb - 1) / rpb;
    const int nb = std::max(1, std::min(nchunk, ctx().num_cu * 16));
    const T* v = static_cast<const T*>(A.values);
    switch (G)
    {
#define DFX_SPMV_CASE(GG)                                                      \
  case GG:                                                                     \
    spmv_k<T, GG><<<nb, kBlock, 0, s>>>(A.nrows, A.row_ptr, A.cols, v, x, y,  \
                                        accumulate);                           \
    break;
      DFX_SPMV_CASE(2)
      DFX_SPMV_CASE(4)
      DFX_SPMV_CASE(8)
      DFX_SPMV_CASE(16)
      DFX_SPMV_CASE(32)
      DFX_SPMV_CASE(64)
#undef DFX_SPMV_CASE
    }
  }
  else
  {
    const T* v = static_cast<const T*>(A.values);
    if (A.nblk > 0 && A.lanes <= 1 && (A.bs == 2 || A.bs == 3))
    {
      const void* kern = A.bs == 3 ? reinterpret_cast<const void*>(bsr_stream_k<T, 3>)
                                   : reinterpret_cast<const void*>(bsr_stream_k<T, 2>);
      const int nb = std::max(1, std::min(A.nblk, resident_blocks(kern, kBlock, 0)));
      if (A.bs == 3)
        bsr_stream_k<T, 3><<<nb, kBlock, 0, s>>>(A.nblk, row_blocks(A), A.row_ptr, A.cols, v, x, y,
                                                 accumulate, nullptr, nullptr, nullptr);
      else
        bsr_stream_k<T, 2><<<nb, kBlock, 0, s>>>(A.nblk, row_blocks(A), A.row_ptr, A.cols, v, x, y,
                                                 accumulate, nullptr, nullptr, nullptr);
      check_launch("spmv");
      return;
    }
    const std::int64_t n = std::int64_t(A.nrows) * A.bs;
    const int nb = static_cast<int>((n + kBlock - 1) / kBlock);
    if (A.bs == 2)
      bsr_spmv_k<T, 2><<<nb, kBlock, 0, s>>>(A.nrows, A.row_ptr, A.cols, v, x, y, accumulate);
    else if (A.bs == 3)
      bsr_spmv_k<T, 3><<<nb, kBlock, 0, s>>>(A.nrows, A.row_ptr, A.cols, v, x, y, accumulate);
    else
      throw Error(DFX_ERR_ARG, "block size must be 1..3");
  }
  check_launch("spmv");
}

// ---------------------------------------------------------------------
// Conjugate gradients, device resident. Three kernels per iteration k,
// no atomics and no host round trip:
//   cg_spmv_k      q = A p, partial sums of p.q
//   cg_update_k    alpha = rz_k / p.q (each block sums the partials);
//                  x += alpha p, r -= alpha q; partials of r.z and r.r
//   cg_direction_k beta = rz_{k+1} / rz_k; stop test ||r||^2/||r0||^2 <
//                  rtol^2; p = z + beta p; block 0 records the scalars.
// rz is kept in a two-slot ring indexed by the iteration parity, so the
// slot being read in an iteration is never the one written. After the
// stop test fires every later kernel returns at once.
// ---------------------------------------------------------------------


struct CGState
{
  double rz[2];
  double rr0, rr, tol2;
  std::int32_t iter, done, max_it;
  std::int32_t kd; // iteration the direction kernel completes (set by the update kernel)
  double partials[3 * kMaxPartials];
};

/// Jacobi inverse diagonal of a (block) CSR matrix: 1 / A[i,i].
template <typename T>
__global__ void __launch_bounds__(kBlock)
    inv_diag_k(std::int32_t nrows, int bs, const std::int32_t* diag, const T* v, T* dinv)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nrows) * bs)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / bs);
  const int k = static_cast<int>(t % bs);
  const std::int32_t p = diag[r];
  const T d = p >= 0 ? v[(std::int64_t(p) * bs + k) * bs + k] : T(0);
  dinv[t] = d != T(0) ? T(1) / d : T(1);
}

/// r = b - y; p = z = dinv r; partial sums of (r.z, r.r)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_start_k(std::int64_t n, const T* b, const T* y, const T* dinv, T* r, T* p, CGState* st)
{
  double s[2] = {0.0, 0.0};
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
  {
    const T ri = b[i] - y[i];
    const T zi = dinv ? dinv[i] * ri : ri;
    r[i] = ri;
    p[i] = zi;
    s[0] += double(ri) * double(zi);
    s[1] += double(ri) * double(ri);
  }
  block_partials<2>(s, st->partials + kMaxPartials);
}

/// One block: rz_0, rr0 from the start partials.
__global__ void __launch_bounds__(kBlock) cg_start_finish_k(int nb, CGState* st)
{
  const double rz = sum_partials(st->partials + kMaxPartials, nb);
  const double rr = sum_partials(st->partials + kMaxPartials + nb, nb);
  if (threadIdx.x == 0)
  {
    st->rz[0] = rz;
    st->rr0 = rr;
    st->rr = rr;
    st->iter = 0;
    st->done = (rr == 0.0 || st->max_it <= 0) ? 1 : 0;
  }
}

/// q = A p and partial sums of p.q, G lanes per row.
template <typename T, int G>
__global__ void __launch_bounds__(kBlock)
    cg_spmv_k(std::int32_t nrows, const std::int32_t* __restrict__ rp,
              const std::int32_t* __restrict__ ci, const T* __restrict__ v,
              const T* __restrict__ p, T* __restrict__ q, CGState* st)
{
  if (block_flag(&st->done))
    return;
  constexpr int RPB = kBlock / G;
  const int nchunk = (nrows + RPB - 1) / RPB;
  const int lane = threadIdx.x % G;
  double d[1] = {0.0};
  for (int chunk = xcd_contiguous(blockIdx.x, gridDim.x); chunk < nchunk;
       chunk += gridDim.x)
  {
    const std::int32_t row = chunk * RPB + threadIdx.x / G;
    T s = 0;
    if (row < nrows)
    {
      const std::int32_t e = rp[row + 1];
      for (std::int32_t j = rp[row] + lane; j < e; j += G)
        s += v[j] * p[ci[j]];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
      s += __shfl_down(s, off, G);
    if (lane == 0 && row < nrows)
    {
      q[row] = s;
      d[0] += double(s) * double(p[row]);
    }
  }
  block_partials<1>(d, st->partials);
}

/// Block CSR variant of cg_spmv_k (one lane per (block row, component)).
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock)
    cg_bsr_spmv_k(std::int32_t nrows, const std::int32_t* __restrict__ rp,
                  const std::int32_t* __restrict__ ci, const T* __restrict__ v,
                  const T* __restrict__ p, T* __restrict__ q, CGState* st)
{
  if (block_flag(&st->done))
    return;
  const std::int64_t n = std::int64_t(nrows) * BS;
  double d[1] = {0.0};
  for (std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; t < n;
       t += std::int64_t(gridDim.x) * kBlock)
  {
    const std::int32_t row = static_cast<std::int32_t>(t / BS);
    const int k0 = static_cast<int>(t % BS);
    T s = 0;
    for (std::int32_t j = rp[row]; j < rp[row + 1]; ++j)
    {
      const T* a = v + (std::int64_t(j) * BS + k0) * BS;
      const T* xx = p + std::int64_t(ci[j]) * BS;
#pragma unroll
      for (int k1 = 0; k1 < BS; ++k1)
        s += a[k1] * xx[k1];
    }
    q[t] = s;
    d[0] += double(s) * double(p[t]);
  }
  block_partials<1>(d, st->partials);
}

/// alpha = rz_k / p.q; x += alpha p; r -= alpha q; partials of r.z, r.r.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_update_k(std::int64_t n, int nb_spmv, T* x, T* r, const T* p, const T* q,
                const T* dinv, CGState* st)
{
  if (block_flag(&st->done))
    return;
  // the iteration index lives in the state (written by the previous
  // direction kernel), so a captured batch of iterations can be replayed
  const int k = st->iter;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    st->kd = k;
  const double pq = sum_partials(st->partials, nb_spmv);
  const T alpha = static_cast<T>(st->rz[k & 1] / pq);
  double s[2] = {0.0, 0.0};
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
  {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv ? dinv[i] * ri : ri;
    s[0] += double(ri) * double(zi);
    s[1] += double(ri) * double(ri);
  }
  block_partials<2>(s, st->partials + kMaxPartials);
}

/// beta = rz_{k+1} / rz_k; stop test; p = dinv r + beta p.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_direction_k(std::int64_t n, int nb_upd, const T* r, const T* dinv, T* p,
                   CGState* st)
{
  if (block_flag(&st->done))
    return;
  const int k = st->kd;
  const double rz = sum_partials(st->partials + kMaxPartials, nb_upd);
  const double rr = sum_partials(st->partials + kMaxPartials + nb_upd, nb_upd);
  const bool stop = rr / st->rr0 < st->tol2 || k + 1 >= st->max_it;
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    st->rz[(k + 1) & 1] = rz;
    st->rr = rr;
    st->iter = k + 1;
    if (stop)
      st->done = 1;
  }
  if (stop)
    return;
  const T beta = static_cast<T>(rz / st->rz[k & 1]);
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    p[i] = (dinv ? dinv[i] * r[i] : r[i]) + beta * p[i];
}

/// Partial sums of (r.z, r.r) with z = dinv r, into the update partials
/// (used after the residual has been replaced).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_rz_partials_k(std::int64_t n, const T* r, const T* dinv, CGState* st)
{
  double s[2] = {0.0, 0.0};
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
  {
    const T ri = r[i];
    const T zi = dinv ? dinv[i] * ri : ri;
    s[0] += double(ri) * double(zi);
    s[1] += double(ri) * double(ri);
  }
  block_partials<2>(s, st->partials + kMaxPartials);
}

/// Re-arms a stopped CG after a residual replacement.
__global__ void cg_resume_k(CGState* st, double rr0, double tol2, int kd)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
  {
    st->rr0 = rr0;
    st->tol2 = tol2;
    st->done = 0;
    if (kd >= 0)
      st->kd = kd;
  }
}

/// Blocks of a reduction kernel: one element per thread up to kMaxPartials
/// blocks (a small system spread over many CUs: each element's loads are
/// one dependent round trip, so per-thread loops would serialise them).
int partial_blocks(std::int64_t n)
{
  std::int64_t b = (n + kBlock - 1) / kBlock;
  return static_cast<int>(std::clamp<std::int64_t>(b, 1, 1024));
}

// ---------------------------------------------------------------------
// CG of a small scalar system in one cooperative launch. A system of a few
// ten thousand rows lives in the L2 caches, and each kernel of the
// three-kernel iteration is a short chain of dependent round trips plus a
// launch; here every block stays resident for the whole solve and the
// phases are separated by grid barriers, two per iteration:
//   phase A  p = z + beta p_old on the block's rows (z = dinv r), written
//            to the other of two p buffers; q = A p for its rows, with p
//            of the neighbour columns formed the same way from z and
//            p_old; partial sums of p.q                         | barrier
//   phase B  alpha = rz / p.q (every block sums the partials in the same
//            order); x += alpha p, r -= alpha q, z = dinv r (written);
//            partials of r.z and r.r                            | barrier
// then every block sums rz, rr identically and takes the same stop
// decision (||r||^2/||r0||^2 < rtol^2 or max_it), so all blocks leave the
// loop at the same iteration. x, r, dinv and the block's own p and q stay
// in registers (at most RPT rows per thread). The barrier counters only
// grow (barrier e completes when the top counter reaches e x groups)
// under agent-scope release/acquire; a wait gives up after a bounded
// number of polls and raises a flag that every block sees, so the grid
// always drains. Same stop test and FP64 reductions as the
// three-kernel path; sums are grouped differently, so the last bits (and
// at most an iteration) may differ from it.
// ---------------------------------------------------------------------

/// Barrier state and results of the cooperative CG (zeroed before each
/// launch): blocks arrive on one of 8 group counters (block b on group
/// b mod 8, each on its own 128-byte line, so arrivals on different
/// counters proceed in parallel); the last arrival of a group advances
/// the top counter, which every block polls.
struct CoopSync
{
  unsigned int top;
  int failed;
  std::int32_t iter;
  std::int32_t pad0;
  double rr0, rr;
  unsigned int pad1[24];
  unsigned int group[8 * 32];
};

constexpr unsigned int kCoopSpinLimit = 1u << 22;
constexpr int kCoopMaxRPT = 4;
constexpr std::int32_t kCoopMaxRows = 65536;

/// Grid barrier of the cooperative CG: false if it timed out (or another
/// block's did), in which case the caller returns. Every wavefront waits
/// for its own memory operations before the block barrier, so thread 0's
/// release (an L2 write-back at agent scope) covers the block's stores;
/// its acquire (an invalidate of the CU's caches) precedes every load of
/// the block after the barrier.
__device__ inline bool coop_barrier(CoopSync* g, unsigned int nb, unsigned int& epoch)
{
  __shared__ int ok;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  epoch += 1;
  if (threadIdx.x == 0)
  {
    const unsigned int grp = blockIdx.x & 7u;
    const unsigned int gsize = (nb - grp + 7u) / 8u;
    const unsigned int ngroups = nb < 8u ? nb : 8u;
    const unsigned int old = __hip_atomic_fetch_add(&g->group[32 * grp], 1u, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    if (old == epoch * gsize - 1u)
      __hip_atomic_fetch_add(&g->top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int target = epoch * ngroups;
    int good = 1;
    unsigned int spins = 0;
    while (__hip_atomic_load(&g->top, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target)
    {
      if (__hip_atomic_load(&g->failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
          || ++spins > kCoopSpinLimit)
      {
        __hip_atomic_store(&g->failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

/// p of row j in iteration k: z_j (k = 0) or z_j + beta p_old_j, one
/// rounding pattern wherever it is formed.
template <typename T>
__device__ inline T coop_dir(T z, T pold, T beta, bool first)
{
  return first ? z : __builtin_fma(beta, pold, z);
}

template <typename T, int RPT>
__global__ void __launch_bounds__(kBlock)
    cg_coop_k(std::int32_t n, const std::int32_t* __restrict__ rp,
              const std::int32_t* __restrict__ ci, const T* __restrict__ v,
              const std::int32_t* __restrict__ diag, int jacobi, const T* __restrict__ b,
              T* __restrict__ x, T* z, T* p0, T* p1, double* parts, CoopSync* g, int max_it,
              double tol2)
{
  const unsigned int nb = gridDim.x;
  const std::int32_t nt = static_cast<std::int32_t>(nb) * kBlock;
  const std::int32_t t0 = static_cast<std::int32_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned int epoch = 0;
  T xr[RPT], rr_[RPT], dr[RPT], pr[RPT], qr[RPT];
  // r = b - A x, z = dinv r; partials of r.z, r.r
  double s[2] = {0.0, 0.0};
#pragma unroll
  for (int u = 0; u < RPT; ++u)
  {
    const std::int32_t i = t0 + u * nt;
    xr[u] = rr_[u] = dr[u] = pr[u] = qr[u] = T(0);
    if (i < n)
    {
      T ax = 0;
      for (std::int32_t e = rp[i]; e < rp[i + 1]; ++e)
        ax += v[e] * x[ci[e]];
      xr[u] = x[i];
      rr_[u] = b[i] - ax;
      T d = T(1);
      if (jacobi)
      {
        const std::int32_t pd = diag[i];
        const T a = pd >= 0 ? v[pd] : T(0);
        d = a != T(0) ? T(1) / a : T(1);
      }
      dr[u] = d;
      const T zi = d * rr_[u];
      z[i] = zi;
      s[0] += double(rr_[u]) * double(zi);
      s[1] += double(rr_[u]) * double(rr_[u]);
    }
  }
  block_partials<2>(s, parts + nb);
  if (!coop_barrier(g, nb, epoch))
    return;
  double rz = sum_partials(parts + nb, nb);
  const double rr0 = sum_partials(parts + 2 * nb, nb);
  double rr = rr0;
  int k = 0;
  bool done = rr0 == 0.0 || max_it <= 0;
  T beta = T(0);
  while (!done)
  {
    const bool first = k == 0;
    T* pn = (k & 1) ? p1 : p0;
    const T* po = (k & 1) ? p0 : p1;
    // phase A
    double d = 0.0;
#pragma unroll
    for (int u = 0; u < RPT; ++u)
    {
      const std::int32_t i = t0 + u * nt;
      if (i < n)
      {
        const T pi = coop_dir<T>(dr[u] * rr_[u], pr[u], beta, first);
        pr[u] = pi;
        pn[i] = pi;
        // the row in chunks of 8 entries: the chunk's columns and values,
        // then all its gathers in flight together, then the sums in order
        T q = 0;
        const std::int32_t re = rp[i + 1];
        for (std::int32_t e0 = rp[i]; e0 < re; e0 += 8)
        {
          std::int32_t jj[8];
          T vv[8], zz[8], pp[8];
#pragma unroll
          for (int h = 0; h < 8; ++h)
          {
            const bool in = e0 + h < re;
            jj[h] = in ? ci[e0 + h] : i;
            vv[h] = in ? v[e0 + h] : T(0);
          }
#pragma unroll
          for (int h = 0; h < 8; ++h)
          {
            zz[h] = z[jj[h]];
            pp[h] = first ? T(0) : po[jj[h]];
          }
#pragma unroll
          for (int h = 0; h < 8; ++h)
            if (e0 + h < re)
              q += vv[h] * coop_dir<T>(zz[h], pp[h], beta, first);
        }
        qr[u] = q;
        d += double(q) * double(pi);
      }
    }
    {
      double dd[1] = {d};
      block_partials<1>(dd, parts);
    }
    if (!coop_barrier(g, nb, epoch))
      return;
    // phase B
    const double pq = sum_partials(parts, nb);
    const T alpha = static_cast<T>(rz / pq);
    s[0] = s[1] = 0.0;
#pragma unroll
    for (int u = 0; u < RPT; ++u)
    {
      const std::int32_t i = t0 + u * nt;
      if (i < n)
      {
        xr[u] += alpha * pr[u];
        rr_[u] = rr_[u] - alpha * qr[u];
        const T zi = dr[u] * rr_[u];
        z[i] = zi;
        s[0] += double(rr_[u]) * double(zi);
        s[1] += double(rr_[u]) * double(rr_[u]);
      }
    }
    block_partials<2>(s, parts + nb);
    if (!coop_barrier(g, nb, epoch))
      return;
    const double rzn = sum_partials(parts + nb, nb);
    rr = sum_partials(parts + 2 * nb, nb);
    done = rr / rr0 < tol2 || k + 1 >= max_it;
    beta = static_cast<T>(rzn / rz);
    rz = rzn;
    ++k;
  }
#pragma unroll
  for (int u = 0; u < RPT; ++u)
  {
    const std::int32_t i = t0 + u * nt;
    if (i < n)
      x[i] = xr[u];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    g->iter = k;
    g->rr0 = rr0;
    g->rr = rr;
  }
}

/// Rows per thread and blocks of the cooperative CG for A, or 0 when the
/// system is not one for it (block size > 1, too large for a co-resident
/// grid, no cooperative launch, or a barrier counter that could wrap).
template <typename T>
std::pair<int, int> coop_shape(const dfx_csr& A, int max_it)
{
  if (A.bs != 1 || A.nrows == 0 || A.nrows > kCoopMaxRows || !ctx().cooperative)
    return {0, 0};
  const int need = (A.nrows + kBlock - 1) / kBlock;
  // a barrier costs more the more blocks arrive (measured on MI355X: 5.8 us
  // at 141 blocks, 3.4 us at 64, 2.5 us at 32): two rows per thread
  // once a system needs more than 64 blocks of one row per thread
  int rpt = need > 64 ? 2 : 1;
  int cap = 0;
  for (; rpt <= kCoopMaxRPT; rpt *= 2)
  {
    const void* kern = rpt == 1   ? reinterpret_cast<const void*>(cg_coop_k<T, 1>)
                       : rpt == 2 ? reinterpret_cast<const void*>(cg_coop_k<T, 2>)
                                  : reinterpret_cast<const void*>(cg_coop_k<T, 4>);
    cap = std::min(resident_blocks(kern, kBlock, 0), kMaxPartials);
    if ((need + rpt - 1) / rpt <= cap)
      break;
  }
  if (rpt > kCoopMaxRPT)
    return {0, 0};
  const int nb = (need + rpt - 1) / rpt;
  // 2 barriers per iteration plus 2: the counter reaches (2 max_it + 2) nb
  if ((2.0 * max_it + 4.0) * nb >= 4294967295.0)
    return {0, 0};
  return {rpt, nb};
}

/// The cooperative CG (coop_shape gave rpt, nb); work vectors from w.
template <typename T>
int cg_coop_run(const dfx_csr& A, const T* b, T* x, int max_it, double rtol, bool jacobi,
                int rpt, int nb, T* z, T* p0, T* p1, double* parts, CoopSync* g,
                bool* converged, double* rr0, double* rr)
{
  hipStream_t s = stream();
  fill_k<unsigned int><<<1, kBlock, 0, s>>>(reinterpret_cast<unsigned int*>(g),
                                            sizeof(CoopSync) / sizeof(unsigned int), 0u);
  check_launch("cg_coop set-up");
  std::int32_t n = A.nrows;
  const std::int32_t* rpp = A.row_ptr;
  const std::int32_t* cip = A.cols;
  const T* vp = static_cast<const T*>(A.values);
  const std::int32_t* dg = A.diag;
  int jac = jacobi ? 1 : 0;
  double tol2 = rtol * rtol;
  void* args[] = {&n, &rpp, &cip, &vp, &dg, &jac, &b, &x, &z, &p0, &p1, &parts, &g, &max_it,
                  &tol2};
  const void* kern = rpt == 1   ? reinterpret_cast<const void*>(cg_coop_k<T, 1>)
                     : rpt == 2 ? reinterpret_cast<const void*>(cg_coop_k<T, 2>)
                                : reinterpret_cast<const void*>(cg_coop_k<T, 4>);
  const hipError_t le = hipLaunchCooperativeKernel(kern, dim3(nb), dim3(kBlock), args, 0, s);
  if (le != hipSuccess)
  {
    // refused (the process cannot launch co-resident grids): the caller
    // runs the three-kernel iteration, here and for later solves
    (void)hipGetLastError();
    ctx().cooperative = false;
    return -1;
  }
  CoopSync h{};
  DFX_HIP(hipMemcpyAsync(&h, g, sizeof(CoopSync), hipMemcpyDeviceToHost, s));
  DFX_HIP(hipStreamSynchronize(s));
  if (h.failed)
    throw Error(DFX_ERR_HIP, "cooperative CG: a grid barrier timed out");
  *rr0 = h.rr0;
  *rr = h.rr;
  *converged = h.rr0 == 0.0 || h.rr / h.rr0 < rtol * rtol;
  return h.iter;
}

/// Per-dtype scratch of a CG solve (device).
template <typename T>
struct CGWork
{
  std::int64_t n = 0;
  T *r = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr;
  bool own = true; // false: the vectors live in caller-provided scratch
  void alloc(std::int64_t nn, bool jacobi)
  {
    n = nn;
    r = static_cast<T*>(dmalloc(sizeof(T) * n));
    p = static_cast<T*>(dmalloc(sizeof(T) * n));
    q = static_cast<T*>(dmalloc(sizeof(T) * n));
    if (jacobi)
      dinv = static_cast<T*>(dmalloc(sizeof(T) * n));
  }
  ~CGWork()
  {
    if (!own)
      return;
    dfree(r);
    dfree(p);
    dfree(q);
    dfree(dinv);
  }
};

struct StateHolder
{
  CGState* st = nullptr;
  StateHolder() { st = static_cast<CGState*>(dmalloc(sizeof(CGState))); }
  ~StateHolder() { dfree(st); }
};

/// Refresh the sliced-ELL copy of A from its CSR values; the slice layout
/// and columns are rebuilt only when the pattern changed. Returns whether
/// the copy is used: scalar CSR with the stream kernel as its SpMV, and
/// padding to the slices' longest rows adding at most 15 % entries (longer
/// and shorter rows mixed in one slice, as in P3 patterns, stay on the
/// stream kernel).
template <typename T>
bool refresh_sell(const dfx_csr& A, bool values = true);

/// The sliced-ELL layout of A (pattern work, allocation) without copying
/// the values; true if the solvers will use the copy.
template <typename T>
bool sell_layout(const dfx_csr& A)
{
  return refresh_sell<T>(A, false);
}

template <typename T>
bool refresh_sell(const dfx_csr& A, bool values)
{
  if (A.bs != 1 || A.nrows == 0 || lanes_per_row(A) != 1)
    return false;
  const std::int32_t ns = (A.nrows + kSlice - 1) / kSlice;
  if (A.sell_pattern != A.pattern_id)
  {
    dfree(A.sell_vals);
    dfree(A.sell_cols);
    dfree(A.sell_off);
    dfree(A.sell_base);
    A.sell_vals = A.sell_cols = nullptr;
    A.sell_off = nullptr;
    A.sell_base = nullptr;
    A.sell_use = false;
    std::int32_t* shape = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * 3 * ns));
    const std::int64_t nt = std::int64_t(ns) * kSlice;
    sell_shape_k<<<static_cast<int>((nt + kBlock - 1) / kBlock), kBlock, 0, stream()>>>(
        A.nrows, ns, A.row_ptr, A.cols, shape);
    check_launch("sell_shape");
    std::vector<std::int32_t> h(3 * std::size_t(ns));
    d2h(h.data(), shape, sizeof(std::int32_t) * h.size());
    dfree(shape);
    // entry offsets, column byte offsets (16-byte aligned), slice bases
    std::vector<std::int64_t> off(2 * (std::size_t(ns) + 1), 0);
    std::vector<std::int32_t> base(ns);
    std::int64_t cb = 0;
    for (std::int32_t sl = 0; sl < ns; ++sl)
    {
      const std::int64_t w = h[3 * sl];
      const std::int32_t lo = w > 0 ? h[3 * sl + 1] : 0, hi = w > 0 ? h[3 * sl + 2] : 0;
      const bool narrow = std::int64_t(hi) - lo < 65536;
      base[sl] = narrow ? lo : -1 - lo; // -1 - first column: 32-bit slice
      off[sl + 1] = off[sl] + w * kSlice;
      off[ns + 1 + sl] = cb;
      cb += (w * kSlice * (narrow ? 2 : 4) + 15) / 16 * 16;
    }
    off[2 * std::size_t(ns) + 1] = cb;
    A.sell_pattern = A.pattern_id;
    if (off[ns] > A.nnz + A.nnz / 20 * 3)
      return false; // padding above 15 %: the stream kernel is used
    A.sell_off = static_cast<std::int64_t*>(dmalloc(sizeof(std::int64_t) * off.size()));
    h2d(A.sell_off, off.data(), sizeof(std::int64_t) * off.size());
    A.sell_base = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max(ns, 1)));
    h2d(A.sell_base, base.data(), sizeof(std::int32_t) * ns);
    A.sell_cols = dmalloc(std::max<std::int64_t>(cb, 16));
    A.sell_vals = dmalloc(sizeof(T) * std::max<std::int64_t>(off[ns], 1));
    A.nslices = ns;
    A.sell_use = true;
    const int nb = static_cast<int>((nt + kBlock - 1) / kBlock);
    // columns (and values unless only the layout is asked for)
    sell_build_k<T><<<nb, kBlock, 0, stream()>>>(
        A.nrows, ns, A.sell_off, A.sell_base, A.row_ptr, A.cols, static_cast<const T*>(A.values),
        static_cast<unsigned char*>(A.sell_cols),
        values ? static_cast<T*>(A.sell_vals) : nullptr);
    check_launch("sell_build");
    return true;
  }
  if (!A.sell_use)
    return false;
  if (!values)
    return true;
  const std::int64_t nt = std::int64_t(ns) * kSlice;
  sell_build_k<T><<<static_cast<int>((nt + kBlock - 1) / kBlock), kBlock, 0, stream()>>>(
      A.nrows, ns, A.sell_off, A.sell_base, A.row_ptr, A.cols, static_cast<const T*>(A.values),
      nullptr, static_cast<T*>(A.sell_vals));
  check_launch("sell_build");
  return true;
}

template <typename T>
int launch_cg_spmv(const dfx_csr& A, const T* p, T* q, CGState* st, bool sell)
{
  hipStream_t s = stream();
  const T* v = static_cast<const T*>(A.values);
  if (sell)
    return launch_sell<T>(A, p, q, 0, p, st->partials, &st->done);
  if (A.bs == 1 && lanes_per_row(A) == 1)
  {
    const int nb = std::max(
        1, std::min({A.nblk,
                     resident_blocks(reinterpret_cast<const void*>(spmv_stream_k<T>), kBlock, 0),
                     kMaxPartials}));
    spmv_stream_k<T><<<nb, kBlock, 0, s>>>(A.nblk, row_blocks(A), A.row_ptr, A.cols, v, p, q, 0, p,
                                           st->partials, &st->done);
    return nb;
  }
  if (A.bs == 1)
  {
    const int G = lanes_per_row(A);
    const int rpb = kBlock / G;
    const int nchunk = (A.nrows + rpb - 1) / rpb;
    const int nb = std::max(1, std::min(nchunk, kMaxPartials));
    switch (G)
    {
#define DFX_CG_CASE(GG)                                                        \
  case GG:                                                                     \
    cg_spmv_k<T, GG><<<nb, kBlock, 0, s>>>(A.nrows, A.row_ptr, A.cols, v, p, q, st); \
    break;
      DFX_CG_CASE(2)
      DFX_CG_CASE(4)
      DFX_CG_CASE(8)
      DFX_CG_CASE(16)
      DFX_CG_CASE(32)
      DFX_CG_CASE(64)
#undef DFX_CG_CASE
    }
    return nb;
  }
  if (A.nblk > 0 && A.lanes <= 1 && (A.bs == 2 || A.bs == 3))
  {
    const void* kern = A.bs == 3 ? reinterpret_cast<const void*>(bsr_stream_k<T, 3>)
                                 : reinterpret_cast<const void*>(bsr_stream_k<T, 2>);
    const int nb = std::max(1, std::min({A.nblk, resident_blocks(kern, kBlock, 0), kMaxPartials}));
    if (A.bs == 3)
      bsr_stream_k<T, 3><<<nb, kBlock, 0, s>>>(A.nblk, row_blocks(A), A.row_ptr, A.cols, v, p, q, 0, p,
                                               st->partials, &st->done);
    else
      bsr_stream_k<T, 2><<<nb, kBlock, 0, s>>>(A.nblk, row_blocks(A), A.row_ptr, A.cols, v, p, q, 0, p,
                                               st->partials, &st->done);
    return nb;
  }
  const int nb = partial_blocks(std::int64_t(A.nrows) * A.bs);
  if (A.bs == 3)
    cg_bsr_spmv_k<T, 3><<<nb, kBlock, 0, s>>>(A.nrows, A.row_ptr, A.cols, v, p, q, st);
  else if (A.bs == 2)
    cg_bsr_spmv_k<T, 2><<<nb, kBlock, 0, s>>>(A.nrows, A.row_ptr, A.cols, v, p, q, st);
  else
    throw Error(DFX_ERR_ARG, "block size must be 1..3");
  return nb;
}

/// A captured sequence of launches on a stream, replayed with
/// hipGraphLaunch (no memset or copy nodes: kernels only).
struct BatchGraph
{
  hipGraphExec_t exec = nullptr;
  template <typename F>
  void capture(hipStream_t s, F&& launches)
  {
    hipGraph_t g = nullptr;
    DFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try
    {
      launches();
    }
    catch (...)
    {
      (void)hipStreamEndCapture(s, &g);
      if (g)
        (void)hipGraphDestroy(g);
      throw;
    }
    DFX_HIP(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
    {
      exec = nullptr;
      throw Error(DFX_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(e));
    }
  }
  ~BatchGraph()
  {
    if (exec)
      (void)hipGraphExecDestroy(exec);
  }
};

/// Runs (preconditioned) CG on A x = b. All arrays are device arrays of T.
/// Returns iterations; sets *converged.
template <typename T>
int cg_run(const dfx_csr& A, const T* b, T* x, int max_it, double rtol, bool jacobi,
           int check_every, CGWork<T>& w, CGState* st, bool* converged, double* rr0,
           double* rr, bool x_is_zero = false, bool allow_sell = true, bool use_graph = false)
{
  hipStream_t s = stream();
  const std::int64_t n = std::int64_t(A.nrows) * A.bs;
  const int nbv = partial_blocks(n);
  // the sliced-ELL copy belongs to the matrix handle and is allocated on
  // its first use; allow_sell = false keeps the CSR kernels
  const bool sell = allow_sell && refresh_sell<T>(A);
  if (jacobi)
  {
    const int nb = static_cast<int>((n + kBlock - 1) / kBlock);
    inv_diag_k<T><<<nb, kBlock, 0, s>>>(A.nrows, A.bs, A.diag,
                                        static_cast<const T*>(A.values), w.dinv);
  }
  // q = A x (initial residual), then r, p and the first reductions
  if (x_is_zero)
    fill_k<T><<<reduce_blocks(n), kBlock, 0, s>>>(w.q, n, T(0));
  else
    launch_spmv<T>(A, x, w.q, false);
  CGState init{};
  init.tol2 = rtol * rtol;
  init.max_it = max_it;
  DFX_HIP(hipMemcpyAsync(st, &init, offsetof(CGState, partials), hipMemcpyHostToDevice, s));
  cg_start_k<T><<<nbv, kBlock, 0, s>>>(n, b, w.q, jacobi ? w.dinv : nullptr, w.r, w.p, st);
  cg_start_finish_k<<<1, kBlock, 0, s>>>(nbv, st);
  check_launch("cg_start");
  struct
  {
    std::int32_t iter, done;
  } h{0, 0};
  const int batch = std::max(1, check_every);
  auto iteration = [&]()
  {
    const int nbs = launch_cg_spmv<T>(A, w.p, w.q, st, sell);
    cg_update_k<T><<<nbv, kBlock, 0, s>>>(n, nbs, x, w.r, w.p, w.q, jacobi ? w.dinv : nullptr,
                                          st);
    cg_direction_k<T><<<nbv, kBlock, 0, s>>>(n, nbv, w.r, jacobi ? w.dinv : nullptr, w.p, st);
  };
  // a batch of iterations captured once and replayed (the kernels read
  // the iteration index from the state and do nothing once it stopped)
  BatchGraph graph;
  if (use_graph)
    graph.capture(s, [&]() { for (int j = 0; j < batch; ++j) iteration(); });
  int k = 0;
  while (true)
  {
    if (graph.exec)
    {
      DFX_HIP(hipGraphLaunch(graph.exec, s));
      k += batch;
    }
    else
      for (int j = 0; j < batch && k < max_it; ++j, ++k)
        iteration();
    check_launch("cg iteration");
    DFX_HIP(hipMemcpyAsync(&h, &st->iter, sizeof(h), hipMemcpyDeviceToHost, s));
    DFX_HIP(hipStreamSynchronize(s));
    if (h.done || k >= max_it)
      break;
  }
  CGState fin;
  DFX_HIP(hipMemcpy(&fin, st, offsetof(CGState, partials), hipMemcpyDeviceToHost));
  *rr0 = fin.rr0;
  *rr = fin.rr;
  *converged = fin.rr0 == 0.0 || fin.rr / fin.rr0 < fin.tol2;
  return fin.iter;
}

/// a . b summed in FP64 through the device buffers partials
/// (kMaxPartials values) and out (one value); returns it on the host.
template <typename T>
double dot_with(double* partials, double* out, const T* a, const T* b, std::int64_t n)
{
  const int nb = partial_blocks(n);
  dot_k<T><<<nb, kBlock, 0, stream()>>>(a, b, n, partials);
  finish_sums_k<<<1, kBlock, 0, stream()>>>(partials, nb, 1, out);
  check_launch("dot");
  double h = 0;
  d2h(&h, out, sizeof(double));
  return h;
}

/// Scratch for host-visible reductions.
struct Reducer
{
  double* partials = nullptr;
  double* out = nullptr;
  Reducer()
  {
    partials = static_cast<double*>(dmalloc(8 * (kMaxPartials + 8)));
    out = partials + kMaxPartials;
  }
  template <typename T>
  double dot(const T* a, const T* b, std::int64_t n)
  {
    return dot_with<T>(partials, out, a, b, n);
  }
};

Reducer& reducer()
{
  static Reducer* r = new Reducer();
  return *r;
}

/// FP64 true residual norm ||b - A x||_2, with y scratch of T; the
/// reduction goes through partials/out when given (else the library's
/// reduction buffers).
template <typename T>
double true_residual(const dfx_csr& A, const T* b, const T* x, T* y,
                     double* partials = nullptr, double* out = nullptr)
{
  hipStream_t s = stream();
  const std::int64_t n = std::int64_t(A.nrows) * A.bs;
  launch_spmv<T>(A, x, y, false);
  axpy_k<T><<<reduce_blocks(n), kBlock, 0, s>>>(y, -1.0, y, b, n);
  check_launch("residual");
  return std::sqrt(partials ? dot_with<T>(partials, out, y, y, n) : reducer().dot(y, y, n));
}

} // namespace dfx

namespace dfx
{
template <typename T>
void cg_solve_t(const dfx_csr& A, const dfx_vec& b, dfx_vec& x,
                       const dfx_cg_opts& o, dfx_cg_result& res)
{
  const std::int64_t n = std::int64_t(A.nrows) * A.bs;
  CGWork<T> w;
  w.alloc(n, o.jacobi != 0);
  StateHolder sh;
  Reducer& red = reducer();
  DFX_HIP(hipMemsetAsync(sh.st, 0, sizeof(CGState), stream()));
  // pattern work of the sliced-ELL copy (as the sparsity pattern, outside
  // the timed window); its values are copied inside
  sell_layout<T>(A);
  DFX_HIP(hipStreamSynchronize(stream()));
  const auto [rpt, nbc] = o.classic ? std::pair<int, int>{0, 0} : coop_shape<T>(A, o.max_it);
  T* p1 = nullptr;
  CoopSync* cs = nullptr;
  if (rpt)
  {
    p1 = static_cast<T*>(dmalloc(sizeof(T) * std::max<std::int64_t>(n, 1)));
    cs = static_cast<CoopSync*>(dmalloc(sizeof(CoopSync)));
  }
  struct Free
  {
    void* p;
    ~Free() { dfree(p); }
  } free_p1{p1}, free_cs{cs};
  DFX_HIP(hipStreamSynchronize(stream()));
  const auto t0 = std::chrono::steady_clock::now();
  bool conv = false;
  double rr0 = 0, rr = 0;
  int it = rpt ? cg_coop_run<T>(A, static_cast<const T*>(b.data), static_cast<T*>(x.data),
                                o.max_it, o.rtol, o.jacobi != 0, rpt, nbc, w.r, w.p, p1,
                                sh.st->partials, cs, &conv, &rr0, &rr)
               : -1;
  res.cooperative = it >= 0 ? 1 : 0;
  if (it < 0)
    it = cg_run<T>(A, static_cast<const T*>(b.data), static_cast<T*>(x.data), o.max_it, o.rtol,
                   o.jacobi != 0, o.check_every, w, sh.st, &conv, &rr0, &rr, false, true,
                   o.graph != 0);
  // the true residual ||b - A x|| / ||b|| that ends the solve
  const double bn = std::sqrt(red.dot(static_cast<const T*>(b.data), static_cast<const T*>(b.data), n));
  const double tr = true_residual<T>(A, static_cast<const T*>(b.data), static_cast<const T*>(x.data),
                                     w.q);
  DFX_HIP(hipStreamSynchronize(stream()));
  const auto t1 = std::chrono::steady_clock::now();
  res.iterations = it;
  res.converged = conv ? 1 : 0;
  res.rnorm0 = std::sqrt(rr0);
  res.rnorm = std::sqrt(rr);
  res.seconds = std::chrono::duration<double>(t1 - t0).count();
  res.true_rel_residual = bn > 0 ? tr / bn : tr;
}

// ---------------------------------------------------------------------
// CG without allocation. The work vectors (r, p, q, diag^-1), the solver
// state and the reduction partials live in one device array, `scratch`,
// of fixed size: kNoAllocMaxN scalar rows at FP64 (32 MiB of vectors plus
// the state). Each device holds its own copy of the array, whose address
// is looked up once per device at library initialisation
// (hipGetSymbolAddress) and used wherever host code hands the array to a
// kernel or a copy. Every call on a device shares that device's array:
// calls that overlap on one device race.
// ---------------------------------------------------------------------

constexpr std::int64_t kNoAllocMaxN = std::int64_t(1) << 20;
constexpr std::size_t kStateBytes = (sizeof(CGState) + 255) / 256 * 256;
constexpr std::size_t kScratchBytes = kStateBytes + 4 * kNoAllocMaxN * sizeof(double);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr; // device -> address of its scratch
} // namespace

/// Looks up the device address of `scratch` on `device` (made current by
/// the caller) once.
void init_noalloc_scratch(int device)
{
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr.count(device))
    return;
  void* p = nullptr;
  DFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p)
    throw Error(DFX_ERR_HIP, "no device address for the CG scratch array");
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
}

unsigned char* noalloc_scratch()
{
  ensure_init();
  int dev = 0;
  DFX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end())
    throw Error(DFX_ERR_HIP, "the CG scratch array was not set up on device "
                                 + std::to_string(dev));
  return it->second;
}

/// cg_solve_t with every device buffer in the scratch array: the CSR SpMV
/// kernels (the sliced-ELL copy would be allocated in the matrix handle)
/// and the reductions through the state's partials.
template <typename T>
void cg_solve_noalloc_t(const dfx_csr& A, const dfx_vec& b, dfx_vec& x, const dfx_cg_opts& o,
                        dfx_cg_result& res)
{
  const std::int64_t n = std::int64_t(A.nrows) * A.bs;
  unsigned char* base = noalloc_scratch();
  CGState* st = reinterpret_cast<CGState*>(base);
  T* v = reinterpret_cast<T*>(base + kStateBytes);
  CGWork<T> w;
  w.own = false;
  w.n = n;
  w.r = v;
  w.p = v + n;
  w.q = v + 2 * n;
  w.dinv = o.jacobi ? v + 3 * n : nullptr;
  // the state starts at zero (the previous call's contents remain)
  DFX_HIP(hipMemsetAsync(st, 0, sizeof(CGState), stream()));
  DFX_HIP(hipStreamSynchronize(stream()));
  const auto t0 = std::chrono::steady_clock::now();
  bool conv = false;
  double rr0 = 0, rr = 0;
  const int it = cg_run<T>(A, static_cast<const T*>(b.data), static_cast<T*>(x.data), o.max_it,
                           o.rtol, o.jacobi != 0, o.check_every, w, st, &conv, &rr0, &rr,
                           false, false);
  DFX_HIP(hipStreamSynchronize(stream()));
  const auto t1 = std::chrono::steady_clock::now();
  res.iterations = it;
  res.converged = conv ? 1 : 0;
  res.rnorm0 = std::sqrt(rr0);
  res.rnorm = std::sqrt(rr);
  res.seconds = std::chrono::duration<double>(t1 - t0).count();
  double* partials = st->partials;
  double* out = st->partials + 2 * kMaxPartials;
  const double bn = std::sqrt(dot_with<T>(partials, out, static_cast<const T*>(b.data),
                                          static_cast<const T*>(b.data), n));
  const double tr = true_residual<T>(A, static_cast<const T*>(b.data),
                                     static_cast<const T*>(x.data), w.q, partials, out);
  res.true_rel_residual = bn > 0 ? tr / bn : tr;
}

/// x += d (FP64 update from an FP32 correction)
__global__ void __launch_bounds__(kBlock) add_correction_k(double* x, const float* d, std::int64_t n)
{
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    x[i] += static_cast<double>(d[i]);
}

/// r = b - y (FP64), r32 = (float) r, partial sums of ||r||^2
__global__ void __launch_bounds__(kBlock)
    residual_split_k(const double* b, const double* y, float* r32, std::int64_t n,
                     double* partials)
{
  double s[1] = {0.0};
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
  {
    const double ri = b[i] - y[i];
    r32[i] = static_cast<float>(ri);
    s[0] += ri * ri;
  }
  block_partials<1>(s, partials);
}

} // namespace dfx

namespace dfx
{
template <typename T>
int autotune_lanes(dfx_csr& A, int reps, float* best_ms)
{
  const std::int64_t nx = std::int64_t(A.ncols) * A.bs, ny = std::int64_t(A.nrows) * A.bs;
  T* x = static_cast<T*>(dmalloc(sizeof(T) * std::max<std::int64_t>(nx, 1)));
  T* y = static_cast<T*>(dmalloc(sizeof(T) * std::max<std::int64_t>(ny, 1)));
  fill_k<T><<<reduce_blocks(nx), kBlock, 0, stream()>>>(x, nx, T(1));
  hipEvent_t e0, e1;
  DFX_HIP(hipEventCreate(&e0));
  DFX_HIP(hipEventCreate(&e1));
  int best = 0;
  float bms = 1e30f;
  // block matrices: 1 stream, 2 one thread per (block row, component)
  const int gmax = A.bs == 1 ? 64 : 2;
  for (int G = 1; G <= gmax; G *= 2)
  {
    if (G == 1 && A.nblk == 0)
      continue;
    A.lanes = G;
    launch_spmv<T>(A, x, y, false); // warm-up
    DFX_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r)
      launch_spmv<T>(A, x, y, false);
    DFX_HIP(hipEventRecord(e1, stream()));
    DFX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    DFX_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < bms)
    {
      bms = ms;
      best = G;
    }
  }
  A.lanes = best;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(x);
  dfree(y);
  *best_ms = bms / reps;
  return best;
}

} // namespace dfx

using namespace dfx;

/// A recorded sequence of library launches, replayed as one graph.
struct dfx_graph
{
  hipGraphExec_t exec = nullptr;
};

extern "C" {

const char* dfx_last_error(void) { return dfx::g_last_error.c_str(); }

const char* dfx_version(void) { return "dolfinx_amd 0.1 (gfx950)"; }

dfx_status dfx_init(void)
{
  return guard([] { ensure_init(); });
}

dfx_status dfx_synchronize(void)
{
  return guard([] { DFX_HIP(hipStreamSynchronize(stream())); });
}

void* dfx_stream(void)
{
  void* s = nullptr;
  guard([&] { s = static_cast<void*>(stream()); });
  return s;
}

dfx_status dfx_event_record(int slot)
{
  return guard(
      [&]
      {
        require(slot >= 0 && slot < 64, "event slot out of range");
        DFX_HIP(hipEventRecord(ctx().events[slot], stream()));
      });
}

dfx_status dfx_event_elapsed(int a, int b, float* ms)
{
  return guard(
      [&]
      {
        require(a >= 0 && a < 64 && b >= 0 && b < 64 && ms, "bad event arguments");
        ensure_init();
        DFX_HIP(hipEventSynchronize(ctx().events[b]));
        DFX_HIP(hipEventElapsedTime(ms, ctx().events[a], ctx().events[b]));
      });
}

// ---- MatrixCSR ----

dfx_status dfx_csr_create(const int64_t* row_ptr, const int32_t* cols, int32_t nrows,
                          int32_t ncols, int bs, int dtype, dfx_csr** out)
{
  return guard(
      [&]
      {
        require(row_ptr && cols && out, "null argument");
        require(nrows >= 0 && ncols >= 0, "negative size");
        require(bs >= 1 && bs <= 3, "block size must be 1..3");
        require(dtype == DFX_F32 || dtype == DFX_F64, "bad dtype");
        const std::int64_t nnz = row_ptr[nrows];
        require(row_ptr[0] == 0 && nnz < INT32_MAX, "row_ptr out of range");
        std::vector<std::int32_t> rp32(std::size_t(nrows) + 1);
        std::int32_t maxrow = 0;
        for (std::int32_t i = 0; i <= nrows; ++i)
        {
          rp32[i] = static_cast<std::int32_t>(row_ptr[i]);
          if (i > 0)
          {
            require(row_ptr[i] >= row_ptr[i - 1], "row_ptr not monotone");
            maxrow = std::max<std::int32_t>(maxrow, rp32[i] - rp32[i - 1]);
          }
        }
        // every row's columns strictly increasing (the finalised rows of
        // SparsityPattern::finalize): the device kernels find entries by
        // binary search in a row
        for (std::int32_t i = 0; i < nrows; ++i)
          for (std::int64_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
          {
            require(cols[j] >= 0 && cols[j] < ncols, "column index out of range");
            require(j == row_ptr[i] || cols[j] > cols[j - 1],
                    "columns of row " + std::to_string(i)
                        + " are not strictly increasing (sorted, no duplicates)");
          }
        auto A = std::make_unique<dfx_csr>();
        A->nrows = nrows;
        A->ncols = ncols;
        A->bs = bs;
        A->dtype = dtype;
        A->nnz = nnz;
        A->max_row = maxrow;
        A->row_ptr = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * (nrows + 1)));
        A->cols = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max<std::int64_t>(nnz, 1)));
        A->diag = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max(nrows, 1)));
        const std::size_t vb = dtype_size(dtype) * std::size_t(nnz) * bs * bs;
        A->values = dmalloc(std::max<std::size_t>(vb, 8));
        h2d(A->row_ptr, rp32.data(), sizeof(std::int32_t) * (nrows + 1));
        const std::int32_t chunk
            = (dtype == DFX_F64 ? stream_chunk<double>() : stream_chunk<float>()) / bs;
        if (bs == 1 || (bs <= 3 && maxrow <= chunk))
        {
          // row blocks of the stream SpMV: greedy, at most kBlock / bs rows
          // and entries fitting the LDS chunk of the matrix precision (a
          // longer scalar row gets a block of its own)
          const std::int32_t max_rows = kBlock / bs;
          // (first row, first entry) of each block, then (nrows, nnz)
          std::vector<std::int32_t> rb{0, 0};
          std::int32_t r = 0;
          while (r < nrows)
          {
            std::int32_t e = r + 1;
            while (e < nrows && e - r < max_rows && rp32[e + 1] - rp32[r] <= chunk)
              ++e;
            rb.push_back(e);
            rb.push_back(rp32[e]);
            r = e;
          }
          A->nblk = static_cast<std::int32_t>(rb.size() / 2) - 1;
          A->rblk = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * rb.size()));
          h2d(A->rblk, rb.data(), sizeof(std::int32_t) * rb.size());
        }
        h2d(A->cols, cols, sizeof(std::int32_t) * nnz);
        DFX_HIP(hipMemsetAsync(A->values, 0, vb, stream()));
        if (nrows > 0)
          diag_pos_k<<<(nrows + kBlock - 1) / kBlock, kBlock, 0, stream()>>>(nrows, A->row_ptr, A->cols, A->diag);
        check_launch("csr_create");
        DFX_HIP(hipStreamSynchronize(stream()));
        prime_stream_grids();
        A->pattern_id = next_pattern_id();
        *out = A.release();
      });
}

dfx_status dfx_csr_create_shared(const dfx_csr* P, int dtype, dfx_csr** out)
{
  return guard(
      [&]
      {
        require(P && out, "null argument");
        require(dtype == DFX_F32 || dtype == DFX_F64, "bad dtype");
        require(!P->pattern_owner, "share the pattern of its owner");
        auto A = std::make_unique<dfx_csr>();
        *A = *P;
        A->dtype = dtype;
        A->pattern_owner = P;
        A->lanes = 0;
        A->sell_vals = nullptr;
        A->sell_cols = nullptr;
        A->sell_off = nullptr;
        A->sell_base = nullptr;
        A->nslices = 0;
        A->sell_pattern = 0;
        A->sell_use = false;
        if (dtype != P->dtype)
        {
          // the stream SpMV row blocks depend on the value size
          A->rblk = nullptr;
          A->nblk = 0;
        }
        const std::size_t vb = dtype_size(dtype) * std::size_t(A->nnz) * A->bs * A->bs;
        A->values = dmalloc(std::max<std::size_t>(vb, 8));
        // zeros stored when first read (an assembly overwrites them all)
        A->zeroed = true;
        A->zero_pending = true;
        *out = A.release();
      });
}

dfx_status dfx_csr_destroy(dfx_csr* A)
{
  if (A)
  {
    if (!A->pattern_owner)
    {
      dfree(A->row_ptr);
      dfree(A->rblk);
      dfree(A->cols);
      dfree(A->diag);
    }
    dfree(A->values);
    dfree(A->sell_vals);
    dfree(A->sell_cols);
    dfree(A->sell_off);
    dfree(A->sell_base);
    delete A;
  }
  return DFX_OK;
}

dfx_status dfx_csr_sizes(const dfx_csr* A, int32_t* nrows, int64_t* nnz, int32_t* bs,
                         int32_t* dtype)
{
  return guard(
      [&]
      {
        require(A, "null matrix");
        if (nrows)
          *nrows = A->nrows;
        if (nnz)
          *nnz = A->nnz;
        if (bs)
          *bs = A->bs;
        if (dtype)
          *dtype = A->dtype;
      });
}

} // extern "C"

namespace dfx
{
namespace
{
/// MatrixCSR::add for staged dense blocks (la/MatrixCSR.h:157-174 via
/// impl::insert_blocked_csr, la/matrix_csr_impl.h:67-109), one thread per
/// (row, component): the row's contributions (block b, local row i) come in
/// the order they were staged (off/item, built on the host by a counting
/// sort over rows), and each adds its n column blocks at the slots found by
/// binary search in the row; a column outside the pattern sets *err.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    add_blocks_k(std::int32_t nrows, int bs, int m, int n, const std::int32_t* __restrict__ off,
                 const std::int64_t* __restrict__ item, const std::int32_t* __restrict__ cols_in,
                 const T* __restrict__ vals_in, const std::int32_t* __restrict__ rp,
                 const std::int32_t* __restrict__ ci, T* __restrict__ values, int* err)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nrows) * bs)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / bs);
  const int cr = static_cast<int>(t - std::int64_t(r) * bs);
  const std::int32_t lo = rp[r], hi = rp[r + 1];
  const std::int64_t ld = std::int64_t(n) * bs; // values per row of a block
  for (std::int32_t p = off[r]; p < off[r + 1]; ++p)
  {
    const std::int64_t bi = item[p]; // b * m + i
    const std::int64_t b = bi / m;
    const T* v = vals_in + (bi * bs + cr) * ld;
    for (int j = 0; j < n; ++j)
    {
      const std::int32_t c = cols_in[b * n + j];
      std::int32_t a = lo, z = hi;
      while (a < z)
      {
        const std::int32_t mid = (a + z) >> 1;
        if (ci[mid] < c)
          a = mid + 1;
        else
          z = mid;
      }
      if (a >= hi || ci[a] != c)
      {
        *err = 1;
        continue;
      }
      T* out = values + (std::int64_t(a) * bs + cr) * bs;
      for (int cc = 0; cc < bs; ++cc)
        out[cc] += v[std::int64_t(j) * bs + cc];
    }
  }
}
/// b[bs r + comp] += the staged block vectors at row r, in staging order,
/// one thread per (row, component) (off/item as for add_blocks_k).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    add_vec_blocks_k(std::int32_t nrows, int bs, const std::int32_t* __restrict__ off,
                     const std::int64_t* __restrict__ item, const T* __restrict__ vals_in,
                     T* __restrict__ b)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nrows) * bs)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / bs);
  const int cr = static_cast<int>(t - std::int64_t(r) * bs);
  T s = b[t];
  for (std::int32_t p = off[r]; p < off[r + 1]; ++p)
    s += vals_in[item[p] * bs + cr];
  b[t] = s;
}

/// The staged rows grouped by matrix row in staging order: off (nrows + 1)
/// and the staged row index of each (a counting sort on the host).
void group_rows(std::int32_t nrows, std::int64_t nitems, const std::int32_t* rows,
                std::vector<std::int32_t>& off, std::vector<std::int64_t>& item)
{
  off.assign(std::size_t(nrows) + 1, 0);
  for (std::int64_t q = 0; q < nitems; ++q)
  {
    require(rows[q] >= 0 && rows[q] < nrows, "row index out of range");
    ++off[std::size_t(rows[q]) + 1];
  }
  for (std::int32_t r = 0; r < nrows; ++r)
    off[std::size_t(r) + 1] += off[r];
  item.resize(nitems);
  std::vector<std::int32_t> pos(off.begin(), off.end() - 1);
  for (std::int64_t q = 0; q < nitems; ++q)
    item[pos[rows[q]]++] = q;
}
} // namespace

void csr_add_blocks(dfx_csr& A, std::int64_t nblocks, int m, int n, const std::int32_t* rows,
                    const std::int32_t* cols, const void* values)
{
  require(nblocks >= 0 && m > 0 && n > 0, "bad block shape");
  if (nblocks == 0)
    return;
  const std::int64_t nitems = nblocks * m;
  require(nitems < (std::int64_t(1) << 40), "too many blocks");
  for (std::int64_t q = 0; q < nblocks * n; ++q)
    require(cols[q] >= 0 && cols[q] < A.ncols, "column index out of range");
  std::vector<std::int32_t> off;
  std::vector<std::int64_t> item;
  group_rows(A.nrows, nitems, rows, off, item);
  const std::size_t ts = dtype_size(A.dtype);
  const std::size_t nval = std::size_t(nblocks) * m * n * A.bs * A.bs;
  std::int32_t* d_off = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * off.size()));
  std::int64_t* d_item = static_cast<std::int64_t*>(dmalloc(sizeof(std::int64_t) * item.size()));
  std::int32_t* d_cols = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * nblocks * n));
  void* d_vals = dmalloc(ts * nval);
  int* err = static_cast<int*>(dmalloc(sizeof(int)));
  struct Free
  {
    std::vector<void*> p;
    ~Free()
    {
      for (void* q : p)
        dfree(q);
    }
  } fr{{d_off, d_item, d_cols, d_vals, err}};
  h2d(d_off, off.data(), sizeof(std::int32_t) * off.size());
  h2d(d_item, item.data(), sizeof(std::int64_t) * item.size());
  h2d(d_cols, cols, sizeof(std::int32_t) * nblocks * n);
  h2d(d_vals, values, ts * nval);
  materialize_zero(A);
  fill_k<int><<<1, kBlock, 0, stream()>>>(err, 1, 0);
  const std::int64_t nt = std::int64_t(A.nrows) * A.bs;
  const int nb = static_cast<int>((nt + kBlock - 1) / kBlock);
  if (nb > 0)
  {
    if (A.dtype == DFX_F64)
      add_blocks_k<double><<<nb, kBlock, 0, stream()>>>(
          A.nrows, A.bs, m, n, d_off, d_item, d_cols, static_cast<const double*>(d_vals),
          A.row_ptr, A.cols, static_cast<double*>(A.values), err);
    else
      add_blocks_k<float><<<nb, kBlock, 0, stream()>>>(
          A.nrows, A.bs, m, n, d_off, d_item, d_cols, static_cast<const float*>(d_vals),
          A.row_ptr, A.cols, static_cast<float*>(A.values), err);
  }
  check_launch("add_blocks");
  A.zeroed = false;
  int h = 0;
  d2h(&h, err, sizeof(int));
  if (h)
    throw Error(DFX_ERR_PATTERN, "Entry not in sparsity");
}

void vec_add_blocks(dfx_vec& b, std::int64_t nblocks, int m, int bs, const std::int32_t* rows,
                    const void* values)
{
  require(nblocks >= 0 && m > 0 && bs > 0, "bad block shape");
  if (nblocks == 0)
    return;
  const std::int32_t nrows = static_cast<std::int32_t>(b.n / bs);
  require(std::int64_t(nrows) * bs == b.n, "vector size is not a multiple of the block size");
  std::vector<std::int32_t> off;
  std::vector<std::int64_t> item;
  group_rows(nrows, nblocks * m, rows, off, item);
  const std::size_t ts = dtype_size(b.dtype);
  const std::size_t nval = std::size_t(nblocks) * m * bs;
  std::int32_t* d_off = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * off.size()));
  std::int64_t* d_item = static_cast<std::int64_t*>(dmalloc(sizeof(std::int64_t) * item.size()));
  void* d_vals = dmalloc(ts * nval);
  struct Free
  {
    std::vector<void*> p;
    ~Free()
    {
      for (void* q : p)
        dfree(q);
    }
  } fr{{d_off, d_item, d_vals}};
  h2d(d_off, off.data(), sizeof(std::int32_t) * off.size());
  h2d(d_item, item.data(), sizeof(std::int64_t) * item.size());
  h2d(d_vals, values, ts * nval);
  const int nb = static_cast<int>((b.n + kBlock - 1) / kBlock);
  if (nb > 0)
  {
    if (b.dtype == DFX_F64)
      add_vec_blocks_k<double><<<nb, kBlock, 0, stream()>>>(
          nrows, bs, d_off, d_item, static_cast<const double*>(d_vals), static_cast<double*>(b.data));
    else
      add_vec_blocks_k<float><<<nb, kBlock, 0, stream()>>>(
          nrows, bs, d_off, d_item, static_cast<const float*>(d_vals), static_cast<float*>(b.data));
  }
  check_launch("add_vec_blocks");
  DFX_HIP(hipStreamSynchronize(stream()));
}
} // namespace dfx

extern "C" {

dfx_status dfx_csr_add_blocks(dfx_csr* A, int64_t nblocks, int32_t m, int32_t n,
                              const int32_t* rows, const int32_t* cols, const void* values)
{
  return guard(
      [&]
      {
        require(A && (nblocks == 0 || (rows && cols && values)), "null argument");
        csr_add_blocks(*A, nblocks, m, n, rows, cols, values);
      });
}

dfx_status dfx_csr_zero(dfx_csr* A)
{
  return guard(
      [&]
      {
        require(A, "null matrix");
        // MatrixCSR::set(0): the zeros are stored when the values are next
        // read, unless an assembly overwrites every entry first
        A->zeroed = true;
        A->zero_pending = true;
      });
}

dfx_status dfx_csr_get_values(const dfx_csr* A, void* values)
{
  return guard(
      [&]
      {
        require(A && values, "null argument");
        materialize_zero(*A);
        d2h(values, A->values, dtype_size(A->dtype) * A->nnz * A->bs * A->bs);
      });
}

dfx_status dfx_csr_set_values(dfx_csr* A, const void* values)
{
  return guard(
      [&]
      {
        require(A && values, "null argument");
        h2d(A->values, values, dtype_size(A->dtype) * A->nnz * A->bs * A->bs);
        A->zeroed = false;
        A->zero_pending = false;
      });
}

dfx_status dfx_csr_get_pattern(const dfx_csr* A, int64_t* row_ptr, int32_t* cols)
{
  return guard(
      [&]
      {
        require(A, "null matrix");
        if (row_ptr)
        {
          std::vector<std::int32_t> rp(std::size_t(A->nrows) + 1);
          d2h(rp.data(), A->row_ptr, sizeof(std::int32_t) * rp.size());
          for (std::size_t i = 0; i < rp.size(); ++i)
            row_ptr[i] = rp[i];
        }
        if (cols)
          d2h(cols, A->cols, sizeof(std::int32_t) * A->nnz);
      });
}

// ---- Vector ----

dfx_status dfx_vec_create(int32_t n, int dtype, dfx_vec** out)
{
  return guard(
      [&]
      {
        require(out && n >= 0, "bad argument");
        require(dtype == DFX_F32 || dtype == DFX_F64, "bad dtype");
        auto v = std::make_unique<dfx_vec>();
        v->n = n;
        v->dtype = dtype;
        v->data = dmalloc(dtype_size(dtype) * std::max(n, 1));
        DFX_HIP(hipMemsetAsync(v->data, 0, dtype_size(dtype) * std::max(n, 1), stream()));
        DFX_HIP(hipStreamSynchronize(stream()));
        *out = v.release();
      });
}

dfx_status dfx_vec_destroy(dfx_vec* v)
{
  if (v)
  {
    dfree(v->data);
    delete v;
  }
  return DFX_OK;
}

dfx_status dfx_vec_size(const dfx_vec* v, int32_t* n, int32_t* dtype)
{
  return guard(
      [&]
      {
        require(v, "null vector");
        if (n)
          *n = v->n;
        if (dtype)
          *dtype = v->dtype;
      });
}

dfx_status dfx_vec_upload(dfx_vec* v, const void* host)
{
  return guard(
      [&]
      {
        require(v && host, "null argument");
        h2d(v->data, host, dtype_size(v->dtype) * v->n);
      });
}

dfx_status dfx_vec_download(const dfx_vec* v, void* host)
{
  return guard(
      [&]
      {
        require(v && host, "null argument");
        d2h(host, v->data, dtype_size(v->dtype) * v->n);
      });
}

dfx_status dfx_vec_set(dfx_vec* v, double value)
{
  return guard(
      [&]
      {
        require(v, "null vector");
        const int nb = reduce_blocks(v->n);
        if (v->dtype == DFX_F64)
          fill_k<double><<<nb, kBlock, 0, stream()>>>(static_cast<double*>(v->data), v->n, value);
        else
          fill_k<float><<<nb, kBlock, 0, stream()>>>(static_cast<float*>(v->data), v->n,
                                                      static_cast<float>(value));
        check_launch("vec_set");
      });
}

dfx_status dfx_vec_copy(dfx_vec* dst, const dfx_vec* src)
{
  return guard(
      [&]
      {
        require(dst && src && dst->n == src->n, "size mismatch");
        const int nb = reduce_blocks(src->n);
        hipStream_t s = stream();
        if (dst->dtype == DFX_F64 && src->dtype == DFX_F64)
          convert_k<double, double><<<nb, kBlock, 0, s>>>(static_cast<double*>(dst->data), static_cast<const double*>(src->data), src->n);
        else if (dst->dtype == DFX_F64)
          convert_k<double, float><<<nb, kBlock, 0, s>>>(static_cast<double*>(dst->data), static_cast<const float*>(src->data), src->n);
        else if (src->dtype == DFX_F64)
          convert_k<float, double><<<nb, kBlock, 0, s>>>(static_cast<float*>(dst->data), static_cast<const double*>(src->data), src->n);
        else
          convert_k<float, float><<<nb, kBlock, 0, s>>>(static_cast<float*>(dst->data), static_cast<const float*>(src->data), src->n);
        check_launch("vec_copy");
      });
}

dfx_status dfx_spmv(const dfx_csr* A, const dfx_vec* x, dfx_vec* y)
{
  return guard(
      [&]
      {
        require(A && x && y, "null argument");
        require(x->dtype == A->dtype && y->dtype == A->dtype, "dtype mismatch");
        require(x->n == A->ncols * A->bs && y->n == A->nrows * A->bs, "size mismatch");
        if (A->dtype == DFX_F64)
          launch_spmv<double>(*A, static_cast<const double*>(x->data), static_cast<double*>(y->data), true);
        else
          launch_spmv<float>(*A, static_cast<const float*>(x->data), static_cast<float*>(y->data), true);
      });
}

dfx_status dfx_csr_autotune(dfx_csr* A, int32_t* lanes, float* ms)
{
  return guard(
      [&]
      {
        require(A, "null matrix");
        materialize_zero(*A);
        float t = 0;
        int g = 0;
        if (A->dtype == DFX_F64)
          g = autotune_lanes<double>(*A, 10, &t);
        else
          g = autotune_lanes<float>(*A, 10, &t);
        if (lanes)
          *lanes = g;
        if (ms)
          *ms = t;
      });
}

dfx_status dfx_csr_set_lanes(dfx_csr* A, int32_t lanes)
{
  return guard(
      [&]
      {
        require(A, "null matrix");
        require(lanes == -1 || lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4
                    || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64,
                "lanes must be -1 (sliced-ELL snapshot), 0 (default), 1 (stream) or a power "
                "of two in [2, 64]");
        if (lanes == -1)
        {
          require(A->bs == 1, "the sliced-ELL copy is for block size 1");
          materialize_zero(*A);
          const int saved = A->lanes;
          A->lanes = 1;
          const bool ok = A->dtype == DFX_F64 ? refresh_sell<double>(*A) : refresh_sell<float>(*A);
          A->lanes = saved;
          require(ok, "no sliced-ELL copy for this matrix (padding above 15 %)");
        }
        A->lanes = lanes;
      });
}

dfx_status dfx_inner_product(const dfx_vec* a, const dfx_vec* b, double* out)
{
  return guard(
      [&]
      {
        require(a && b && out && a->n == b->n && a->dtype == b->dtype, "mismatch");
        if (a->dtype == DFX_F64)
          *out = reducer().dot(static_cast<const double*>(a->data), static_cast<const double*>(b->data), a->n);
        else
          *out = reducer().dot(static_cast<const float*>(a->data), static_cast<const float*>(b->data), a->n);
      });
}

dfx_status dfx_squared_norm(const dfx_vec* a, double* out)
{
  return dfx_inner_product(a, a, out);
}

dfx_status dfx_axpy(dfx_vec* r, double alpha, const dfx_vec* x, const dfx_vec* y)
{
  return guard(
      [&]
      {
        require(r && x && y && r->n == x->n && x->n == y->n, "size mismatch");
        require(r->dtype == x->dtype && x->dtype == y->dtype, "dtype mismatch");
        const int nb = reduce_blocks(r->n);
        if (r->dtype == DFX_F64)
          axpy_k<double><<<nb, kBlock, 0, stream()>>>(static_cast<double*>(r->data), alpha, static_cast<const double*>(x->data), static_cast<const double*>(y->data), r->n);
        else
          axpy_k<float><<<nb, kBlock, 0, stream()>>>(static_cast<float*>(r->data), alpha, static_cast<const float*>(x->data), static_cast<const float*>(y->data), r->n);
        check_launch("axpy");
      });
}

// ---- Solvers ----

dfx_status dfx_cg_solve(const dfx_csr* A, const dfx_vec* b, dfx_vec* x,
                        const dfx_cg_opts* opts, dfx_cg_result* res)
{
  return guard(
      [&]
      {
        require(A && b && x && opts && res, "null argument");
        require(b->dtype == A->dtype && x->dtype == A->dtype, "dtype mismatch");
        require(A->nrows == A->ncols, "matrix not square");
        require(b->n == A->nrows * A->bs && x->n == b->n, "size mismatch");
        require(opts->max_it >= 0 && opts->rtol >= 0, "bad options");
        materialize_zero(*A);
        *res = dfx_cg_result{};
        if (A->dtype == DFX_F64)
          cg_solve_t<double>(*A, *b, *x, *opts, *res);
        else
          cg_solve_t<float>(*A, *b, *x, *opts, *res);
        if (!res->converged)
          throw Error(DFX_ERR_NOCONV, "CG exceeded max iterations ("
                                          + std::to_string(opts->max_it)
                                          + "). Relative residual="
                                          + std::to_string(res->rnorm / std::max(res->rnorm0, 1e-300)));
      });
}

dfx_status dfx_cg_solve_noalloc(const dfx_csr* A, const dfx_vec* b, dfx_vec* x,
                                const dfx_cg_opts* opts, dfx_cg_result* res)
{
  return guard(
      [&]
      {
        require(A && b && x && opts && res, "null argument");
        require(b->dtype == A->dtype && x->dtype == A->dtype, "dtype mismatch");
        require(A->nrows == A->ncols, "matrix not square");
        require(b->n == A->nrows * A->bs && x->n == b->n, "size mismatch");
        require(opts->max_it >= 0 && opts->rtol >= 0, "bad options");
        require(std::int64_t(A->nrows) * A->bs <= kNoAllocMaxN,
                "dfx_cg_solve_noalloc: " + std::to_string(std::int64_t(A->nrows) * A->bs)
                    + " rows exceed the scratch maximum of " + std::to_string(kNoAllocMaxN)
                    + " (use dfx_cg_solve)");
        materialize_zero(*A);
        *res = dfx_cg_result{};
        if (A->dtype == DFX_F64)
          cg_solve_noalloc_t<double>(*A, *b, *x, *opts, *res);
        else
          cg_solve_noalloc_t<float>(*A, *b, *x, *opts, *res);
        if (!res->converged)
          throw Error(DFX_ERR_NOCONV, "CG exceeded max iterations ("
                                          + std::to_string(opts->max_it)
                                          + "). Relative residual="
                                          + std::to_string(res->rnorm / std::max(res->rnorm0, 1e-300)));
      });
}

dfx_status dfx_mixed_solve(const dfx_csr* A64, const dfx_csr* A32, const dfx_vec* b,
                           dfx_vec* x, const dfx_mixed_opts* o, dfx_mixed_result* res)
{
  return guard(
      [&]
      {
        require(A64 && A32 && b && x && o && res, "null argument");
        require(A64->dtype == DFX_F64 && A32->dtype == DFX_F32, "need FP64 and FP32 matrices");
        require(b->dtype == DFX_F64 && x->dtype == DFX_F64, "b and x must be FP64");
        require(A64->nrows == A32->nrows && A64->nnz == A32->nnz && A64->bs == A32->bs,
                "matrices differ in shape");
        require(A64->nrows == A64->ncols, "matrix not square");
        require(b->n == A64->nrows * A64->bs && x->n == b->n, "size mismatch");
        materialize_zero(*A64);
        materialize_zero(*A32);
        // pattern work of the FP32 sliced-ELL copy outside the timed window
        // (as for the FP64 solve); its values are copied inside
        sell_layout<float>(*A32);
        *res = dfx_mixed_result{};
        hipStream_t s = stream();
        const std::int64_t n = b->n;
        const double* bd = static_cast<const double*>(b->data);
        double* xd = static_cast<double*>(x->data);
        CGWork<float> w;
        w.alloc(n, true);
        StateHolder sh;
        CGState* st = sh.st;
        float* d32 = static_cast<float*>(dmalloc(sizeof(float) * n));
        double* y = static_cast<double*>(dmalloc(sizeof(double) * n));
        Reducer& red = reducer();
        DFX_HIP(hipMemsetAsync(st, 0, sizeof(CGState), s));
        DFX_HIP(hipStreamSynchronize(s));
        const int nb = reduce_blocks(n);
        const int nbv = partial_blocks(n);
        const auto t0 = std::chrono::steady_clock::now();
        const bool sell = refresh_sell<float>(*A32);
        const double bn2 = red.dot(bd, bd, n);
        const double tol2 = o->rtol * o->rtol * bn2;
        // r64 = b - A64 x, r32 = (float) r64 (written into the CG residual)
        auto true_residual_split = [&]() -> double
        {
          launch_spmv<double>(*A64, xd, y, false);
          const int nbr = partial_blocks(n);
          residual_split_k<<<nbr, kBlock, 0, s>>>(bd, y, w.r, n, red.partials);
          finish_sums_k<<<1, kBlock, 0, s>>>(red.partials, nbr, 1, red.out);
          check_launch("residual");
          double h = 0;
          d2h(&h, red.out, sizeof(double));
          return h;
        };
        int outer = 0, inner = 0, inner_base = 0;
        double rr = true_residual_split();
        // FP64 residual before the current refinement step (stall test)
        double rr_prev = rr;
        // an inner solve stops when its residual fell by inner_rtol, or
        // earlier once it meets half the outer target (the last refinement
        // is not solved past what the outer test needs)
        auto inner_tol2 = [&](double rr_outer)
        {
          const double fin = 0.25 * tol2 / std::max(rr_outer, 1e-300);
          return std::max(o->inner_rtol * o->inner_rtol, fin);
        };
        bool ok = rr <= tol2;
        const bool restart = o->restart != 0;
        if (!ok)
        {
          inv_diag_k<float><<<static_cast<int>((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
              A32->nrows, A32->bs, A32->diag, static_cast<const float*>(A32->values), w.dinv);
          fill_k<float><<<nb, kBlock, 0, s>>>(d32, n, 0.0f);
          fill_k<float><<<nb, kBlock, 0, s>>>(w.q, n, 0.0f);
          CGState init{};
          init.tol2 = o->inner_rtol * o->inner_rtol;
          init.max_it = o->max_inner;
          DFX_HIP(hipMemcpyAsync(st, &init, offsetof(CGState, partials), hipMemcpyHostToDevice, s));
          // r = r32 - A32 * 0, p = z: reuse cg_start with y = 0
          cg_start_k<float><<<nbv, kBlock, 0, s>>>(n, w.r, w.q, w.dinv, w.r, w.p, st);
          cg_start_finish_k<<<1, kBlock, 0, s>>>(nbv, st);
          cg_resume_k<<<1, 1, 0, s>>>(st, rr, inner_tol2(rr), -1);
          check_launch("mixed start");
        }
        int k = 0;
        struct
        {
          std::int32_t iter, done;
        } h{0, 0};
        const int batch = std::max(1, o->check_every);
        auto iteration = [&]()
        {
          const int nbs = launch_cg_spmv<float>(*A32, w.p, w.q, st, sell);
          cg_update_k<float><<<nbv, kBlock, 0, s>>>(n, nbs, d32, w.r, w.p, w.q, w.dinv, st);
          cg_direction_k<float><<<nbv, kBlock, 0, s>>>(n, nbv, w.r, w.dinv, w.p, st);
        };
        // inner iterations replayed from a captured batch (the kernels read
        // the iteration index from the state)
        BatchGraph graph;
        if (o->graph && !ok)
          graph.capture(s, [&]() { for (int j = 0; j < batch; ++j) iteration(); });
        while (!ok && k < o->max_inner)
        {
          if (graph.exec)
          {
            DFX_HIP(hipGraphLaunch(graph.exec, s));
            k += batch;
          }
          else
            for (int j = 0; j < batch && k < o->max_inner; ++j, ++k)
              iteration();
          check_launch("mixed iteration");
          DFX_HIP(hipMemcpyAsync(&h, &st->iter, sizeof(h), hipMemcpyDeviceToHost, s));
          DFX_HIP(hipStreamSynchronize(s));
          if (!h.done)
            continue;
          // FP32 residual dropped by inner_rtol (or the inner budget ran out):
          // FP64 correction and true residual
          inner = inner_base + h.iter;
          add_correction_k<<<nb, kBlock, 0, s>>>(xd, d32, n);
          fill_k<float><<<nb, kBlock, 0, s>>>(d32, n, 0.0f);
          rr = true_residual_split();
          ++outer;
          if (rr <= tol2)
          {
            ok = true;
            break;
          }
          // stalled: the step did not halve the FP64 residual norm, or the
          // budget ran out
          if (rr > 0.25 * rr_prev || outer >= o->max_outer || h.iter >= o->max_inner)
            break;
          rr_prev = rr;
          if (restart)
          {
            // classic refinement: a new inner CG from d = 0 on the new r
            inner_base += h.iter;
            fill_k<float><<<nb, kBlock, 0, s>>>(w.q, n, 0.0f);
            cg_start_k<float><<<nbv, kBlock, 0, s>>>(n, w.r, w.q, w.dinv, w.r, w.p, st);
            cg_start_finish_k<<<1, kBlock, 0, s>>>(nbv, st);
            cg_resume_k<<<1, 1, 0, s>>>(st, rr, inner_tol2(rr), -1);
            k = 0;
          }
          else
          {
            // reliable update: keep the search direction; beta is formed
            // from the replaced residual, rz_new / rz_old
            const int kk = h.iter - 1;
            cg_resume_k<<<1, 1, 0, s>>>(st, rr, inner_tol2(rr), kk);
            cg_rz_partials_k<float><<<nbv, kBlock, 0, s>>>(n, w.r, w.dinv, st);
            cg_direction_k<float><<<nbv, kBlock, 0, s>>>(n, nbv, w.r, w.dinv, w.p, st);
            k = kk + 1;
          }
          check_launch("mixed refinement");
        }
        // FP64 fallback after a stall: Jacobi-PCG on A64 from the current
        // x, stopping when its residual norm reaches rtol ||b||
        int fb_it = 0;
        if (!ok && o->fallback && o->max_inner > 0)
        {
          CGWork<double> w64;
          w64.alloc(n, true);
          bool conv = false;
          double rr0f = 0, rrf = 0;
          const double rt = std::sqrt(tol2 / std::max(rr, 1e-300));
          fb_it = cg_run<double>(*A64, bd, xd, o->max_inner, rt, true, batch, w64, st, &conv,
                                 &rr0f, &rrf);
          rr = true_residual_split();
          ok = rr <= tol2;
        }
        DFX_HIP(hipStreamSynchronize(s));
        const auto t1 = std::chrono::steady_clock::now();
        dfree(d32);
        dfree(y);
        res->outer = outer;
        res->inner_total = std::max(inner, inner_base + h.iter);
        res->fallback_iterations = fb_it;
        res->converged = ok ? 1 : 0;
        res->true_rel_residual = bn2 > 0 ? std::sqrt(rr / bn2) : std::sqrt(rr);
        res->seconds = std::chrono::duration<double>(t1 - t0).count();
        if (!ok)
          throw Error(DFX_ERR_NOCONV, "mixed-precision refinement did not converge");
      });
}

// ---------------------------------------------------------------------
// Graph capture of library calls: the library stream records instead of
// running between dfx_graph_begin and dfx_graph_end (thread-local mode, so
// other threads keep launching), and the recorded work replays as one
// graph launch.
// ---------------------------------------------------------------------

dfx_status dfx_graph_begin(void)
{
  return guard(
      [&]
      {
        ensure_init();
        require(!capturing(), "a graph capture is already open on this thread");
        DFX_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
        g_capturing = true;
        g_capture_error.clear();
      });
}

dfx_status dfx_graph_end(dfx_graph** out)
{
  return guard(
      [&]
      {
        require(capturing(), "no graph capture is open on this thread");
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream(), &g);
        g_capturing = false;
        const std::string first = g_capture_error;
        g_capture_error.clear();
        if (!first.empty() || e != hipSuccess || !out)
        {
          if (g)
            (void)hipGraphDestroy(g);
          (void)hipGetLastError();
          if (!out)
            throw Error(DFX_ERR_ARG, "null output (the capture was discarded)");
          throw Error(DFX_ERR_CAPTURE,
                      "graph capture discarded: "
                          + (first.empty() ? std::string(hipGetErrorString(e)) : first));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess)
          throw Error(DFX_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(ei));
        *out = new dfx_graph{exec};
      });
}

dfx_status dfx_graph_launch(const dfx_graph* g)
{
  return guard(
      [&]
      {
        require(g && g->exec, "null graph");
        no_capture("dfx_graph_launch");
        DFX_HIP(hipGraphLaunch(g->exec, stream()));
      });
}

dfx_status dfx_graph_destroy(dfx_graph* g)
{
  if (g)
  {
    if (g->exec)
      (void)hipGraphExecDestroy(g->exec);
    delete g;
  }
  return DFX_OK;
}

} // extern "C"
