// Cell-loop assembly on the device: hand-written element kernels for the
// config forms, batched over all cells, scattering into a device CSR with
// FP atomics through a precomputed cell -> CSR-slot map. Also Dirichlet
// lifting, set_diagonal and DirichletBC::set. gfx950 (MI355X) only.
//
// Replaces fem::impl::assemble_cells_matrix (fem/assemble_matrix_impl.h:
// 92-200), fem::impl::assemble_cells (fem/assemble_vector_impl.h:72-116),
// fem::impl::lift_bc (fem/assemble_vector_impl.h:361-414) and the
// MatrixCSR::add binary-search insert (la/matrix_csr_impl.h:67-109).
#include "device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dfx
{
std::uint64_t next_pattern_id();

namespace
{

constexpr int kBlock = 256;
// LDS slice of one elasticity row group (6 rows at most)
constexpr std::size_t kTeamSlice = 16 * 1024;

/// Two values of T as one aligned vector access (double2 or float2).
template <typename T>
struct Vec2;
template <>
struct Vec2<double>
{
  using type = double2;
};
template <>
struct Vec2<float>
{
  using type = float2;
};
template <typename T>
using vec2_t = typename Vec2<T>::type;

struct KernelInfo
{
  int cell;         // DFX_TET / DFX_HEX
  int npc;          // geometry nodes per cell
  int k;            // dofs (nodes) per cell
  int bs;           // block size
  bool bilinear;    // matrix (true) or vector (false) form
};

bool kernel_info(int id, KernelInfo& ki)
{
  switch (id)
  {
  case DFX_K_P1TET_A:
    ki = {DFX_TET, 4, 4, 1, true};
    return true;
  case DFX_K_P1TET_L:
    ki = {DFX_TET, 4, 4, 1, false};
    return true;
  case DFX_K_Q1HEX_A:
    ki = {DFX_HEX, 8, 8, 1, true};
    return true;
  case DFX_K_Q1HEX_L:
    ki = {DFX_HEX, 8, 8, 1, false};
    return true;
  case DFX_K_P2TET_ELAST_A:
    ki = {DFX_TET, 4, 10, 3, true};
    return true;
  case DFX_K_P2TET_ELAST_L:
    ki = {DFX_TET, 4, 10, 3, false};
    return true;
  case DFX_K_P1TET_NEUMANN:
    ki = {DFX_TET, 4, 4, 1, false};
    return true;
  case DFX_K_Q1HEX_NEUMANN:
    ki = {DFX_HEX, 8, 8, 1, false};
    return true;
  case DFX_K_P1TET_ROBIN_A:
    ki = {DFX_TET, 4, 4, 1, true};
    return true;
  case DFX_K_Q1HEX_ROBIN_A:
    ki = {DFX_HEX, 8, 8, 1, true};
    return true;
  case DFX_K_P1TET_GRADJUMP_A:
    ki = {DFX_TET, 4, 4, 1, true};
    return true;
  case DFX_K_ELEMENT_A:
    ki = {-1, 0, 0, 1, true}; // any element: k from the descriptor
    return true;
  case DFX_K_ELEMENT_L:
    ki = {-1, 0, 0, 1, false};
    return true;
  case DFX_K_P3TET_A:
    ki = {DFX_TET, 4, 20, 1, true};
    return true;
  case DFX_K_P3TET_L:
    ki = {DFX_TET, 4, 20, 1, false};
    return true;
  default:
    return false;
  }
}

/// Kernels of exterior facet matrices (beta u v ds).
bool is_facet_matrix_kernel(int id)
{
  return id == DFX_K_P1TET_ROBIN_A || id == DFX_K_Q1HEX_ROBIN_A;
}

/// Kernels of interior facet integrals.
bool is_interior_kernel(int id) { return id == DFX_K_P1TET_GRADJUMP_A; }

// ---------------------------------------------------------------------
// Geometry helpers
// ---------------------------------------------------------------------

/// adj = det(J) J^{-1} (row-major, adj[3a + r]) for J[3r + a] = dx_r/dxi_a;
/// returns det(J).
template <typename T>
__device__ inline T adjugate(const T J[9], T adj[9])
{
  adj[0] = J[4] * J[8] - J[5] * J[7];
  adj[1] = J[2] * J[7] - J[1] * J[8];
  adj[2] = J[1] * J[5] - J[2] * J[4];
  adj[3] = J[5] * J[6] - J[3] * J[8];
  adj[4] = J[0] * J[8] - J[2] * J[6];
  adj[5] = J[2] * J[3] - J[0] * J[5];
  adj[6] = J[3] * J[7] - J[4] * J[6];
  adj[7] = J[1] * J[6] - J[0] * J[7];
  adj[8] = J[0] * J[4] - J[1] * J[3];
  return J[0] * adj[0] + J[1] * adj[3] + J[2] * adj[6];
}

/// Slot of (row dof i, column dof j) of each cell as an offset inside row i.
template <typename O>
__global__ void __launch_bounds__(kBlock)
    slot_map_k(std::int32_t nc, int k, const std::int32_t* __restrict__ dm,
               const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
               O* __restrict__ slots, int* err)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  const std::int64_t kk = std::int64_t(k) * k;
  if (t >= std::int64_t(nc) * kk)
    return;
  const std::int64_t c = t / kk;
  const int e = static_cast<int>(t % kk);
  const std::int32_t row = dm[c * k + e / k];
  const std::int32_t col = dm[c * k + e % k];
  std::int32_t lo = rp[row], hi = rp[row + 1];
  const std::int32_t r0 = lo, r1 = hi;
  while (lo < hi)
  {
    const std::int32_t mid = (lo + hi) >> 1;
    if (ci[mid] < col)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= r1 || ci[lo] != col)
  {
    atomicOr(err, 1);
    slots[t] = 0;
    return;
  }
  slots[t] = static_cast<O>(lo - r0);
}

// ---------------------------------------------------------------------
// P1 tetrahedra: a = kappa grad u . grad v + c u v, 16 lanes per cell,
// lane (i, j) computes one entry. MODE 0: A += Ae with BC rows/cols
// zeroed; MODE 1: b -= alpha Ae[:, bc] (g - x0) (lifting).
// ---------------------------------------------------------------------
template <typename T, typename O, int MODE>
__global__ void __launch_bounds__(kBlock)
    p1tet_a_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
              const std::int32_t* __restrict__ dm, const T* __restrict__ coeffs, int cstride,
              T creac, const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
              const O* __restrict__ slots, T* __restrict__ vals, const double* __restrict__ g,
              const T* __restrict__ x0, double alpha, T* __restrict__ b)
{
  const std::int64_t gid = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  const std::int64_t c = gid >> 4;
  const int e = static_cast<int>(gid & 15), i = e >> 2, j = e & 3;
  if (c >= nc)
    return;
  const int4 gv = *reinterpret_cast<const int4*>(xdm + 4 * c);
  const int4 dv = *reinterpret_cast<const int4*>(dm + 4 * c);
  const std::int32_t di = (&dv.x)[i], dj = (&dv.x)[j];
  if constexpr (MODE == 0)
  {
    if (bc && (bc[di] | bc[dj]))
      return;
  }
  else
  {
    if (!bc[dj])
      return;
  }
  const T* p0 = x + 3 * std::int64_t(gv.x);
  const T* p1 = x + 3 * std::int64_t(gv.y);
  const T* p2 = x + 3 * std::int64_t(gv.z);
  const T* p3 = x + 3 * std::int64_t(gv.w);
  T J[9], adj[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    J[3 * r + 0] = p1[r] - p0[r];
    J[3 * r + 1] = p2[r] - p0[r];
    J[3 * r + 2] = p3[r] - p0[r];
  }
  const T det = adjugate(J, adj);
  const T adet = det < T(0) ? -det : det;
  // h_i = det * grad phi_i: rows of adj (phi_0: minus their sum)
  T hi[3], hj[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    const T s = adj[r] + adj[3 + r] + adj[6 + r];
    hi[r] = i == 0 ? -s : adj[3 * (i - 1) + r];
    hj[r] = j == 0 ? -s : adj[3 * (j - 1) + r];
  }
  const T kap = coeffs[c * cstride];
  T a = kap * (hi[0] * hj[0] + hi[1] * hj[1] + hi[2] * hj[2]) / (T(6) * adet);
  a += creac * adet * (i == j ? T(2) : T(1)) / T(120);
  if constexpr (MODE == 0)
  {
    atomicAdd(vals + rp[di] + slots[gid], a);
  }
  else
  {
    const double xb = g ? g[dj] : 0.0;
    const double xo = x0 ? double(x0[dj]) : 0.0;
    atomicAdd(b + di, static_cast<T>(-double(a) * alpha * (xb - xo)));
  }
}

/// P1 tetrahedra: L = f v with a P1 coefficient f (w = f at the vertices):
/// b_i += |K|/20 (f_i + sum_j f_j). 4 lanes per cell.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    p1tet_l_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
              const std::int32_t* __restrict__ dm, const T* __restrict__ coeffs, int cstride,
              T* __restrict__ b)
{
  const std::int64_t gid = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  const std::int64_t c = gid >> 2;
  const int i = static_cast<int>(gid & 3);
  if (c >= nc)
    return;
  const int4 gv = *reinterpret_cast<const int4*>(xdm + 4 * c);
  const T* p0 = x + 3 * std::int64_t(gv.x);
  const T* p1 = x + 3 * std::int64_t(gv.y);
  const T* p2 = x + 3 * std::int64_t(gv.z);
  const T* p3 = x + 3 * std::int64_t(gv.w);
  T J[9], adj[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    J[3 * r + 0] = p1[r] - p0[r];
    J[3 * r + 1] = p2[r] - p0[r];
    J[3 * r + 2] = p3[r] - p0[r];
  }
  const T det = adjugate(J, adj);
  const T adet = det < T(0) ? -det : det;
  const T* f = coeffs + c * cstride;
  const T v = adet / T(120) * (f[0] + f[1] + f[2] + f[3] + f[i]);
  atomicAdd(b + dm[4 * c + i], v);
}

// ---------------------------------------------------------------------
// Q1 hexahedra. Reference cube [0,1]^3, vertex i at bits (i&1, i>>1&1,
// i>>2&1). One-dimensional integrals of the linear shape functions give
// every reference tensor exactly.
// ---------------------------------------------------------------------

/// Symmetrised reference stiffness S6 (order 00, 11, 22, 01+10, 02+20,
/// 12+21) and reference mass of entry (i, j).
template <typename T>
__device__ inline void q1_reference(int i, int j, T S6[6], T& M)
{
  T m[3], k[3], d[3], e[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
  {
    const int p = (i >> a) & 1, q = (j >> a) & 1;
    m[a] = p == q ? T(1) / T(3) : T(1) / T(6);
    k[a] = p == q ? T(1) : T(-1);
    d[a] = p ? T(0.5) : T(-0.5); // int f_p' f_q
    e[a] = q ? T(0.5) : T(-0.5); // int f_p f_q'
  }
  M = m[0] * m[1] * m[2];
  S6[0] = k[0] * m[1] * m[2];
  S6[1] = m[0] * k[1] * m[2];
  S6[2] = m[0] * m[1] * k[2];
  S6[3] = (d[0] * e[1] + e[0] * d[1]) * m[2];
  S6[4] = (d[0] * e[2] + e[0] * d[2]) * m[1];
  S6[5] = (d[1] * e[2] + e[1] * d[2]) * m[0];
}

/// Affine hexahedra, a = kappa grad u.grad v + c u v. Each wavefront
/// takes 64 cells: phase 1 (lane = cell) forms the 6 geometric factors
/// G = kappa adj adj^T / |det|, the mass scale c|det|, the dofs, their row
/// starts and BC flags in LDS; phase 2 (lane = entry (i, j)) evaluates
/// A_ij = sum G S6 + c|det| M for each of the 64 cells and scatters.
template <typename T, typename O, int MODE>
__global__ void __launch_bounds__(kBlock)
    q1hex_a_affine_k(std::int32_t nc, const T* __restrict__ x,
                     const std::int32_t* __restrict__ xdm, const std::int32_t* __restrict__ dm,
                     const T* __restrict__ coeffs, int cstride, T creac,
                     const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                     const O* __restrict__ slots, T* __restrict__ vals,
                     const double* __restrict__ g, const T* __restrict__ x0, double alpha,
                     T* __restrict__ b)
{
  __shared__ T sG[4][64][7];
  __shared__ std::int32_t sDof[4][64][8];
  __shared__ std::int32_t sRow[4][64][8];
  __shared__ unsigned sBC[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const std::int64_t base = (std::int64_t(blockIdx.x) * 4 + wave) * 64;
  {
    const std::int64_t c = base + lane;
    if (c < nc)
    {
      const int4 g0 = *reinterpret_cast<const int4*>(xdm + 8 * c);
      const int4 g1 = *reinterpret_cast<const int4*>(xdm + 8 * c + 4);
      const T* q0 = x + 3 * std::int64_t(g0.x);
      const T* q1 = x + 3 * std::int64_t(g0.y);
      const T* q2 = x + 3 * std::int64_t(g0.z);
      const T* q4 = x + 3 * std::int64_t(g1.x);
      T J[9], adj[9];
#pragma unroll
      for (int r = 0; r < 3; ++r)
      {
        J[3 * r + 0] = q1[r] - q0[r];
        J[3 * r + 1] = q2[r] - q0[r];
        J[3 * r + 2] = q4[r] - q0[r];
      }
      const T det = adjugate(J, adj);
      const T adet = det < T(0) ? -det : det;
      const T s = coeffs[c * cstride] / adet;
      T* G = sG[wave][lane];
      G[0] = s * (adj[0] * adj[0] + adj[1] * adj[1] + adj[2] * adj[2]);
      G[1] = s * (adj[3] * adj[3] + adj[4] * adj[4] + adj[5] * adj[5]);
      G[2] = s * (adj[6] * adj[6] + adj[7] * adj[7] + adj[8] * adj[8]);
      G[3] = s * (adj[0] * adj[3] + adj[1] * adj[4] + adj[2] * adj[5]);
      G[4] = s * (adj[0] * adj[6] + adj[1] * adj[7] + adj[2] * adj[8]);
      G[5] = s * (adj[3] * adj[6] + adj[4] * adj[7] + adj[5] * adj[8]);
      G[6] = creac * adet;
      const int4 d0 = *reinterpret_cast<const int4*>(dm + 8 * c);
      const int4 d1 = *reinterpret_cast<const int4*>(dm + 8 * c + 4);
      const std::int32_t dd[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
      unsigned mask = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        sDof[wave][lane][k] = dd[k];
        if constexpr (MODE == 0)
          sRow[wave][lane][k] = rp[dd[k]];
        if (bc && bc[dd[k]])
          mask |= 1u << k;
      }
      sBC[wave][lane] = mask;
    }
  }
  __syncthreads();
  const int i = lane >> 3, j = lane & 7;
  T S6[6], M;
  q1_reference<T>(i, j, S6, M);
  for (int t = 0; t < 64; ++t)
  {
    const std::int64_t c = base + t;
    if (c >= nc)
      break;
    const unsigned mask = sBC[wave][t];
    if constexpr (MODE == 0)
    {
      if ((mask >> i | mask >> j) & 1u)
        continue;
    }
    else
    {
      if (!((mask >> j) & 1u))
        continue;
    }
    const T* G = sG[wave][t];
    const T a = G[0] * S6[0] + G[1] * S6[1] + G[2] * S6[2] + G[3] * S6[3] + G[4] * S6[4]
                + G[5] * S6[5] + G[6] * M;
    if constexpr (MODE == 0)
    {
      atomicAdd(vals + sRow[wave][t][i] + slots[c * 64 + lane], a);
    }
    else
    {
      const std::int32_t dj = sDof[wave][t][j];
      const double xb = g ? g[dj] : 0.0;
      const double xo = x0 ? double(x0[dj]) : 0.0;
      atomicAdd(b + sDof[wave][t][i], static_cast<T>(-double(a) * alpha * (xb - xo)));
    }
  }
}

/// Q1 shape functions and reference gradients at point xi.
template <typename T>
__device__ inline void q1_tab(const T xi[3], T phi[8], T dphi[8][3])
{
#pragma unroll
  for (int i = 0; i < 8; ++i)
  {
    T f[3], df[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
      const int bit = (i >> a) & 1;
      f[a] = bit ? xi[a] : T(1) - xi[a];
      df[a] = bit ? T(1) : T(-1);
    }
    phi[i] = f[0] * f[1] * f[2];
    dphi[i][0] = df[0] * f[1] * f[2];
    dphi[i][1] = f[0] * df[1] * f[2];
    dphi[i][2] = f[0] * f[1] * df[2];
  }
}

/// Element matrices of the Q1 stiffness + reaction form on general
/// (non-affine) hexahedra: one thread per cell, 2x2x2 Gauss, the 8x8
/// matrix summed in registers over the quadrature points and stored
/// (E: cells x 8 x 8, row-major); the row-owner element path then adds
/// them in cell order, so the result does not depend on the schedule.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    q1hex_elem_general_k(std::int32_t nc, const T* __restrict__ x,
                         const std::int32_t* __restrict__ xdm, const T* __restrict__ coeffs,
                         int cstride, T creac, T* __restrict__ E)
{
  const std::int64_t c = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  T cd[8][3];
  for (int v = 0; v < 8; ++v)
    for (int r = 0; r < 3; ++r)
      cd[v][r] = x[3 * std::int64_t(xdm[8 * c + v]) + r];
  const T kap = coeffs[c * cstride];
  T A[8][8];
  for (int p = 0; p < 8; ++p)
    for (int q = 0; q < 8; ++q)
      A[p][q] = 0;
  const T ga = T(0.5) - T(0.5) / sqrt(T(3)), gb = T(0.5) + T(0.5) / sqrt(T(3));
  for (int qp = 0; qp < 8; ++qp)
  {
    const T xi[3] = {(qp & 1) ? gb : ga, ((qp >> 1) & 1) ? gb : ga, ((qp >> 2) & 1) ? gb : ga};
    T phi[8], dphi[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, adj[9];
    q1_tab(xi, phi, dphi);
    for (int v = 0; v < 8; ++v)
      for (int r = 0; r < 3; ++r)
        for (int a = 0; a < 3; ++a)
          J[3 * r + a] += cd[v][r] * dphi[v][a];
    const T det = adjugate(J, adj);
    const T adet = det < T(0) ? -det : det;
    T h[8][3];
    for (int v = 0; v < 8; ++v)
      for (int r = 0; r < 3; ++r)
        h[v][r] = dphi[v][0] * adj[r] + dphi[v][1] * adj[3 + r] + dphi[v][2] * adj[6 + r];
    const T ws = kap / (T(8) * adet), wm = creac * adet / T(8);
    for (int p = 0; p < 8; ++p)
      for (int q = 0; q < 8; ++q)
        A[p][q] += ws * (h[p][0] * h[q][0] + h[p][1] * h[q][1] + h[p][2] * h[q][2])
                   + wm * phi[p] * phi[q];
  }
  T* out = E + 64 * c;
  for (int p = 0; p < 8; ++p)
    for (int q = 0; q < 8; ++q)
      out[8 * p + q] = A[p][q];
}

/// Q1 hexahedra: L = f v with a Q1 coefficient (w = f at the vertices).
/// 8 lanes per cell; affine cells use b_i = |det| sum_j M_ij f_j, others
/// 2x2x2 Gauss.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    q1hex_l_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
              const std::int32_t* __restrict__ dm, const T* __restrict__ coeffs, int cstride,
              int affine, T* __restrict__ b)
{
  const std::int64_t gid = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  const std::int64_t c = gid >> 3;
  const int i = static_cast<int>(gid & 7);
  if (c >= nc)
    return;
  const T* f = coeffs + c * cstride;
  T v = 0;
  if (affine)
  {
    const T* q0 = x + 3 * std::int64_t(xdm[8 * c + 0]);
    const T* q1 = x + 3 * std::int64_t(xdm[8 * c + 1]);
    const T* q2 = x + 3 * std::int64_t(xdm[8 * c + 2]);
    const T* q4 = x + 3 * std::int64_t(xdm[8 * c + 4]);
    T J[9], adj[9];
    for (int r = 0; r < 3; ++r)
    {
      J[3 * r + 0] = q1[r] - q0[r];
      J[3 * r + 1] = q2[r] - q0[r];
      J[3 * r + 2] = q4[r] - q0[r];
    }
    const T det = adjugate(J, adj);
    const T adet = det < T(0) ? -det : det;
    for (int j = 0; j < 8; ++j)
    {
      T m = 1;
      for (int a = 0; a < 3; ++a)
        m *= (((i >> a) & 1) == ((j >> a) & 1)) ? T(1) / T(3) : T(1) / T(6);
      v += m * f[j];
    }
    v *= adet;
  }
  else
  {
    T cd[8][3];
    for (int w = 0; w < 8; ++w)
      for (int r = 0; r < 3; ++r)
        cd[w][r] = x[3 * std::int64_t(xdm[8 * c + w]) + r];
    const T ga = T(0.5) - T(0.5) / sqrt(T(3)), gb = T(0.5) + T(0.5) / sqrt(T(3));
    for (int qp = 0; qp < 8; ++qp)
    {
      const T xi[3] = {(qp & 1) ? gb : ga, ((qp >> 1) & 1) ? gb : ga, ((qp >> 2) & 1) ? gb : ga};
      T phi[8], dphi[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, adj[9];
      q1_tab(xi, phi, dphi);
      for (int w = 0; w < 8; ++w)
        for (int r = 0; r < 3; ++r)
          for (int a = 0; a < 3; ++a)
            J[3 * r + a] += cd[w][r] * dphi[w][a];
      const T det = adjugate(J, adj);
      const T adet = det < T(0) ? -det : det;
      T fq = 0;
      for (int w = 0; w < 8; ++w)
        fq += f[w] * phi[w];
      v += adet / T(8) * fq * phi[i];
    }
  }
  atomicAdd(b + dm[8 * c + i], v);
}

// ---------------------------------------------------------------------
// Dirichlet helpers
// ---------------------------------------------------------------------

/// A[d, d] = diag for each BC dof d (unrolled index; block row d / bs).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    set_diagonal_k(std::int32_t n, const std::int32_t* dofs, int bs, const std::int32_t* diag,
                   T* vals, T value, int* err)
{
  const std::int32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n)
    return;
  const std::int32_t d = dofs[t];
  const std::int32_t p = diag[d / bs];
  if (p < 0)
  {
    if (err)
      atomicOr(err, 1);
    return;
  }
  const int k = d % bs;
  vals[(std::int64_t(p) * bs + k) * bs + k] = value;
}

/// b[d] = alpha (g[d] - x0[d]) for each BC dof d.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bc_set_k(std::int32_t n, const std::int32_t* dofs, const double* g, const T* x0,
             double alpha, T* b)
{
  const std::int32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n)
    return;
  const std::int32_t d = dofs[t];
  const double gv = g ? g[d] : 0.0;
  const double xv = x0 ? double(x0[d]) : 0.0;
  b[d] = static_cast<T>(alpha * (gv - xv));
}

template <typename T>
__global__ void __launch_bounds__(kBlock) to_dtype_k(T* dst, const double* src, std::int64_t n)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t < n)
    dst[t] = static_cast<T>(src[t]);
}

int blocks_for(std::int64_t n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

/// Device copy of a host FP64 array converted to dtype.
/// coeffs[t] = values[map[t] * ncomp + t % ncomp] (cell layout, map per
/// cell) or values[map[t]] (vertex layout, map per position), in T.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    pack_gather_k(std::int64_t n, int ncomp, int per_cell, const std::int32_t* __restrict__ map,
                  const double* __restrict__ values, T* __restrict__ coeffs)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= n)
    return;
  const std::int64_t src
      = per_cell ? std::int64_t(map[t / ncomp]) * ncomp + t % ncomp : std::int64_t(map[t]);
  coeffs[t] = static_cast<T>(values[src]);
}

void* upload_as(const double* src, std::int64_t n, int dtype)
{
  void* d = dmalloc(dtype_size(dtype) * std::max<std::int64_t>(n, 1));
  if (n == 0)
    return d;
  if (dtype == DFX_F64)
  {
    h2d(d, src, sizeof(double) * n);
    return d;
  }
  double* tmp = static_cast<double*>(dmalloc(sizeof(double) * n));
  h2d(tmp, src, sizeof(double) * n);
  to_dtype_k<float><<<blocks_for(n), kBlock, 0, stream()>>>(static_cast<float*>(d), tmp, n);
  check_launch("upload_as");
  DFX_HIP(hipStreamSynchronize(stream()));
  dfree(tmp);
  return d;
}

/// Builds (or reuses) the cell -> CSR slot map of form f for matrix A.
void ensure_slots(dfx_form& f, const dfx_csr& A)
{
  if (f.slots && f.slot_pattern == A.pattern_id)
    return;
  dfree(f.slots);
  f.slots = nullptr;
  const int k = f.d.dofs_per_cell;
  const std::int64_t n = std::int64_t(f.d.num_cells) * k * k;
  f.slot_bytes = A.max_row <= 256 ? 1 : 2;
  require(A.max_row <= 65536, "row too long for the slot map");
  f.slots = dmalloc(std::max<std::int64_t>(n, 1) * f.slot_bytes);
  int* err = static_cast<int*>(dmalloc(sizeof(int)));
  DFX_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
  if (n > 0)
  {
    if (f.slot_bytes == 1)
      slot_map_k<std::uint8_t><<<blocks_for(n), kBlock, 0, stream()>>>(
          f.d.num_cells, k, f.dofmap, A.row_ptr, A.cols, static_cast<std::uint8_t*>(f.slots), err);
    else
      slot_map_k<std::uint16_t><<<blocks_for(n), kBlock, 0, stream()>>>(
          f.d.num_cells, k, f.dofmap, A.row_ptr, A.cols, static_cast<std::uint16_t*>(f.slots), err);
    check_launch("slot_map");
  }
  int h = 0;
  d2h(&h, err, sizeof(int));
  dfree(err);
  if (h)
  {
    dfree(f.slots);
    f.slots = nullptr;
    throw Error(DFX_ERR_PATTERN, "Entry not in sparsity");
  }
  f.slot_pattern = A.pattern_id;
}

// ---------------------------------------------------------------------
// Row-owner assembly of affine cells (no atomics, deterministic).
//
// Pass 1, cell_factors_k (one thread per cell): the geometric factors
//   G = kappa |det J| J^{-1} J^{-T}   (6 values, symmetric)
// and the reaction scale c |det J|, plus the bit mask of the cell's dofs
// that carry a Dirichlet condition.
// Pass 2, row_assemble_k (one thread per matrix row): loop over the row's
// incidences (cell c, local index i), in cell order, and add
//   A_ij = sum_g G_g(c) S_g(i, j) + c|det J| M(i, j)
// for every column dof j of c (S: symmetrised reference stiffness
// integrals, M: reference mass) into an LDS copy of the block's slice of
// the CSR values; the slice is then stored with coalesced writes. Every
// entry is summed in the same order on every run.
// ---------------------------------------------------------------------

/// A 32-bit mask carried bit for bit in a value of T.
template <typename T>
__device__ inline T mask_bits(std::uint32_t m)
{
  if constexpr (sizeof(T) == 8)
    return __longlong_as_double(static_cast<long long>(m));
  else
    return __uint_as_float(m);
}

/// The affine factors of one cell from its Jacobian vertices q0..q3
/// (x1-x0, x2-x0 and x3-x0 for a tetrahedron, x4-x0 for a hexahedron with
/// lexicographic vertices): G = kappa |det J| J^-1 J^-T (6 values, through
/// the adjugate) and c |det J|.
template <typename T>
__device__ inline void affine_factors(const T* q0, const T* q1, const T* q2, const T* q3,
                                      T kappa, T creac, T (&G)[7])
{
  T J[9], adj[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    J[3 * r + 0] = q1[r] - q0[r];
    J[3 * r + 1] = q2[r] - q0[r];
    J[3 * r + 2] = q3[r] - q0[r];
  }
  const T det = adjugate(J, adj);
  const T adet = det < T(0) ? -det : det;
  const T s = kappa / adet;
  G[0] = s * (adj[0] * adj[0] + adj[1] * adj[1] + adj[2] * adj[2]);
  G[1] = s * (adj[3] * adj[3] + adj[4] * adj[4] + adj[5] * adj[5]);
  G[2] = s * (adj[6] * adj[6] + adj[7] * adj[7] + adj[8] * adj[8]);
  G[3] = s * (adj[0] * adj[3] + adj[1] * adj[4] + adj[2] * adj[5]);
  G[4] = s * (adj[0] * adj[6] + adj[1] * adj[7] + adj[2] * adj[8]);
  G[5] = s * (adj[3] * adj[6] + adj[4] * adj[7] + adj[5] * adj[8]);
  G[6] = creac * adet;
}

/// Compact record of a cell whose Jacobian is diagonal (G is then
/// diagonal): G00, G11, G22 and c |det J|, 16-byte aligned.
template <typename T>
__device__ inline void store_compact(T* __restrict__ cd, std::int64_t c, const T (&g)[7])
{
  if constexpr (sizeof(T) == 8)
  {
    double2* G = reinterpret_cast<double2*>(cd + 4 * c);
    G[0] = make_double2(g[0], g[1]);
    G[1] = make_double2(g[2], g[6]);
  }
  else
    *reinterpret_cast<float4*>(cd + 4 * c) = make_float4(g[0], g[1], g[2], g[6]);
}

/// The factors of a compact record (off-diagonal factors zero).
template <typename T>
__device__ inline void load_compact(const T* __restrict__ cd, std::int64_t c, T (&g)[7])
{
  if constexpr (sizeof(T) == 8)
  {
    const double2* p = reinterpret_cast<const double2*>(cd + 4 * c);
    const double2 a = p[0], b = p[1];
    g[0] = a.x; g[1] = a.y; g[2] = b.x; g[6] = b.y;
  }
  else
  {
    const float4 a = *reinterpret_cast<const float4*>(cd + 4 * c);
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[6] = a.w;
  }
  g[3] = g[4] = g[5] = T(0);
}

/// Pass 1: the factor record of every cell (one thread per cell). With SAME
/// the dofmap is the vertex map (P1/Q1 on the mesh's numbering): the
/// vertex row, loaded with vector loads, also addresses the BC markers, so
/// the marker and coordinate gathers are in flight together.
template <typename T, int NV, bool SAME, bool COMPACT = false>
__global__ void __launch_bounds__(kBlock)
    cell_factors_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                   const std::int32_t* __restrict__ dm, int k, const T* __restrict__ coeffs,
                   int cstride, T creac, const std::int8_t* __restrict__ bc, T* __restrict__ cd,
                   std::uint32_t* __restrict__ mask)
{
  const std::int64_t c
      = xcd_contiguous(blockIdx.x, gridDim.x) * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  std::int32_t v[NV];
  {
    const int4 a = *reinterpret_cast<const int4*>(xdm + NV * c);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    if constexpr (NV == 8)
    {
      const int4 b = *reinterpret_cast<const int4*>(xdm + NV * c + 4);
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
  }
  std::uint32_t m = 0;
  if constexpr (SAME)
  {
    if (mask)
    {
#pragma unroll
      for (int j = 0; j < NV; ++j)
        if (bc[v[j]])
          m |= 1u << j;
    }
  }
  T g[7];
  affine_factors<T>(x + 3 * std::int64_t(v[0]), x + 3 * std::int64_t(v[1]),
                    x + 3 * std::int64_t(v[2]), x + 3 * std::int64_t(v[NV == 4 ? 3 : 4]),
                    coeffs[c * cstride], creac, g);
  if constexpr (!SAME)
  {
    if (mask)
      for (int j = 0; j < k; ++j)
        if (bc[dm[c * k + j]])
          m |= 1u << j;
  }
  if (mask)
    mask[c] = m;
  if constexpr (COMPACT)
  {
    store_compact<T>(cd, c, g);
    return;
  }
  // the eighth slot carries the BC column mask bits (moved, never computed
  // with), so a row kernel reads factors and mask with one load; the record
  // is stored with full-width vector stores
  const T g7 = mask_bits<T>(m);
  if constexpr (sizeof(T) == 8)
  {
    double2* G = reinterpret_cast<double2*>(cd + 8 * c);
    G[0] = make_double2(g[0], g[1]);
    G[1] = make_double2(g[2], g[3]);
    G[2] = make_double2(g[4], g[5]);
    G[3] = make_double2(g[6], g7);
  }
  else
  {
    float4* G = reinterpret_cast<float4*>(cd + 8 * c);
    G[0] = make_float4(g[0], g[1], g[2], g[3]);
    G[1] = make_float4(g[4], g[5], g[6], g7);
  }
}

/// Incidence slot offsets: for incidence p = (cell c, local row i) and
/// column j, the position of dof(c, j) inside row dof(c, i).
template <typename O>
__global__ void __launch_bounds__(kBlock)
    incidence_slots_k(std::int64_t ninc, int k, const std::int32_t* __restrict__ inc,
                      const std::int32_t* __restrict__ dm, const std::int32_t* __restrict__ rp,
                      const std::int32_t* __restrict__ ci, O* __restrict__ islots, int* err)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= ninc * k)
    return;
  const std::int64_t p = t / k;
  const int j = static_cast<int>(t % k);
  const std::int32_t cik = inc[p];
  const std::int32_t c = cik / k;
  const std::int32_t row = dm[cik];
  const std::int32_t col = dm[std::int64_t(c) * k + j];
  std::int32_t lo = rp[row], hi = rp[row + 1];
  const std::int32_t r0 = lo, r1 = hi;
  while (lo < hi)
  {
    const std::int32_t mid = (lo + hi) >> 1;
    if (ci[mid] < col)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= r1 || ci[lo] != col)
  {
    atomicOr(err, 1);
    islots[t] = 0;
    return;
  }
  islots[t] = static_cast<O>(lo - r0);
}

/// The 7 cell factors of cell c and its BC column mask (stored as 8 values,
/// 16-byte aligned, the mask bits in the eighth).
template <typename T>
__device__ inline void load_factors(const T* __restrict__ cd, std::int64_t c, T (&g)[7],
                                    std::uint32_t& m)
{
  if constexpr (sizeof(T) == 8)
  {
    const double2* p = reinterpret_cast<const double2*>(cd + 8 * c);
    const double2 a = p[0], b = p[1], d = p[2], e = p[3];
    g[0] = a.x; g[1] = a.y; g[2] = b.x; g[3] = b.y; g[4] = d.x; g[5] = d.y; g[6] = e.x;
    m = static_cast<std::uint32_t>(__double_as_longlong(e.y));
  }
  else
  {
    const float4* p = reinterpret_cast<const float4*>(cd + 8 * c);
    const float4 a = p[0], b = p[1];
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w; g[4] = b.x; g[5] = b.y; g[6] = b.z;
    m = __float_as_uint(b.w);
  }
}

/// Orders the LDS accesses of the lanes of one wavefront around it
/// (release and acquire fences at wavefront scope and a wave barrier): a
/// value one lane writes before it is seen by the others reading after it.
__device__ inline void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/// a[sl[j]] += v[j] for the unmasked j of one incidence. The K slots of an
/// incidence are distinct columns of its row, so every value is read
/// before any is written: the K LDS reads are in flight together instead
/// of K dependent read-add-write round trips, and each entry still gets
/// one addition, in the same order as element by element.
template <typename T, typename O, int K>
__device__ inline void add_slots(T* a, const O (&sl)[K], const T (&v)[K], std::uint32_t m)
{
  T cur[K];
#pragma unroll
  for (int j = 0; j < K; ++j)
    cur[j] = a[sl[j]];
#pragma unroll
  for (int j = 0; j < K; ++j)
    if (!((m >> j) & 1u))
      a[sl[j]] = cur[j] + v[j];
}

/// The K slot offsets of incidence p with one vector load where the width
/// allows (K bytes or 2K bytes, naturally aligned).
template <typename O, int K>
__device__ inline void load_slots(const O* __restrict__ islots, std::int64_t p, O (&sl)[K])
{
  const O* src = islots + p * K;
  if constexpr (sizeof(O) * K == 8)
  {
    const std::uint64_t w = *reinterpret_cast<const std::uint64_t*>(src);
#pragma unroll
    for (int j = 0; j < K; ++j)
      sl[j] = static_cast<O>(w >> (8 * sizeof(O) * j));
  }
  else if constexpr (sizeof(O) * K == 4)
  {
    const std::uint32_t w = *reinterpret_cast<const std::uint32_t*>(src);
#pragma unroll
    for (int j = 0; j < K; ++j)
      sl[j] = static_cast<O>(w >> (8 * sizeof(O) * j));
  }
  else if constexpr (sizeof(O) * K == 16)
  {
    const uint4 w = *reinterpret_cast<const uint4*>(src);
    const std::uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < K; ++j)
      sl[j] = static_cast<O>(q[(j * sizeof(O)) / 4] >> (8 * ((j * sizeof(O)) % 4)));
  }
  else
  {
#pragma unroll
    for (int j = 0; j < K; ++j)
      sl[j] = src[j];
  }
}

/// Entry (i, j) of the element matrix from the cell factors g and the
/// reference tensors ref[7][K][K].
template <typename T, int K>
__device__ inline T element_entry(const T (&g)[7], const T* __restrict__ ref, int i, int j)
{
  const T* Si = ref + i * K + j;
  return g[0] * Si[0] + g[1] * Si[K * K] + g[2] * Si[2 * K * K] + g[3] * Si[3 * K * K]
         + g[4] * Si[4 * K * K] + g[5] * Si[5 * K * K] + g[6] * Si[6 * K * K];
}

/// Entry (i, j) of the affine Q1 element matrix in closed form, from the
/// cell factors g (coefficients of S00, S11, S22, S01+S10, S02+S20,
/// S12+S21 and M). On the reference cube, with vertex bits (b0, b1, b2) =
/// lexicographic vertex index, the 1D factors are m = int phi_a phi_b (1/3
/// if a = b, else 1/6), k = int phi_a' phi_b' (+1 if a = b, else -1) and
/// int phi_a' phi_b = s_a / 2 (s_0 = -1, s_1 = +1), so with m_d, k_d taken
/// on bit d of i and j:
///   S00 = k0 m1 m2, S11 = m0 k1 m2, S22 = m0 m1 k2, M = m0 m1 m2,
///   S01 + S10 = m2 s_i0 s_i1 (k0 + k1) / 4   (and cyclically).
/// gs holds g[3..5] already multiplied by s_i s_i' / 4.
template <typename T>
__device__ inline T q1_entry(const T (&g)[7], const T (&gs)[3], int i, int j)
{
  const int e = i ^ j;
  const T third = T(1) / T(3), sixth = T(1) / T(6);
  const T m0 = (e & 1) ? sixth : third, k0 = (e & 1) ? T(-1) : T(1);
  const T m1 = (e & 2) ? sixth : third, k1 = (e & 2) ? T(-1) : T(1);
  const T m2 = (e & 4) ? sixth : third, k2 = (e & 4) ? T(-1) : T(1);
  const T m01 = m0 * m1;
  return g[0] * (k0 * m1 * m2) + g[1] * (m0 * k1 * m2) + g[2] * (m01 * k2)
         + gs[0] * (m2 * (k0 + k1)) + gs[1] * (m1 * (k0 + k2)) + gs[2] * (m0 * (k1 + k2))
         + g[6] * (m01 * m2);
}

/// The mixed-derivative factors of q1_entry for row vertex i.
template <typename T>
__device__ inline void q1_signed(const T (&g)[7], int i, T (&gs)[3])
{
  const T s0 = (i & 1) ? T(1) : T(-1);
  const T s1 = (i & 2) ? T(1) : T(-1);
  const T s2 = (i & 4) ? T(1) : T(-1);
  gs[0] = T(0.25) * s0 * s1 * g[3];
  gs[1] = T(0.25) * s0 * s2 * g[4];
  gs[2] = T(0.25) * s1 * s2 * g[5];
}

/// Variant "rows": one thread per matrix row, R rows per workgroup. The
/// thread walks the row's incidences U at a time (all global loads of a
/// group issued before use) and accumulates into the workgroup's LDS copy
/// of its CSR slice, which is then stored with coalesced writes. With CF
/// (affine Q1 hexahedra) the element entries come from q1_entry instead of the
/// reference tensors, and LDS holds only the slice. The body assembles rows
/// [r0, r0 + blockDim.x) of one matrix; smem holds the tensors (unless CF)
/// and the slice.
template <typename T, typename O, int K, int U, bool CF>
__device__ inline void row_block(std::int32_t r0, std::int32_t nrows, std::int32_t nplan,
                                 const std::int32_t* __restrict__ inc_off,
                                 const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
                                 const T* __restrict__ cd, const std::int8_t* __restrict__ bc,
                                 const std::int32_t* __restrict__ rp, const T* __restrict__ ref,
                                 T* __restrict__ vals, int accumulate, T* smem)
{
  static_assert(!CF || K == 8, "closed form is the Q1 hexahedron");
  T* S = smem; // reference tensors [7][K][K]
  T* acc = CF ? S : S + 7 * K * K;
  const int R = blockDim.x;
  const std::int32_t r1 = min(r0 + R, nrows);
  const std::int32_t base = rp[r0];
  const std::int32_t len = rp[r1] - base;
  if constexpr (!CF)
    for (int t = threadIdx.x; t < 7 * K * K; t += R)
      S[t] = ref[t];
  for (int t = threadIdx.x; t < len; t += R)
    acc[t] = T(0);
  __syncthreads();
  const std::int32_t r = r0 + threadIdx.x;
  if (r < r1 && r < nplan && !(bc && bc[r]))
  {
    T* a = acc + (rp[r] - base);
    const std::int32_t pe = inc_off[r + 1];
    for (std::int32_t pb = inc_off[r]; pb < pe; pb += U)
    {
      // loads at clamped positions, none behind a branch, so that all of a
      // group's loads are in flight together; incidences past pe are
      // skipped below
      std::int32_t cik[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        cik[u] = inc[min(pb + u, pe - 1)];
      T g[U][7];
      std::uint32_t m[U];
      O sl[U][K];
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
        const std::int64_t c = cik[u] / K;
        load_factors<T>(cd, c, g[u], m[u]);
        load_slots<O, K>(islots, min(pb + u, pe - 1), sl[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
        if (pb + u >= pe)
          continue;
        const int i = cik[u] % K;
        if constexpr (CF)
        {
          T gs[3];
          q1_signed<T>(g[u], i, gs);
          T v[8];
#pragma unroll
          for (int j = 0; j < 8; ++j)
            v[j] = q1_entry<T>(g[u], gs, i, j);
          add_slots<T, O, 8>(a, sl[u], v, m[u]);
        }
        else
        {
          T v[K];
#pragma unroll
          for (int j = 0; j < K; ++j)
            v[j] = element_entry<T, K>(g[u], S, i, j);
          add_slots<T, O, K>(a, sl[u], v, m[u]);
        }
      }
    }
  }
  __syncthreads();
  if (accumulate)
    for (int t = threadIdx.x; t < len; t += R)
      vals[base + t] += acc[t];
  else
    for (int t = threadIdx.x; t < len; t += R)
      vals[base + t] = acc[t];
}

template <typename T, typename O, int K, int U, bool CF>
__global__ void __launch_bounds__(kBlock)
    row_assemble_k(std::int32_t nrows, std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
                   const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
                   const T* __restrict__ cd, const std::uint32_t* __restrict__ mask,
                   const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                   const T* __restrict__ ref, T* __restrict__ vals, int accumulate)
{
  (void)mask; // carried in the factor records
  extern __shared__ unsigned char smem_raw[];
  row_block<T, O, K, U, CF>(xcd_contiguous(blockIdx.x, gridDim.x) * blockDim.x, nrows, nplan,
                            inc_off, inc, islots, cd, bc, rp, ref, vals, accumulate,
                            reinterpret_cast<T*>(smem_raw));
}

// ---------------------------------------------------------------------
// Element tensors of any size and block size (the reference's
// la::MatrixCSR::mat_add_values<BS, BS> for the element matrices of a
// kernel, la/MatrixCSR.h:157-174): one thread per row dof (node r,
// component a) walks the row's incidences (c, i) in cell order and adds
// element row BS i + a into the blocks of its row (the thread owns
// component row a of every block of block row r: no atomics, each entry
// summed in cell order as MatrixCSR::add sums the cells' contributions).
// ---------------------------------------------------------------------

template <typename T, typename O, int BS>
__global__ void __launch_bounds__(kBlock)
    element_rows_bs_k(std::int32_t nplan, int k, const std::int32_t* __restrict__ inc_off,
                      const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
                      const std::int32_t* __restrict__ dm, const T* __restrict__ E,
                      const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                      T* __restrict__ vals)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nplan) * BS)
    return;
  if (bc && bc[t])
    return; // BC row: no contribution
  const std::int32_t r = static_cast<std::int32_t>(t / BS);
  const int a = static_cast<int>(t - std::int64_t(r) * BS);
  const std::int64_t nd = std::int64_t(BS) * k;
  T* row = vals + std::int64_t(rp[r]) * BS * BS + BS * a;
  for (std::int32_t p = inc_off[r]; p < inc_off[r + 1]; ++p)
  {
    const std::int32_t cik = inc[p];
    const std::int32_t c = cik / k;
    const int i = cik - c * k;
    const T* Ei = E + std::int64_t(c) * nd * nd + (std::int64_t(BS) * i + a) * nd;
    for (int j = 0; j < k; ++j)
    {
      const std::int64_t node = dm[std::int64_t(c) * k + j];
      T* blk = row + std::int64_t(islots[std::int64_t(p) * k + j]) * BS * BS;
#pragma unroll
      for (int b = 0; b < BS; ++b)
        if (!(bc && bc[BS * node + b]))
          blk[b] += Ei[BS * j + b];
    }
  }
}

/// b[(r, a)] += sum over the row's incidences of e[c][BS i + a].
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock)
    element_vector_bs_k(std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
                        const std::int32_t* __restrict__ inc, const T* __restrict__ e,
                        T* __restrict__ b)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nplan) * BS)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / BS);
  const int a = static_cast<int>(t - std::int64_t(r) * BS);
  const std::int32_t p0 = inc_off[r], p1 = inc_off[r + 1];
  if (p1 == p0)
    return;
  T s = b[t];
  for (std::int32_t p = p0; p < p1; ++p)
    s += e[std::int64_t(inc[p]) * BS + a];
  b[t] = s;
}

/// Lifting from element matrices: b[(r, a)] -= alpha sum over the row's
/// incidences (c, i), in cell order, of sum over the cell's Dirichlet
/// columns (j, d) of E_c[BS i + a][BS j + d] (g - x0)_(j, d).
template <typename T, int BS>
__global__ void __launch_bounds__(kBlock)
    element_lift_bs_k(std::int32_t nplan, int k, const std::int32_t* __restrict__ inc_off,
                      const std::int32_t* __restrict__ inc, const std::int32_t* __restrict__ dm,
                      const T* __restrict__ E, const std::int8_t* __restrict__ bc,
                      const double* __restrict__ g, const T* __restrict__ x0, double alpha,
                      T* __restrict__ b)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nplan) * BS)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / BS);
  const int a = static_cast<int>(t - std::int64_t(r) * BS);
  const std::int64_t nd = std::int64_t(BS) * k;
  double s = 0.0;
  for (std::int32_t p = inc_off[r]; p < inc_off[r + 1]; ++p)
  {
    const std::int32_t cik = inc[p];
    const std::int32_t c = cik / k;
    const int i = cik - c * k;
    const T* Ei = E + std::int64_t(c) * nd * nd + (std::int64_t(BS) * i + a) * nd;
    for (int j = 0; j < k; ++j)
    {
      const std::int64_t node = dm[std::int64_t(c) * k + j];
      for (int d = 0; d < BS; ++d)
      {
        const std::int64_t col = BS * node + d;
        if (!bc[col])
          continue;
        const double gv = g ? g[col] : 0.0;
        const double xv = x0 ? double(x0[col]) : 0.0;
        s += double(Ei[BS * j + d]) * (gv - xv);
      }
    }
  }
  if (s != 0.0)
    b[t] = static_cast<T>(double(b[t]) - alpha * s);
}

typedef int dfx_i32x4 __attribute__((ext_vector_type(4)));

/// Variant 12-14 (Q1 hexahedra whose vertices lie in at most 8 cells): one
/// thread per row with closed-form entries as variants 4-6, but the row's
/// plan sits at fixed positions -- its incidences c * 8 + i in cell order,
/// padded with -1 to 8 (inc8), and their 64 slot offsets (islot8) -- so it
/// is loaded together with the row bounds and the cell records of all 8
/// incidences are the second and last dependent loads, all in flight at
/// once. The plan streams are read and the slice is written with
/// non-temporal accesses (each byte is used once), which leaves the L2 to
/// the cell records that neighbouring rows share. Entries are summed in
/// cell order as in the other variants.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    row_fixed_k(std::int32_t nrows, std::int32_t nplan, const std::int32_t* __restrict__ inc8,
                const std::uint8_t* __restrict__ islot8, const T* __restrict__ cd,
                const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                T* __restrict__ vals, int accumulate)
{
  extern __shared__ unsigned char smem_raw[];
  T* acc = reinterpret_cast<T*>(smem_raw);
  const int R = blockDim.x;
  const std::int32_t r0 = xcd_contiguous(blockIdx.x, gridDim.x) * R;
  const std::int32_t r1 = min(r0 + R, nrows);
  const std::int32_t base = rp[r0];
  const std::int32_t len = rp[r1] - base;
  const std::int32_t r = r0 + threadIdx.x;
  const bool live = r < r1 && r < nplan;
  const std::int32_t rr = live ? r : 0;
  // the row's plan and bounds: independent loads
  const dfx_i32x4* ip = reinterpret_cast<const dfx_i32x4*>(inc8 + 8 * std::int64_t(rr));
  const dfx_i32x4 ia = __builtin_nontemporal_load(ip);
  const dfx_i32x4 ib = __builtin_nontemporal_load(ip + 1);
  const dfx_i32x4* sp = reinterpret_cast<const dfx_i32x4*>(islot8 + 64 * std::int64_t(rr));
  dfx_i32x4 sw[4];
#pragma unroll
  for (int h = 0; h < 4; ++h)
    sw[h] = __builtin_nontemporal_load(sp + h);
  const std::int32_t roff = (r < r1 ? rp[r] : base) - base;
  const bool skip = !live || (bc && bc[rr]);
  const std::int32_t cik[8] = {ia.x, ia.y, ia.z, ia.w, ib.x, ib.y, ib.z, ib.w};
  T g[8][7];
  std::uint32_t m[8];
#pragma unroll
  for (int u = 0; u < 8; ++u)
  {
    load_factors<T>(cd, cik[u] < 0 ? 0 : cik[u] >> 3, g[u], m[u]);
    if (cik[u] < 0 || skip)
      m[u] = 0xffu; // padding (or a skipped row): every column masked
  }
  for (int t = threadIdx.x; t < len; t += R)
    acc[t] = T(0);
  __syncthreads();
  T* a = acc + roff;
#pragma unroll
  for (int u = 0; u < 8; ++u)
  {
    const int i = cik[u] & 7;
    T gs[3];
    q1_signed<T>(g[u], i, gs);
    const std::uint64_t w = static_cast<std::uint64_t>(static_cast<std::uint32_t>(sw[u / 2][2 * (u % 2)]))
                            | (static_cast<std::uint64_t>(static_cast<std::uint32_t>(sw[u / 2][2 * (u % 2) + 1])) << 32);
    // column j = i ^ e: the 1D factors of q1_entry depend on e only, so
    // with e fixed at compile time an entry is seven products; the column
    // order differs from j order only in which slot is added first
    std::uint32_t sle[8];
    T v[8];
    std::uint32_t me = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e)
    {
      const int j = i ^ e;
      sle[e] = static_cast<std::uint32_t>(w >> (8 * j)) & 0xffu;
      me |= ((m[u] >> j) & 1u) << e;
      v[e] = q1_entry<T>(g[u], gs, 0, e);
    }
    add_slots<T, std::uint32_t, 8>(a, sle, v, me);
  }
  __syncthreads();
  if (accumulate)
    for (int t = threadIdx.x; t < len; t += R)
      vals[base + t] += acc[t];
  else
    for (int t = threadIdx.x; t < len; t += R)
      __builtin_nontemporal_store(acc[t], vals + base + t);
}

/// The fixed-width plan from the row-owner plan: for each row its
/// incidences (padded with -1 to 8) and their 8-byte slot offsets.
__global__ void __launch_bounds__(kBlock)
    fixed_plan_k(std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
                 const std::int32_t* __restrict__ inc, const std::uint8_t* __restrict__ islots,
                 std::int32_t* __restrict__ inc8, std::uint8_t* __restrict__ islot8)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  const std::int32_t p0 = inc_off[r], n = inc_off[r + 1] - p0;
  for (int u = 0; u < 8; ++u)
  {
    inc8[8 * std::int64_t(r) + u] = u < n ? inc[p0 + u] : -1;
    for (int j = 0; j < 8; ++j)
      islot8[64 * std::int64_t(r) + 8 * u + j] = u < n ? islots[8 * std::int64_t(p0 + u) + j] : 0;
  }
}

// ---------------------------------------------------------------------
// Batched assembly of many independent systems of one affine kernel (no
// counterpart in the reference): one launch forms the cell factors of all
// systems and one assembles the rows of all systems, from device tables of
// per-system pointers and of (system, first cell / first row) work items.
// ---------------------------------------------------------------------

template <typename T>
struct BatchSys
{
  const T* x;
  const std::int32_t* xdm;
  const std::int32_t* dm;
  const T* coeffs;
  const std::int8_t* bc;     // markers or nullptr
  T* cd;                     // cell factor records (8 per cell)
  const std::int32_t* inc_off;
  const std::int32_t* inc;
  const std::uint8_t* islots;
  const std::int32_t* rp;
  T* vals;
  T creac;
  std::int32_t cstride, nc, nrows, nplan, accumulate;
};

template <typename T, int NV>
__global__ void __launch_bounds__(kBlock)
    batch_factors_k(const BatchSys<T>* __restrict__ sys, const int2* __restrict__ work)
{
  const int2 wi = work[blockIdx.x];
  const BatchSys<T>& S = sys[wi.x];
  const std::int64_t c = wi.y + static_cast<std::int64_t>(threadIdx.x);
  if (c >= S.nc)
    return;
  const std::int32_t* v = S.xdm + NV * c;
  T g[7];
  affine_factors<T>(S.x + 3 * std::int64_t(v[0]), S.x + 3 * std::int64_t(v[1]),
                    S.x + 3 * std::int64_t(v[2]), S.x + 3 * std::int64_t(v[NV == 4 ? 3 : 4]),
                    S.coeffs[c * S.cstride], S.creac, g);
  std::uint32_t m = 0;
  if (S.bc)
    for (int j = 0; j < NV; ++j)
      if (S.bc[S.dm[c * NV + j]])
        m |= 1u << j;
  T* G = S.cd + 8 * c;
#pragma unroll
  for (int i = 0; i < 7; ++i)
    G[i] = g[i];
  G[7] = mask_bits<T>(m);
}

template <typename T, int K, bool CF>
__global__ void __launch_bounds__(kBlock)
    batch_rows_k(const BatchSys<T>* __restrict__ sys, const int2* __restrict__ work,
                 const T* __restrict__ ref)
{
  extern __shared__ unsigned char smem_raw[];
  const int2 wi = work[blockIdx.x];
  const BatchSys<T>& S = sys[wi.x];
  row_block<T, std::uint8_t, K, 4, CF>(wi.y, S.nrows, S.nplan, S.inc_off, S.inc, S.islots, S.cd,
                                       S.bc, S.rp, ref, S.vals, S.accumulate,
                                       reinterpret_cast<T*>(smem_raw));
}

/// Variant "fused geometry": as row_assemble_k, but every incidence forms
/// its cell's affine factors from the cell's vertices, coefficient and BC
/// markers (no cell pass and no factor array): per incidence the cell's
/// vertex row and kappa are read, then its four Jacobian vertices
/// (neighbouring rows share them in cache), instead of a 64-byte factor
/// record written by a separate pass. Loads are staged for eight
/// incidences at a time (indices, then vertex rows, then coordinates four
/// cells at a time) at clamped positions, so each stage is one round trip.
/// With SAME the dofmap is the vertex map (P1/Q1 on the mesh's own
/// numbering) and the BC markers are read through the vertex row.
template <typename T, typename O, int K, bool CF, bool SAME>
__global__ void __launch_bounds__(kBlock)
    row_fused_k(std::int32_t nrows, std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
                const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
                const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                const std::int32_t* __restrict__ dm, const T* __restrict__ coeffs, int cstride,
                T creac, const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                const T* __restrict__ ref, T* __restrict__ vals, int accumulate)
{
  static_assert(!CF || K == 8, "closed form is the Q1 hexahedron");
  static_assert(K == 4 || K == 8, "affine P1 tetrahedra or Q1 hexahedra");
  constexpr int NV = K;               // vertices per cell
  constexpr int JV = K == 4 ? 3 : 4;  // third Jacobian vertex
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw); // reference tensors [7][K][K]
  T* acc = CF ? S : S + 7 * K * K;
  const int R = blockDim.x;
  const std::int32_t r0 = xcd_contiguous(blockIdx.x, gridDim.x) * R;
  const std::int32_t r1 = min(r0 + R, nrows);
  const std::int32_t base = rp[r0];
  const std::int32_t len = rp[r1] - base;
  if constexpr (!CF)
    for (int t = threadIdx.x; t < 7 * K * K; t += R)
      S[t] = ref[t];
  for (int t = threadIdx.x; t < len; t += R)
    acc[t] = T(0);
  __syncthreads();
  const std::int32_t r = r0 + threadIdx.x;
  if (r < r1 && r < nplan && !(bc && bc[r]))
  {
    T* a = acc + (rp[r] - base);
    const std::int32_t p0 = inc_off[r], pe = inc_off[r + 1];
    for (std::int32_t pb = p0; pb < pe; pb += 8)
    {
      // stage 1: incidences and slot offsets
      std::int32_t cik[8];
      O sl[8][K];
#pragma unroll
      for (int u = 0; u < 8; ++u)
      {
        const std::int32_t p = min(pb + u, pe - 1);
        cik[u] = inc[p];
        load_slots<O, K>(islots, p, sl[u]);
      }
      // stage 2: vertex rows, kappa, BC column masks
      std::int32_t vid[8][4];
      T kap[8];
      std::uint32_t m[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
      {
        const std::int64_t c = cik[u] / K;
        const int4 va = *reinterpret_cast<const int4*>(xdm + NV * c);
        int row[K];
        row[0] = va.x; row[1] = va.y; row[2] = va.z; row[3] = va.w;
        if constexpr (K == 8)
        {
          const int4 vb = *reinterpret_cast<const int4*>(xdm + NV * c + 4);
          row[4] = vb.x; row[5] = vb.y; row[6] = vb.z; row[7] = vb.w;
        }
        vid[u][0] = row[0];
        vid[u][1] = row[1];
        vid[u][2] = row[2];
        vid[u][3] = row[JV];
        kap[u] = coeffs[c * cstride];
        std::uint32_t mm = 0;
        if (bc)
        {
#pragma unroll
          for (int j = 0; j < K; ++j)
            if (bc[SAME ? row[j] : dm[c * K + j]])
              mm |= 1u << j;
        }
        m[u] = mm;
      }
      // stage 3: coordinates and factors, four cells at a time
#pragma unroll
      for (int h = 0; h < 2; ++h)
      {
        T q[4][4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int w = 0; w < 4; ++w)
#pragma unroll
            for (int d = 0; d < 3; ++d)
              q[u][w][d] = x[3 * std::int64_t(vid[4 * h + u][w]) + d];
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
          const int uu = 4 * h + u;
          if (pb + uu >= pe)
            continue;
          T g[7];
          affine_factors<T>(q[u][0], q[u][1], q[u][2], q[u][3], kap[uu], creac, g);
          const int i = cik[uu] % K;
          if constexpr (CF)
          {
            T gs[3];
            q1_signed<T>(g, i, gs);
            T v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
              v[j] = q1_entry<T>(g, gs, i, j);
            add_slots<T, O, 8>(a, sl[uu], v, m[uu]);
          }
          else
          {
            T v[K];
#pragma unroll
            for (int j = 0; j < K; ++j)
              v[j] = element_entry<T, K>(g, S, i, j);
            add_slots<T, O, K>(a, sl[uu], v, m[uu]);
          }
        }
      }
    }
  }
  __syncthreads();
  if (accumulate)
    for (int t = threadIdx.x; t < len; t += R)
      vals[base + t] += acc[t];
  else
    for (int t = threadIdx.x; t < len; t += R)
      vals[base + t] = acc[t];
}

/// Variant "lanes": L lanes per row, one incidence per lane, so every lane
/// has a single short chain of loads; the lanes of a row add their entries
/// into the row's LDS slice with LDS atomics.
template <typename T, typename O, int K, int L>
__global__ void __launch_bounds__(kBlock)
    row_assemble_lanes_k(std::int32_t nrows, std::int32_t nplan,
                         const std::int32_t* __restrict__ inc_off,
                         const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
                         const T* __restrict__ cd, const std::uint32_t* __restrict__ mask,
                         const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                         const T* __restrict__ ref, T* __restrict__ vals, int accumulate)
{
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw); // reference tensors [7][K][K]
  T* acc = S + 7 * K * K;
  constexpr int RPB = kBlock / L;
  const std::int32_t r0 = xcd_contiguous(blockIdx.x, gridDim.x) * RPB;
  const std::int32_t r1 = min(r0 + RPB, nrows);
  const std::int32_t base = rp[r0];
  const std::int32_t len = rp[r1] - base;
  for (int t = threadIdx.x; t < 7 * K * K; t += kBlock)
    S[t] = ref[t];
  for (int t = threadIdx.x; t < len; t += kBlock)
    acc[t] = T(0);
  __syncthreads();
  const std::int32_t r = r0 + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (r < r1 && r < nplan && !(bc && bc[r]))
  {
    T* a = acc + (rp[r] - base);
    const std::int32_t pe = inc_off[r + 1];
    for (std::int32_t p = inc_off[r] + lane; p < pe; p += L)
    {
      const std::int32_t cik = inc[p];
      const std::int64_t c = cik / K;
      const int i = cik - static_cast<int>(c) * K;
      T g[7];
      std::uint32_t m;
      load_factors<T>(cd, c, g, m);
      O sl[K];
      load_slots<O, K>(islots, p, sl);
#pragma unroll
      for (int j = 0; j < K; ++j)
        if (!((m >> j) & 1u))
          atomicAdd(a + sl[j], element_entry<T, K>(g, S, i, j));
    }
  }
  __syncthreads();
  if (accumulate)
    for (int t = threadIdx.x; t < len; t += kBlock)
      vals[base + t] += acc[t];
  else
    for (int t = threadIdx.x; t < len; t += kBlock)
      vals[base + t] = acc[t];
}

/// Lifting b[r] -= alpha sum_{(c,i) at r} sum_{j in BC} A_ij (g_j - x0_j),
/// one thread per row, over cells whose mask marks a BC column.
template <typename T, int K, bool COMPACT = false>
__global__ void __launch_bounds__(kBlock)
    row_lift_k(std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
               const std::int32_t* __restrict__ inc, const std::int32_t* __restrict__ dm,
               const T* __restrict__ cd, const std::uint32_t* __restrict__ mask,
               const T* __restrict__ ref, const double* __restrict__ g,
               const T* __restrict__ x0, double alpha, T* __restrict__ b)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  double s = 0.0;
  const std::int32_t p0 = inc_off[r], p1 = inc_off[r + 1];
  // eight incidences at a time: their cells and masks loaded together
  // (clamped positions), then the cells with a Dirichlet column in order
  for (std::int32_t pb = p0; pb < p1; pb += 8)
  {
    std::int32_t cq[8];
    std::uint32_t mq[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      cq[u] = inc[min(pb + u, p1 - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      mq[u] = mask[cq[u] / K];
    for (int u = 0; u < 8 && pb + u < p1; ++u)
    {
      const std::uint32_t m = mq[u];
      if (!m)
        continue;
      const std::int32_t cik = cq[u];
      const std::int32_t c = cik / K;
      const int i = cik - c * K;
      T G[7];
      if constexpr (COMPACT)
      {
        const T* rc = cd + 4 * std::int64_t(c);
        G[0] = rc[0]; G[1] = rc[1]; G[2] = rc[2]; G[6] = rc[3];
        G[3] = G[4] = G[5] = T(0);
      }
      else
      {
        const T* rc = cd + 8 * std::int64_t(c);
#pragma unroll
        for (int q = 0; q < 7; ++q)
          G[q] = rc[q];
      }
      const T* Si = ref + i * K;
      for (int j = 0; j < K; ++j)
      {
        if (!((m >> j) & 1u))
          continue;
        const T v = G[0] * Si[j] + G[1] * Si[K * K + j] + G[2] * Si[2 * K * K + j]
                    + G[3] * Si[3 * K * K + j] + G[4] * Si[4 * K * K + j]
                    + G[5] * Si[5 * K * K + j] + G[6] * Si[6 * K * K + j];
        const std::int32_t col = dm[std::int64_t(c) * K + j];
        const double gv = g ? g[col] : 0.0;
        const double xv = x0 ? double(x0[col]) : 0.0;
        s += double(v) * (gv - xv);
      }
    }
  }
  if (s != 0.0)
    b[r] = static_cast<T>(double(b[r]) - alpha * s);
}

typedef double dfx_double4 __attribute__((ext_vector_type(4)));
typedef float dfx_float4 __attribute__((ext_vector_type(4)));

/// One 16x16x4 matrix-core product-accumulate, acc += A B, in FP64 or
/// FP32. Operand layout (one wavefront, lane l): a = A[l % 16][l / 16],
/// b = B[l / 16][l % 16]; acc[r] = D[mfma16_row<T>(l / 16, r)][l % 16]:
/// row 4 (l / 16) + r in FP32, row 4 r + l / 16 in FP64.
__device__ inline dfx_double4 mfma16(double a, double b, dfx_double4 acc)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ inline dfx_float4 mfma16(float a, float b, dfx_float4 acc)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}
template <typename T>
__device__ inline int mfma16_row(int kq, int r)
{
  return sizeof(T) == 8 ? 4 * r + kq : 4 * kq + r;
}

/// Element matrices of all cells as one batched product on the matrix
/// cores: E (cells x K*K) = F (cells x 8) * T (8 x K*K), where F holds the
/// cell factors (G_0..G_5, c|det J|, 0) and T the reference tensors
/// [7][K][K] with a zero eighth row. One wavefront forms 16 cells: two
/// 16x16x4 products (k = 0..3, 4..7) per 16 columns of E.
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
    elem_mfma_k(std::int32_t nc, const T* __restrict__ cd, const T* __restrict__ ref,
                T* __restrict__ elem)
{
  constexpr int KK = K * K;
  constexpr int NT = (KK + 15) / 16;
  using V4 = decltype(mfma16(T(), T(), {}));
  __shared__ T tab[8 * NT * 16];
  for (int t = threadIdx.x; t < 8 * NT * 16; t += kBlock)
  {
    const int g = t / (NT * 16), e = t % (NT * 16);
    tab[t] = (g < 7 && e < KK) ? ref[g * KK + e] : T(0);
  }
  __syncthreads();
  const int l = threadIdx.x & 63;
  const std::int64_t c0
      = (std::int64_t(xcd_contiguous(blockIdx.x, gridDim.x)) * (kBlock / 64) + (threadIdx.x >> 6))
        * 16;
  if (c0 >= nc)
    return;
  const std::int64_t ca = c0 + (l & 15);
  const int kq = l >> 4;
  // column 7 of F is zero (the factor data keeps the BC mask bits there)
  const T a0 = ca < nc ? cd[8 * ca + kq] : T(0);
  const T a1 = (ca < nc && kq < 3) ? cd[8 * ca + 4 + kq] : T(0);
#pragma unroll 5
  for (int t = 0; t < NT; ++t)
  {
    const int e = 16 * t + (l & 15);
    V4 acc = {0, 0, 0, 0};
    acc = mfma16(a0, tab[kq * NT * 16 + e], acc);
    acc = mfma16(a1, tab[(4 + kq) * NT * 16 + e], acc);
    if (e < KK)
    {
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const std::int64_t c = c0 + mfma16_row<T>(kq, r);
        if (c < nc)
          elem[c * KK + e] = acc[r];
      }
    }
  }
}

/// Incidence-parallel row assembly: a workgroup owns a block of consecutive
/// rows holding at most one incidence per thread. Every thread loads its
/// incidence's element row (from the element matrices E when MAT, else from
/// the cell factors and the reference tensors in LDS), then the rows'
/// incidences are added into the LDS slice in rounds by their rank in the
/// row, so each entry is summed in cell order as in the other variants.
template <typename T, typename O, int K, bool MAT>
__global__ void __launch_bounds__(kBlock)
    row_gather_k(const std::int32_t* __restrict__ gblk, const std::int32_t* __restrict__ inc_off,
                 const std::int32_t* __restrict__ inc, const std::int32_t* __restrict__ inc_row,
                 const O* __restrict__ islots, const T* __restrict__ cd,
                 const T* __restrict__ elem, const std::uint32_t* __restrict__ mask,
                 const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                 const T* __restrict__ ref, T* __restrict__ vals, int accumulate, int ngblk,
                 std::int32_t nplan)
{
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw); // reference tensors [7][K][K] unless MAT
  T* acc = MAT ? S : S + 7 * K * K;
  const int b = xcd_contiguous(blockIdx.x, gridDim.x);
  const std::int32_t R0 = gblk[b], R1 = gblk[b + 1];
  const std::int32_t base = rp[R0];
  const std::int32_t len = rp[R1] - base;
  if constexpr (!MAT)
    for (int t = threadIdx.x; t < 7 * K * K; t += kBlock)
      S[t] = ref[t];
  for (int t = threadIdx.x; t < len; t += kBlock)
    acc[t] = T(0);
  // rows from nplan on have no incidences (inc_off has nplan + 1 entries)
  const std::int32_t pb = inc_off[min(R0, nplan)], pe = inc_off[min(R1, nplan)];
  const std::int32_t p = pb + threadIdx.x;
  bool live = p < pe;
  std::int32_t r = 0, rank = -1;
  T e[K];
  O sl[K];
  std::uint32_t m = 0;
  T g[7];
  int i = 0;
  if (pe > pb)
  {
    // every load at a clamped position and none behind a branch: the
    // incidence's row, cell, slots and element row (or cell factors) are
    // in flight together
    const std::int32_t pc = min(p, pe - 1);
    r = inc_row[pc];
    const std::int32_t cik = inc[pc];
    load_slots<O, K>(islots, pc, sl);
    const std::int64_t c = cik / K;
    i = cik - static_cast<int>(c) * K;
    m = mask ? mask[c] : 0u;
    if constexpr (MAT)
    {
      const T* er = elem + std::int64_t(cik) * K;
      if constexpr ((K * sizeof(T)) % 16 == 0)
      {
        // element rows start 16-byte aligned: two values per load
        using T2 = vec2_t<T>;
        const T2* er2 = reinterpret_cast<const T2*>(er);
#pragma unroll
        for (int j = 0; j < K / 2; ++j)
        {
          const T2 v2 = er2[j];
          e[2 * j] = v2.x;
          e[2 * j + 1] = v2.y;
        }
      }
      else
      {
#pragma unroll
        for (int j = 0; j < K; ++j)
          e[j] = er[j];
      }
    }
    else
    {
      std::uint32_t mf;
      load_factors<T>(cd, c, g, mf);
    }
    rank = p - inc_off[r];
    live = live && !(bc && bc[r]);
  }
  __syncthreads();
  if constexpr (!MAT)
    if (pe > pb)
    {
      // entries from the factors and the tensors staged in LDS above
#pragma unroll
      for (int j = 0; j < K; ++j)
        e[j] = element_entry<T, K>(g, S, i, j);
    }
  T* a = acc + (live ? rp[r] - base : 0);
  const int rounds = gblk[ngblk + 1 + b];
  for (int k = 0; k < rounds; ++k)
  {
    if (live && rank == k)
      add_slots<T, O, K>(a, sl, e, m);
    __syncthreads();
  }
  if (accumulate)
    for (int t = threadIdx.x; t < len; t += kBlock)
      vals[base + t] += acc[t];
  else
    for (int t = threadIdx.x; t < len; t += kBlock)
      vals[base + t] = acc[t];
}

/// Variant 16 (forms with reference tensors, K <= 32 nodes per cell): a
/// team of K lanes per row, 64 / K teams per wavefront, no element matrix
/// stored. The team walks its row's incidences (cell c, local row i) in
/// cell order; lane j forms the entry of column j from the incidence's
/// cell factors and the reference tensors in LDS and adds it into the
/// row's LDS slice at the slot of column j. The K slots of an incidence are
/// distinct, so the team's adds need no atomics and every entry is summed
/// in cell order, with the additions of variant 8. The cell factors of the
/// row's incidences are loaded lane-parallel, K incidences at a time (lane
/// j holds incidence bK + j of batch b), and handed to the team through an
/// 8-value LDS slot per team; each step's slots are loaded one step ahead. The wavefronts of a
/// workgroup (blockDim.x / 64 of them) share one copy of the tensors. Rows
/// come ordered by incidence count inside chunks of 4096 consecutive rows
/// (trow), so the teams of a wavefront take similar numbers of steps.
template <typename T, typename O, int K>
__global__ void __launch_bounds__(1024, 8)
    row_team_k(std::int32_t ntrow, const int4* __restrict__ trow,
               const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
               const T* __restrict__ cd, const std::uint32_t* __restrict__ mask,
               const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
               const T* __restrict__ ref, T* __restrict__ vals, int accumulate,
               std::int32_t maxrow)
{
  constexpr int G = 64 / K;
  static_assert(K <= 24, "the BC mask bits share a word with the local row");
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw); // reference tensors [7][K][K]
  for (int t = threadIdx.x; t < 7 * K * K; t += blockDim.x)
    S[t] = ref[t];
  const int nw = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int team = lane / K, j = lane - team * K;
  const bool member = team < G;
  const std::int64_t tq
      = (std::int64_t(xcd_contiguous(blockIdx.x, gridDim.x)) * nw + wave) * G + team;
  int4 ro = make_int4(-1, 0, 0, 0);
  if (member && tq < ntrow)
    ro = trow[tq];
  const std::int32_t r = ro.x, p0 = ro.y, n = ro.z;
  const std::size_t slot = std::size_t(wave) * G + (member ? team : 0);
  T* a = S + 7 * K * K + slot * maxrow;
  // the team's broadcast slot (8 values, 16-byte aligned) after the slices
  const std::size_t slices = (std::size_t(nw) * G * maxrow + 3) & ~std::size_t(3);
  T* bs = S + 7 * K * K + slices + slot * 8;
  std::int32_t rbase = 0, rlen = 0;
  bool skip = true;
  if (r >= 0)
  {
    rbase = rp[r];
    rlen = rp[r + 1] - rbase;
    skip = (bc && bc[r]) || n == 0;
  }
  // one batch of the row's incidences: lane j holds incidence q = b K + j
  // (its cell factors, and its local row i above the BC mask bits)
  T bg[7] = {};
  std::uint32_t bmi = 0;
  auto load_batch = [&](int b)
  {
    const int q = b * K + j;
    if (q < n)
    {
      const std::int32_t cik = inc[p0 + q];
      const std::int32_t c = cik / K;
      std::uint32_t mf;
      load_factors<T>(cd, c, bg, mf);
      bmi = (mask ? mask[c] : 0u) | (static_cast<std::uint32_t>(cik - c * K) << 24);
    }
  };
  O sl_next = 0;
  if (!skip)
  {
    load_batch(0);
    sl_next = islots[std::int64_t(p0) * K + j];
  }
  for (int t = j; t < rlen; t += K)
    a[t] = T(0);
  __syncthreads();
  if (!skip)
  {
    const int tb = team * K;
    for (int st = 0; st < n; ++st)
    {
      const int bq = st / K, src = tb + (st - bq * K);
      if (st > 0 && st == bq * K)
        load_batch(bq);
      const O sl = sl_next;
      if (st + 1 < n)
        sl_next = islots[std::int64_t(p0 + st + 1) * K + j];
      // the lane holding incidence st writes its factors into the team's
      // slot and the team reads them back (LDS accesses of one wavefront
      // complete in order): four 16-byte writes and four broadcast reads
      // instead of fifteen shuffles
      (void)src;
      using T2 = vec2_t<T>;
      T2* b2 = reinterpret_cast<T2*>(bs);
      // the slot passes values between lanes: wavefront-scope fences and a
      // wave barrier keep every lane's read of the previous step before
      // this write, and this write before every lane's read below (without
      // them the compiler may move a lane's read across another lane's
      // write)
      wave_sync();
      if (j == st - bq * K)
      {
        T w8[8] = {bg[0], bg[1], bg[2], bg[3], bg[4], bg[5], bg[6], T(0)};
        if constexpr (sizeof(T) == 8)
          w8[7] = __longlong_as_double(static_cast<long long>(bmi));
        else
          w8[7] = __uint_as_float(bmi);
#pragma unroll
        for (int h = 0; h < 4; ++h)
          b2[h] = T2{w8[2 * h], w8[2 * h + 1]};
      }
      wave_sync();
      T g[7];
      std::uint32_t mi;
      {
        T r8[8];
#pragma unroll
        for (int h = 0; h < 4; ++h)
        {
          const T2 v = b2[h];
          r8[2 * h] = v.x;
          r8[2 * h + 1] = v.y;
        }
#pragma unroll
        for (int u = 0; u < 7; ++u)
          g[u] = r8[u];
        if constexpr (sizeof(T) == 8)
          mi = static_cast<std::uint32_t>(__double_as_longlong(r8[7]));
        else
          mi = __float_as_uint(r8[7]);
      }
      const T e = element_entry<T, K>(g, S, static_cast<int>(mi >> 24), j);
      if (!((mi >> j) & 1u))
        a[sl] += e;
    }
  }
  // the team stores its own row (the whole slice: zeros for skipped rows),
  // whose entries other lanes of the team added
  wave_sync();
  if (r >= 0)
  {
    if (accumulate)
      for (int t = j; t < rlen; t += K)
        vals[rbase + t] += a[t];
    else
      for (int t = j; t < rlen; t += K)
        __builtin_nontemporal_store(a[t], vals + rbase + t);
  }
}

/// Variant 17: the rows of variant 16 (a team of K lanes per row, rows in
/// trow order) gathering the element matrices E of variant 7 instead of
/// forming entries: lane j of the team adds E[c][i][j] of each incidence
/// (c, i) into the row's slice at the slot of column j, in cell order (the
/// additions of variant 7). A step's element value, slot and mask bit are
/// loaded four steps ahead; the first 2K incidences of the row are loaded
/// lane-parallel and handed out by shuffles.
template <typename T, typename O, int K>
__global__ void __launch_bounds__(1024, 8)
    row_team_gather_k(std::int32_t ntrow, const int4* __restrict__ trow,
                      const std::int32_t* __restrict__ inc, const O* __restrict__ islots,
                      const T* __restrict__ elem, const std::uint32_t* __restrict__ mask,
                      const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                      T* __restrict__ vals, int accumulate, std::int32_t maxrow)
{
  constexpr int G = 64 / K;
  extern __shared__ unsigned char smem_raw[];
  const int nw = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int team = lane / K, j = lane - team * K;
  const bool member = team < G;
  const std::int64_t tq
      = (std::int64_t(xcd_contiguous(blockIdx.x, gridDim.x)) * nw + wave) * G + team;
  int4 ro = make_int4(-1, 0, 0, 0);
  if (member && tq < ntrow)
    ro = trow[tq];
  const std::int32_t r = ro.x, p0 = ro.y, n = ro.z;
  T* a = reinterpret_cast<T*>(smem_raw) + (std::size_t(wave) * G + (member ? team : 0)) * maxrow;
  std::int32_t rbase = 0, rlen = 0;
  bool skip = true;
  std::int32_t pre0 = 0, pre1 = 0;
  if (r >= 0)
  {
    rbase = rp[r];
    rlen = rp[r + 1] - rbase;
    skip = (bc && bc[r]) || n == 0;
    if (!skip)
    {
      pre0 = inc[p0 + min(j, n - 1)];
      pre1 = inc[p0 + min(K + j, n - 1)];
    }
  }
  for (int t = j; t < rlen; t += K)
    a[t] = T(0);
  __syncthreads();
  if (!skip)
  {
    const int tb = team * K;
    // incidence st of the row (every lane of the team calls it with the
    // same st, so the shuffles read lanes of the team only)
    auto cik_at = [&](int st) -> std::int32_t
    {
      if (st < 2 * K)
      {
        const int src = tb + (st < K ? st : st - K);
        const std::int32_t v0 = __shfl(pre0, src, 64);
        const std::int32_t v1 = __shfl(pre1, src, 64);
        return st < K ? v0 : v1;
      }
      return inc[p0 + st];
    };
    struct Step
    {
      T e;
      O sl;
      std::uint32_t off; // 1: column j of this incidence is a Dirichlet dof
    };
    auto fetch = [&](Step& q, int st)
    {
      if (st < n)
      {
        const std::int32_t cik = cik_at(st);
        q.e = elem[std::int64_t(cik) * K + j];
        q.sl = islots[std::int64_t(p0 + st) * K + j];
        q.off = mask ? (mask[cik / K] >> j) & 1u : 0u;
      }
    };
    auto add = [&](const Step& q)
    {
      if (!q.off)
        a[q.sl] += q.e;
    };
    Step q0{}, q1{}, q2{}, q3{};
    fetch(q0, 0);
    fetch(q1, 1);
    fetch(q2, 2);
    fetch(q3, 3);
    for (int st = 0; st < n; st += 4)
    {
      add(q0);
      fetch(q0, st + 4);
      if (st + 1 >= n)
        break;
      add(q1);
      fetch(q1, st + 5);
      if (st + 2 >= n)
        break;
      add(q2);
      fetch(q2, st + 6);
      if (st + 3 >= n)
        break;
      add(q3);
      fetch(q3, st + 7);
    }
  }
  // the row's entries were added by other lanes of the team
  wave_sync();
  if (r >= 0)
  {
    if (accumulate)
      for (int t = j; t < rlen; t += K)
        vals[rbase + t] += a[t];
    else
      for (int t = j; t < rlen; t += K)
        __builtin_nontemporal_store(a[t], vals + rbase + t);
  }
}

/// BC column masks of element-tensor forms: bit j of mask[c] is set when
/// node j of cell c is a Dirichlet dof.
__global__ void __launch_bounds__(kBlock)
    element_mask_k(std::int32_t nc, int k, const std::int32_t* __restrict__ dm,
                   const std::int8_t* __restrict__ bc, std::uint32_t* __restrict__ mask)
{
  const std::int64_t c = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  std::uint32_t m = 0;
  for (int j = 0; j < k; ++j)
    if (bc[dm[c * k + j]])
      m |= 1u << j;
  mask[c] = m;
}

/// b += element vectors supplied by the caller (e: cells x k, cell-major),
/// one thread per row adding its incidences in cell order onto b[r], the
/// order of the reference's cell loop.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    element_vector_k(std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
                     const std::int32_t* __restrict__ inc, const T* __restrict__ e,
                     T* __restrict__ b)
{
  const std::int32_t r = xcd_contiguous(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  const std::int32_t p1 = inc_off[r + 1];
  T s = b[r];
  for (std::int32_t p = inc_off[r]; p < p1; ++p)
    s += e[inc[p]];
  b[r] = s;
}

/// Lifting with caller-supplied element matrices (E: cells x k x k):
/// b[r] -= alpha sum_{(c,i) at r} sum_{j in BC} E_c[i][j] (g_j - x0_j), one
/// thread per row over the cells whose mask marks a BC column.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    element_lift_k(std::int32_t nplan, int k, const std::int32_t* __restrict__ inc_off,
                   const std::int32_t* __restrict__ inc, const std::int32_t* __restrict__ dm,
                   const T* __restrict__ E, const std::uint32_t* __restrict__ mask,
                   const double* __restrict__ g, const T* __restrict__ x0, double alpha,
                   T* __restrict__ b)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  double s = 0.0;
  for (std::int32_t p = inc_off[r]; p < inc_off[r + 1]; ++p)
  {
    const std::int32_t cik = inc[p];
    const std::int32_t c = cik / k;
    const std::uint32_t m = mask[c];
    if (!m)
      continue;
    const T* er = E + std::int64_t(cik) * k;
    for (int j = 0; j < k; ++j)
    {
      if (!((m >> j) & 1u))
        continue;
      const std::int32_t col = dm[std::int64_t(c) * k + j];
      const double gv = g ? g[col] : 0.0;
      const double xv = x0 ? double(x0[col]) : 0.0;
      s += double(er[j]) * (gv - xv);
    }
  }
  if (s != 0.0)
    b[r] = static_cast<T>(double(b[r]) - alpha * s);
}

// ---------------------------------------------------------------------
// Exterior facet loads g v ds with g given at the cell's vertices (P1 on a
// tetrahedron's facet triangle, Q1 on a hexahedron's facet
// parallelogram). Facet vertices in DOLFINx sub-entity order.
// ---------------------------------------------------------------------

__constant__ int kTetFacetV[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
__constant__ int kHexFacetV[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                     {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};

/// One thread per (cell, local facet) pair e: the facet's contribution to
/// each of its vertices, out[4 e + a] = int_F g phi_a ds: the facet mass
/// matrix times g, |F| (1 + d_ab) / 12 on triangles and |F| M_ab / 36 with
/// M the tensor product of [2 1; 1 2] on parallelograms.
template <typename T, bool HEX>
__global__ void __launch_bounds__(kBlock)
    facet_load_k(std::int32_t npairs, const std::int32_t* __restrict__ pairs,
                 const T* __restrict__ x, const std::int32_t* __restrict__ xdm, int npc,
                 const T* __restrict__ coeffs, int cstride, T* __restrict__ out)
{
  const std::int32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= npairs)
    return;
  constexpr int NF = HEX ? 4 : 3;
  const std::int32_t c = pairs[2 * e];
  const int lf = pairs[2 * e + 1];
  T X[NF][3], g[NF];
#pragma unroll
  for (int a = 0; a < NF; ++a)
  {
    const int lv = HEX ? kHexFacetV[lf][a] : kTetFacetV[lf][a];
    const std::int64_t v = xdm[std::int64_t(c) * npc + lv];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      X[a][d] = x[3 * v + d];
    g[a] = coeffs[std::int64_t(c) * cstride + lv];
  }
  T t0[3], t1[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    t0[d] = X[1][d] - X[0][d];
    t1[d] = X[2][d] - X[0][d];
  }
  const T n0 = t0[1] * t1[2] - t0[2] * t1[1];
  const T n1 = t0[2] * t1[0] - t0[0] * t1[2];
  const T n2 = t0[0] * t1[1] - t0[1] * t1[0];
  const T nrm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if constexpr (HEX)
  {
    constexpr T M[4][4] = {{4, 2, 2, 1}, {2, 4, 1, 2}, {2, 1, 4, 2}, {1, 2, 2, 4}};
    const T w = nrm / T(36);
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
      T s = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        s += w * M[a][b] * g[b];
      out[4 * std::int64_t(e) + a] = s;
    }
  }
  else
  {
    const T area = T(0.5) * nrm;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
      T s = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b)
        s += area * (a == b ? T(2) : T(1)) / T(12) * g[b];
      out[4 * std::int64_t(e) + a] = s;
    }
    out[4 * std::int64_t(e) + 3] = T(0);
  }
}

/// One thread per (cell, local facet) pair e: the facet matrix of
/// beta u v ds, out[16 e + 4 a + b] = beta |F| (1 + d_ab) / 12 on triangles
/// (a, b < 3, the rest zero) and beta |F| M_ab / 36 on parallelograms, with
/// beta the cell's first packed coefficient.
template <typename T, bool HEX>
__global__ void __launch_bounds__(kBlock)
    facet_matrix_k(std::int32_t npairs, const std::int32_t* __restrict__ pairs,
                   const T* __restrict__ x, const std::int32_t* __restrict__ xdm, int npc,
                   const T* __restrict__ coeffs, int cstride, T* __restrict__ out)
{
  const std::int32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= npairs)
    return;
  const std::int32_t c = pairs[2 * e];
  const int lf = pairs[2 * e + 1];
  T X[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
  {
    const int lv = HEX ? kHexFacetV[lf][a] : kTetFacetV[lf][a];
    const std::int64_t v = xdm[std::int64_t(c) * npc + lv];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      X[a][d] = x[3 * v + d];
  }
  const T beta = coeffs[std::int64_t(c) * cstride];
  T t0[3], t1[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    t0[d] = X[1][d] - X[0][d];
    t1[d] = X[2][d] - X[0][d];
  }
  const T n0 = t0[1] * t1[2] - t0[2] * t1[1];
  const T n1 = t0[2] * t1[0] - t0[0] * t1[2];
  const T n2 = t0[0] * t1[1] - t0[1] * t1[0];
  const T nrm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  T* o = out + 16 * std::int64_t(e);
  if constexpr (HEX)
  {
    constexpr T M[4][4] = {{4, 2, 2, 1}, {2, 4, 1, 2}, {2, 1, 4, 2}, {1, 2, 2, 4}};
    const T w = beta * (nrm / T(36));
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        o[4 * a + b] = w * M[a][b];
  }
  else
  {
    const T area = T(0.5) * nrm;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        o[4 * a + b] = (a < 3 && b < 3) ? beta * area * (a == b ? T(2) : T(1)) / T(12) : T(0);
  }
}

/// A[r, :] += the facet matrices' rows at row r, facet by facet in facet
/// order (one thread per row with facet incidences, rows[i], over its
/// incidences finc = 4 e + facet vertex); entries in a BC row or column add zero. The column's slot is
/// found by binary search in the row; an absent one sets *err (err may be
/// null once the pattern is known to hold every entry).
template <typename T, bool HEX>
__global__ void __launch_bounds__(kBlock)
    facet_matrix_rows_k(std::int32_t nrows, const std::int32_t* __restrict__ rows,
                        const std::int32_t* __restrict__ off,
                        const std::int32_t* __restrict__ finc,
                        const std::int32_t* __restrict__ pairs,
                        const std::int32_t* __restrict__ dofmap, int k,
                        const T* __restrict__ fm, const std::int8_t* __restrict__ mark,
                        const std::int32_t* __restrict__ row_ptr,
                        const std::int32_t* __restrict__ cols, T* __restrict__ values,
                        int* __restrict__ err, std::uint8_t* __restrict__ fslot)
{
  const std::int32_t ir = blockIdx.x * kBlock + threadIdx.x;
  if (ir >= nrows)
    return;
  const std::int32_t r = rows[ir];
  constexpr int NF = HEX ? 4 : 3;
  const std::int32_t p1 = off[r + 1];
  if (off[r] == p1)
    return;
  const std::int32_t rb = row_ptr[r], re = row_ptr[r + 1];
  const bool rmask = mark && mark[r];
  for (std::int32_t p = off[r]; p < p1; ++p)
  {
    const std::int32_t e = finc[p] >> 2;
    const int a = finc[p] & 3;
    const std::int32_t c = pairs[2 * e];
    const int lf = pairs[2 * e + 1];
#pragma unroll
    for (int b = 0; b < NF; ++b)
    {
      const int lv = HEX ? kHexFacetV[lf][b] : kTetFacetV[lf][b];
      const std::int32_t col = dofmap[std::int64_t(c) * k + lv];
      std::int32_t lo = rb, hi = re;
      while (lo < hi)
      {
        const std::int32_t mid = (lo + hi) >> 1;
        if (cols[mid] < col)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo == re || cols[lo] != col)
      {
        if (err)
          *err = 1;
        continue;
      }
      if (fslot)
        fslot[4 * std::int64_t(p) + b] = static_cast<std::uint8_t>(lo - rb);
      const bool zero = rmask || (mark && mark[col]);
      values[lo] += zero ? T(0) : fm[16 * std::int64_t(e) + 4 * a + b];
    }
  }
}

/// Lifting by the facet matrices: b[r] -= alpha sum over the row's facet
/// incidences and the facet's BC columns j of F_rj (g_j - x0_j).
template <typename T, bool HEX>
__global__ void __launch_bounds__(kBlock)
    facet_lift_rows_k(std::int32_t nplan, const std::int32_t* __restrict__ off,
                      const std::int32_t* __restrict__ finc,
                      const std::int32_t* __restrict__ pairs,
                      const std::int32_t* __restrict__ dofmap, int k,
                      const T* __restrict__ fm, const std::int8_t* __restrict__ mark,
                      const double* __restrict__ g, const T* __restrict__ x0, double alpha,
                      T* __restrict__ bvec)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  constexpr int NF = HEX ? 4 : 3;
  double s = 0;
  for (std::int32_t p = off[r]; p < off[r + 1]; ++p)
  {
    const std::int32_t e = finc[p] >> 2;
    const int a = finc[p] & 3;
    const std::int32_t c = pairs[2 * e];
    const int lf = pairs[2 * e + 1];
#pragma unroll
    for (int b = 0; b < NF; ++b)
    {
      const int lv = HEX ? kHexFacetV[lf][b] : kTetFacetV[lf][b];
      const std::int32_t col = dofmap[std::int64_t(c) * k + lv];
      if (!mark[col])
        continue;
      const double gv = g ? g[col] : 0.0;
      const double xv = x0 ? double(x0[col]) : 0.0;
      s += double(fm[16 * std::int64_t(e) + 4 * a + b]) * (gv - xv);
    }
  }
  if (s != 0.0)
    bvec[r] = static_cast<T>(double(bvec[r]) - alpha * s);
}

// ---------------------------------------------------------------------
// Interior facet integrals on P1 tetrahedra: the macro element of facet e
// has the 4 dofs of its first cell then the 4 of its second; its 8 x 8
// matrix is added entry by entry like mat_set(dmapjoint0, dmapjoint1, Ae)
// (shared dofs appear twice). Incidences finc = 8 e + macro index.
// ---------------------------------------------------------------------

/// P1 gradients of a tetrahedron (rows of the inverse Jacobian, lambda_0 =
/// 1 - lambda_1 - lambda_2 - lambda_3).
template <typename T>
__device__ inline void p1_gradients(const T X[4][3], T g[4][3])
{
  T J[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int a = 0; a < 3; ++a)
      J[r][a] = X[a + 1][r] - X[0][r];
  const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const T det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const T id = T(1) / det;
  // K = J^-1: row a is grad lambda_{a+1}
  const T K[3][3] = {{c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                      (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
                     {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
                      (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
                     {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                      (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    g[1][d] = K[0][d];
    g[2][d] = K[1][d];
    g[3][d] = K[2][d];
    g[0][d] = -(K[0][d] + K[1][d] + K[2][d]);
  }
}

/// One thread per interior facet e: out[64 e + 8 a + b] = gamma |F| j_a j_b
/// with j the jump of the normal derivative of the macro basis functions,
/// grad phi.n on the first cell and -grad phi.n on the second (n the unit
/// normal of the facet).
template <typename T>
__global__ void __launch_bounds__(64)
    interior_matrix_k(std::int32_t nfacets, const std::int32_t* __restrict__ quads,
                      const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                      T gamma, T* __restrict__ out)
{
  // the 64 facets of a workgroup stage their matrices in LDS (row padded
  // to 65 entries against bank conflicts) and store them as one
  // contiguous, coalesced block
  __shared__ T tile[64 * 65];
  const int t = threadIdx.x;
  const std::int32_t e0 = blockIdx.x * 64;
  const std::int32_t e = min(e0 + t, nfacets - 1);
  T X[2][4][3];
#pragma unroll
  for (int s = 0; s < 2; ++s)
  {
    const std::int32_t c = quads[4 * e + 2 * s];
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
      const std::int64_t v = xdm[std::int64_t(c) * 4 + a];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        X[s][a][d] = x[3 * v + d];
    }
  }
  const int lf = quads[4 * e + 1];
  const int f0 = kTetFacetV[lf][0], f1 = kTetFacetV[lf][1], f2 = kTetFacetV[lf][2];
  T t0[3], t1[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    t0[d] = X[0][f1][d] - X[0][f0][d];
    t1[d] = X[0][f2][d] - X[0][f0][d];
  }
  T n[3] = {t0[1] * t1[2] - t0[2] * t1[1], t0[2] * t1[0] - t0[0] * t1[2],
            t0[0] * t1[1] - t0[1] * t1[0]};
  const T nrm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d)
    n[d] /= nrm;
  const T w = gamma * (T(0.5) * nrm);
  T jv[8];
#pragma unroll
  for (int s = 0; s < 2; ++s)
  {
    T g[4][3];
    p1_gradients<T>(X[s], g);
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
      const T dn = g[a][0] * n[0] + g[a][1] * n[1] + g[a][2] * n[2];
      jv[4 * s + a] = s == 0 ? dn : -dn;
    }
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b)
      tile[t * 65 + 8 * a + b] = w * jv[a] * jv[b];
  __syncthreads();
  const std::int64_t nv = std::int64_t(min(64, nfacets - e0)) * 64;
  T* o = out + 64 * std::int64_t(e0);
  for (int k = 0; k < 64; ++k)
  {
    const int idx = k * 64 + t;
    if (idx < nv)
      o[idx] = tile[(idx >> 6) * 65 + (idx & 63)];
  }
}

/// A[r, :] += the macro-element rows at row r, facet by facet and macro
/// index by macro index (one thread per row); BC rows and columns add zero;
/// slots by binary search, an absent one sets *err (if err is not null).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    interior_rows_k(std::int32_t nplan, const std::int32_t* __restrict__ off,
                    const std::int32_t* __restrict__ finc, const std::int32_t* __restrict__ quads,
                    const std::int32_t* __restrict__ dofmap, const T* __restrict__ fm,
                    const std::int8_t* __restrict__ mark, const std::int32_t* __restrict__ row_ptr,
                    const std::int32_t* __restrict__ cols, T* __restrict__ values,
                    int* __restrict__ err, std::uint8_t* __restrict__ fslot)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  const std::int32_t p1 = off[r + 1];
  if (off[r] == p1)
    return;
  const std::int32_t rb = row_ptr[r], re = row_ptr[r + 1];
  const bool rmask = mark && mark[r];
  for (std::int32_t p = off[r]; p < p1; ++p)
  {
    const std::int32_t e = finc[p] >> 3;
    const int i = finc[p] & 7;
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
      const std::int32_t c = quads[4 * e + 2 * (j >> 2)];
      const std::int32_t col = dofmap[std::int64_t(c) * 4 + (j & 3)];
      std::int32_t lo = rb, hi = re;
      while (lo < hi)
      {
        const std::int32_t mid = (lo + hi) >> 1;
        if (cols[mid] < col)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo == re || cols[lo] != col)
      {
        if (err)
          *err = 1;
        continue;
      }
      if (fslot)
        fslot[8 * std::int64_t(p) + j] = static_cast<std::uint8_t>(lo - rb);
      const bool zero = rmask || (mark && mark[col]);
      values[lo] += zero ? T(0) : fm[64 * std::int64_t(e) + 8 * i + j];
    }
  }
}

/// Facet matrix rows once their slots are known (fslot): one thread per
/// row with facet incidences, the row's values staged in LDS, the
/// contributions added there in the same order as the search kernels add
/// them (facet, then entry), the row written back. W entries per
/// incidence (NF used), EW values per facet matrix row, rows of at most
/// kFacetRowMax entries.
constexpr int kFacetRowMax = 64;
template <typename T, int NF, int W, int SH>
__global__ void __launch_bounds__(64)
    facet_rows_fast_k(std::int32_t nfrows, const std::int32_t* __restrict__ frows,
                      const std::int32_t* __restrict__ off, const std::int32_t* __restrict__ finc,
                      const std::uint8_t* __restrict__ fslot, const T* __restrict__ fm,
                      const std::int8_t* __restrict__ mark, const std::int32_t* __restrict__ row_ptr,
                      const std::int32_t* __restrict__ cols, T* __restrict__ values)
{
  __shared__ T acc[kFacetRowMax][64];
  const int t = threadIdx.x;
  const std::int32_t i = blockIdx.x * 64 + t;
  if (i >= nfrows)
    return;
  const std::int32_t r = frows[i];
  const std::int32_t rb = row_ptr[r], len = row_ptr[r + 1] - rb;
  for (std::int32_t s = 0; s < len; ++s)
    acc[s][t] = values[rb + s];
  const bool rmask = mark && mark[r];
  for (std::int32_t p = off[r]; p < off[r + 1]; ++p)
  {
    const std::int32_t e = finc[p] >> SH;
    const int a = finc[p] & ((1 << SH) - 1);
    const T* fr = fm + std::int64_t(e) * (W * W) + W * a;
    const std::uint8_t* sl = fslot + std::int64_t(p) * W;
#pragma unroll
    for (int b = 0; b < NF; ++b)
    {
      const int s = sl[b];
      const bool zero = rmask || (mark && mark[cols[rb + s]]);
      acc[s][t] += zero ? T(0) : fr[b];
    }
  }
  for (std::int32_t s = 0; s < len; ++s)
    values[rb + s] = acc[s][t];
}

/// Lifting by the macro-element matrices: b[r] -= alpha sum over the row's
/// incidences and the macro element's BC columns j of F_ij (g_j - x0_j).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    interior_lift_rows_k(std::int32_t nplan, const std::int32_t* __restrict__ off,
                         const std::int32_t* __restrict__ finc,
                         const std::int32_t* __restrict__ quads,
                         const std::int32_t* __restrict__ dofmap, const T* __restrict__ fm,
                         const std::int8_t* __restrict__ mark, const double* __restrict__ g,
                         const T* __restrict__ x0, double alpha, T* __restrict__ bvec)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nplan)
    return;
  double s = 0;
  for (std::int32_t p = off[r]; p < off[r + 1]; ++p)
  {
    const std::int32_t e = finc[p] >> 3;
    const int i = finc[p] & 7;
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
      const std::int32_t c = quads[4 * e + 2 * (j >> 2)];
      const std::int32_t col = dofmap[std::int64_t(c) * 4 + (j & 3)];
      if (!mark[col])
        continue;
      const double gv = g ? g[col] : 0.0;
      const double xv = x0 ? double(x0[col]) : 0.0;
      s += double(fm[64 * std::int64_t(e) + 8 * i + j]) * (gv - xv);
    }
  }
  if (s != 0.0)
    bvec[r] = static_cast<T>(double(bvec[r]) - alpha * s);
}

/// b[r] += the facet contributions at row r in facet order (one thread per
/// row with facet incidences, rows[i], over its incidences finc = 4 e +
/// facet vertex).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    facet_rows_k(std::int32_t nrows, const std::int32_t* __restrict__ rows,
                 const std::int32_t* __restrict__ off, const std::int32_t* __restrict__ finc,
                 const T* __restrict__ fv, T* __restrict__ b)
{
  const std::int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nrows)
    return;
  const std::int32_t r = rows[i];
  const std::int32_t p1 = off[r + 1];
  T s = 0;
  for (std::int32_t p = off[r]; p < p1; ++p)
    s += fv[finc[p]];
  if (p1 > off[r])
    b[r] += s;
}

// ---------------------------------------------------------------------
// Structured Q1 rows. On a box of hexahedra with lexicographic vertex order
// a row vertex is corner I of at most one cell for each I = 0..7, and
// column j of that cell is the neighbour at offset bit(j) - bit(I) in each
// direction, one of the 27 neighbour slots q(I, j). With the corner cells
// and the CSR position of each slot precomputed, a thread accumulates its
// row in 27 registers with indices and reference values fixed at compile
// time; entries are summed in corner order.
// ---------------------------------------------------------------------

__host__ __device__ constexpr int q1_slot(int i, int j)
{
  return (1 + ((j >> 0) & 1) - ((i >> 0) & 1)) + 3 * (1 + ((j >> 1) & 1) - ((i >> 1) & 1))
         + 9 * (1 + ((j >> 2) & 1) - ((i >> 2) & 1));
}

__global__ void __launch_bounds__(kBlock) fill_i32_k(std::int32_t* v, std::int64_t n, std::int32_t a)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t < n)
    v[t] = a;
}

/// corner_cell[row * 8 + i] = c for dof(c, i) = row; a second cell at the
/// same corner of a row means the mesh is not a structured box.
__global__ void __launch_bounds__(kBlock)
    struct_corner_k(std::int32_t nc, const std::int32_t* __restrict__ dm,
                    std::int32_t* __restrict__ corner, int* err)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nc) * 8)
    return;
  const std::int32_t c = static_cast<std::int32_t>(t / 8);
  const int i = static_cast<int>(t % 8);
  const std::int32_t row = dm[t];
  if (atomicCAS(corner + std::int64_t(row) * 8 + i, -1, c) != -1)
    atomicOr(err, 1);
}

/// tmp[row * 27 + q(i, j)] = position of dof(c, j) in row dof(c, i); all
/// cells must agree on a slot's position.
__global__ void __launch_bounds__(kBlock)
    struct_slots_k(std::int32_t nc, const std::int32_t* __restrict__ dm,
                   const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                   std::int32_t* __restrict__ tmp, int* err)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nc) * 64)
    return;
  const std::int64_t c = t / 64;
  const int i = static_cast<int>((t / 8) % 8), j = static_cast<int>(t % 8);
  const std::int32_t row = dm[c * 8 + i], col = dm[c * 8 + j];
  std::int32_t lo = rp[row], hi = rp[row + 1];
  const std::int32_t r0 = lo, r1 = hi;
  while (lo < hi)
  {
    const std::int32_t mid = (lo + hi) >> 1;
    if (ci[mid] < col)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= r1 || ci[lo] != col)
  {
    atomicOr(err, 2);
    return;
  }
  const std::int32_t pos = lo - r0;
  const std::int32_t old = atomicCAS(tmp + std::int64_t(row) * 27 + q1_slot(i, j), -1, pos);
  if (old != -1 && old != pos)
    atomicOr(err, 4);
}

/// Pack the slot positions as bytes (255 absent) and check that the slots
/// of each row cover every entry of the row exactly once.
__global__ void __launch_bounds__(kBlock)
    struct_pack_k(std::int32_t nrows, const std::int32_t* __restrict__ tmp,
                  const std::int32_t* __restrict__ rp, std::uint8_t* __restrict__ slotmap,
                  int* err)
{
  const std::int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrows)
    return;
  const std::int32_t len = rp[r + 1] - rp[r];
  std::uint32_t seen = 0;
  int count = 0;
  for (int q = 0; q < 27; ++q)
  {
    const std::int32_t v = tmp[std::int64_t(r) * 27 + q];
    if (v >= 0)
    {
      if (v >= 32 || ((seen >> v) & 1u))
        atomicOr(err, 8);
      else
        seen |= 1u << v;
      ++count;
    }
    slotmap[std::int64_t(r) * 32 + q] = v < 0 ? std::uint8_t(255) : static_cast<std::uint8_t>(v);
  }
  for (int q = 27; q < 32; ++q)
    slotmap[std::int64_t(r) * 32 + q] = 255;
  if (count != len)
    atomicOr(err, 8);
}

/// One thread per row of a structured Q1 matrix (see above). Factors of the
/// row's corner cells are loaded four corners at a time at clamped indices
/// (absent corners masked), so no load waits behind a branch.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    row_struct_q1_k(std::int32_t nrows, const std::int32_t* __restrict__ corner,
                    const std::uint8_t* __restrict__ slotmap, const T* __restrict__ cd,
                    const std::int8_t* __restrict__ bc, const std::int32_t* __restrict__ rp,
                    T* __restrict__ vals, int accumulate)
{
  const std::int32_t r = xcd_contiguous(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (r >= nrows)
    return;
  T acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q)
    acc[q] = T(0);
  if (!(bc && bc[r]))
  {
    const int4 ca = *reinterpret_cast<const int4*>(corner + 8 * std::int64_t(r));
    const int4 cb = *reinterpret_cast<const int4*>(corner + 8 * std::int64_t(r) + 4);
    const std::int32_t cc[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
      T g[4][7];
      std::uint32_t m[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
      {
        const std::int32_t c = cc[4 * h + u];
        load_factors<T>(cd, c < 0 ? 0 : c, g[u], m[u]);
        if (c < 0)
          m[u] = 0xffu; // absent corner: every column masked
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
      {
        const int I = 4 * h + u;
        T gs[3];
        q1_signed<T>(g[u], I, gs);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (!((m[u] >> j) & 1u))
            acc[q1_slot(I, j)] += q1_entry<T>(g[u], gs, I, j);
      }
    }
  }
  const std::uint32_t* sm = reinterpret_cast<const std::uint32_t*>(slotmap + 32 * std::int64_t(r));
  std::uint32_t w[7];
#pragma unroll
  for (int k = 0; k < 7; ++k)
    w[k] = sm[k];
  T* out = vals + rp[r];
#pragma unroll
  for (int q = 0; q < 27; ++q)
  {
    const std::uint32_t s = (w[q / 4] >> (8 * (q % 4))) & 0xffu;
    if (s != 0xffu)
      out[s] = accumulate ? out[s] + acc[q] : acc[q];
  }
}

/// Variant 15 (Q1 on a structured box of hexahedra): one thread per row as
/// variant 9 -- the row vertex's 8 corner cells, its 27 entries summed in
/// registers with compile-time reference values, corner order -- but the
/// rows of a workgroup are consecutive and their entries go through an LDS
/// copy of the workgroup's CSR slice, so the slice is stored with coalesced
/// non-temporal writes. Every entry of a row is one of its 27 slots
/// (checked when the plan is built), so the slice needs no zero fill.
template <typename T, bool COMPACT = false>
__global__ void __launch_bounds__(kBlock)
    row_struct_lds_k(std::int32_t nrows, const std::int32_t* __restrict__ corner,
                     const std::uint8_t* __restrict__ slotmap, const T* __restrict__ cd,
                     const std::uint32_t* __restrict__ mask, const std::int8_t* __restrict__ bc,
                     const std::int32_t* __restrict__ rp, T* __restrict__ vals, int accumulate)
{
  extern __shared__ unsigned char smem_raw[];
  T* slice = reinterpret_cast<T*>(smem_raw);
  const int R = blockDim.x;
  const std::int32_t r0 = xcd_contiguous(blockIdx.x, gridDim.x) * R;
  const std::int32_t r1 = min(r0 + R, nrows);
  const std::int32_t base = rp[r0];
  const std::int32_t len = rp[r1] - base;
  const std::int32_t r = r0 + threadIdx.x;
  if (r < r1)
  {
    const dfx_i32x4* cp = reinterpret_cast<const dfx_i32x4*>(corner + 8 * std::int64_t(r));
    const dfx_i32x4 ca = __builtin_nontemporal_load(cp);
    const dfx_i32x4 cb = __builtin_nontemporal_load(cp + 1);
    const std::int32_t cc[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
    // the row's 27 slot offsets, padded to 32 bytes: two 16-byte loads
    const dfx_i32x4* smp = reinterpret_cast<const dfx_i32x4*>(slotmap + 32 * std::int64_t(r));
    const dfx_i32x4 sa = __builtin_nontemporal_load(smp);
    const dfx_i32x4 sb = __builtin_nontemporal_load(smp + 1);
    const std::uint32_t w[7] = {std::uint32_t(sa.x), std::uint32_t(sa.y), std::uint32_t(sa.z),
                                std::uint32_t(sa.w), std::uint32_t(sb.x), std::uint32_t(sb.y),
                                std::uint32_t(sb.z)};
    const bool rowbc = bc && bc[r];
    const std::int32_t roff = rp[r] - base;
    T acc[27];
#pragma unroll
    for (int q = 0; q < 27; ++q)
      acc[q] = T(0);
    T g[8][7];
    std::uint32_t m[8];
#pragma unroll
    for (int I = 0; I < 8; ++I)
    {
      const std::int32_t c = cc[I] < 0 ? 0 : cc[I];
      if constexpr (COMPACT)
      {
        load_compact<T>(cd, c, g[I]);
        m[I] = mask ? mask[c] : 0u;
      }
      else
        load_factors<T>(cd, c, g[I], m[I]);
      if (cc[I] < 0 || rowbc)
        m[I] = 0xffu; // absent corner (or a Dirichlet row): nothing added
    }
#pragma unroll
    for (int I = 0; I < 8; ++I)
    {
      T gs[3];
      q1_signed<T>(g[I], I, gs);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (!((m[I] >> j) & 1u))
          acc[q1_slot(I, j)] += q1_entry<T>(g[I], gs, I, j);
    }
    T* a = slice + roff;
#pragma unroll
    for (int q = 0; q < 27; ++q)
    {
      const std::uint32_t sl = (w[q / 4] >> (8 * (q % 4))) & 0xffu;
      if (sl != 0xffu)
        a[sl] = acc[q];
    }
  }
  __syncthreads();
  if (accumulate)
    for (int t = threadIdx.x; t < len; t += R)
      vals[base + t] += slice[t];
  else
    for (int t = threadIdx.x; t < len; t += R)
      __builtin_nontemporal_store(slice[t], vals + base + t);
}

/// Element source vector of an affine cell from its vertices q0..q3 (P1:
/// the four vertices; Q1: vertices 0, 1, 2, 4) and the source values at its
/// NV nodes: P1 b_i = |K|/20 (f_i + sum f), Q1 b_i = |det J| sum_j M_ij f_j.
template <typename T, int NV>
__device__ inline void affine_source(const T* q0, const T* q1, const T* q2, const T* q3,
                                     const T (&f)[NV], T (&out)[NV])
{
  T J[9], adj[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    J[3 * r + 0] = q1[r] - q0[r];
    J[3 * r + 1] = q2[r] - q0[r];
    J[3 * r + 2] = q3[r] - q0[r];
  }
  const T det = adjugate(J, adj);
  const T adet = det < T(0) ? -det : det;
  if constexpr (NV == 4)
  {
    const T sum = f[0] + f[1] + f[2] + f[3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      out[i] = adet / T(120) * (sum + f[i]);
  }
  else
  {
#pragma unroll
    for (int i = 0; i < 8; ++i)
    {
      T a = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j)
      {
        T m = 1;
#pragma unroll
        for (int d = 0; d < 3; ++d)
          m *= (((i >> d) & 1) == ((j >> d) & 1)) ? T(1) / T(3) : T(1) / T(6);
        a += m * f[j];
      }
      out[i] = adet * a;
    }
  }
}

/// The NV source values of cell c, with vector loads when the coefficient
/// rows are 16-byte aligned.
template <typename T, int NV>
__device__ inline void load_source(const T* __restrict__ coeffs, int cstride, std::int64_t c,
                                   T (&f)[NV])
{
  const T* fp = coeffs + c * cstride;
  if ((cstride * sizeof(T)) % 16 == 0)
  {
    using T2 = vec2_t<T>;
    const T2* f2 = reinterpret_cast<const T2*>(fp);
#pragma unroll
    for (int h = 0; h < NV / 2; ++h)
    {
      const T2 w = f2[h];
      f[2 * h] = w.x;
      f[2 * h + 1] = w.y;
    }
  }
  else
  {
#pragma unroll
    for (int h = 0; h < NV; ++h)
      f[h] = fp[h];
  }
}

/// Element vectors of the source forms, one thread per cell:
/// P1 tetrahedra b_i = |K|/20 (f_i + sum f); affine Q1 hexahedra
/// b_i = |det J| sum_j M_ij f_j; other hexahedra 2x2x2 Gauss.
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock)
    cell_source_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                  const T* __restrict__ coeffs, int cstride, int affine, T* __restrict__ be)
{
  const std::int64_t c = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  // the cell's vertex row (NV ints, 16-byte aligned) and its NV source
  // values with vector loads when the coefficient rows are 16-byte aligned
  std::int32_t v[NV];
  {
    const int4 a4 = *reinterpret_cast<const int4*>(xdm + NV * c);
    v[0] = a4.x; v[1] = a4.y; v[2] = a4.z; v[3] = a4.w;
    if constexpr (NV == 8)
    {
      const int4 b4 = *reinterpret_cast<const int4*>(xdm + NV * c + 4);
      v[4] = b4.x; v[5] = b4.y; v[6] = b4.z; v[7] = b4.w;
    }
  }
  T f[NV];
  load_source<T, NV>(coeffs, cstride, c, f);
  T out[NV];
  if (affine)
    affine_source<T, NV>(x + 3 * std::int64_t(v[0]), x + 3 * std::int64_t(v[1]),
                         x + 3 * std::int64_t(v[2]), x + 3 * std::int64_t(v[NV == 4 ? 3 : 4]),
                         f, out);
  else
  {
    T cdv[8][3];
    for (int w = 0; w < 8; ++w)
      for (int r = 0; r < 3; ++r)
        cdv[w][r] = x[3 * std::int64_t(v[w]) + r];
    for (int i = 0; i < NV; ++i)
      out[i] = 0;
    const T ga = T(0.5) - T(0.5) / sqrt(T(3)), gb = T(0.5) + T(0.5) / sqrt(T(3));
    for (int qp = 0; qp < 8; ++qp)
    {
      const T xi[3] = {(qp & 1) ? gb : ga, ((qp >> 1) & 1) ? gb : ga, ((qp >> 2) & 1) ? gb : ga};
      T phi[8], dphi[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, adj[9];
      q1_tab(xi, phi, dphi);
      for (int w = 0; w < 8; ++w)
        for (int r = 0; r < 3; ++r)
          for (int a = 0; a < 3; ++a)
            J[3 * r + a] += cdv[w][r] * dphi[w][a];
      const T det = adjugate(J, adj);
      const T adet = det < T(0) ? -det : det;
      T fq = 0;
      for (int w = 0; w < 8; ++w)
        fq += f[w] * phi[w];
      for (int i = 0; i < NV; ++i)
        out[i] += adet / T(8) * fq * phi[i];
    }
  }
  // corner-major element vectors, be[i][c]: consecutive cells of a
  // wavefront store contiguously, and consecutive rows, whose cells are
  // consecutive at a given corner, later read the same lines
#pragma unroll
  for (int i = 0; i < NV; ++i)
    be[std::int64_t(i) * nc + c] = out[i];
}

/// One pass over the cells of a system whose bilinear and linear forms
/// share their geometry (dfx_assemble_system with an affine P1/Q1
/// stiffness form and the P1/Q1 source form on the same mesh): the cell
/// factor record and BC mask of the matrix, as cell_factors_k, and the
/// element source vector, as cell_source_k, from one load of the cell's
/// vertex row and coordinates.
template <typename T, int NV, bool COMPACT = false>
__global__ void __launch_bounds__(kBlock)
    cell_system_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                  const T* __restrict__ coeffs, int cstride, T creac,
                  const std::int8_t* __restrict__ bc, T* __restrict__ cd,
                  std::uint32_t* __restrict__ mask, const T* __restrict__ src, int sstride,
                  T* __restrict__ be)
{
  const std::int64_t c
      = xcd_contiguous(blockIdx.x, gridDim.x) * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  std::int32_t v[NV];
  {
    const int4 a = *reinterpret_cast<const int4*>(xdm + NV * c);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    if constexpr (NV == 8)
    {
      const int4 b = *reinterpret_cast<const int4*>(xdm + NV * c + 4);
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
  }
  T f[NV];
  load_source<T, NV>(src, sstride, c, f);
  std::uint32_t m = 0;
  if (mask)
  {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (bc[v[j]])
        m |= 1u << j;
  }
  const T* q0 = x + 3 * std::int64_t(v[0]);
  const T* q1 = x + 3 * std::int64_t(v[1]);
  const T* q2 = x + 3 * std::int64_t(v[2]);
  const T* q3 = x + 3 * std::int64_t(v[NV == 4 ? 3 : 4]);
  T g[7];
  affine_factors<T>(q0, q1, q2, q3, coeffs[c * cstride], creac, g);
  T out[NV];
  affine_source<T, NV>(q0, q1, q2, q3, f, out);
  if (mask)
    mask[c] = m;
  if constexpr (COMPACT)
    store_compact<T>(cd, c, g);
  else
  {
    const T g7 = mask_bits<T>(m);
    if constexpr (sizeof(T) == 8)
    {
      double2* G = reinterpret_cast<double2*>(cd + 8 * c);
      G[0] = make_double2(g[0], g[1]);
      G[1] = make_double2(g[2], g[3]);
      G[2] = make_double2(g[4], g[5]);
      G[3] = make_double2(g[6], g7);
    }
    else
    {
      float4* G = reinterpret_cast<float4*>(cd + 8 * c);
      G[0] = make_float4(g[0], g[1], g[2], g[3]);
      G[1] = make_float4(g[4], g[5], g[6], g7);
    }
  }
#pragma unroll
  for (int i = 0; i < NV; ++i)
    be[std::int64_t(i) * nc + c] = out[i];
}

/// *diff = 1 if a[0..n) and b[0..n) differ in any byte (32-bit words).
__global__ void __launch_bounds__(kBlock)
    words_differ_k(std::int64_t n, const std::uint32_t* __restrict__ a,
                   const std::uint32_t* __restrict__ b, int* diff)
{
  for (std::int64_t i = blockIdx.x * std::int64_t(kBlock) + threadIdx.x; i < n;
       i += std::int64_t(gridDim.x) * kBlock)
    if (a[i] != b[i])
    {
      *diff = 1;
      return;
    }
}

/// b[BS r + comp] += sum of the element vector entries of the row's
/// incidences, in cell order. inc holds c * K + i; the cell pass stored
/// entry (i, comp) of cell c at be[(i BS + comp) nc + c] (corner-major).
/// Eight incidences per group, loads at clamped positions so that all
/// sixteen loads of a group are in flight together.
template <typename T, int K, int BS>
__global__ void __launch_bounds__(kBlock)
    row_vector_k(std::int32_t nplan, std::int32_t nc, const std::int32_t* __restrict__ inc_off,
                 const std::int32_t* __restrict__ inc, const T* __restrict__ be,
                 T* __restrict__ b)
{
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nplan) * BS)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / BS);
  const int comp = static_cast<int>(t % BS);
  const std::int32_t p0 = inc_off[r], p1 = inc_off[r + 1];
  T s = 0;
  for (std::int32_t pb = p0; pb < p1; pb += 8)
  {
    std::int32_t q[8];
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      q[u] = inc[min(pb + u, p1 - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = be[std::int64_t((q[u] % K) * BS + comp) * nc + q[u] / K];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (pb + u < p1)
        s += v[u];
  }
  b[t] += s;
}

// ---------------------------------------------------------------------
// Lagrange elements of degree 2 and 3 on the reference tetrahedron, for the
// reference tensors of the P2 elasticity and P3 forms. Nodes in the Basix
// order (vertices; edge interiors from the edge's first vertex to its
// second, at GLL points for degree 3, Basix's default gll_warped variant;
// face centroids for degree 3); basis in monomial coefficients from the
// Vandermonde matrix; integrals exact from int xi^a eta^b zeta^c =
// a! b! c! / (a+b+c+3)! over the reference cell.
// ---------------------------------------------------------------------

struct TetBasis
{
  int n = 0;
  std::vector<std::array<int, 3>> mono; // exponents of the n monomials
  std::vector<double> C;                // C[m * n + i]: coefficient of mono m in phi_i
};

double factorial_d(int k)
{
  double f = 1;
  for (int i = 2; i <= k; ++i)
    f *= i;
  return f;
}

double mono_integral(int a, int b, int c)
{
  return factorial_d(a) * factorial_d(b) * factorial_d(c) / factorial_d(a + b + c + 3);
}

TetBasis tet_basis(int degree)
{
  static constexpr int E[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
  static constexpr int F[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  std::vector<std::array<double, 3>> pts;
  for (auto& v : V)
    pts.push_back({v[0], v[1], v[2]});
  std::vector<double> t;
  if (degree == 2)
    t = {0.5};
  if (degree == 3)
    t = {0.5 - 0.5 / std::sqrt(5.0), 0.5 + 0.5 / std::sqrt(5.0)};
  for (auto& e : E)
    for (double s : t)
      pts.push_back({V[e[0]][0] + s * (V[e[1]][0] - V[e[0]][0]),
                     V[e[0]][1] + s * (V[e[1]][1] - V[e[0]][1]),
                     V[e[0]][2] + s * (V[e[1]][2] - V[e[0]][2])});
  if (degree == 3)
    for (auto& f : F)
      pts.push_back({(V[f[0]][0] + V[f[1]][0] + V[f[2]][0]) / 3,
                     (V[f[0]][1] + V[f[1]][1] + V[f[2]][1]) / 3,
                     (V[f[0]][2] + V[f[1]][2] + V[f[2]][2]) / 3});
  TetBasis B;
  B.n = static_cast<int>(pts.size());
  for (int s = 0; s <= degree; ++s)
    for (int a = s; a >= 0; --a)
      for (int b = s - a; b >= 0; --b)
        B.mono.push_back({a, b, s - a - b});
  const int n = B.n;
  // Solve V C = I with V[p][m] = mono_m(point p) (Gauss-Jordan, pivoting)
  std::vector<double> A(n * n), X(n * n, 0.0);
  for (int p = 0; p < n; ++p)
    for (int m = 0; m < n; ++m)
      A[p * n + m] = std::pow(pts[p][0], B.mono[m][0]) * std::pow(pts[p][1], B.mono[m][1])
                     * std::pow(pts[p][2], B.mono[m][2]);
  // we need C with sum_m C[m][i] mono_m(p) = delta_ip, i.e. V C = I
  for (int i = 0; i < n; ++i)
    X[i * n + i] = 1.0;
  for (int col = 0; col < n; ++col)
  {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::abs(A[r * n + col]) > std::abs(A[piv * n + col]))
        piv = r;
    for (int k = 0; k < n; ++k)
    {
      std::swap(A[col * n + k], A[piv * n + k]);
      std::swap(X[col * n + k], X[piv * n + k]);
    }
    const double d = A[col * n + col];
    for (int k = 0; k < n; ++k)
    {
      A[col * n + k] /= d;
      X[col * n + k] /= d;
    }
    for (int r = 0; r < n; ++r)
      if (r != col && A[r * n + col] != 0.0)
      {
        const double f = A[r * n + col];
        for (int k = 0; k < n; ++k)
        {
          A[r * n + k] -= f * A[col * n + k];
          X[r * n + k] -= f * X[col * n + k];
        }
      }
  }
  B.C = X; // X = V^{-1}: X[m * n + i]
  return B;
}

/// int (d_a phi_i)(d_b phi_j) (a, b = -1: no derivative) over the
/// reference tetrahedron.
double basis_product_integral(const TetBasis& B, int i, int a, int j, int b)
{
  const int n = B.n;
  double s = 0;
  for (int m = 0; m < n; ++m)
  {
    std::array<int, 3> e = B.mono[m];
    double cm = B.C[m * n + i];
    if (cm == 0.0)
      continue;
    if (a >= 0)
    {
      if (e[a] == 0)
        continue;
      cm *= e[a];
      e[a] -= 1;
    }
    for (int q = 0; q < n; ++q)
    {
      std::array<int, 3> f = B.mono[q];
      double cq = B.C[q * n + j];
      if (cq == 0.0)
        continue;
      if (b >= 0)
      {
        if (f[b] == 0)
          continue;
        cq *= f[b];
        f[b] -= 1;
      }
      s += cm * cq * mono_integral(e[0] + f[0], e[1] + f[1], e[2] + f[2]);
    }
  }
  return s;
}

double basis_integral(const TetBasis& B, int i)
{
  double s = 0;
  for (int m = 0; m < B.n; ++m)
    s += B.C[m * B.n + i] * mono_integral(B.mono[m][0], B.mono[m][1], B.mono[m][2]);
  return s;
}

/// [7][k][k] tensors of the affine stiffness + reaction form of a basis.
std::vector<double> scalar_tensors(const TetBasis& B)
{
  const int k = B.n;
  std::vector<double> t(7 * k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j)
    {
      const int e = i * k + j;
      auto S = [&](int a, int b) { return basis_product_integral(B, i, a, j, b); };
      t[0 * k * k + e] = S(0, 0);
      t[1 * k * k + e] = S(1, 1);
      t[2 * k * k + e] = S(2, 2);
      t[3 * k * k + e] = S(0, 1) + S(1, 0);
      t[4 * k * k + e] = S(0, 2) + S(2, 0);
      t[5 * k * k + e] = S(1, 2) + S(2, 1);
      t[6 * k * k + e] = basis_product_integral(B, i, -1, j, -1);
    }
  return t;
}

/// Tensors S_ab(i, j) = int d_a phi_i d_b phi_j (elasticity), stored per
/// (i, j) as the 3 x 3 matrix over (a, b) padded to 10 values:
/// t[(i k + j) 10 + 3 a + b], t[(i k + j) 10 + 9] = 0.
std::vector<double> gradient_tensors(const TetBasis& B)
{
  const int k = B.n;
  std::vector<double> t(10 * k * k, 0.0);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          t[(i * k + j) * 10 + 3 * a + b] = basis_product_integral(B, i, a, j, b);
  return t;
}

// ---------------------------------------------------------------------
// P2 vector elasticity, a = sigma(u):eps(v) with DG0 lambda, mu. Per cell
// a record of 16 values: J^{-1} (9), 0, |det J| lambda, |det J| mu, 0...
// Entry block (i, j) (3 x 3, row component c, column component d) with
// P = J^{-T}-transformed reference gradient products,
// P_cd = sum_ab Kinv[a][c] S_ab(i,j) Kinv[b][d]:
//   A_(i,c)(j,d) = |det J| mu (delta_cd tr P + P_dc) + |det J| lambda P_cd.
// ---------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kBlock)
    elast_factors_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                    const std::int32_t* __restrict__ dm, const T* __restrict__ coeffs, int cstride,
                    const std::int8_t* __restrict__ bc, T* __restrict__ cd,
                    std::uint32_t* __restrict__ mask)
{
  using T2 = vec2_t<T>;
  const std::int64_t c
      = xcd_contiguous(blockIdx.x, gridDim.x) * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  const int4 v = *reinterpret_cast<const int4*>(xdm + 4 * c);
  const T* q0 = x + 3 * std::int64_t(v.x);
  const T* q1 = x + 3 * std::int64_t(v.y);
  const T* q2 = x + 3 * std::int64_t(v.z);
  const T* q3 = x + 3 * std::int64_t(v.w);
  T J[9], adj[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    J[3 * r + 0] = q1[r] - q0[r];
    J[3 * r + 1] = q2[r] - q0[r];
    J[3 * r + 2] = q3[r] - q0[r];
  }
  const T det = adjugate(J, adj);
  const T adet = det < T(0) ? -det : det;
  T o[16];
#pragma unroll
  for (int q = 0; q < 9; ++q)
    o[q] = adj[q] / det;
  o[9] = T(0);
  o[10] = adet * coeffs[c * cstride];     // |det J| lambda
  o[11] = adet * coeffs[c * cstride + 1]; // |det J| mu
#pragma unroll
  for (int q = 12; q < 16; ++q)
    o[q] = T(0);
  T2* out = reinterpret_cast<T2*>(cd + 16 * c);
#pragma unroll
  for (int h = 0; h < 8; ++h)
    out[h] = T2{o[2 * h], o[2 * h + 1]};
  if (mask)
  {
    std::uint32_t m = 0;
    for (int j = 0; j < 10; ++j)
    {
      const std::int32_t node = dm[c * 10 + j];
      for (int b = 0; b < 3; ++b)
        if (bc[3 * std::int64_t(node) + b])
          m |= 1u << (3 * j + b);
    }
    mask[c] = m;
  }
}

/// Row teams for the elasticity rows: a wavefront holds 6 teams of 10
/// lanes; team t takes one block row, lane j the column node j of every
/// incidence of that row. The team walks the row's incidences in cell
/// order, lane j adding the 3 x 3 block (i, j) into the row's LDS slice at
/// the block of column slot j (the 10 slots of one incidence are distinct,
/// so the additions need no atomics and every entry is summed in cell
/// order). Rows come in groups of up to 6 with similar incidence counts
/// (egrp: per group its first position in erow and its LDS slice length,
/// then a terminating entry; erow: per row (row, slice offset, first
/// incidence, incidences)); blocks are stored in LDS padded to 10 values.
template <typename T, typename O>
__global__ void __launch_bounds__(64)
    elast_team_k(const int2* __restrict__ egrp, const int4* __restrict__ erow, const std::int32_t* __restrict__ inc,
                 const O* __restrict__ islots, const T* __restrict__ cd,
                 const std::uint32_t* __restrict__ mask, const std::int8_t* __restrict__ bc,
                 const std::int32_t* __restrict__ rp, const T* __restrict__ ref,
                 T* __restrict__ vals, int accumulate)
{
  using T2 = vec2_t<T>;
  constexpr int K = 10;
  extern __shared__ unsigned char smem_raw[];
  T* acc = reinterpret_cast<T*>(smem_raw);
  const int lane = threadIdx.x, team = lane / K, j = lane - K * team;
  // A group's description as seen by this lane's team: the group's rows
  // [q0, q1) and slice length, the team's (row, slice offset, first
  // incidence, incidences), the row's CSR position and length, its
  // Dirichlet components and its first 30 incidences (3 per lane).
  struct Meta
  {
    std::int32_t q0, q1, total;
    bool has;
    int4 ro;
    std::int32_t rbase, rlen;
    std::uint32_t rowbc;
    std::int32_t pre0, pre1, pre2;
  };
  // first stage: the group's bounds and the team's row entry
  auto meta_rows = [&](Meta& m, int g)
  {
    const int2 gb = egrp[g];
    m.q0 = gb.x;
    m.total = gb.y;
    m.q1 = egrp[g + 1].x;
    m.has = lane < 6 * K && m.q0 + team < m.q1;
    m.ro = m.has ? erow[m.q0 + team] : make_int4(0, 0, 0, 0);
  };
  // second stage: the loads that need the row entry
  auto meta_data = [&](Meta& m)
  {
    m.rbase = m.rlen = 0;
    m.rowbc = 0;
    m.pre0 = m.pre1 = m.pre2 = 0;
    if (m.has)
    {
      const std::int32_t r = m.ro.x, p0 = m.ro.z, n = m.ro.w;
      m.rbase = rp[r];
      m.rlen = rp[r + 1] - m.rbase;
      if (bc)
        for (int cc = 0; cc < 3; ++cc)
          if (bc[3 * std::int64_t(r) + cc])
            m.rowbc |= 1u << cc;
      if (n > 0)
      {
        m.pre0 = inc[p0 + min(j, n - 1)];
        m.pre1 = inc[p0 + min(K + j, n - 1)];
        m.pre2 = inc[p0 + min(2 * K + j, n - 1)];
      }
    }
  };
  // the team adds its row's incidences into the row's slice
  auto rows = [&](const Meta m)
  {
    const std::uint32_t rowbc = m.rowbc;
    const std::int32_t p0 = m.ro.z, n = m.ro.w;
    if (!m.has || rowbc == 7u || n == 0)
      return;
    T* a = acc + m.ro.y;
    // the first 30 incidences are handed out by shuffles within the team
    // (the team's lanes run the same steps)
    const std::int32_t pre0 = m.pre0, pre1 = m.pre1, pre2 = m.pre2;
    auto cik_at = [pre0, pre1, pre2, p0, inc, team](int st) -> std::int32_t
    {
      if (st < 3 * K)
      {
        const int h = st >= 2 * K ? 2 : (st >= K ? 1 : 0);
        // (masks rather than a select chain, which the compiler turns into
        // a table in scratch memory)
        const std::int32_t m1 = -static_cast<std::int32_t>(h == 1);
        const std::int32_t m2 = -static_cast<std::int32_t>(h == 2);
        const std::int32_t v = (pre0 & ~(m1 | m2)) | (pre1 & m1) | (pre2 & m2);
        return __shfl(v, team * K + (st - K * h), 64);
      }
      return inc[p0 + st];
    };
    // each step's cell record, mask, slot and reference block are loaded
    // two steps ahead
    struct Stage
    {
      T2 rec[6];
      T2 tm[5];
      std::uint32_t m;
      std::int32_t sl;
    };
    // 32-bit byte offsets from the (uniform) array bases; the launcher
    // checks that every offset fits
    const char* cdb = reinterpret_cast<const char*>(cd);
    const char* refb = reinterpret_cast<const char*>(ref);
    const char* maskb = reinterpret_cast<const char*>(mask);
    const char* slb = reinterpret_cast<const char*>(islots);
    auto fetch = [&](Stage& sg, int st)
    {
      const std::int32_t cik = cik_at(st);
      const std::int32_t c = cik / K;
      const int i = cik - c * K;
      const T2* src = reinterpret_cast<const T2*>(
          cdb + static_cast<std::uint32_t>(c) * std::uint32_t(16 * sizeof(T)));
#pragma unroll
      for (int h = 0; h < 6; ++h)
        sg.rec[h] = src[h];
      const T2* s2 = reinterpret_cast<const T2*>(
          refb + static_cast<std::uint32_t>(i * K + j) * std::uint32_t(10 * sizeof(T)));
#pragma unroll
      for (int h = 0; h < 5; ++h)
        sg.tm[h] = s2[h];
      sg.m = mask ? *reinterpret_cast<const std::uint32_t*>(
                 maskb + static_cast<std::uint32_t>(c) * std::uint32_t(sizeof(std::uint32_t)))
                  : 0u;
      sg.sl = *reinterpret_cast<const O*>(
          slb + static_cast<std::uint32_t>((p0 + st) * K + j) * std::uint32_t(sizeof(O)));
    };
    auto add = [&](const Stage& sg)
    {
      T Ki[9], Tm[10];
#pragma unroll
      for (int h = 0; h < 4; ++h)
      {
        Ki[2 * h] = sg.rec[h].x;
        Ki[2 * h + 1] = sg.rec[h].y;
      }
      Ki[8] = sg.rec[4].x;
      const T lam = sg.rec[5].x, mu = sg.rec[5].y;
#pragma unroll
      for (int h = 0; h < 5; ++h)
      {
        Tm[2 * h] = sg.tm[h].x;
        Tm[2 * h + 1] = sg.tm[h].y;
      }
      // W = S Kinv, P = Kinv^T W
      T W[9], P[9];
#pragma unroll
      for (int aa = 0; aa < 3; ++aa)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          W[aa * 3 + d] = Tm[aa * 3 + 0] * Ki[0 * 3 + d] + Tm[aa * 3 + 1] * Ki[1 * 3 + d]
                          + Tm[aa * 3 + 2] * Ki[2 * 3 + d];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          P[cc * 3 + d] = Ki[0 * 3 + cc] * W[0 * 3 + d] + Ki[1 * 3 + cc] * W[1 * 3 + d]
                          + Ki[2 * 3 + cc] * W[2 * 3 + d];
      const T tr = P[0] + P[4] + P[8];
      T out[9];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          out[cc * 3 + d] = cc == d ? mu * (tr + P[cc * 4]) + lam * P[cc * 4]
                                    : mu * P[d * 3 + cc] + lam * P[cc * 3 + d];
      // Dirichlet components of the row and of column node j add nothing
      // (rare: branched around)
      const std::uint32_t colb = (sg.m >> (3 * j)) & 7u;
      if (rowbc | colb)
      {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
          for (int d = 0; d < 3; ++d)
            if (((rowbc >> cc) | (colb >> d)) & 1u)
              out[cc * 3 + d] = T(0);
      }
      T2* blk = reinterpret_cast<T2*>(a + K * sg.sl);
#pragma unroll
      for (int h = 0; h < 4; ++h)
      {
        T2 w = blk[h];
        w.x += out[2 * h];
        w.y += out[2 * h + 1];
        blk[h] = w;
      }
      a[K * sg.sl + 8] += out[8];
    };
    Stage s0, s1, s2;
    fetch(s0, 0);
    fetch(s1, min(1, n - 1));
    for (int st = 0; st < n; st += 3)
    {
      fetch(s2, min(st + 2, n - 1));
      add(s0);
      if (st + 1 >= n)
        break;
      fetch(s0, min(st + 3, n - 1));
      add(s1);
      if (st + 2 >= n)
        break;
      fetch(s1, min(st + 4, n - 1));
      add(s2);
    }
  };
  // the group's rows are stored one at a time by the whole wavefront (row
  // t's position, length and slice offset come from lane K t)
  auto store = [&](const Meta m)
  {
    for (std::int32_t t = 0; t < m.q1 - m.q0; ++t)
    {
      const std::int32_t base = __builtin_amdgcn_readlane(m.rbase, K * t);
      const std::int32_t n9 = 9 * __builtin_amdgcn_readlane(m.rlen, K * t);
      const T* src = acc + __builtin_amdgcn_readlane(m.ro.y, K * t);
      T* dst = vals + std::int64_t(base) * 9;
      for (int e = lane; e < n9; e += 64)
      {
        const int blk = e / 9;
        const T v = src[blk * K + (e - 9 * blk)];
        if (accumulate)
          dst[e] += v;
        else
          dst[e] = v;
      }
    }
  };
  Meta cur;
  meta_rows(cur, xcd_contiguous(blockIdx.x, gridDim.x));
  for (int t = lane; t < cur.total; t += 64)
    acc[t] = T(0);
  meta_data(cur);
  __syncthreads();
  rows(cur);
  __syncthreads();
  store(cur);
}

/// Lifting of the P2 elasticity form, b[(r, c)] -= alpha sum over the
/// row's incidences (cell, i), in cell order, of sum over the cell's
/// Dirichlet columns (j, d) of A_(i,c)(j,d) (g - x0)_(j,d); one thread per
/// row dof (node r, component c). The 3 x 3 blocks are formed from the
/// cell records and reference tensors as in the row kernel, for the cells
/// whose mask marks a Dirichlet column only.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    elast_lift_k(std::int32_t nplan, const std::int32_t* __restrict__ inc_off,
                 const std::int32_t* __restrict__ inc, const std::int32_t* __restrict__ dm,
                 const T* __restrict__ cd, const std::uint32_t* __restrict__ mask,
                 const T* __restrict__ ref, const double* __restrict__ g,
                 const T* __restrict__ x0, double alpha, T* __restrict__ b)
{
  constexpr int K = 10;
  const std::int64_t t = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (t >= std::int64_t(nplan) * 3)
    return;
  const std::int32_t r = static_cast<std::int32_t>(t / 3);
  const int comp = static_cast<int>(t - 3 * std::int64_t(r));
  double s = 0.0;
  const std::int32_t p0 = inc_off[r], p1 = inc_off[r + 1];
  // eight incidences at a time: their cells and masks loaded together
  // (clamped positions), then the cells with a Dirichlet column in order
  for (std::int32_t pb = p0; pb < p1; pb += 8)
  {
    std::int32_t cq[8];
    std::uint32_t mq[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      cq[u] = inc[min(pb + u, p1 - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      mq[u] = mask[cq[u] / K];
    for (int u = 0; u < 8 && pb + u < p1; ++u)
    {
      const std::int32_t cik = cq[u];
      const std::int32_t c = cik / K;
      const std::uint32_t m = mq[u];
      if (!m)
        continue;
      const int i = cik - c * K;
      const T* rec = cd + 16 * std::int64_t(c);
      T Ki[9];
#pragma unroll
      for (int q = 0; q < 9; ++q)
        Ki[q] = rec[q];
      const T lam = rec[10], mu = rec[11];
      for (int j = 0; j < K; ++j)
      {
        const std::uint32_t colb = (m >> (3 * j)) & 7u;
        if (!colb)
          continue;
        const T* Tm = ref + (i * K + j) * 10;
        T W[9], P[9];
#pragma unroll
        for (int aa = 0; aa < 3; ++aa)
#pragma unroll
          for (int d = 0; d < 3; ++d)
            W[aa * 3 + d] = Tm[aa * 3 + 0] * Ki[0 * 3 + d] + Tm[aa * 3 + 1] * Ki[1 * 3 + d]
                            + Tm[aa * 3 + 2] * Ki[2 * 3 + d];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
          for (int d = 0; d < 3; ++d)
            P[cc * 3 + d] = Ki[0 * 3 + cc] * W[0 * 3 + d] + Ki[1 * 3 + cc] * W[1 * 3 + d]
                            + Ki[2 * 3 + cc] * W[2 * 3 + d];
        const T tr = P[0] + P[4] + P[8];
        const std::int64_t node = dm[std::int64_t(c) * K + j];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
          if (!((colb >> d) & 1u))
            continue;
          const T v = comp == d ? mu * (tr + P[comp * 4]) + lam * P[comp * 4]
                                : mu * P[d * 3 + comp] + lam * P[comp * 3 + d];
          const std::int64_t col = 3 * node + d;
          const double gv = g ? g[col] : 0.0;
          const double xv = x0 ? double(x0[col]) : 0.0;
          s += double(v) * (gv - xv);
        }
      }
    }
  }
  if (s != 0.0)
    b[t] = static_cast<T>(double(b[t]) - alpha * s);
}

/// Element vectors of the tetrahedral source forms with constant data,
/// one thread per cell, stored corner-major (be[(i BS + comp) nc + c]):
/// P2 body force b_(i,c) = |det J| w_i f_c; P3 source b_i = |det J| w_i f
/// (w_i = int phi_i over the reference cell).
template <typename T, int K, int BS>
__global__ void __launch_bounds__(kBlock)
    cell_const_source_k(std::int32_t nc, const T* __restrict__ x, const std::int32_t* __restrict__ xdm,
                        const T* __restrict__ w, T f0, T f1, T f2, T* __restrict__ be)
{
  const std::int64_t c = blockIdx.x * std::int64_t(kBlock) + threadIdx.x;
  if (c >= nc)
    return;
  const std::int32_t* v = xdm + 4 * c;
  const T* q0 = x + 3 * std::int64_t(v[0]);
  const T* q1 = x + 3 * std::int64_t(v[1]);
  const T* q2 = x + 3 * std::int64_t(v[2]);
  const T* q3 = x + 3 * std::int64_t(v[3]);
  T J[9], adj[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    J[3 * r + 0] = q1[r] - q0[r];
    J[3 * r + 1] = q2[r] - q0[r];
    J[3 * r + 2] = q3[r] - q0[r];
  }
  const T det = adjugate(J, adj);
  const T adet = det < T(0) ? -det : det;
  const T f[3] = {f0, f1, f2};
  for (int i = 0; i < K; ++i)
  {
#pragma unroll
    for (int b = 0; b < BS; ++b)
      be[std::int64_t(i * BS + b) * nc + c] = adet * w[i] * f[b];
  }
}


/// Reference tensors of the affine stiffness + reaction forms, as
/// [7][k][k]: the symmetrised stiffness integrals S00, S11, S22, S01+S10,
/// S02+S20, S12+S21 (S_ab(i,j) = int d_a phi_i d_b phi_j over the reference
/// cell) and the mass M(i,j) = int phi_i phi_j.
std::vector<double> reference_tensors(int kernel)
{
  if (kernel == DFX_K_P1TET_A)
  {
    const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::vector<double> t(7 * 16);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
      {
        const int e = 4 * i + j;
        auto S = [&](int a, int b) { return d[i][a] * d[j][b] / 6.0; };
        t[0 * 16 + e] = S(0, 0);
        t[1 * 16 + e] = S(1, 1);
        t[2 * 16 + e] = S(2, 2);
        t[3 * 16 + e] = S(0, 1) + S(1, 0);
        t[4 * 16 + e] = S(0, 2) + S(2, 0);
        t[5 * 16 + e] = S(1, 2) + S(2, 1);
        t[6 * 16 + e] = (i == j ? 2.0 : 1.0) / 120.0;
      }
    return t;
  }
  if (kernel == DFX_K_Q1HEX_A)
  {
    std::vector<double> t(7 * 64);
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 8; ++j)
      {
        double m[3], kk[3], dd[3], ee[3];
        for (int a = 0; a < 3; ++a)
        {
          const int p = (i >> a) & 1, q = (j >> a) & 1;
          m[a] = p == q ? 1.0 / 3.0 : 1.0 / 6.0;
          kk[a] = p == q ? 1.0 : -1.0;
          dd[a] = p ? 0.5 : -0.5;
          ee[a] = q ? 0.5 : -0.5;
        }
        const int e = 8 * i + j;
        t[0 * 64 + e] = kk[0] * m[1] * m[2];
        t[1 * 64 + e] = m[0] * kk[1] * m[2];
        t[2 * 64 + e] = m[0] * m[1] * kk[2];
        t[3 * 64 + e] = (dd[0] * ee[1] + ee[0] * dd[1]) * m[2];
        t[4 * 64 + e] = (dd[0] * ee[2] + ee[0] * dd[2]) * m[1];
        t[5 * 64 + e] = (dd[1] * ee[2] + ee[1] * dd[2]) * m[0];
        t[6 * 64 + e] = m[0] * m[1] * m[2];
      }
    return t;
  }
  if (kernel == DFX_K_P3TET_A)
    return scalar_tensors(tet_basis(3));
  if (kernel == DFX_K_P2TET_ELAST_A)
    return gradient_tensors(tet_basis(2));
  if (kernel == DFX_K_P2TET_ELAST_L || kernel == DFX_K_P3TET_L)
  {
    const TetBasis B = tet_basis(kernel == DFX_K_P3TET_L ? 3 : 2);
    std::vector<double> w(B.n);
    for (int i = 0; i < B.n; ++i)
      w[i] = basis_integral(B, i);
    return w;
  }
  return {};
}

bool row_path(const dfx_form& f)
{
  return f.affine && f.ref_tensors
         && (f.d.kernel == DFX_K_P1TET_A || f.d.kernel == DFX_K_Q1HEX_A
             || f.d.kernel == DFX_K_P3TET_A || f.d.kernel == DFX_K_P2TET_ELAST_A);
}

/// Incidence slot offsets of form f for the pattern of A (built once).
void ensure_plan(dfx_form& f, const dfx_csr& A)
{
  if (f.islots && f.plan_pattern == A.pattern_id)
    return;
  dfree(f.islots);
  f.islots = nullptr;
  const int k = f.d.dofs_per_cell;
  const std::int64_t ninc = std::int64_t(f.d.num_cells) * k;
  require(A.max_row <= 65536, "row too long for the slot map");
  f.islot_bytes = A.max_row <= 256 ? 1 : 2;
  f.islots = dmalloc(std::max<std::int64_t>(ninc * k, 1) * f.islot_bytes);
  int* err = static_cast<int*>(dmalloc(sizeof(int)));
  DFX_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
  if (ninc > 0)
  {
    if (f.islot_bytes == 1)
      incidence_slots_k<std::uint8_t><<<blocks_for(ninc * k), kBlock, 0, stream()>>>(
          ninc, k, f.inc, f.dofmap, A.row_ptr, A.cols, static_cast<std::uint8_t*>(f.islots), err);
    else
      incidence_slots_k<std::uint16_t><<<blocks_for(ninc * k), kBlock, 0, stream()>>>(
          ninc, k, f.inc, f.dofmap, A.row_ptr, A.cols, static_cast<std::uint16_t*>(f.islots), err);
    check_launch("incidence_slots");
  }
  int h = 0;
  d2h(&h, err, sizeof(int));
  dfree(err);
  if (h)
  {
    dfree(f.islots);
    f.islots = nullptr;
    throw Error(DFX_ERR_PATTERN, "Entry not in sparsity");
  }
  f.plan_pattern = A.pattern_id;
}

template <typename T>
void launch_cell_factors(const dfx_form& f, const dfx_bc* bc, bool compact = false)
{
  // done by the system cell pass of this assembly
  if (f.cell_data_ready)
  {
    f.cell_data_ready = false;
    return;
  }
  const int nv = f.d.nodes_per_cell;
  const std::int32_t nc = f.d.num_cells;
  const T creac = static_cast<T>(f.constants[0]);
  std::uint32_t* mask = bc ? f.cell_mask : nullptr;
  const bool same = f.dofmap == f.x_dofmap && f.d.dofs_per_cell == nv;
  auto go = [&](auto kern)
  {
    kern<<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, f.dofmap, f.d.dofs_per_cell,
        static_cast<const T*>(f.coeffs), f.d.cstride, creac, bc ? bc->marker : nullptr,
        static_cast<T*>(f.cell_data), mask);
  };
  if (nc > 0)
  {
    if (compact)
    {
      require(nv == 8 && same, "compact records need Q1 on the vertex map");
      go(cell_factors_k<T, 8, true, true>);
    }
    else if (nv == 4)
      same ? go(cell_factors_k<T, 4, true>) : go(cell_factors_k<T, 4, false>);
    else
      same ? go(cell_factors_k<T, 8, true>) : go(cell_factors_k<T, 8, false>);
  }
  check_launch("cell_factors");
}

/// Lanes per row of the "lanes" variant: the power of two covering the
/// most incidences of one row (8 for Q1 hexahedra, 32 for P1 tetrahedra).
int incidence_lanes(const dfx_form& f)
{
  int L = 8;
  while (L < 64 && L < f.max_inc)
    L *= 2;
  return L;
}

/// Row-kernel variants, picked per form by dfx_form_autotune: 0..2 "rows"
/// with 64/128/256 rows per workgroup; 3 "lanes"; 4..6 closed-form Q1 rows;
/// 7, 8 incidence-parallel gathers; 9 structured Q1; 10 fused geometry;
/// 11 as 1 with eight incidences in flight; 12..14 fixed-width Q1 rows;
/// 15 structured Q1 rows in registers stored through an LDS slice; 16 row
/// teams forming the entries; 17 row teams gathering the element matrices.
constexpr int kRowVariants = 18;

template <typename T, typename O, int K, int L>
void launch_lanes(dfx_form& f, dfx_csr& A, const dfx_bc* bc)
{
  constexpr int RPB = kBlock / L;
  const std::size_t lds = (7 * K * K + std::size_t(RPB) * A.max_row) * sizeof(T);
  require(lds <= 64 * 1024, "row too long for the LDS slice");
  const int nb = (A.nrows + RPB - 1) / RPB;
  if (nb > 0)
    row_assemble_lanes_k<T, O, K, L><<<nb, kBlock, lds, stream()>>>(
        A.nrows, f.max_dof + 1, f.inc_off, f.inc, static_cast<const O*>(f.islots),
        static_cast<const T*>(f.cell_data), bc ? f.cell_mask : nullptr, bc ? bc->marker : nullptr,
        A.row_ptr, static_cast<const T*>(f.ref_tensors), static_cast<T*>(A.values),
        A.zeroed ? 0 : 1);
}

/// Row groups of the elasticity team kernel. Rows are ordered by their
/// incidence count (longest first) inside chunks of 4096 consecutive rows,
/// so the 6 rows of a wavefront take similar numbers of steps while a
/// chunk's cells stay together in the L2; a group is cut at 6 rows or when
/// its padded LDS slice would exceed kTeamSlice. Groups are stored as
/// (first position, slice length) pairs ended by (number of rows, 0), rows
/// as (row, slice offset, first incidence, incidences).
void ensure_teams(dfx_form& f, const dfx_csr& A, std::size_t tsize)
{
  const std::int64_t cap = static_cast<std::int64_t>(kTeamSlice / tsize);
  if (f.team_grp && f.team_pattern == A.pattern_id && f.team_cap == cap)
    return;
  dfree(f.team_grp);
  f.team_grp = nullptr;
  f.nteam = 0;
  if (A.nrows == 0)
    return;
  require(10 * std::int64_t(A.max_row) <= cap, "row too long for the LDS slice");
  const std::int32_t nplan = f.max_dof + 1;
  std::vector<std::int32_t> off(nplan + 1), rp(A.nrows + 1);
  d2h(off.data(), f.inc_off, sizeof(std::int32_t) * off.size());
  d2h(rp.data(), A.row_ptr, sizeof(std::int32_t) * rp.size());
  auto count = [&](std::int32_t r) { return r < nplan ? off[r + 1] - off[r] : 0; };
  std::vector<std::int32_t> order(A.nrows);
  for (std::int32_t r = 0; r < A.nrows; ++r)
    order[r] = r;
  std::sort(order.begin(), order.end(),
            [&](std::int32_t x, std::int32_t y)
            {
              if ((x >> 12) != (y >> 12))
                return x < y;
              const std::int32_t cx = count(x), cy = count(y);
              return cx != cy ? cx > cy : x < y;
            });
  std::vector<std::int32_t> grp{0, 0}, rows;
  rows.reserve(4 * std::size_t(A.nrows));
  std::int64_t slice = 0;
  int members = 0;
  for (std::int32_t q = 0; q < A.nrows; ++q)
  {
    const std::int32_t r = order[q];
    const std::int64_t len = 10 * std::int64_t(rp[r + 1] - rp[r]);
    if (members == 6 || slice + len > cap)
    {
      grp.back() = static_cast<std::int32_t>(slice);
      grp.push_back(q);
      grp.push_back(0);
      slice = 0;
      members = 0;
    }
    rows.push_back(r);
    rows.push_back(static_cast<std::int32_t>(slice));
    rows.push_back(r < nplan ? off[r] : 0);
    rows.push_back(count(r));
    slice += len;
    ++members;
  }
  grp.back() = static_cast<std::int32_t>(slice);
  grp.push_back(A.nrows);
  grp.push_back(0);
  f.nteam = static_cast<std::int32_t>(grp.size() / 2) - 1;
  // the row quads start at a multiple of 4 (16-byte aligned)
  const std::size_t at = (grp.size() + 3) & ~std::size_t(3);
  std::vector<std::int32_t> buf(at + rows.size(), 0);
  std::copy(grp.begin(), grp.end(), buf.begin());
  std::copy(rows.begin(), rows.end(), buf.begin() + at);
  f.team_grp = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * buf.size()));
  h2d(f.team_grp, buf.data(), sizeof(std::int32_t) * buf.size());
  f.team_rows_at = static_cast<std::int32_t>(at);
  f.team_pattern = A.pattern_id;
  f.team_cap = cap;
}

/// Row order of variant 16 (built once per pattern): rows sorted by their
/// incidence count, longest first, inside chunks of 4096 consecutive rows
/// (so a chunk's cells stay together in the L2), each as (row, first
/// incidence, incidences, 0).
void ensure_trow(dfx_form& f, const dfx_csr& A)
{
  if (f.trow && f.trow_pattern == A.pattern_id)
    return;
  dfree(f.trow);
  f.trow = nullptr;
  f.ntrow = 0;
  const std::int32_t nplan = f.max_dof + 1;
  std::vector<std::int32_t> off(nplan + 1);
  d2h(off.data(), f.inc_off, sizeof(std::int32_t) * off.size());
  auto count = [&](std::int32_t r) { return r < nplan ? off[r + 1] - off[r] : 0; };
  std::vector<std::int32_t> order(A.nrows);
  for (std::int32_t r = 0; r < A.nrows; ++r)
    order[r] = r;
  std::sort(order.begin(), order.end(),
            [&](std::int32_t x, std::int32_t y)
            {
              if ((x >> 12) != (y >> 12))
                return x < y;
              const std::int32_t cx = count(x), cy = count(y);
              return cx != cy ? cx > cy : x < y;
            });
  std::vector<std::int32_t> rows(4 * std::size_t(A.nrows));
  for (std::int32_t q = 0; q < A.nrows; ++q)
  {
    const std::int32_t r = order[q];
    rows[4 * std::size_t(q)] = r;
    rows[4 * std::size_t(q) + 1] = r < nplan ? off[r] : 0;
    rows[4 * std::size_t(q) + 2] = count(r);
    rows[4 * std::size_t(q) + 3] = 0;
  }
  f.trow = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max<std::size_t>(rows.size(), 4)));
  h2d(f.trow, rows.data(), sizeof(std::int32_t) * rows.size());
  f.ntrow = A.nrows;
  f.trow_pattern = A.pattern_id;
}

template <typename T, typename O>
void launch_elast(dfx_form& f, dfx_csr& A, const dfx_bc* bc)
{
  const std::int32_t nc = f.d.num_cells;
  if (nc > 0)
    elast_factors_k<T><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, f.dofmap, static_cast<const T*>(f.coeffs),
        f.d.cstride, bc ? bc->marker : nullptr, static_cast<T*>(f.cell_data),
        bc ? f.cell_mask : nullptr);
  // the team kernel addresses the cell records and slot offsets with 32-bit
  // byte offsets
  require(std::int64_t(nc) * 16 * std::int64_t(sizeof(T)) <= 0xffffffffLL
              && std::int64_t(nc) * 100 * std::int64_t(sizeof(O)) <= 0xffffffffLL,
          "mesh too large for the elasticity row kernel (cell records above 4 GiB)");
  ensure_teams(f, A, sizeof(T));
  if (f.nteam > 0)
    elast_team_k<T, O><<<f.nteam, 64, kTeamSlice, stream()>>>(
        reinterpret_cast<const int2*>(f.team_grp),
        reinterpret_cast<const int4*>(f.team_grp + f.team_rows_at), f.inc,
        static_cast<const O*>(f.islots),
        static_cast<const T*>(f.cell_data), bc ? f.cell_mask : nullptr, bc ? bc->marker : nullptr,
        A.row_ptr, static_cast<const T*>(f.ref_tensors), static_cast<T*>(A.values),
        A.zeroed ? 0 : 1);
  check_launch("elasticity assembly");
}

/// Structured-box plan of a Q1 form (corner cells and slot positions);
/// returns false when the mesh is not a structured box of hexahedra.
bool ensure_struct(dfx_form& f, const dfx_csr& A)
{
  if (f.struct_pattern == A.pattern_id && f.struct_ok >= 0)
    return f.struct_ok == 1;
  dfree(f.corner_cell);
  dfree(f.slotmap);
  f.corner_cell = nullptr;
  f.slotmap = nullptr;
  f.struct_pattern = A.pattern_id;
  f.struct_ok = 0;
  const std::int32_t nrows = A.nrows, nc = f.d.num_cells;
  if (f.d.kernel != DFX_K_Q1HEX_A || !f.affine || f.max_dof + 1 != nrows || A.max_row > 27)
    return false;
  f.corner_cell = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * 8 * std::max(nrows, 1)));
  f.slotmap = static_cast<std::uint8_t*>(dmalloc(32 * std::size_t(std::max(nrows, 1))));
  std::int32_t* tmp = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * 27 * std::size_t(std::max(nrows, 1))));
  int* err = static_cast<int*>(dmalloc(sizeof(int)));
  DFX_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
  fill_i32_k<<<blocks_for(std::int64_t(nrows) * 8), kBlock, 0, stream()>>>(f.corner_cell, std::int64_t(nrows) * 8, -1);
  fill_i32_k<<<blocks_for(std::int64_t(nrows) * 27), kBlock, 0, stream()>>>(tmp, std::int64_t(nrows) * 27, -1);
  if (nc > 0)
  {
    struct_corner_k<<<blocks_for(std::int64_t(nc) * 8), kBlock, 0, stream()>>>(nc, f.dofmap, f.corner_cell, err);
    struct_slots_k<<<blocks_for(std::int64_t(nc) * 64), kBlock, 0, stream()>>>(nc, f.dofmap, A.row_ptr, A.cols, tmp, err);
  }
  if (nrows > 0)
    struct_pack_k<<<blocks_for(nrows), kBlock, 0, stream()>>>(nrows, tmp, A.row_ptr, f.slotmap, err);
  check_launch("struct_plan");
  int h = 0;
  d2h(&h, err, sizeof(int));
  dfree(err);
  dfree(tmp);
  if (h)
  {
    dfree(f.corner_cell);
    dfree(f.slotmap);
    f.corner_cell = nullptr;
    f.slotmap = nullptr;
    return false;
  }
  f.struct_ok = 1;
  return true;
}

/// Row blocks of the incidence-parallel variants: consecutive rows with at
/// most one incidence per thread and an LDS slice of at most 32 KB, plus the
/// row of every incidence. Built once per pattern.
void ensure_gather(dfx_form& f, const dfx_csr& A, std::size_t tsize)
{
  if (f.gblk && f.gather_pattern == A.pattern_id)
    return;
  dfree(f.gblk);
  dfree(f.inc_row);
  f.gblk = nullptr;
  f.inc_row = nullptr;
  const std::int32_t nplan = f.max_dof + 1;
  std::vector<std::int32_t> off(nplan + 1), rp(A.nrows + 1);
  d2h(off.data(), f.inc_off, sizeof(std::int32_t) * off.size());
  d2h(rp.data(), A.row_ptr, sizeof(std::int32_t) * rp.size());
  std::vector<std::int32_t> irow(std::max<std::int32_t>(off[nplan], 1));
  for (std::int32_t r = 0; r < nplan; ++r)
    for (std::int32_t q = off[r]; q < off[r + 1]; ++q)
      irow[q] = r;
  const std::int64_t cap_inc = kBlock;
  const std::int64_t cap_slice = (32 * 1024) / static_cast<std::int64_t>(tsize);
  std::vector<std::int32_t> gb{0}, ranks;
  std::int64_t slice = 0;
  std::int32_t r = 0;
  while (r < A.nrows)
  {
    std::int32_t e = r;
    std::int64_t ninc = 0, nent = 0, most = 0;
    while (e < A.nrows)
    {
      const std::int64_t ci = e < nplan ? off[e + 1] - off[e] : 0;
      const std::int64_t ce = rp[e + 1] - rp[e];
      if (ninc + ci > cap_inc || nent + ce > cap_slice)
      {
        require(e > r, "row too long for the incidence-parallel variants");
        break;
      }
      ninc += ci;
      nent += ce;
      most = std::max(most, ci);
      ++e;
    }
    gb.push_back(e);
    ranks.push_back(static_cast<std::int32_t>(most));
    slice = std::max(slice, nent);
    r = e;
  }
  f.ngblk = static_cast<std::int32_t>(gb.size()) - 1;
  // after the ngblk + 1 row boundaries: each block's most incidences of a
  // row (its number of rounds)
  gb.insert(gb.end(), ranks.begin(), ranks.end());
  f.gblk_slice = static_cast<std::int32_t>(slice);
  f.gblk = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * gb.size()));
  h2d(f.gblk, gb.data(), sizeof(std::int32_t) * gb.size());
  f.inc_row = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * irow.size()));
  h2d(f.inc_row, irow.data(), sizeof(std::int32_t) * irow.size());
  f.gather_pattern = A.pattern_id;
}

/// Variants 7 (element matrices on the matrix cores, then gathered by rows)
/// and 8 (entries formed from the cell factors while gathering).
template <typename T, typename O, int K>
void launch_gather(dfx_form& f, dfx_csr& A, const dfx_bc* bc, bool mat)
{
  ensure_gather(f, A, sizeof(T));
  const std::size_t lds = ((mat ? 0 : 7 * K * K) + std::size_t(f.gblk_slice)) * sizeof(T);
  require(lds <= 64 * 1024, "row too long for the LDS slice");
  launch_cell_factors<T>(f, bc);
  const std::int32_t nc = f.d.num_cells;
  if (mat)
  {
    if (!f.cell_elem)
      f.cell_elem = dmalloc(sizeof(T) * std::max<std::int64_t>(std::int64_t(nc) * K * K, 1));
    const int nb = static_cast<int>((std::int64_t(nc) + 63) / 64);
    if (nb > 0)
      elem_mfma_k<T, K><<<nb, kBlock, 0, stream()>>>(
          nc, static_cast<const T*>(f.cell_data), static_cast<const T*>(f.ref_tensors),
          static_cast<T*>(f.cell_elem));
  }
  if (f.ngblk == 0)
    return;
  auto go = [&](auto kern)
  {
    kern<<<f.ngblk, kBlock, lds, stream()>>>(
        f.gblk, f.inc_off, f.inc, f.inc_row, static_cast<const O*>(f.islots),
        static_cast<const T*>(f.cell_data), static_cast<const T*>(f.cell_elem),
        bc ? f.cell_mask : nullptr, bc ? bc->marker : nullptr, A.row_ptr,
        static_cast<const T*>(f.ref_tensors), static_cast<T*>(A.values), A.zeroed ? 0 : 1,
        f.ngblk, f.max_dof + 1);
  };
  if (mat)
    go(row_gather_k<T, O, K, true>);
  else
    go(row_gather_k<T, O, K, false>);
}

/// Element matrices supplied by the caller (E: cells x K x K), gathered by
/// rows with the incidence-parallel kernel.
template <typename T, typename O, int K>
void launch_elements(dfx_form& f, dfx_csr& A, const dfx_bc* bc, const T* E)
{
  ensure_gather(f, A, sizeof(T));
  const std::size_t lds = std::size_t(f.gblk_slice) * sizeof(T);
  require(lds <= 64 * 1024, "row too long for the LDS slice");
  const std::int32_t nc = f.d.num_cells;
  if (bc && nc > 0)
    element_mask_k<<<blocks_for(nc), kBlock, 0, stream()>>>(nc, K, f.dofmap, bc->marker,
                                                            f.cell_mask);
  if (f.ngblk > 0)
    row_gather_k<T, O, K, true><<<f.ngblk, kBlock, lds, stream()>>>(
        f.gblk, f.inc_off, f.inc, f.inc_row, static_cast<const O*>(f.islots), nullptr, E,
        bc ? f.cell_mask : nullptr, bc ? bc->marker : nullptr, A.row_ptr, nullptr,
        static_cast<T*>(A.values), A.zeroed ? 0 : 1, f.ngblk, f.max_dof + 1);
  check_launch("element matrix assembly");
}

template <typename T>
void assemble_elements(dfx_form& f, dfx_csr& A, const dfx_bc* bc, const T* E)
{
  ensure_plan(f, A);
  auto go = [&](auto o)
  {
    using O = decltype(o);
    switch (f.d.dofs_per_cell)
    {
    case 4:
      launch_elements<T, O, 4>(f, A, bc, E);
      break;
    case 8:
      launch_elements<T, O, 8>(f, A, bc, E);
      break;
    case 10:
      launch_elements<T, O, 10>(f, A, bc, E);
      break;
    default:
      launch_elements<T, O, 20>(f, A, bc, E);
    }
  };
  if (f.islot_bytes == 1)
    go(std::uint8_t{});
  else
    go(std::uint16_t{});
}

/// Fixed-width plan of a Q1 form (built once per pattern); false if a row
/// has more than 8 incidences or the slot offsets are wider than a byte.
bool ensure_fixed(dfx_form& f, const dfx_csr& A)
{
  if (f.max_inc > 8 || f.islot_bytes != 1)
    return false;
  if (f.inc8 && f.fixed_pattern == A.pattern_id)
    return true;
  dfree(f.inc8);
  dfree(f.islot8);
  f.inc8 = nullptr;
  f.islot8 = nullptr;
  const std::int32_t nplan = f.max_dof + 1;
  f.inc8 = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * 8 * std::max(nplan, 1)));
  f.islot8 = static_cast<std::uint8_t*>(dmalloc(64 * std::size_t(std::max(nplan, 1))));
  if (nplan > 0)
    fixed_plan_k<<<blocks_for(nplan), kBlock, 0, stream()>>>(
        nplan, f.inc_off, f.inc, static_cast<const std::uint8_t*>(f.islots), f.inc8, f.islot8);
  check_launch("fixed plan");
  f.fixed_pattern = A.pattern_id;
  return true;
}

template <typename T, typename O, int K>
void launch_row_assemble(dfx_form& f, dfx_csr& A, const dfx_bc* bc)
{
  const int variant = f.row_variant;
  if (variant == 15)
  {
    if constexpr (K == 8)
    {
      require(ensure_struct(f, A), "variant not available for this mesh");
      const bool compact = f.compact_records;
      launch_cell_factors<T>(f, bc, compact);
      const int R = 128;
      const std::size_t lds = std::size_t(R) * A.max_row * sizeof(T);
      const int nb = (A.nrows + R - 1) / R;
      const std::uint32_t* mk = bc ? f.cell_mask : nullptr;
      if (nb > 0 && compact)
        row_struct_lds_k<T, true><<<nb, R, lds, stream()>>>(
            A.nrows, f.corner_cell, f.slotmap, static_cast<const T*>(f.cell_data), mk,
            bc ? bc->marker : nullptr, A.row_ptr, static_cast<T*>(A.values), A.zeroed ? 0 : 1);
      else if (nb > 0)
        row_struct_lds_k<T><<<nb, R, lds, stream()>>>(
            A.nrows, f.corner_cell, f.slotmap, static_cast<const T*>(f.cell_data), mk,
            bc ? bc->marker : nullptr, A.row_ptr, static_cast<T*>(A.values), A.zeroed ? 0 : 1);
      check_launch("row_assemble");
      return;
    }
    else
      throw Error(DFX_ERR_ARG, "variant not available for this element");
  }
  if (variant >= 12 && variant <= 14)
  {
    // fixed-width rows: 64, 128 or 256 rows per workgroup
    require(K == 8 && f.d.kernel == DFX_K_Q1HEX_A && f.affine && ensure_fixed(f, A),
            "variant not available for this element");
    launch_cell_factors<T>(f, bc);
    const int R = 64 << (variant - 12);
    const std::size_t lds = std::size_t(R) * A.max_row * sizeof(T);
    require(lds <= 64 * 1024, "row too long for the LDS slice");
    const int nb = (A.nrows + R - 1) / R;
    if (nb > 0)
      row_fixed_k<T><<<nb, R, lds, stream()>>>(
          A.nrows, f.max_dof + 1, f.inc8, f.islot8, static_cast<const T*>(f.cell_data),
          bc ? bc->marker : nullptr, A.row_ptr, static_cast<T*>(A.values), A.zeroed ? 0 : 1);
    check_launch("row_assemble");
    return;
  }
  if (variant == 10)
  {
    // fused geometry: no cell pass; 64 rows per workgroup
    require(f.affine && (K == 4 || K == 8) && f.d.kernel != DFX_K_P3TET_A,
            "variant not available for this element");
    if constexpr (K == 4 || K == 8)
    {
      constexpr bool cf = K == 8;
      const int R = 64;
      const std::size_t lds = ((cf ? 0 : 7 * K * K) + std::size_t(R) * A.max_row) * sizeof(T);
      require(lds <= 64 * 1024, "row too long for the LDS slice");
      const int nb = (A.nrows + R - 1) / R;
      auto go = [&](auto kern)
      {
        kern<<<nb, R, lds, stream()>>>(
            A.nrows, f.max_dof + 1, f.inc_off, f.inc, static_cast<const O*>(f.islots),
            static_cast<const T*>(f.x), f.x_dofmap, f.dofmap, static_cast<const T*>(f.coeffs),
            f.d.cstride, static_cast<T>(f.constants[0]), bc ? bc->marker : nullptr, A.row_ptr,
            static_cast<const T*>(f.ref_tensors), static_cast<T*>(A.values), A.zeroed ? 0 : 1);
      };
      if (nb > 0)
      {
        if (f.x_dofmap == f.dofmap)
          go(row_fused_k<T, O, K, cf, true>);
        else
          go(row_fused_k<T, O, K, cf, false>);
      }
      check_launch("row_assemble");
    }
    return;
  }
  if (variant == 9)
  {
    if constexpr (K == 8)
    {
      require(ensure_struct(f, A), "variant not available for this mesh");
      launch_cell_factors<T>(f, bc);
      const int nb = (A.nrows + kBlock - 1) / kBlock;
      if (nb > 0)
        row_struct_q1_k<T><<<nb, kBlock, 0, stream()>>>(
            A.nrows, f.corner_cell, f.slotmap, static_cast<const T*>(f.cell_data),
            bc ? bc->marker : nullptr, A.row_ptr, static_cast<T*>(A.values), A.zeroed ? 0 : 1);
      check_launch("row_assemble");
      return;
    }
    else
      throw Error(DFX_ERR_ARG, "variant not available for this element");
  }
  if (variant == 7 || variant == 8)
  {
    launch_gather<T, O, K>(f, A, bc, variant == 7);
    check_launch("row_assemble");
    return;
  }
  if (variant == 17)
  {
    // teams of K lanes per row gathering the element matrices of the
    // matrix-core product
    ensure_trow(f, A);
    constexpr int G = 64 / K;
    int nw = 16;
    auto lds_for = [&](int w) { return std::size_t(w) * G * A.max_row * sizeof(T); };
    while (nw > 4 && lds_for(nw) > 64 * 1024)
      nw /= 2;
    const std::size_t lds = lds_for(nw);
    require(lds <= 64 * 1024, "row too long for the LDS slice");
    launch_cell_factors<T>(f, bc);
    const std::int32_t nc = f.d.num_cells;
    if (!f.cell_elem)
      f.cell_elem = dmalloc(sizeof(T) * std::max<std::int64_t>(std::int64_t(nc) * K * K, 1));
    const int nbe = static_cast<int>((std::int64_t(nc) + 63) / 64);
    if (nbe > 0)
      elem_mfma_k<T, K><<<nbe, kBlock, 0, stream()>>>(
          nc, static_cast<const T*>(f.cell_data), static_cast<const T*>(f.ref_tensors),
          static_cast<T*>(f.cell_elem));
    const std::int64_t per_block = std::int64_t(nw) * G;
    const int nb = static_cast<int>((std::int64_t(f.ntrow) + per_block - 1) / per_block);
    if (nb > 0)
      row_team_gather_k<T, O, K><<<nb, 64 * nw, lds, stream()>>>(
          f.ntrow, reinterpret_cast<const int4*>(f.trow), f.inc, static_cast<const O*>(f.islots),
          static_cast<const T*>(f.cell_elem), bc ? f.cell_mask : nullptr,
          bc ? bc->marker : nullptr, A.row_ptr, static_cast<T*>(A.values), A.zeroed ? 0 : 1,
          A.max_row);
    check_launch("row_assemble");
    return;
  }
  if (variant == 16)
  {
    // teams of K lanes per row; the wavefronts of a workgroup (16, 8 or 4,
    // the most whose slices fit 64 KB of LDS) share the reference tensors
    ensure_trow(f, A);
    constexpr int G = 64 / K;
    int nw = 16;
    auto lds_for = [&](int w)
    {
      const std::size_t slices = (std::size_t(w) * G * A.max_row + 3) & ~std::size_t(3);
      return (7 * std::size_t(K) * K + slices + std::size_t(w) * G * 8) * sizeof(T);
    };
    while (nw > 4 && lds_for(nw) > 64 * 1024)
      nw /= 2;
    const std::size_t lds = lds_for(nw);
    require(lds <= 64 * 1024, "row too long for the LDS slice");
    launch_cell_factors<T>(f, bc);
    const std::int64_t per_block = std::int64_t(nw) * G;
    const int nb = static_cast<int>((std::int64_t(f.ntrow) + per_block - 1) / per_block);
    if (nb > 0)
      row_team_k<T, O, K><<<nb, 64 * nw, lds, stream()>>>(
          f.ntrow, reinterpret_cast<const int4*>(f.trow), f.inc, static_cast<const O*>(f.islots),
          static_cast<const T*>(f.cell_data), bc ? f.cell_mask : nullptr,
          bc ? bc->marker : nullptr, A.row_ptr, static_cast<const T*>(f.ref_tensors),
          static_cast<T*>(A.values), A.zeroed ? 0 : 1, A.max_row);
    check_launch("row_assemble");
    return;
  }
  // wide elements: the LDS slice of a thread-per-row block does not fit
  require(K <= 8 || variant == 3, "variant not available for this element");
  launch_cell_factors<T>(f, bc);
  if (variant != 3)
  {
    // 0-2: four incidences per group, reference tensors in LDS;
    // 4-6: closed-form entries for Q1 hexahedra (four per group), else
    // eight incidences per group; 11: as 1 with eight per group
    const bool wide = variant >= 4 && variant != 11;
    const bool cf = wide && K == 8 && f.d.kernel == DFX_K_Q1HEX_A;
    const int R = variant == 11 ? 128 : 64 << (wide ? variant - 4 : variant);
    const std::size_t lds
        = ((cf ? 0 : 7 * K * K) + std::size_t(R) * A.max_row) * sizeof(T);
    require(lds <= 64 * 1024, "row too long for the LDS slice");
    const int nb = (A.nrows + R - 1) / R;
    auto go = [&](auto kern)
    {
      kern<<<nb, R, lds, stream()>>>(
          A.nrows, f.max_dof + 1, f.inc_off, f.inc, static_cast<const O*>(f.islots),
          static_cast<const T*>(f.cell_data), bc ? f.cell_mask : nullptr,
          bc ? bc->marker : nullptr, A.row_ptr, static_cast<const T*>(f.ref_tensors),
          static_cast<T*>(A.values), A.zeroed ? 0 : 1);
    };
    if (nb > 0)
    {
      if (variant == 11)
        go(row_assemble_k<T, O, K, 8, false>);
      else if (!wide)
        go(row_assemble_k<T, O, K, 4, false>);
      else if constexpr (K == 8)
      {
        if (cf)
          go(row_assemble_k<T, O, K, 4, true>);
        else
          go(row_assemble_k<T, O, K, 8, false>);
      }
      else
        go(row_assemble_k<T, O, K, 8, false>);
    }
  }
  else
  {
    switch (incidence_lanes(f))
    {
    case 8:
      launch_lanes<T, O, K, 8>(f, A, bc);
      break;
    case 16:
      launch_lanes<T, O, K, 16>(f, A, bc);
      break;
    case 32:
      launch_lanes<T, O, K, 32>(f, A, bc);
      break;
    default:
      launch_lanes<T, O, K, 64>(f, A, bc);
    }
  }
  check_launch("row_assemble");
}

template <typename T>
void assemble_rows(dfx_form& f, dfx_csr& A, const dfx_bc* bc)
{
  ensure_plan(f, A);
  const int K = f.d.dofs_per_cell;
  if (f.d.kernel == DFX_K_P2TET_ELAST_A)
  {
    if (f.islot_bytes == 1)
      launch_elast<T, std::uint8_t>(f, A, bc);
    else
      launch_elast<T, std::uint16_t>(f, A, bc);
  }
  else if (K == 4)
  {
    if (f.islot_bytes == 1)
      launch_row_assemble<T, std::uint8_t, 4>(f, A, bc);
    else
      launch_row_assemble<T, std::uint16_t, 4>(f, A, bc);
  }
  else if (K == 8)
  {
    if (f.islot_bytes == 1)
      launch_row_assemble<T, std::uint8_t, 8>(f, A, bc);
    else
      launch_row_assemble<T, std::uint16_t, 8>(f, A, bc);
  }
  else
  {
    if (f.islot_bytes == 1)
      launch_row_assemble<T, std::uint8_t, 20>(f, A, bc);
    else
      launch_row_assemble<T, std::uint16_t, 20>(f, A, bc);
  }
}

template <typename T>
void lift_rows(dfx_form& f, const dfx_bc& bc, const T* x0, double alpha, T* b)
{
  // axis-aligned Q1 boxes: the 4-value records (the masks are always
  // written here, a BC being given)
  const bool compact = f.compact_records && f.d.dofs_per_cell == 8;
  launch_cell_factors<T>(f, &bc, compact);
  const std::int32_t nplan = f.max_dof + 1;
  const T* cd = static_cast<const T*>(f.cell_data);
  const T* ref = static_cast<const T*>(f.ref_tensors);
  if (f.d.dofs_per_cell == 4)
    row_lift_k<T, 4><<<blocks_for(nplan), kBlock, 0, stream()>>>(
        nplan, f.inc_off, f.inc, f.dofmap, cd, f.cell_mask, ref, bc.g, x0, alpha, b);
  else if (compact)
    row_lift_k<T, 8, true><<<blocks_for(nplan), kBlock, 0, stream()>>>(
        nplan, f.inc_off, f.inc, f.dofmap, cd, f.cell_mask, ref, bc.g, x0, alpha, b);
  else if (f.d.dofs_per_cell == 8)
    row_lift_k<T, 8><<<blocks_for(nplan), kBlock, 0, stream()>>>(
        nplan, f.inc_off, f.inc, f.dofmap, cd, f.cell_mask, ref, bc.g, x0, alpha, b);
  else
    row_lift_k<T, 20><<<blocks_for(nplan), kBlock, 0, stream()>>>(
        nplan, f.inc_off, f.inc, f.dofmap, cd, f.cell_mask, ref, bc.g, x0, alpha, b);
  check_launch("row_lift");
}

/// apply_lifting of the P2 elasticity form: cell records and the cells'
/// Dirichlet column masks for this BC, then one thread per row dof.
template <typename T>
void elast_lift_rows(dfx_form& f, const dfx_bc& bc, const T* x0, double alpha, T* b)
{
  const std::int32_t nc = f.d.num_cells;
  if (nc > 0)
    elast_factors_k<T><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, f.dofmap, static_cast<const T*>(f.coeffs),
        f.d.cstride, bc.marker, static_cast<T*>(f.cell_data), f.cell_mask);
  const std::int32_t nplan = f.max_dof + 1;
  if (nplan > 0)
    elast_lift_k<T><<<blocks_for(std::int64_t(nplan) * 3), kBlock, 0, stream()>>>(
        nplan, f.inc_off, f.inc, f.dofmap, static_cast<const T*>(f.cell_data), f.cell_mask,
        static_cast<const T*>(f.ref_tensors), bc.g, x0, alpha, b);
  check_launch("elasticity lifting");
}

/// Exterior facet loads: the facets' vertex contributions, then one
/// thread per row adding its facet incidences.
/// The facet matrices of a DFX_K_*_ROBIN_A form into f.fvals.
template <typename T>
void facet_matrices(const dfx_form& f)
{
  const auto* x = static_cast<const T*>(f.x);
  const auto* co = static_cast<const T*>(f.coeffs);
  T* fm = static_cast<T*>(f.fvals);
  if (f.d.kernel == DFX_K_Q1HEX_ROBIN_A)
    facet_matrix_k<T, true><<<blocks_for(f.nfacet), kBlock, 0, stream()>>>(
        f.nfacet, f.facets, x, f.x_dofmap, f.d.nodes_per_cell, co, f.d.cstride, fm);
  else
    facet_matrix_k<T, false><<<blocks_for(f.nfacet), kBlock, 0, stream()>>>(
        f.nfacet, f.facets, x, f.x_dofmap, f.d.nodes_per_cell, co, f.d.cstride, fm);
  check_launch("facet matrices");
}

/// Exterior facet matrices added into A row by row (facet order per row).
template <typename T>
void facet_matrix_rows(const dfx_form& f, dfx_csr& A, const dfx_bc* bc)
{
  if (f.nfacet == 0)
    return;
  facet_matrices<T>(f);
  const std::int8_t* mark = bc ? bc->marker : nullptr;
  auto* vals = static_cast<T*>(A.values);
  const auto* fm = static_cast<const T*>(f.fvals);
  const bool hex = f.d.kernel == DFX_K_Q1HEX_ROBIN_A;
  if (f.facet_pattern == A.pattern_id && A.max_row <= kFacetRowMax)
  {
    const int nb = static_cast<int>((f.nfrows + 63) / 64);
    if (hex)
      facet_rows_fast_k<T, 4, 4, 2><<<nb, 64, 0, stream()>>>(
          f.nfrows, f.frows, f.finc_off, f.finc, f.fslot, fm, mark, A.row_ptr, A.cols, vals);
    else
      facet_rows_fast_k<T, 3, 4, 2><<<nb, 64, 0, stream()>>>(
          f.nfrows, f.frows, f.finc_off, f.finc, f.fslot, fm, mark, A.row_ptr, A.cols, vals);
    check_launch("facet matrix rows");
    return;
  }
  const bool check = f.facet_pattern != A.pattern_id;
  int* err = check ? static_cast<int*>(dmalloc(sizeof(int))) : nullptr;
  if (check)
    DFX_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
  const std::int32_t np = f.nfrows;
  std::uint8_t* fslot = check && A.max_row <= 256 ? f.fslot : nullptr;
  if (hex)
    facet_matrix_rows_k<T, true><<<blocks_for(np), kBlock, 0, stream()>>>(
        np, f.frows, f.finc_off, f.finc, f.facets, f.dofmap, f.d.dofs_per_cell, fm, mark,
        A.row_ptr, A.cols, vals, err, fslot);
  else
    facet_matrix_rows_k<T, false><<<blocks_for(np), kBlock, 0, stream()>>>(
        np, f.frows, f.finc_off, f.finc, f.facets, f.dofmap, f.d.dofs_per_cell, fm, mark,
        A.row_ptr, A.cols, vals, err, fslot);
  check_launch("facet matrix rows");
  if (!check)
    return;
  int h = 0;
  d2h(&h, err, sizeof(int));
  dfree(err);
  if (h)
    throw Error(DFX_ERR_PATTERN, "Entry not in sparsity");
  const_cast<dfx_form&>(f).facet_pattern = A.max_row <= 256 ? A.pattern_id : 0;
}

/// Interior facet macro-element matrices into f.fvals.
template <typename T>
void interior_matrices(const dfx_form& f)
{
  interior_matrix_k<T><<<(f.nfacet + 63) / 64, 64, 0, stream()>>>(
      f.nfacet, f.facets, static_cast<const T*>(f.x), f.x_dofmap, static_cast<T>(f.constants[0]),
      static_cast<T*>(f.fvals));
  check_launch("interior facet matrices");
}

/// Interior facet integral added into A row by row.
template <typename T>
void interior_rows(const dfx_form& f, dfx_csr& A, const dfx_bc* bc)
{
  if (f.nfacet == 0)
    return;
  interior_matrices<T>(f);
  if (f.facet_pattern == A.pattern_id && A.max_row <= kFacetRowMax)
  {
    facet_rows_fast_k<T, 8, 8, 3><<<static_cast<int>((f.nfrows + 63) / 64), 64, 0, stream()>>>(
        f.nfrows, f.frows, f.finc_off, f.finc, f.fslot, static_cast<const T*>(f.fvals),
        bc ? bc->marker : nullptr, A.row_ptr, A.cols, static_cast<T*>(A.values));
    check_launch("interior facet rows");
    return;
  }
  const bool check = f.facet_pattern != A.pattern_id;
  int* err = check ? static_cast<int*>(dmalloc(sizeof(int))) : nullptr;
  if (check)
    DFX_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
  const std::int32_t np = f.max_dof + 1;
  interior_rows_k<T><<<blocks_for(np), kBlock, 0, stream()>>>(
      np, f.finc_off, f.finc, f.facets, f.dofmap, static_cast<const T*>(f.fvals),
      bc ? bc->marker : nullptr, A.row_ptr, A.cols, static_cast<T*>(A.values), err,
      check && A.max_row <= 256 ? f.fslot : nullptr);
  check_launch("interior facet rows");
  if (!check)
    return;
  int h = 0;
  d2h(&h, err, sizeof(int));
  dfree(err);
  if (h)
    throw Error(DFX_ERR_PATTERN, "Entry not in sparsity");
  const_cast<dfx_form&>(f).facet_pattern = A.max_row <= 256 ? A.pattern_id : 0;
}

/// Lifting by the interior facet matrices.
template <typename T>
void interior_lift_rows(const dfx_form& f, const dfx_bc& bc, const T* x0, double alpha, T* b)
{
  if (f.nfacet == 0)
    return;
  interior_matrices<T>(f);
  const std::int32_t np = f.max_dof + 1;
  interior_lift_rows_k<T><<<blocks_for(np), kBlock, 0, stream()>>>(
      np, f.finc_off, f.finc, f.facets, f.dofmap, static_cast<const T*>(f.fvals), bc.marker, bc.g,
      x0, alpha, b);
  check_launch("interior facet lifting");
}

/// Lifting by the exterior facet matrices.
template <typename T>
void facet_lift_rows(const dfx_form& f, const dfx_bc& bc, const T* x0, double alpha, T* b)
{
  if (f.nfacet == 0)
    return;
  facet_matrices<T>(f);
  const std::int32_t np = f.max_dof + 1;
  const auto* fm = static_cast<const T*>(f.fvals);
  if (f.d.kernel == DFX_K_Q1HEX_ROBIN_A)
    facet_lift_rows_k<T, true><<<blocks_for(np), kBlock, 0, stream()>>>(
        np, f.finc_off, f.finc, f.facets, f.dofmap, f.d.dofs_per_cell, fm, bc.marker, bc.g, x0,
        alpha, b);
  else
    facet_lift_rows_k<T, false><<<blocks_for(np), kBlock, 0, stream()>>>(
        np, f.finc_off, f.finc, f.facets, f.dofmap, f.d.dofs_per_cell, fm, bc.marker, bc.g, x0,
        alpha, b);
  check_launch("facet lifting");
}

template <typename T>
void facet_rows(const dfx_form& f, T* b)
{
  if (f.nfacet == 0)
    return;
  T* fv = static_cast<T*>(f.fvals);
  const auto* x = static_cast<const T*>(f.x);
  const auto* co = static_cast<const T*>(f.coeffs);
  if (f.d.kernel == DFX_K_Q1HEX_NEUMANN)
    facet_load_k<T, true><<<blocks_for(f.nfacet), kBlock, 0, stream()>>>(
        f.nfacet, f.facets, x, f.x_dofmap, f.d.nodes_per_cell, co, f.d.cstride, fv);
  else
    facet_load_k<T, false><<<blocks_for(f.nfacet), kBlock, 0, stream()>>>(
        f.nfacet, f.facets, x, f.x_dofmap, f.d.nodes_per_cell, co, f.d.cstride, fv);
  facet_rows_k<T><<<blocks_for(f.nfrows), kBlock, 0, stream()>>>(f.nfrows, f.frows, f.finc_off,
                                                                  f.finc, fv, b);
  check_launch("facet loads");
}

template <typename T>
void vector_rows(const dfx_form& f, T* b)
{
  const std::int32_t nc = f.d.num_cells;
  const int nv = f.d.nodes_per_cell;
  T* be = static_cast<T*>(f.cell_vec);
  const std::int32_t np = f.max_dof + 1;
  if (f.d.kernel == DFX_K_P2TET_ELAST_L)
  {
    cell_const_source_k<T, 10, 3><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, static_cast<const T*>(f.ref_tensors),
        static_cast<T>(f.constants[0]), static_cast<T>(f.constants[1]),
        static_cast<T>(f.constants[2]), be);
    row_vector_k<T, 10, 3><<<blocks_for(std::int64_t(np) * 3), kBlock, 0, stream()>>>(
        np, nc, f.inc_off, f.inc, be, b);
    check_launch("assemble_vector");
    return;
  }
  if (f.d.kernel == DFX_K_P3TET_L)
  {
    cell_const_source_k<T, 20, 1><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, static_cast<const T*>(f.ref_tensors),
        static_cast<T>(f.constants[0]), T(0), T(0), be);
    row_vector_k<T, 20, 1><<<blocks_for(np), kBlock, 0, stream()>>>(np, nc, f.inc_off, f.inc, be, b);
    check_launch("assemble_vector");
    return;
  }
  if (f.cell_vec_ready)
    f.cell_vec_ready = false; // element vectors from the system cell pass
  else if (nv == 4)
    cell_source_k<T, 4><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, static_cast<const T*>(f.coeffs),
        f.d.cstride, 1, be);
  else
    cell_source_k<T, 8><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, static_cast<const T*>(f.coeffs),
        f.d.cstride, f.affine ? 1 : 0, be);
  const std::int32_t nplan = f.max_dof + 1;
  if (nv == 4)
    row_vector_k<T, 4, 1><<<blocks_for(nplan), kBlock, 0, stream()>>>(nplan, nc, f.inc_off, f.inc,
                                                                       be, b);
  else
    row_vector_k<T, 8, 1><<<blocks_for(nplan), kBlock, 0, stream()>>>(nplan, nc, f.inc_off, f.inc,
                                                                       be, b);
  check_launch("assemble_vector");
}

/// The element matrices of a Q1 form on general hexahedra, computed into
/// the form's element-matrix array (device).
template <typename T>
const T* general_hex_elements(dfx_form& f)
{
  const std::int32_t nc = f.d.num_cells;
  if (!f.cell_elem)
    f.cell_elem = dmalloc(sizeof(T) * std::max<std::int64_t>(std::int64_t(nc) * 64, 1));
  if (nc > 0)
    q1hex_elem_general_k<T><<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(f.x), f.x_dofmap, static_cast<const T*>(f.coeffs),
        f.d.cstride, static_cast<T>(f.constants[0]), static_cast<T*>(f.cell_elem));
  check_launch("general hexahedra");
  return static_cast<const T*>(f.cell_elem);
}

/// Q1 stiffness + reaction on general hexahedra.
bool general_hex(const dfx_form& f)
{
  return f.d.kernel == DFX_K_Q1HEX_A && !f.affine && !f.host_kernel;
}

template <typename T, typename O, int MODE>
void launch_bilinear(const dfx_form& f, const dfx_csr* A, const dfx_bc* bc, const T* x0,
                     double alpha, T* b)
{
  hipStream_t s = stream();
  const std::int32_t nc = f.d.num_cells;
  if (nc == 0)
    return;
  const T* x = static_cast<const T*>(f.x);
  const T* co = static_cast<const T*>(f.coeffs);
  const T creac = static_cast<T>(f.constants[0]);
  const std::int8_t* mark = bc ? bc->marker : nullptr;
  const double* g = bc ? bc->g : nullptr;
  const std::int32_t* rp = A ? A->row_ptr : nullptr;
  const O* sl = static_cast<const O*>(f.slots);
  T* vals = A ? static_cast<T*>(A->values) : nullptr;
  switch (f.d.kernel)
  {
  case DFX_K_P1TET_A:
    p1tet_a_k<T, O, MODE><<<blocks_for(std::int64_t(nc) * 16), kBlock, 0, s>>>(
        nc, x, f.x_dofmap, f.dofmap, co, f.d.cstride, creac, mark, rp, sl, vals, g, x0, alpha, b);
    break;
  case DFX_K_Q1HEX_A:
    // general hexahedra go through their element matrices (general_hex_elements)
    require(f.affine, "general hexahedra are assembled from their element matrices");
    q1hex_a_affine_k<T, O, MODE><<<static_cast<int>((nc + 255) / 256), kBlock, 0, s>>>(
        nc, x, f.x_dofmap, f.dofmap, co, f.d.cstride, creac, mark, rp, sl, vals, g, x0, alpha, b);
    break;
  default:
    throw Error(DFX_ERR_ARG, "no device kernel for this bilinear form yet");
  }
  check_launch("assemble_matrix");
}

template <typename T>
void launch_linear(const dfx_form& f, T* b)
{
  hipStream_t s = stream();
  const std::int32_t nc = f.d.num_cells;
  if (nc == 0)
    return;
  const T* x = static_cast<const T*>(f.x);
  const T* co = static_cast<const T*>(f.coeffs);
  switch (f.d.kernel)
  {
  case DFX_K_P1TET_L:
    p1tet_l_k<T><<<blocks_for(std::int64_t(nc) * 4), kBlock, 0, s>>>(nc, x, f.x_dofmap, f.dofmap,
                                                                      co, f.d.cstride, b);
    break;
  case DFX_K_Q1HEX_L:
    q1hex_l_k<T><<<blocks_for(std::int64_t(nc) * 8), kBlock, 0, s>>>(
        nc, x, f.x_dofmap, f.dofmap, co, f.d.cstride, f.affine ? 1 : 0, b);
    break;
  default:
    throw Error(DFX_ERR_ARG, "no device kernel for this linear form yet");
  }
  check_launch("assemble_vector");
}

/// True if every hexahedron is a parallelepiped (x3 = x1 + x2 - x0 and
/// likewise for x5, x6, x7), so that its Jacobian is constant.
bool is_element_kernel(int id) { return id == DFX_K_ELEMENT_A || id == DFX_K_ELEMENT_L; }

/// Kernels of exterior facet integrals.
bool is_facet_kernel(int id)
{
  return id == DFX_K_P1TET_NEUMANN || id == DFX_K_Q1HEX_NEUMANN || id == DFX_K_P1TET_ROBIN_A
         || id == DFX_K_Q1HEX_ROBIN_A;
}

/// Element tensors of these sizes k (k x k matrices) have row kernels.
bool element_matrix_size(int k) { return k == 4 || k == 8 || k == 10 || k == 20; }

/// Incidences (cell, local index) of every dofmap row in cell order: the
/// row-owner plan of a form (counting sort of the dofmap entries).
void build_incidences(dfx_form& f, const std::int32_t* dofmap)
{
  const std::int64_t nc = f.d.num_cells;
  const int k = f.d.dofs_per_cell;
  const std::int32_t nplan = f.max_dof + 1;
  std::vector<std::int32_t> off(std::size_t(nplan) + 1, 0);
  for (std::int64_t i = 0; i < nc * k; ++i)
    ++off[dofmap[i] + 1];
  std::int32_t mx = 0;
  for (std::int32_t r = 0; r < nplan; ++r)
  {
    mx = std::max(mx, off[r + 1]);
    off[r + 1] += off[r];
  }
  std::vector<std::int32_t> inc(std::max<std::int64_t>(nc * k, 1));
  std::vector<std::int32_t> pos(off.begin(), off.end() - 1);
  for (std::int64_t i = 0; i < nc * k; ++i)
    inc[pos[dofmap[i]]++] = static_cast<std::int32_t>(i);
  f.max_inc = mx;
  f.inc_off = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * off.size()));
  h2d(f.inc_off, off.data(), sizeof(std::int32_t) * off.size());
  f.inc = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * inc.size()));
  h2d(f.inc, inc.data(), sizeof(std::int32_t) * inc.size());
}

/// A form whose element tensors the caller supplies: the dofmap and its
/// row-owner plan only.
std::unique_ptr<dfx_form> create_element_form(const dfx_form_desc& d, const KernelInfo& ki)
{
  (void)ki;
  require(d.bs == 1 || d.bs == 3, "element tensors are supported for block sizes 1 and 3");
  require(d.dofs_per_cell >= 1 && d.dofs_per_cell <= 32, "dofs per cell must be 1..32");
  require(d.dofmap, "null dofmap");
  const std::int64_t nc = d.num_cells;
  std::int32_t max_dof = -1;
  for (std::int64_t i = 0; i < nc * d.dofs_per_cell; ++i)
  {
    require(d.dofmap[i] >= 0, "negative dof");
    max_dof = std::max(max_dof, d.dofmap[i]);
  }
  auto f = std::make_unique<dfx_form>();
  f->max_dof = max_dof;
  f->d = d;
  f->d.x = nullptr;
  f->d.x_dofmap = nullptr;
  f->d.dofmap = nullptr;
  f->d.coeffs = nullptr;
  f->d.constants = nullptr;
  f->d.num_nodes = 0;
  f->d.nodes_per_cell = 0;
  f->d.cstride = 0;
  f->d.num_constants = 0;
  f->dofmap = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max<std::int64_t>(nc * d.dofs_per_cell, 1)));
  h2d(f->dofmap, d.dofmap, sizeof(std::int32_t) * nc * d.dofs_per_cell);
  build_incidences(*f, d.dofmap);
  if (ki.bilinear)
    f->cell_mask = static_cast<std::uint32_t*>(dmalloc(sizeof(std::uint32_t) * std::max<std::int64_t>(nc, 1)));
  return f;
}

bool hex_all_affine(const dfx_form_desc& d)
{
  for (std::int64_t c = 0; c < d.num_cells; ++c)
  {
    const std::int32_t* v = d.x_dofmap + 8 * c;
    const double* X[8];
    for (int k = 0; k < 8; ++k)
      X[k] = d.x + 3 * std::int64_t(v[k]);
    for (int r = 0; r < 3; ++r)
    {
      const double h = std::abs(X[1][r] - X[0][r]) + std::abs(X[2][r] - X[0][r])
                       + std::abs(X[4][r] - X[0][r]) + 1e-300;
      const double e3 = X[3][r] - (X[1][r] + X[2][r] - X[0][r]);
      const double e5 = X[5][r] - (X[1][r] + X[4][r] - X[0][r]);
      const double e6 = X[6][r] - (X[2][r] + X[4][r] - X[0][r]);
      const double e7 = X[7][r] - (X[1][r] + X[2][r] + X[4][r] - 2 * X[0][r]);
      const double tol = 1e-13 * h;
      if (std::abs(e3) > tol || std::abs(e5) > tol || std::abs(e6) > tol || std::abs(e7) > tol)
        return false;
    }
  }
  return true;
}

/// Whether every hexahedron's Jacobian (x1 - x0, x2 - x0, x4 - x0 as
/// columns) is exactly diagonal: each edge from vertex 0 moves along one
/// axis only.
bool hex_all_axis_aligned(const dfx_form_desc& d)
{
  for (std::int64_t c = 0; c < d.num_cells; ++c)
  {
    const std::int32_t* v = d.x_dofmap + 8 * c;
    const double* X0 = d.x + 3 * std::int64_t(v[0]);
    const int col[3] = {1, 2, 4};
    for (int a = 0; a < 3; ++a)
    {
      const double* Xa = d.x + 3 * std::int64_t(v[col[a]]);
      for (int r = 0; r < 3; ++r)
        if (r != a && Xa[r] - X0[r] != 0.0)
          return false;
    }
  }
  return true;
}

/// Device copies of the batch tables, grown as needed and reused.
struct BatchTables
{
  void* sys = nullptr;
  std::size_t sys_bytes = 0;
  int2* cwork = nullptr;
  std::size_t cwork_n = 0;
  int2* rwork = nullptr;
  std::size_t rwork_n = 0;
  // host copies of what the device tables hold: an assembly whose tables
  // are unchanged uploads nothing (and so can be recorded into a graph)
  std::vector<unsigned char> h_sys;
  std::vector<int2> h_cw, h_rw;
};
BatchTables& batch_tables()
{
  static BatchTables t;
  return t;
}

template <typename X>
void upload_grow(X*& dst, std::size_t& cap, const X* src, std::size_t n)
{
  if (n > cap)
  {
    dfree(dst);
    dst = static_cast<X*>(dmalloc(sizeof(X) * n));
    cap = n;
  }
  h2d(dst, src, sizeof(X) * n);
}

template <typename T>
void batch_assemble(int n, dfx_csr* const* A, const dfx_form* const* a, const dfx_bc* const* bc)
{
  const int kernel = a[0]->d.kernel;
  const int K = a[0]->d.dofs_per_cell;
  constexpr int R = 64;
  std::vector<BatchSys<T>> sys(n);
  std::vector<int2> cw, rw;
  std::int32_t max_row = 0;
  for (int s = 0; s < n; ++s)
  {
    dfx_form& f = const_cast<dfx_form&>(*a[s]);
    dfx_csr& M = *A[s];
    require(f.d.kernel == kernel && row_path(f), "batched forms must share one affine kernel");
    require(M.dtype == f.d.dtype && M.bs == 1 && f.d.bs == 1, "matrix and form do not match");
    require(!bc[s] || bc[s]->n_total == M.nrows, "BC size does not match the matrix");
    require(f.max_dof < M.nrows, "dofmap exceeds the matrix size");
    ensure_plan(f, M);
    require(f.islot_bytes == 1, "row too long for the batched kernel");
    max_row = std::max(max_row, M.max_row);
    BatchSys<T>& d = sys[s];
    d.x = static_cast<const T*>(f.x);
    d.xdm = f.x_dofmap;
    d.dm = f.dofmap;
    d.coeffs = static_cast<const T*>(f.coeffs);
    d.bc = bc[s] ? bc[s]->marker : nullptr;
    d.cd = static_cast<T*>(f.cell_data);
    d.inc_off = f.inc_off;
    d.inc = f.inc;
    d.islots = static_cast<const std::uint8_t*>(f.islots);
    d.rp = M.row_ptr;
    d.vals = static_cast<T*>(M.values);
    d.creac = static_cast<T>(f.constants[0]);
    d.cstride = f.d.cstride;
    d.nc = f.d.num_cells;
    d.nrows = M.nrows;
    d.nplan = f.max_dof + 1;
    d.accumulate = M.zeroed ? 0 : 1;
    for (std::int32_t c0 = 0; c0 < d.nc; c0 += kBlock)
      cw.push_back(make_int2(s, c0));
    for (std::int32_t r0 = 0; r0 < d.nrows; r0 += R)
      rw.push_back(make_int2(s, r0));
  }
  const bool cf = kernel == DFX_K_Q1HEX_A;
  const std::size_t lds = ((cf ? 0 : 7 * K * K) + std::size_t(R) * max_row) * sizeof(T);
  require(lds <= 64 * 1024, "row too long for the LDS slice");
  BatchTables& t = batch_tables();
  const auto* sb = reinterpret_cast<const unsigned char*>(sys.data());
  const std::size_t sn = sizeof(BatchSys<T>) * sys.size();
  if (t.h_sys.size() != sn || std::memcmp(t.h_sys.data(), sb, sn) != 0)
  {
    std::size_t cap = t.sys_bytes / sizeof(BatchSys<T>);
    BatchSys<T>* p = static_cast<BatchSys<T>*>(t.sys);
    t.h_sys.clear();
    upload_grow(p, cap, sys.data(), sys.size());
    t.sys = p;
    t.sys_bytes = cap * sizeof(BatchSys<T>);
    t.h_sys.assign(sb, sb + sn);
  }
  auto same = [](const std::vector<int2>& h, const std::vector<int2>& v)
  {
    return h.size() == v.size()
           && std::memcmp(h.data(), v.data(), sizeof(int2) * v.size()) == 0;
  };
  if (!cw.empty() && !same(t.h_cw, cw))
  {
    t.h_cw.clear();
    upload_grow(t.cwork, t.cwork_n, cw.data(), cw.size());
    t.h_cw = cw;
  }
  if (!rw.empty() && !same(t.h_rw, rw))
  {
    t.h_rw.clear();
    upload_grow(t.rwork, t.rwork_n, rw.data(), rw.size());
    t.h_rw = rw;
  }
  const BatchSys<T>* dsys = static_cast<const BatchSys<T>*>(t.sys);
  const T* ref = static_cast<const T*>(a[0]->ref_tensors);
  if (!cw.empty())
  {
    if (K == 4)
      batch_factors_k<T, 4><<<static_cast<int>(cw.size()), kBlock, 0, stream()>>>(dsys, t.cwork);
    else
      batch_factors_k<T, 8><<<static_cast<int>(cw.size()), kBlock, 0, stream()>>>(dsys, t.cwork);
  }
  if (!rw.empty())
  {
    if (K == 4)
      batch_rows_k<T, 4, false><<<static_cast<int>(rw.size()), R, lds, stream()>>>(dsys, t.rwork, ref);
    else
      batch_rows_k<T, 8, true><<<static_cast<int>(rw.size()), R, lds, stream()>>>(dsys, t.rwork, ref);
  }
  check_launch("assemble_matrix_batched");
  for (int s = 0; s < n; ++s)
  {
    A[s]->zeroed = false;
    A[s]->zero_pending = false;
  }
}

} // namespace
} // namespace dfx

using namespace dfx;

/// Tensors of the fast element path (block size 1, k in {4, 8, 10, 20});
/// other sizes take the generic row-dof kernels.
bool fast_elements(const dfx_form& f)
{
  return f.d.bs == 1 && element_matrix_size(f.d.dofs_per_cell);
}

/// A += E (element matrices on the device) with BC rows/columns skipped.
template <typename T>
void elements_into_matrix(dfx_form& f, dfx_csr& A, const dfx_bc* bc, const T* E)
{
  if (fast_elements(f))
  {
    assemble_elements<T>(f, A, bc, E);
    return;
  }
  ensure_plan(f, A);
  materialize_zero(A); // the generic kernels add into stored values
  const std::int32_t nplan = f.max_dof + 1;
  const std::int64_t nt = std::int64_t(nplan) * f.d.bs;
  if (nt > 0)
  {
    auto go = [&](auto o)
    {
      using O = decltype(o);
      const O* sl = static_cast<const O*>(f.islots);
      const std::int8_t* mk = bc ? bc->marker : nullptr;
      if (f.d.bs == 3)
        element_rows_bs_k<T, O, 3><<<blocks_for(nt), kBlock, 0, stream()>>>(
            nplan, f.d.dofs_per_cell, f.inc_off, f.inc, sl, f.dofmap, E, mk, A.row_ptr,
            static_cast<T*>(A.values));
      else
        element_rows_bs_k<T, O, 1><<<blocks_for(nt), kBlock, 0, stream()>>>(
            nplan, f.d.dofs_per_cell, f.inc_off, f.inc, sl, f.dofmap, E, mk, A.row_ptr,
            static_cast<T*>(A.values));
    };
    if (f.islot_bytes == 1)
      go(std::uint8_t{});
    else
      go(std::uint16_t{});
  }
  check_launch("element matrix assembly");
}

/// b += e (element vectors on the device).
template <typename T>
void elements_into_vector(const dfx_form& L, T* b, const T* e)
{
  const std::int32_t nplan = L.max_dof + 1;
  if (nplan <= 0)
    return;
  if (L.d.bs == 3)
    element_vector_bs_k<T, 3><<<blocks_for(std::int64_t(nplan) * 3), kBlock, 0, stream()>>>(
        nplan, L.inc_off, L.inc, e, b);
  else
    element_vector_k<T><<<blocks_for(nplan), kBlock, 0, stream()>>>(nplan, L.inc_off, L.inc, e,
                                                                      b);
  check_launch("element vector assembly");
}

/// b -= alpha E[:, bc] (g - x0) from element matrices on the device.
template <typename T>
void elements_lift(const dfx_form& a, T* b, const dfx_bc& bc, const T* E, const T* x0,
                   double alpha)
{
  const std::int32_t nc = a.d.num_cells;
  const std::int32_t nplan = a.max_dof + 1;
  if (nplan <= 0)
    return;
  if (!fast_elements(a))
  {
    const std::int64_t nt = std::int64_t(nplan) * a.d.bs;
    if (a.d.bs == 3)
      element_lift_bs_k<T, 3><<<blocks_for(nt), kBlock, 0, stream()>>>(
          nplan, a.d.dofs_per_cell, a.inc_off, a.inc, a.dofmap, E, bc.marker, bc.g, x0, alpha,
          b);
    else
      element_lift_bs_k<T, 1><<<blocks_for(nt), kBlock, 0, stream()>>>(
          nplan, a.d.dofs_per_cell, a.inc_off, a.inc, a.dofmap, E, bc.marker, bc.g, x0, alpha,
          b);
    check_launch("element lifting");
    return;
  }
  if (nc > 0)
    element_mask_k<<<blocks_for(nc), kBlock, 0, stream()>>>(nc, a.d.dofs_per_cell, a.dofmap,
                                                            bc.marker, a.cell_mask);
  element_lift_k<T><<<blocks_for(nplan), kBlock, 0, stream()>>>(
      nplan, a.d.dofs_per_cell, a.inc_off, a.inc, a.dofmap, E, a.cell_mask, bc.g, x0, alpha, b);
  check_launch("element lifting");
}

/// Runs the form's host element kernel (the reference's kernelptr_t,
/// fem/kernel.h:18-20) on every cell in cell order, as
/// assemble_cells_matrix / assemble_cells call it
/// (fem/assemble_matrix_impl.h:144-152, fem/assemble_vector_impl.h:99-107):
/// A zeroed, w = the cell's packed coefficients, c = the constants,
/// coordinate_dofs = the cell's geometry nodes (3 values each), no entity
/// index or permutation. The tensors go to the device in one copy; returns
/// their device array.
template <typename T>
const T* run_host_kernel(dfx_form& f)
{
  using Fn = void (*)(T*, const T*, const T*, const T*, const int*, const std::uint8_t*, void*);
  const Fn fn = reinterpret_cast<Fn>(f.host_kernel);
  const std::int64_t nc = f.d.num_cells;
  const int npc = f.d.nodes_per_cell;
  const std::int64_t nd = std::int64_t(f.d.bs) * f.d.dofs_per_cell;
  const std::int64_t sz = f.d.kernel == DFX_K_ELEMENT_A ? nd * nd : nd;
  const int cs = f.d.cstride;
  std::vector<T> E(std::max<std::int64_t>(nc * sz, 1), T(0));
  std::vector<T> cdofs(3 * std::size_t(std::max(npc, 1)));
  std::vector<T> w(std::max(cs, 1), T(0));
  std::vector<T> cst(std::max<std::size_t>(f.h_constants.size(), 1), T(0));
  for (std::size_t q = 0; q < f.h_constants.size(); ++q)
    cst[q] = static_cast<T>(f.h_constants[q]);
  for (std::int64_t c = 0; c < nc; ++c)
  {
    for (int i = 0; i < npc; ++i)
    {
      const double* xi = f.h_x.data() + 3 * std::int64_t(f.h_x_dofmap[c * npc + i]);
      for (int r = 0; r < 3; ++r)
        cdofs[3 * i + r] = static_cast<T>(xi[r]);
    }
    for (int q = 0; q < cs; ++q)
      w[q] = static_cast<T>(f.h_coeffs[c * cs + q]);
    fn(E.data() + c * sz, w.data(), cst.data(), cdofs.data(), nullptr, nullptr,
       f.host_kernel_data);
  }
  if (!f.elem_tensors)
    f.elem_tensors = dmalloc(sizeof(T) * E.size());
  h2d(f.elem_tensors, E.data(), sizeof(T) * E.size());
  return static_cast<const T*>(f.elem_tensors);
}

/// Facet integrals with a host kernel of the reference's ABI, as
/// assemble_exterior_facets / assemble_interior_facets call theirs
/// (fem/assemble_matrix_impl.h:264-383, :442-667;
/// fem/assemble_vector_impl.h:118-174): per facet, in the given order, the
/// kernel gets the cell's (interior: both cells') geometry nodes, packed
/// coefficients and local facet indices (quadrature permutations 0); BC
/// rows and columns of the tensor are zeroed as there; the tensors of all
/// facets are then added on the device, every entry in facet order
/// (mat_set(dmapjoint0, dmapjoint1, Ae)).
template <typename T>
void host_facet_tensors(const dfx_form& f, const std::int8_t* mark, std::vector<std::int32_t>& dofs,
                        std::vector<T>& vals)
{
  using Fn = void (*)(T*, const T*, const T*, const T*, const int*, const std::uint8_t*, void*);
  const Fn fn = reinterpret_cast<Fn>(f.host_kernel);
  const bool interior = f.host_facets == 2;
  const int sides = interior ? 2 : 1;
  const int k = f.d.dofs_per_cell, bs = f.d.bs, npc = f.d.nodes_per_cell;
  const int m = sides * k;          // block rows of a facet tensor
  const int nd = m * bs;            // scalar rows
  const bool mat = f.d.kernel == DFX_K_ELEMENT_A;
  const std::int64_t sz = mat ? std::int64_t(nd) * nd : nd;
  const int per = interior ? 4 : 2;
  const std::int64_t ne = std::int64_t(f.h_entities.size()) / per;
  const int cs = f.d.cstride;
  dofs.resize(std::size_t(ne) * m);
  vals.assign(std::size_t(std::max<std::int64_t>(ne * sz, 1)), T(0));
  std::vector<T> cdofs(3 * std::size_t(sides) * npc);
  std::vector<T> w(std::max(1, sides * cs), T(0));
  std::vector<T> cst(std::max<std::size_t>(f.h_constants.size(), 1), T(0));
  for (std::size_t q = 0; q < f.h_constants.size(); ++q)
    cst[q] = static_cast<T>(f.h_constants[q]);
  for (std::int64_t e = 0; e < ne; ++e)
  {
    int lf[2] = {0, 0};
    const std::uint8_t perm[2] = {0, 0};
    std::int32_t* de = dofs.data() + e * m;
    for (int s = 0; s < sides; ++s)
    {
      const std::int32_t c = f.h_entities[per * e + 2 * s];
      lf[s] = f.h_entities[per * e + 2 * s + 1];
      for (int i = 0; i < npc; ++i)
      {
        const double* xi = f.h_x.data() + 3 * std::int64_t(f.h_x_dofmap[std::int64_t(c) * npc + i]);
        for (int r = 0; r < 3; ++r)
          cdofs[3 * (std::size_t(s) * npc + i) + r] = static_cast<T>(xi[r]);
      }
      for (int q = 0; q < cs; ++q)
        w[std::size_t(s) * cs + q] = static_cast<T>(f.h_coeffs[std::int64_t(c) * cs + q]);
      for (int i = 0; i < k; ++i)
        de[s * k + i] = f.h_dofmap[std::int64_t(c) * k + i];
    }
    T* Ae = vals.data() + e * sz;
    fn(Ae, w.data(), cst.data(), cdofs.data(), lf, perm, f.host_kernel_data);
    if (mat && mark)
    {
      // zero the rows and columns of Dirichlet dofs
      for (int i = 0; i < m; ++i)
        for (int c2 = 0; c2 < bs; ++c2)
          if (mark[std::int64_t(bs) * de[i] + c2])
          {
            for (int j = 0; j < nd; ++j)
              Ae[std::int64_t(bs * i + c2) * nd + j] = T(0);
            for (int j = 0; j < nd; ++j)
              Ae[std::int64_t(j) * nd + bs * i + c2] = T(0);
          }
    }
  }
}

/// Host copy of a BC marker (nullptr without BCs).
std::vector<std::int8_t> host_marker(const dfx_bc* bc)
{
  std::vector<std::int8_t> h;
  if (bc)
  {
    h.resize(bc->n_total);
    d2h(h.data(), bc->marker, bc->n_total);
  }
  return h;
}

template <typename T>
void host_facets_matrix(dfx_csr& A, const dfx_form& f, const dfx_bc* bc)
{
  const std::vector<std::int8_t> mark = host_marker(bc);
  std::vector<std::int32_t> dofs;
  std::vector<T> vals;
  host_facet_tensors<T>(f, bc ? mark.data() : nullptr, dofs, vals);
  const int m = (f.host_facets == 2 ? 2 : 1) * f.d.dofs_per_cell;
  const std::int64_t ne = std::int64_t(dofs.size()) / m;
  csr_add_blocks(A, ne, m, m, dofs.data(), dofs.data(), vals.data());
}

template <typename T>
void host_facets_vector(dfx_vec& b, const dfx_form& f)
{
  std::vector<std::int32_t> dofs;
  std::vector<T> vals;
  host_facet_tensors<T>(f, nullptr, dofs, vals);
  const int m = (f.host_facets == 2 ? 2 : 1) * f.d.dofs_per_cell;
  const std::int64_t ne = std::int64_t(dofs.size()) / m;
  vec_add_blocks(b, ne, m, f.d.bs, dofs.data(), vals.data());
}

extern "C" {

dfx_status dfx_form_create(const dfx_form_desc* desc, dfx_form** out)
{
  return guard(
      [&]
      {
        require(desc && out, "null argument");
        const dfx_form_desc& d = *desc;
        KernelInfo ki;
        require(kernel_info(d.kernel, ki), "unknown kernel id");
        require(d.dtype == DFX_F32 || d.dtype == DFX_F64, "bad dtype");
        require(d.num_cells >= 0 && d.num_nodes >= 0, "negative size");
        if (is_element_kernel(d.kernel))
        {
          *out = create_element_form(d, ki).release();
          return;
        }
        require(d.nodes_per_cell == ki.npc, "geometry nodes per cell do not match the kernel");
        require(d.dofs_per_cell == ki.k && d.bs == ki.bs, "dofmap does not match the kernel");
        require(d.x && d.x_dofmap && d.dofmap, "null mesh data");
        require(d.num_constants >= 0 && d.num_constants <= 8, "at most 8 constants");
        require(d.cstride >= 0 && (d.cstride == 0 || d.coeffs), "missing coefficients");
        // Required coefficient layouts of the kernels
        int need = 0;
        if (d.kernel == DFX_K_P1TET_A || d.kernel == DFX_K_Q1HEX_A || d.kernel == DFX_K_P3TET_A)
          need = 1;
        else if (d.kernel == DFX_K_P1TET_L)
          need = 4;
        else if (d.kernel == DFX_K_Q1HEX_L)
          need = 8;
        else if (d.kernel == DFX_K_P2TET_ELAST_A)
          need = 2;
        else if (d.kernel == DFX_K_P1TET_NEUMANN)
          need = 4;
        else if (d.kernel == DFX_K_Q1HEX_NEUMANN)
          need = 8;
        else if (is_facet_matrix_kernel(d.kernel))
          need = 1;
        require(d.cstride >= need, "too few packed coefficients for this kernel");
        const std::int64_t nc = d.num_cells;
        for (std::int64_t i = 0; i < nc * d.nodes_per_cell; ++i)
          require(d.x_dofmap[i] >= 0 && d.x_dofmap[i] < d.num_nodes, "geometry dofmap out of range");
        std::int32_t max_dof = -1;
        for (std::int64_t i = 0; i < nc * d.dofs_per_cell; ++i)
        {
          require(d.dofmap[i] >= 0, "negative dof");
          max_dof = std::max(max_dof, d.dofmap[i]);
        }
        auto f = std::make_unique<dfx_form>();
        f->max_dof = max_dof;
        f->d = d;
        f->d.x = nullptr;
        f->d.x_dofmap = nullptr;
        f->d.dofmap = nullptr;
        f->d.coeffs = nullptr;
        f->d.constants = nullptr;
        for (int i = 0; i < d.num_constants; ++i)
          f->constants[i] = d.constants[i];
        f->affine = ki.cell == DFX_TET ? true : hex_all_affine(d);
        f->compact_records = ki.cell != DFX_TET && f->affine && d.kernel == DFX_K_Q1HEX_A
                             && d.dofs_per_cell == 8
                             && (d.dofmap == d.x_dofmap
                                 || std::memcmp(d.dofmap, d.x_dofmap,
                                                sizeof(std::int32_t) * 8 * d.num_cells) == 0)
                             && hex_all_axis_aligned(d);
        f->x = upload_as(d.x, std::int64_t(d.num_nodes) * 3, d.dtype);
        f->x_dofmap = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * nc * d.nodes_per_cell));
        h2d(f->x_dofmap, d.x_dofmap, sizeof(std::int32_t) * nc * d.nodes_per_cell);
        if (d.dofmap == d.x_dofmap
            || (d.dofs_per_cell == d.nodes_per_cell
                && std::memcmp(d.dofmap, d.x_dofmap, sizeof(std::int32_t) * nc * d.nodes_per_cell) == 0))
          f->dofmap = f->x_dofmap;
        else
        {
          f->dofmap = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * nc * d.dofs_per_cell));
          h2d(f->dofmap, d.dofmap, sizeof(std::int32_t) * nc * d.dofs_per_cell);
        }
        f->coeffs = upload_as(d.coeffs, nc * d.cstride, d.dtype);
        // Row-owner plan: incidences (cell, local index) of every row, in
        // cell order
        {
          const int k = d.dofs_per_cell;
          build_incidences(*f, d.dofmap);
          // thread-per-row slices do not fit LDS for wide elements
          f->row_variant = k > 8 ? 3 : 1;
          const std::size_t ts = dtype_size(d.dtype);
          if (is_facet_kernel(d.kernel) || is_interior_kernel(d.kernel))
          {
            // facet forms keep per-facet data only (dfx_form_create_facets)
          }
          else if (ki.bilinear)
          {
            const int per_cell = d.kernel == DFX_K_P2TET_ELAST_A ? 16 : 8;
            f->cell_data = dmalloc(ts * per_cell * std::max<std::int64_t>(nc, 1));
            f->cell_mask = static_cast<std::uint32_t*>(dmalloc(sizeof(std::uint32_t) * std::max<std::int64_t>(nc, 1)));
            const std::vector<double> ref = reference_tensors(d.kernel);
            if (!ref.empty())
              f->ref_tensors = upload_as(ref.data(), static_cast<std::int64_t>(ref.size()), d.dtype);
          }
          else
          {
            f->cell_vec = dmalloc(ts * std::max<std::int64_t>(nc * k * d.bs, 1));
            const std::vector<double> w = reference_tensors(d.kernel);
            if (!w.empty())
              f->ref_tensors = upload_as(w.data(), static_cast<std::int64_t>(w.size()), d.dtype);
          }
        }
        f->id = next_pattern_id();
        *out = f.release();
      });
}

/// The rows with facet incidences and room for their slots (nslot bytes).
void set_facet_rows(dfx_form& f, const std::vector<std::int32_t>& off, std::int64_t nslot)
{
  std::vector<std::int32_t> rows;
  for (std::size_t r = 0; r + 1 < off.size(); ++r)
    if (off[r + 1] > off[r])
      rows.push_back(static_cast<std::int32_t>(r));
  f.nfrows = static_cast<std::int32_t>(rows.size());
  f.frows = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max<std::size_t>(rows.size(), 1)));
  h2d(f.frows, rows.data(), sizeof(std::int32_t) * rows.size());
  f.fslot = static_cast<std::uint8_t*>(dmalloc(std::max<std::int64_t>(nslot, 1)));
}

dfx_status dfx_form_create_facets(const dfx_form_desc* desc, int32_t npairs,
                                  const int32_t* pairs, dfx_form** out)
{
  return guard(
      [&]
      {
        require(desc && out && (npairs == 0 || pairs), "null argument");
        require(is_facet_kernel(desc->kernel), "the kernel has no exterior facet integral");
        require(npairs >= 0, "negative size");
        // incidences are packed as 4 e + facet vertex in an int32
        require(npairs < (1 << 29), "too many facets (at most 2^29 - 1)");
        const bool hex = desc->kernel == DFX_K_Q1HEX_NEUMANN || desc->kernel == DFX_K_Q1HEX_ROBIN_A;
        const bool matrix = is_facet_matrix_kernel(desc->kernel);
        const int nfc = hex ? 6 : 4, nfv = hex ? 4 : 3, k = desc->dofs_per_cell;
        static constexpr int tetv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
        static constexpr int hexv[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                           {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};
        for (std::int32_t e = 0; e < npairs; ++e)
          require(pairs[2 * e] >= 0 && pairs[2 * e] < desc->num_cells && pairs[2 * e + 1] >= 0
                      && pairs[2 * e + 1] < nfc,
                  "facet pair out of range");
        dfx_form* raw = nullptr;
        const dfx_status st = dfx_form_create(desc, &raw);
        if (st != DFX_OK)
          throw Error(st, dfx_last_error());
        std::unique_ptr<dfx_form, dfx_status (*)(dfx_form*)> f(raw, dfx_form_destroy);
        if (hex)
          require(f->affine, "facet integrals need affine hexahedra");
        // facet incidences of every row in facet order (counting sort)
        const std::int32_t nplan = f->max_dof + 1;
        std::vector<std::int32_t> off(std::size_t(nplan) + 1, 0);
        auto dof_of = [&](std::int32_t e, int a)
        {
          const int lv = hex ? hexv[pairs[2 * e + 1]][a] : tetv[pairs[2 * e + 1]][a];
          return desc->dofmap[std::int64_t(pairs[2 * e]) * k + lv];
        };
        for (std::int32_t e = 0; e < npairs; ++e)
          for (int a = 0; a < nfv; ++a)
            ++off[dof_of(e, a) + 1];
        for (std::int32_t r = 0; r < nplan; ++r)
          off[r + 1] += off[r];
        std::vector<std::int32_t> finc(std::max<std::int64_t>(std::int64_t(npairs) * nfv, 1));
        std::vector<std::int32_t> pos(off.begin(), off.end() - 1);
        for (std::int32_t e = 0; e < npairs; ++e)
          for (int a = 0; a < nfv; ++a)
            finc[pos[dof_of(e, a)]++] = 4 * e + a;
        f->nfacet = npairs;
        f->facets = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max(2 * npairs, 1)));
        h2d(f->facets, pairs, sizeof(std::int32_t) * 2 * npairs);
        f->finc_off = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * off.size()));
        h2d(f->finc_off, off.data(), sizeof(std::int32_t) * off.size());
        f->finc = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * finc.size()));
        h2d(f->finc, finc.data(), sizeof(std::int32_t) * finc.size());
        f->fvals = dmalloc(dtype_size(desc->dtype) * (matrix ? 16 : 4) * std::max(npairs, 1));
        set_facet_rows(*f, off, matrix ? 4 * std::int64_t(npairs) * nfv : 0);
        *out = f.release();
      });
}

dfx_status dfx_form_create_interior_facets(const dfx_form_desc* desc, int32_t nfacets,
                                           const int32_t* quads, dfx_form** out)
{
  return guard(
      [&]
      {
        require(desc && out && (nfacets == 0 || quads), "null argument");
        require(is_interior_kernel(desc->kernel), "the kernel has no interior facet integral");
        require(nfacets >= 0 && nfacets < (1 << 28), "bad interior facet count");
        for (std::int64_t e = 0; e < 2 * std::int64_t(nfacets); ++e)
          require(quads[2 * e] >= 0 && quads[2 * e] < desc->num_cells && quads[2 * e + 1] >= 0
                      && quads[2 * e + 1] < 4,
                  "interior facet out of range");
        dfx_form* raw = nullptr;
        const dfx_status st = dfx_form_create(desc, &raw);
        if (st != DFX_OK)
          throw Error(st, dfx_last_error());
        std::unique_ptr<dfx_form, dfx_status (*)(dfx_form*)> f(raw, dfx_form_destroy);
        // macro-element incidences of every row, facet by facet (counting
        // sort of 8 e + macro index)
        const std::int32_t nplan = f->max_dof + 1;
        std::vector<std::int32_t> off(std::size_t(nplan) + 1, 0);
        auto dof_of = [&](std::int64_t e, int i)
        { return desc->dofmap[std::int64_t(quads[4 * e + 2 * (i >> 2)]) * 4 + (i & 3)]; };
        for (std::int64_t e = 0; e < nfacets; ++e)
          for (int i = 0; i < 8; ++i)
            ++off[dof_of(e, i) + 1];
        for (std::int32_t r = 0; r < nplan; ++r)
          off[r + 1] += off[r];
        std::vector<std::int32_t> finc(std::max<std::int64_t>(8 * std::int64_t(nfacets), 1));
        std::vector<std::int32_t> pos(off.begin(), off.end() - 1);
        for (std::int64_t e = 0; e < nfacets; ++e)
          for (int i = 0; i < 8; ++i)
            finc[pos[dof_of(e, i)]++] = static_cast<std::int32_t>(8 * e + i);
        f->nfacet = nfacets;
        f->facets = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max(4 * nfacets, 1)));
        h2d(f->facets, quads, sizeof(std::int32_t) * 4 * nfacets);
        f->finc_off = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * off.size()));
        h2d(f->finc_off, off.data(), sizeof(std::int32_t) * off.size());
        f->finc = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * finc.size()));
        h2d(f->finc, finc.data(), sizeof(std::int32_t) * finc.size());
        f->fvals = dmalloc(dtype_size(desc->dtype) * 64 * std::max(nfacets, 1));
        set_facet_rows(*f, off, 8 * 8 * std::int64_t(nfacets));
        *out = f.release();
      });
}

dfx_status dfx_form_create_kernel(const dfx_form_desc* desc, void* kernel, void* custom_data,
                                  dfx_form** out)
{
  return guard(
      [&]
      {
        require(desc && kernel && out, "null argument");
        require(is_element_kernel(desc->kernel),
                "a kernel form is DFX_K_ELEMENT_A (matrix) or DFX_K_ELEMENT_L (vector)");
        require(desc->dtype == DFX_F64 || desc->dtype == DFX_F32, "bad dtype");
        require(desc->num_cells >= 0 && desc->num_nodes >= 0, "negative size");
        require(desc->nodes_per_cell >= 1 && desc->x && desc->x_dofmap,
                "the kernel needs the cell geometry (x, x_dofmap)");
        require(desc->cstride >= 0 && (desc->cstride == 0 || desc->coeffs),
                "null coefficients");
        require(desc->num_constants >= 0 && (desc->num_constants == 0 || desc->constants),
                "null constants");
        const std::int64_t nc = desc->num_cells;
        for (std::int64_t i = 0; i < nc * desc->nodes_per_cell; ++i)
          require(desc->x_dofmap[i] >= 0 && desc->x_dofmap[i] < desc->num_nodes,
                  "geometry dofmap entry out of range");
        KernelInfo ki;
        kernel_info(desc->kernel, ki);
        auto f = create_element_form(*desc, ki);
        f->d.nodes_per_cell = desc->nodes_per_cell;
        f->d.num_nodes = desc->num_nodes;
        f->d.cstride = desc->cstride;
        f->d.num_constants = desc->num_constants;
        f->h_x.assign(desc->x, desc->x + 3 * std::int64_t(desc->num_nodes));
        f->h_x_dofmap.assign(desc->x_dofmap, desc->x_dofmap + nc * desc->nodes_per_cell);
        if (desc->cstride > 0)
          f->h_coeffs.assign(desc->coeffs, desc->coeffs + nc * desc->cstride);
        if (desc->num_constants > 0)
          f->h_constants.assign(desc->constants, desc->constants + desc->num_constants);
        f->host_kernel = kernel;
        f->host_kernel_data = custom_data;
        *out = f.release();
      });
}

dfx_status dfx_form_create_kernel_facets(const dfx_form_desc* desc, void* kernel,
                                         void* custom_data, int32_t type, int32_t nentities,
                                         const int32_t* entities, dfx_form** out)
{
  return guard(
      [&]
      {
        require(desc && kernel && out && (nentities == 0 || entities), "null argument");
        require(type == 1 || type == 2, "type is 1 (exterior facets) or 2 (interior facets)");
        require(is_element_kernel(desc->kernel),
                "a kernel form is DFX_K_ELEMENT_A (matrix) or DFX_K_ELEMENT_L (vector)");
        require(type == 1 || desc->kernel == DFX_K_ELEMENT_A,
                "interior facet integrals are bilinear here");
        require(desc->dtype == DFX_F64 || desc->dtype == DFX_F32, "bad dtype");
        require(desc->bs == 1 || desc->bs == 3, "block size 1 or 3");
        require(desc->num_cells >= 0 && desc->num_nodes >= 0 && nentities >= 0, "negative size");
        require(desc->nodes_per_cell == 4 || desc->nodes_per_cell == 8,
                "tetrahedra (4 nodes) or hexahedra (8 nodes)");
        require(desc->x && desc->x_dofmap && desc->dofmap, "null mesh data");
        require(desc->dofs_per_cell >= 1 && desc->dofs_per_cell <= 32, "1 to 32 dofs per cell");
        require(desc->cstride >= 0 && (desc->cstride == 0 || desc->coeffs), "null coefficients");
        require(desc->num_constants >= 0 && (desc->num_constants == 0 || desc->constants),
                "null constants");
        const std::int64_t nc = desc->num_cells;
        const int nfac = desc->nodes_per_cell == 4 ? 4 : 6;
        const int per = type == 2 ? 4 : 2;
        for (std::int64_t e = 0; e < std::int64_t(nentities); ++e)
          for (int s = 0; s < per / 2; ++s)
          {
            const std::int32_t c = entities[per * e + 2 * s];
            const std::int32_t l = entities[per * e + 2 * s + 1];
            require(c >= 0 && c < nc && l >= 0 && l < nfac, "facet entity out of range");
          }
        for (std::int64_t i = 0; i < nc * desc->nodes_per_cell; ++i)
          require(desc->x_dofmap[i] >= 0 && desc->x_dofmap[i] < desc->num_nodes,
                  "geometry dofmap entry out of range");
        auto f = std::make_unique<dfx_form>();
        f->d = *desc;
        f->d.x = nullptr;
        f->d.x_dofmap = nullptr;
        f->d.dofmap = nullptr;
        f->d.coeffs = nullptr;
        f->d.constants = nullptr;
        std::int32_t maxd = -1;
        f->h_dofmap.assign(desc->dofmap, desc->dofmap + nc * desc->dofs_per_cell);
        for (std::int32_t v : f->h_dofmap)
        {
          require(v >= 0, "negative dof");
          maxd = std::max(maxd, v);
        }
        f->max_dof = maxd;
        f->h_x.assign(desc->x, desc->x + 3 * std::int64_t(desc->num_nodes));
        f->h_x_dofmap.assign(desc->x_dofmap, desc->x_dofmap + nc * desc->nodes_per_cell);
        if (desc->cstride > 0)
          f->h_coeffs.assign(desc->coeffs, desc->coeffs + nc * desc->cstride);
        if (desc->num_constants > 0)
          f->h_constants.assign(desc->constants, desc->constants + desc->num_constants);
        f->h_entities.assign(entities, entities + std::int64_t(nentities) * per);
        f->host_kernel = kernel;
        f->host_kernel_data = custom_data;
        f->host_facets = type;
        f->id = next_pattern_id();
        *out = f.release();
      });
}

dfx_status dfx_form_destroy(dfx_form* f)
{
  if (f)
  {
    if (f->dofmap != f->x_dofmap)
      dfree(f->dofmap);
    dfree(f->x_dofmap);
    dfree(f->x);
    dfree(f->coeffs);
    dfree(f->slots);
    dfree(f->inc_off);
    dfree(f->inc);
    dfree(f->islots);
    dfree(f->cell_data);
    dfree(f->cell_mask);
    dfree(f->ref_tensors);
    dfree(f->cell_vec);
    dfree(f->inc_row);
    dfree(f->gblk);
    dfree(f->team_grp);
    dfree(f->trow);
    dfree(f->facets);
    dfree(f->finc_off);
    dfree(f->finc);
    dfree(f->fvals);
    dfree(f->fslot);
    dfree(f->frows);
    dfree(f->cell_elem);
    dfree(f->corner_cell);
    dfree(f->slotmap);
    dfree(f->inc8);
    dfree(f->islot8);
    dfree(f->elem_tensors);
    dfree(f->pack_map);
    dfree(f->pack_stage);
    delete f;
  }
  return DFX_OK;
}

dfx_status dfx_form_set_coefficients(dfx_form* f, const double* coeffs, const double* constants)
{
  return guard(
      [&]
      {
        require(f, "null form");
        const std::int64_t n = std::int64_t(f->d.num_cells) * f->d.cstride;
        if (f->host_kernel)
        {
          // a host kernel reads host copies
          if (coeffs && n > 0)
            f->h_coeffs.assign(coeffs, coeffs + n);
          if (constants)
            f->h_constants.assign(constants, constants + f->d.num_constants);
          return;
        }
        if (coeffs && n > 0)
        {
          if (f->d.dtype == DFX_F64 && f->coeffs)
            h2d(f->coeffs, coeffs, sizeof(double) * n); // same size: reuse the buffer
          else
          {
            void* nc = upload_as(coeffs, n, f->d.dtype);
            dfree(f->coeffs);
            f->coeffs = nc;
          }
        }
        if (constants)
          for (int i = 0; i < f->d.num_constants; ++i)
            f->constants[i] = constants[i];
      });
}

dfx_status dfx_form_pack_coefficients(dfx_form* f, const dfx_mesh* m, int layout,
                                      const double* values, int64_t nvalues)
{
  return guard(
      [&]
      {
        require(f && m && values, "null argument");
        require(m->num_cells == f->d.num_cells, "mesh and form differ in cells");
        const int cs = f->d.cstride;
        require(cs > 0, "form has no coefficients");
        std::int64_t need = 0;
        if (layout == DFX_PACK_CELL)
          need = std::int64_t(m->num_cells) * cs;
        else if (layout == DFX_PACK_VERTEX)
        {
          require(cs == m->nvpc, "vertex layout needs one coefficient per cell vertex");
          need = m->num_vertices;
        }
        else
          throw Error(DFX_ERR_ARG, "unknown coefficient layout");
        require(nvalues == need, "wrong number of values for the layout");
        if (f->pack_mesh != m || f->pack_layout != layout || !f->pack_map)
        {
          dfree(f->pack_map);
          f->pack_map = nullptr;
          std::vector<std::int32_t> map;
          if (layout == DFX_PACK_CELL)
          {
            map.resize(m->num_cells);
            for (std::int32_t c = 0; c < m->num_cells; ++c)
              map[c] = static_cast<std::int32_t>(m->original_cell[c]);
          }
          else
          {
            map.resize(std::size_t(m->num_cells) * m->nvpc);
            for (std::size_t t = 0; t < map.size(); ++t)
              map[t] = static_cast<std::int32_t>(m->input_vertex[m->cells[t]]);
          }
          f->pack_map = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max<std::size_t>(map.size(), 1)));
          h2d(f->pack_map, map.data(), sizeof(std::int32_t) * map.size());
          f->pack_mesh = m;
          f->pack_layout = layout;
        }
        if (f->pack_stage_n < nvalues)
        {
          dfree(f->pack_stage);
          f->pack_stage = static_cast<double*>(dmalloc(sizeof(double) * std::max<std::int64_t>(nvalues, 1)));
          f->pack_stage_n = nvalues;
        }
        h2d(f->pack_stage, values, sizeof(double) * nvalues);
        const std::int64_t n = std::int64_t(f->d.num_cells) * cs;
        if (!f->coeffs)
          f->coeffs = dmalloc(dtype_size(f->d.dtype) * std::max<std::int64_t>(n, 1));
        const int per_cell = layout == DFX_PACK_CELL ? 1 : 0;
        if (n > 0)
        {
          if (f->d.dtype == DFX_F64)
            pack_gather_k<double><<<blocks_for(n), kBlock, 0, stream()>>>(
                n, cs, per_cell, f->pack_map, f->pack_stage, static_cast<double*>(f->coeffs));
          else
            pack_gather_k<float><<<blocks_for(n), kBlock, 0, stream()>>>(
                n, cs, per_cell, f->pack_map, f->pack_stage, static_cast<float*>(f->coeffs));
          check_launch("pack_coefficients");
        }
      });
}

dfx_status dfx_bc_create(const int32_t* dofs, int32_t ndofs, int32_t n_total, const double* g,
                         dfx_bc** out)
{
  return guard(
      [&]
      {
        require(out && ndofs >= 0 && n_total >= 0 && (ndofs == 0 || dofs), "bad argument");
        std::vector<std::int8_t> mark(std::max(n_total, 1), 0);
        for (std::int32_t i = 0; i < ndofs; ++i)
        {
          require(dofs[i] >= 0 && dofs[i] < n_total, "BC dof out of range");
          mark[dofs[i]] = 1;
        }
        auto bc = std::make_unique<dfx_bc>();
        bc->n_total = n_total;
        bc->ndofs = ndofs;
        bc->dofs = static_cast<std::int32_t*>(dmalloc(sizeof(std::int32_t) * std::max(ndofs, 1)));
        h2d(bc->dofs, dofs, sizeof(std::int32_t) * ndofs);
        bc->marker = static_cast<std::int8_t*>(dmalloc(mark.size()));
        h2d(bc->marker, mark.data(), mark.size());
        if (g)
        {
          bc->g = static_cast<double*>(dmalloc(sizeof(double) * std::max(n_total, 1)));
          h2d(bc->g, g, sizeof(double) * n_total);
        }
        *out = bc.release();
      });
}

dfx_status dfx_bc_destroy(dfx_bc* bc)
{
  if (bc)
  {
    dfree(bc->dofs);
    dfree(bc->marker);
    dfree(bc->g);
    delete bc;
  }
  return DFX_OK;
}

dfx_status dfx_assemble_matrix(dfx_csr* A, const dfx_form* a, const dfx_bc* bc)
{
  return guard(
      [&]
      {
        require(A && a, "null argument");
        KernelInfo ki;
        kernel_info(a->d.kernel, ki);
        require(ki.bilinear, "assemble_matrix needs a bilinear form");
        require(!is_element_kernel(a->d.kernel) || a->host_kernel,
                "element-tensor forms are assembled with dfx_assemble_matrix_elements");
        require(A->dtype == a->d.dtype, "matrix and form dtypes differ");
        require(A->bs == a->d.bs, "matrix and form block sizes differ");
        require(!bc || bc->n_total == A->nrows * A->bs, "BC size does not match the matrix");
        require(a->max_dof < A->nrows && a->max_dof < A->ncols, "dofmap exceeds the matrix size");
        dfx_form& f = const_cast<dfx_form&>(*a);
        if (f.host_facets)
        {
          // facet tensors from the host kernel, added on the device
          if (A->dtype == DFX_F64)
            host_facets_matrix<double>(*A, f, bc);
          else
            host_facets_matrix<float>(*A, f, bc);
          return;
        }
        if (f.host_kernel)
        {
          // element matrices from the host kernel, inserted on the device
          if (A->dtype == DFX_F64)
            elements_into_matrix<double>(f, *A, bc, run_host_kernel<double>(f));
          else
            elements_into_matrix<float>(f, *A, bc, run_host_kernel<float>(f));
          A->zeroed = false;
          A->zero_pending = false;
          return;
        }
        if (is_facet_matrix_kernel(f.d.kernel) || is_interior_kernel(f.d.kernel))
        {
          materialize_zero(*A);
          A->zeroed = false;
          const bool in = is_interior_kernel(f.d.kernel);
          if (A->dtype == DFX_F64)
            in ? interior_rows<double>(f, *A, bc) : facet_matrix_rows<double>(f, *A, bc);
          else
            in ? interior_rows<float>(f, *A, bc) : facet_matrix_rows<float>(f, *A, bc);
          return;
        }
        if (general_hex(f))
        {
          // element matrices by quadrature, added by rows in cell order
          if (A->dtype == DFX_F64)
            elements_into_matrix<double>(f, *A, bc, general_hex_elements<double>(f));
          else
            elements_into_matrix<float>(f, *A, bc, general_hex_elements<float>(f));
          A->zeroed = false;
          A->zero_pending = false;
          return;
        }
        if (row_path(f))
        {
          // store mode (A zeroed) overwrites every entry; add mode needs
          // A's values, which are then already stored
          if (A->dtype == DFX_F64)
            assemble_rows<double>(f, *A, bc);
          else
            assemble_rows<float>(f, *A, bc);
          A->zeroed = false;
          A->zero_pending = false;
          return;
        }
        materialize_zero(*A);
        A->zeroed = false;
        ensure_slots(f, *A);
        if (A->dtype == DFX_F64)
        {
          if (f.slot_bytes == 1)
            launch_bilinear<double, std::uint8_t, 0>(f, A, bc, nullptr, 0.0, nullptr);
          else
            launch_bilinear<double, std::uint16_t, 0>(f, A, bc, nullptr, 0.0, nullptr);
        }
        else
        {
          if (f.slot_bytes == 1)
            launch_bilinear<float, std::uint8_t, 0>(f, A, bc, nullptr, 0.0, nullptr);
          else
            launch_bilinear<float, std::uint16_t, 0>(f, A, bc, nullptr, 0.0, nullptr);
        }
      });
}

dfx_status dfx_assemble_matrix_batched(int32_t n, dfx_csr* const* A, const dfx_form* const* a,
                                       const dfx_bc* const* bc)
{
  return guard(
      [&]
      {
        require(n >= 0 && (n == 0 || (A && a && bc)), "null argument");
        if (n == 0)
          return;
        for (int s = 0; s < n; ++s)
          require(A[s] && a[s], "null matrix or form");
        const int dt = a[0]->d.dtype;
        const int k = a[0]->d.kernel;
        require(k == DFX_K_P1TET_A || k == DFX_K_Q1HEX_A,
                "batched assembly supports the P1 and Q1 stiffness forms");
        if (dt == DFX_F64)
          batch_assemble<double>(n, A, a, bc);
        else
          batch_assemble<float>(n, A, a, bc);
      });
}

dfx_status dfx_form_autotune(dfx_form* a, dfx_csr* A, const dfx_bc* bc, int32_t* variant,
                             float* ms)
{
  return guard(
      [&]
      {
        require(a && A, "null argument");
        KernelInfo ki;
        kernel_info(a->d.kernel, ki);
        require(ki.bilinear && A->dtype == a->d.dtype && A->bs == a->d.bs,
                "form and matrix do not match");
        require(a->max_dof < A->nrows, "dofmap exceeds the matrix size");
        require(!is_facet_matrix_kernel(a->d.kernel) && !is_interior_kernel(a->d.kernel),
                "facet forms have no kernel variants");
        dfx_form& f = *a;
        int best = f.row_variant;
        float bms = -1;
        // the elasticity form has one row kernel
        if (row_path(f) && f.d.kernel != DFX_K_P2TET_ELAST_A)
        {
          hipEvent_t e0, e1;
          DFX_HIP(hipEventCreate(&e0));
          DFX_HIP(hipEventCreate(&e1));
          for (int v = 0; v < kRowVariants; ++v)
          {
            f.row_variant = v;
            float t = 0;
            try
            {
              // timed in store mode (A zeroed), the mode of an assembly
              // after MatrixCSR::set(0)
              for (int rep = 0; rep < 4; ++rep)
              {
                if (rep == 1)
                  DFX_HIP(hipEventRecord(e0, stream()));
                A->zeroed = true;
                A->zero_pending = true;
                if (A->dtype == DFX_F64)
                  assemble_rows<double>(f, *A, bc);
                else
                  assemble_rows<float>(f, *A, bc);
                A->zeroed = false;
                A->zero_pending = false;
              }
              DFX_HIP(hipEventRecord(e1, stream()));
              DFX_HIP(hipEventSynchronize(e1));
              DFX_HIP(hipEventElapsedTime(&t, e0, e1));
              t /= 3;
            }
            catch (const Error&)
            {
              continue; // variant not applicable (LDS slice too large)
            }
            if (bms < 0 || t < bms)
            {
              bms = t;
              best = v;
            }
          }
          (void)hipEventDestroy(e0);
          (void)hipEventDestroy(e1);
        }
        f.row_variant = best;
        // release the plans of the variants not kept (the element matrices
        // of variant 7 alone are k^2 values per cell)
        if (best != 7 && best != 17)
        {
          dfree(f.cell_elem);
          f.cell_elem = nullptr;
        }
        if (best != 7 && best != 8)
        {
          dfree(f.gblk);
          dfree(f.inc_row);
          f.gblk = nullptr;
          f.inc_row = nullptr;
          f.ngblk = 0;
          f.gather_pattern = 0;
        }
        if (best < 12 || best > 14)
        {
          dfree(f.inc8);
          dfree(f.islot8);
          f.inc8 = nullptr;
          f.islot8 = nullptr;
          f.fixed_pattern = 0;
        }
        if (best != 9 && best != 15)
        {
          dfree(f.corner_cell);
          dfree(f.slotmap);
          f.corner_cell = nullptr;
          f.slotmap = nullptr;
          f.struct_pattern = 0;
          f.struct_ok = -1;
        }
        if (best != 16 && best != 17)
        {
          dfree(f.trow);
          f.trow = nullptr;
          f.ntrow = 0;
          f.trow_pattern = 0;
        }
        // leave A zeroed (stored when next read)
        A->zeroed = true;
        A->zero_pending = true;
        if (variant)
          *variant = best;
        if (ms)
          *ms = bms;
      });
}

dfx_status dfx_form_set_variant(dfx_form* a, int32_t variant)
{
  return guard(
      [&]
      {
        require(a, "null form");
        require(variant >= 0 && variant < kRowVariants, "variant out of range");
        a->row_variant = variant;
      });
}

dfx_status dfx_set_diagonal(dfx_csr* A, const dfx_bc* bc, double diag)
{
  return guard(
      [&]
      {
        require(A && bc, "null argument");
        require(bc->n_total == A->nrows * A->bs, "BC size does not match the matrix");
        if (bc->ndofs == 0)
          return;
        materialize_zero(*A);
        // the first call on a pattern checks that every BC dof has a
        // diagonal slot (one host round trip); later calls only launch
        const bool check = bc->diag_checked != A->pattern_id;
        int* err = nullptr;
        if (check)
        {
          err = static_cast<int*>(dmalloc(sizeof(int)));
          DFX_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
        }
        if (A->dtype == DFX_F64)
          set_diagonal_k<double><<<blocks_for(bc->ndofs), kBlock, 0, stream()>>>(
              bc->ndofs, bc->dofs, A->bs, A->diag, static_cast<double*>(A->values), diag, err);
        else
          set_diagonal_k<float><<<blocks_for(bc->ndofs), kBlock, 0, stream()>>>(
              bc->ndofs, bc->dofs, A->bs, A->diag, static_cast<float*>(A->values),
              static_cast<float>(diag), err);
        check_launch("set_diagonal");
        A->zeroed = false;
        if (!check)
          return;
        int h = 0;
        d2h(&h, err, sizeof(int));
        dfree(err);
        if (h)
          throw Error(DFX_ERR_PATTERN, "Entry not in sparsity");
        bc->diag_checked = A->pattern_id;
      });
}

dfx_status dfx_assemble_vector(dfx_vec* b, const dfx_form* L)
{
  return guard(
      [&]
      {
        require(b && L, "null argument");
        KernelInfo ki;
        kernel_info(L->d.kernel, ki);
        require(!ki.bilinear, "assemble_vector needs a linear form");
        require(!is_element_kernel(L->d.kernel) || L->host_kernel,
                "element-tensor forms are assembled with dfx_assemble_vector_elements");
        require(b->dtype == L->d.dtype, "vector and form dtypes differ");
        require(std::int64_t(L->max_dof + 1) * L->d.bs <= b->n, "dofmap exceeds the vector size");
        if (L->host_facets)
        {
          if (b->dtype == DFX_F64)
            host_facets_vector<double>(*b, *L);
          else
            host_facets_vector<float>(*b, *L);
          return;
        }
        if (L->host_kernel)
        {
          dfx_form& f = const_cast<dfx_form&>(*L);
          if (b->dtype == DFX_F64)
            elements_into_vector<double>(f, static_cast<double*>(b->data),
                                         run_host_kernel<double>(f));
          else
            elements_into_vector<float>(f, static_cast<float*>(b->data),
                                        run_host_kernel<float>(f));
          return;
        }
        if (is_facet_kernel(L->d.kernel))
        {
          if (b->dtype == DFX_F64)
            facet_rows<double>(*L, static_cast<double*>(b->data));
          else
            facet_rows<float>(*L, static_cast<float*>(b->data));
          return;
        }
        if (L->d.kernel == DFX_K_P1TET_L || L->d.kernel == DFX_K_Q1HEX_L
            || L->d.kernel == DFX_K_P2TET_ELAST_L || L->d.kernel == DFX_K_P3TET_L)
        {
          if (b->dtype == DFX_F64)
            vector_rows<double>(*L, static_cast<double*>(b->data));
          else
            vector_rows<float>(*L, static_cast<float*>(b->data));
          return;
        }
        if (b->dtype == DFX_F64)
          launch_linear<double>(*L, static_cast<double*>(b->data));
        else
          launch_linear<float>(*L, static_cast<float*>(b->data));
      });
}

dfx_status dfx_assemble_matrix_elements(dfx_csr* A, const dfx_form* a, const dfx_bc* bc,
                                        const dfx_vec* E)
{
  return guard(
      [&]
      {
        require(A && a && E, "null argument");
        require(a->d.kernel == DFX_K_ELEMENT_A, "the form does not take element matrices");
        require(A->dtype == a->d.dtype && E->dtype == a->d.dtype,
                "matrix, form and element matrices dtypes differ");
        require(A->bs == a->d.bs, "matrix and form block sizes differ");
        const std::int64_t nd = std::int64_t(a->d.dofs_per_cell) * a->d.bs;
        require(std::int64_t(E->n) == std::int64_t(a->d.num_cells) * nd * nd,
                "element matrices: num_cells x (bs k) x (bs k) values expected");
        require(!bc || bc->n_total == A->nrows * A->bs, "BC size does not match the matrix");
        require(a->max_dof < A->nrows && a->max_dof < A->ncols, "dofmap exceeds the matrix size");
        dfx_form& f = const_cast<dfx_form&>(*a);
        // store mode (A zeroed) overwrites every entry
        if (A->dtype == DFX_F64)
          elements_into_matrix<double>(f, *A, bc, static_cast<const double*>(E->data));
        else
          elements_into_matrix<float>(f, *A, bc, static_cast<const float*>(E->data));
        A->zeroed = false;
        A->zero_pending = false;
      });
}

dfx_status dfx_assemble_vector_elements(dfx_vec* b, const dfx_form* L, const dfx_vec* e)
{
  return guard(
      [&]
      {
        require(b && L && e, "null argument");
        require(L->d.kernel == DFX_K_ELEMENT_L, "the form does not take element vectors");
        require(b->dtype == L->d.dtype && e->dtype == L->d.dtype,
                "vector, form and element vectors dtypes differ");
        require(std::int64_t(e->n)
                    == std::int64_t(L->d.num_cells) * L->d.dofs_per_cell * L->d.bs,
                "element vectors: num_cells x (bs k) values expected");
        require(std::int64_t(L->max_dof + 1) * L->d.bs <= b->n, "dofmap exceeds the vector size");
        if (b->dtype == DFX_F64)
          elements_into_vector<double>(*L, static_cast<double*>(b->data),
                                       static_cast<const double*>(e->data));
        else
          elements_into_vector<float>(*L, static_cast<float*>(b->data),
                                      static_cast<const float*>(e->data));
      });
}

dfx_status dfx_apply_lifting_elements(dfx_vec* b, const dfx_form* a, const dfx_bc* bc,
                                      const dfx_vec* E, const dfx_vec* x0, double alpha)
{
  return guard(
      [&]
      {
        require(b && a && E, "null argument");
        require(a->d.kernel == DFX_K_ELEMENT_A, "the form does not take element matrices");
        require(b->dtype == a->d.dtype && E->dtype == a->d.dtype,
                "vector, form and element matrices dtypes differ");
        const std::int64_t nd = std::int64_t(a->d.dofs_per_cell) * a->d.bs;
        require(std::int64_t(E->n) == std::int64_t(a->d.num_cells) * nd * nd,
                "element matrices: num_cells x (bs k) x (bs k) values expected");
        if (!bc || bc->ndofs == 0)
          return;
        require(bc->n_total == b->n, "BC size does not match the vector");
        require(!x0 || (x0->n == b->n && x0->dtype == b->dtype), "x0 does not match b");
        require(std::int64_t(a->max_dof + 1) * a->d.bs <= b->n, "dofmap exceeds the vector size");
        // g = 0 and x0 = 0: every term alpha E_ij (g_j - x0_j) is zero
        if (!bc->g && !x0)
          return;
        if (b->dtype == DFX_F64)
          elements_lift<double>(*a, static_cast<double*>(b->data), *bc,
                                static_cast<const double*>(E->data),
                                x0 ? static_cast<const double*>(x0->data) : nullptr, alpha);
        else
          elements_lift<float>(*a, static_cast<float*>(b->data), *bc,
                               static_cast<const float*>(E->data),
                               x0 ? static_cast<const float*>(x0->data) : nullptr, alpha);
      });
}

dfx_status dfx_apply_lifting(dfx_vec* b, const dfx_form* a, const dfx_bc* bc, const dfx_vec* x0,
                             double alpha)
{
  return guard(
      [&]
      {
        require(b && a, "null argument");
        KernelInfo ki;
        kernel_info(a->d.kernel, ki);
        require(ki.bilinear, "apply_lifting needs a bilinear form");
        require(!is_element_kernel(a->d.kernel) || a->host_kernel,
                "apply_lifting needs a form with a kernel (element-tensor forms have none)");
        require(b->dtype == a->d.dtype, "vector and form dtypes differ");
        if (!bc || bc->ndofs == 0)
          return;
        require(bc->n_total == b->n, "BC size does not match the vector");
        require(!x0 || (x0->n == b->n && x0->dtype == b->dtype), "x0 does not match b");
        require(std::int64_t(a->max_dof + 1) * a->d.bs <= b->n, "dofmap exceeds the vector size");
        // g = 0 and x0 = 0: every term alpha A_ij (g_j - x0_j) is zero
        if (!bc->g && !x0)
          return;
        dfx_form& f = const_cast<dfx_form&>(*a);
        require(!f.host_facets,
                "apply_lifting with nonzero Dirichlet data through a facet integral given by "
                "a host kernel is not supported");
        if (f.host_kernel)
        {
          // the kernel's element matrices (as lift_bc recomputes them)
          if (b->dtype == DFX_F64)
            elements_lift<double>(f, static_cast<double*>(b->data), *bc,
                                  run_host_kernel<double>(f),
                                  x0 ? static_cast<const double*>(x0->data) : nullptr, alpha);
          else
            elements_lift<float>(f, static_cast<float*>(b->data), *bc,
                                 run_host_kernel<float>(f),
                                 x0 ? static_cast<const float*>(x0->data) : nullptr, alpha);
          return;
        }
        if (f.d.kernel == DFX_K_P2TET_ELAST_A)
        {
          if (b->dtype == DFX_F64)
            elast_lift_rows<double>(f, *bc, x0 ? static_cast<const double*>(x0->data) : nullptr,
                                    alpha, static_cast<double*>(b->data));
          else
            elast_lift_rows<float>(f, *bc, x0 ? static_cast<const float*>(x0->data) : nullptr,
                                   alpha, static_cast<float*>(b->data));
          return;
        }
        if (is_interior_kernel(f.d.kernel))
        {
          if (b->dtype == DFX_F64)
            interior_lift_rows<double>(f, *bc, x0 ? static_cast<const double*>(x0->data) : nullptr,
                                       alpha, static_cast<double*>(b->data));
          else
            interior_lift_rows<float>(f, *bc, x0 ? static_cast<const float*>(x0->data) : nullptr,
                                      alpha, static_cast<float*>(b->data));
          return;
        }
        if (is_facet_matrix_kernel(f.d.kernel))
        {
          if (b->dtype == DFX_F64)
            facet_lift_rows<double>(f, *bc, x0 ? static_cast<const double*>(x0->data) : nullptr,
                                    alpha, static_cast<double*>(b->data));
          else
            facet_lift_rows<float>(f, *bc, x0 ? static_cast<const float*>(x0->data) : nullptr,
                                   alpha, static_cast<float*>(b->data));
          return;
        }
        if (general_hex(f))
        {
          if (b->dtype == DFX_F64)
            elements_lift<double>(f, static_cast<double*>(b->data), *bc,
                                  general_hex_elements<double>(f),
                                  x0 ? static_cast<const double*>(x0->data) : nullptr, alpha);
          else
            elements_lift<float>(f, static_cast<float*>(b->data), *bc,
                                 general_hex_elements<float>(f),
                                 x0 ? static_cast<const float*>(x0->data) : nullptr, alpha);
          return;
        }
        if (row_path(f))
        {
          if (b->dtype == DFX_F64)
            lift_rows<double>(f, *bc, x0 ? static_cast<const double*>(x0->data) : nullptr, alpha,
                              static_cast<double*>(b->data));
          else
            lift_rows<float>(f, *bc, x0 ? static_cast<const float*>(x0->data) : nullptr, alpha,
                             static_cast<float*>(b->data));
          return;
        }
        if (b->dtype == DFX_F64)
          launch_bilinear<double, std::uint8_t, 1>(
              f, nullptr, bc, x0 ? static_cast<const double*>(x0->data) : nullptr, alpha,
              static_cast<double*>(b->data));
        else
          launch_bilinear<float, std::uint8_t, 1>(
              f, nullptr, bc, x0 ? static_cast<const float*>(x0->data) : nullptr, alpha,
              static_cast<float*>(b->data));
      });
}

dfx_status dfx_bc_set(dfx_vec* b, const dfx_bc* bc, const dfx_vec* x0, double alpha)
{
  return guard(
      [&]
      {
        require(b && bc, "null argument");
        require(bc->n_total == b->n, "BC size does not match the vector");
        require(!x0 || (x0->n == b->n && x0->dtype == b->dtype), "x0 does not match b");
        if (bc->ndofs == 0)
          return;
        if (b->dtype == DFX_F64)
          bc_set_k<double><<<blocks_for(bc->ndofs), kBlock, 0, stream()>>>(
              bc->ndofs, bc->dofs, bc->g, x0 ? static_cast<const double*>(x0->data) : nullptr,
              alpha, static_cast<double*>(b->data));
        else
          bc_set_k<float><<<blocks_for(bc->ndofs), kBlock, 0, stream()>>>(
              bc->ndofs, bc->dofs, bc->g, x0 ? static_cast<const float*>(x0->data) : nullptr,
              alpha, static_cast<float*>(b->data));
        check_launch("bc_set");
      });
}

} // extern "C"

namespace dfx
{
namespace
{
/// The system cell pass (cell_system_k) for a stiffness form a and source
/// form L that share their geometry: returns false (nothing launched) when
/// the pair does not qualify. Whether the device geometry of the two forms
/// is the same is checked once per pair (one comparison pass and a host
/// round trip), then remembered by form id.
template <typename T>
void launch_cell_system(const dfx_form& a, const dfx_form& L, const dfx_bc* bc)
{
  const std::int32_t nc = a.d.num_cells;
  auto go = [&](auto kern)
  {
    kern<<<blocks_for(nc), kBlock, 0, stream()>>>(
        nc, static_cast<const T*>(a.x), a.x_dofmap, static_cast<const T*>(a.coeffs),
        a.d.cstride, static_cast<T>(a.constants[0]), bc ? bc->marker : nullptr,
        static_cast<T*>(a.cell_data), bc ? a.cell_mask : nullptr,
        static_cast<const T*>(L.coeffs), L.d.cstride, static_cast<T*>(L.cell_vec));
  };
  // the records variant 15 reads when the form keeps compact ones
  if (a.d.nodes_per_cell == 8 && a.row_variant == 15 && a.compact_records)
    go(cell_system_k<T, 8, true>);
  else if (a.d.nodes_per_cell == 4)
    go(cell_system_k<T, 4>);
  else
    go(cell_system_k<T, 8>);
  check_launch("cell_system");
}

bool system_cell_pass(dfx_form& a, dfx_form& L, const dfx_bc* bc)
{
  const int ka = a.d.kernel, kl = L.d.kernel;
  const bool pair = (ka == DFX_K_P1TET_A && kl == DFX_K_P1TET_L)
                    || (ka == DFX_K_Q1HEX_A && kl == DFX_K_Q1HEX_L);
  if (!pair || !a.affine || !L.affine || a.d.dtype != L.d.dtype || a.d.bs != 1 || L.d.bs != 1
      || a.d.num_cells != L.d.num_cells || a.d.num_nodes != L.d.num_nodes
      || a.d.nodes_per_cell != L.d.nodes_per_cell || a.dofmap != a.x_dofmap
      || !row_path(a) || a.row_variant == 10 || a.d.num_cells == 0 || !a.cell_data
      || !L.cell_vec || (bc && !a.cell_mask))
    return false;
  if (a.same_geometry_as != L.id)
  {
    if (capturing())
      return false;
    const std::size_t ts = dtype_size(a.d.dtype);
    int* diff = static_cast<int*>(dmalloc(sizeof(int)));
    fill_i32_k<<<1, kBlock, 0, stream()>>>(diff, 1, 0);
    const std::int64_t nx = std::int64_t(a.d.num_nodes) * 3 * ts / 4;
    const std::int64_t nd = std::int64_t(a.d.num_cells) * a.d.nodes_per_cell;
    if (nx > 0)
      words_differ_k<<<blocks_for(nx), kBlock, 0, stream()>>>(
          nx, static_cast<const std::uint32_t*>(a.x), static_cast<const std::uint32_t*>(L.x), diff);
    words_differ_k<<<blocks_for(nd), kBlock, 0, stream()>>>(
        nd, reinterpret_cast<const std::uint32_t*>(a.x_dofmap),
        reinterpret_cast<const std::uint32_t*>(L.x_dofmap), diff);
    check_launch("geometry comparison");
    int h = 1;
    d2h(&h, diff, sizeof(int));
    dfree(diff);
    if (h)
      return false;
    a.same_geometry_as = L.id;
  }
  if (a.d.dtype == DFX_F64)
    launch_cell_system<double>(a, L, bc);
  else
    launch_cell_system<float>(a, L, bc);
  a.cell_data_ready = true;
  L.cell_vec_ready = true;
  return true;
}
} // namespace
} // namespace dfx

extern "C" {

dfx_status dfx_assemble_system(dfx_csr* A, const dfx_form* a, const dfx_bc* bc, double diag,
                               dfx_vec* b, const dfx_form* L, int32_t zero_b)
{
  return guard(
      [&]
      {
        require(A && a && b && L, "null argument");
        auto ok = [](dfx_status st)
        {
          if (st != DFX_OK)
            throw Error(st, dfx_last_error());
        };
        Context& c = ctx();
        ensure_init();
        // a stiffness and a source form on the same geometry share one
        // pass over the cells (records for the matrix, element vectors for
        // b); the flags that tell the two assemblies so are cleared on
        // every way out
        dfx_form& af = const_cast<dfx_form&>(*a);
        dfx_form& Lf = const_cast<dfx_form&>(*L);
        struct ReadyGuard
        {
          dfx_form& a;
          dfx_form& L;
          ~ReadyGuard()
          {
            a.cell_data_ready = false;
            L.cell_vec_ready = false;
          }
        } rg{af, Lf};
        const bool shared = system_cell_pass(af, Lf, bc);
        // the vector assembly reads only L's data and writes only b: on
        // large meshes it runs on the second stream while the matrix is
        // assembled (below ~1M cells the fork and join cost more than the
        // overlap gains). After a shared cell pass only the vector's row
        // pass is left, and beside the matrix's row kernel both slow down
        // more than they overlap, so it runs after it on one stream.
        const bool fork = !shared && L->d.num_cells >= (1 << 20);
        // once forked, the library stream waits for the second stream on
        // every way out (an error in either part included), so no later
        // work on the library stream runs beside writes to b
        struct JoinGuard
        {
          Context& c;
          bool pending = false;
          void join()
          {
            pending = false;
            DFX_HIP(hipStreamWaitEvent(c.stream, c.join, 0));
          }
          ~JoinGuard()
          {
            if (pending)
            {
              (void)hipEventRecord(c.join, c.aux);
              (void)hipStreamWaitEvent(c.stream, c.join, 0);
            }
          }
        } jg{c};
        if (fork)
        {
          DFX_HIP(hipEventRecord(c.fork, c.stream));
          DFX_HIP(hipStreamWaitEvent(c.aux, c.fork, 0));
          jg.pending = true;
        }
        {
          StreamScope on(fork ? c.aux : c.stream);
          if (zero_b)
            ok(dfx_vec_set(b, 0.0));
          ok(dfx_assemble_vector(b, L));
        }
        if (fork)
          DFX_HIP(hipEventRecord(c.join, c.aux));
        ok(dfx_assemble_matrix(A, a, bc));
        if (bc)
          ok(dfx_set_diagonal(A, bc, diag));
        if (fork)
          jg.join();
        if (bc)
        {
          ok(dfx_apply_lifting(b, a, bc, nullptr, 1.0));
          ok(dfx_bc_set(b, bc, nullptr, 1.0));
        }
      });
}

} // extern "C"
