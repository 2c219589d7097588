// Shared internals of libdfx: error plumbing and handle definitions.
#pragma once

#include "../../include/dfx.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dfx
{

/// Error carrying a C ABI status code.
struct Error : std::runtime_error
{
  dfx_status code;
  Error(dfx_status c, const std::string& msg) : std::runtime_error(msg), code(c)
  {
  }
};

void set_last_error(const std::string& msg);

/// Run f, converting exceptions into a status code and the message of
/// dfx_last_error().
template <typename F>
dfx_status guard(F&& f) noexcept
{
  try
  {
    f();
    return DFX_OK;
  }
  catch (const Error& e)
  {
    set_last_error(e.what());
    return e.code;
  }
  catch (const std::exception& e)
  {
    set_last_error(e.what());
    return DFX_ERR_ARG;
  }
  catch (...)
  {
    set_last_error("unknown error");
    return DFX_ERR_ARG;
  }
}

inline void require(bool ok, const std::string& msg)
{
  if (!ok)
    throw Error(DFX_ERR_ARG, msg);
}

} // namespace dfx

/// Mesh after create_box + RCM cell reordering + vertex numbering.
struct dfx_mesh
{
  int cell = DFX_TET;
  int nvpc = 4;
  std::int32_t num_cells = 0;
  std::int32_t num_vertices = 0;
  std::vector<double> x;                  // num_vertices x 3
  std::vector<std::int32_t> cells;        // num_cells x nvpc
  std::vector<std::int64_t> original_cell;
  std::vector<std::int64_t> input_vertex;
  // Lazily computed entities of dimension 1 and 2 (cell->entity and
  // entity->sorted vertices) and their counts.
  std::vector<std::int32_t> c2e[3];
  std::vector<std::int32_t> e2v[3];
  std::int32_t num_entities[3] = {0, 0, 0};
  bool have_entities[3] = {false, false, false};
};

/// Lagrange space: (node) dofmap with block size.
struct dfx_space
{
  dfx_mesh* mesh = nullptr;
  int degree = 1;
  int bs = 1;
  int dofs_per_cell = 4;
  std::int32_t num_nodes = 0;
  std::vector<std::int32_t> dofmap; // num_cells x dofs_per_cell
};

struct dfx_sparsity
{
  std::int32_t nrows = 0;
  std::int32_t ncols = 0;
  int bs = 1;
  std::vector<std::int64_t> row_ptr;
  std::vector<std::int32_t> cols;
};
