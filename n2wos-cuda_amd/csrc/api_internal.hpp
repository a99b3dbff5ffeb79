// api_internal.hpp -- error plumbing and handle types behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/n2wos_amd.h"
#include "bvh.hpp"

namespace n2 {

// Raised inside the library; converted to a return code at the ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define N2_HIP(expr)                                                                                 \
  do {                                                                                               \
    hipError_t e__ = (expr);                                                                         \
    if (e__ != hipSuccess)                                                                           \
      throw ::n2::Error(N2WOS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e__) + " at " + \
                                           __FILE__ + ":" + std::to_string(__LINE__));               \
  } while (0)

#define N2_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) throw ::n2::Error((code), (msg)); \
  } while (0)

// Runs f and maps exceptions to return codes (no exception crosses the ABI).
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return N2WOS_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return N2WOS_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return N2WOS_ERR_INTERNAL;
  }
}

}  // namespace n2

// Device-resident geometry handle.
struct n2wos_bvh {
  n2::Bvh4Node* d_nodes = nullptr;
  n2::TriRec* d_tris = nullptr;
  int device = 0;
  int64_t triangle_count = 0;
  int64_t node_count = 0;
  int64_t binary_nodes = 0;
  int leaf_size = 8;
  int max_depth = 0;       // binary tree depth
  int depth = 0;           // 4-wide tree depth
  float delta = 0.0f;
  double build_ms = 0.0;
  float bounds_lo[3] = {0, 0, 0};
  float bounds_hi[3] = {0, 0, 0};
  std::string build_method = "median";
};
