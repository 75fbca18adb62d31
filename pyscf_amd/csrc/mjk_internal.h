// Context services shared by the library's host-side translation units (the
// J/K driver, the device-resident SCF, the request server).  Not part of the
// C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <mutex>
#include <shared_mutex>

#include "mjk.h"

namespace mjk_internal {

hipStream_t ctx_stream(mjk_ctx* ctx);
int ctx_device(const mjk_ctx* ctx);
size_t ctx_budget(const mjk_ctx* ctx);
// sets the context's last error (printf format) and returns code
int ctx_fail(mjk_ctx* ctx, int code, const char* fmt, ...);
// serialises the calls on one context
std::recursive_mutex& ctx_mutex(mjk_ctx* ctx);
// Held shared by every stream capture in the process: an operation that may
// synchronise the device (hipFree, a library handle's set-up or tear-down)
// takes it exclusively, so it never runs while another thread captures.
std::shared_mutex& capture_mutex();
// The context's dense-algebra library handle (opaque), created by its user
// under capture_mutex() and destroyed with the context by `destroy`.
void* ctx_blas(mjk_ctx* ctx);
void ctx_set_blas(mjk_ctx* ctx, void* handle, void (*destroy)(void*));
// the context's rocBLAS handle (a rocblas_handle), made on first use; null on failure
void* ctx_rocblas(mjk_ctx* ctx);
// Rys root table of the context (device)
const double* ctx_rys(const mjk_ctx* ctx);
// density-fitted J and K of one density (device, nao x nao) with K from its
// occupied orbitals X (device, column-major nao x nocc, D = X X^T)
int df_get_jk_dev(mjk_df* df, const double* d_dm, const double* d_orbo_colmajor, int nocc, double* d_vj,
                  double* d_vk);
int df_nao(const mjk_df* df);
mjk_ctx* df_ctx(const mjk_df* df);

}  // namespace mjk_internal
