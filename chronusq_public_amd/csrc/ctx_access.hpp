// Internal accessors of the opaque context for the library's own host
// modules (the SCF driver and the serving loop).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "basis.hpp"
#include "cqx.h"

namespace cqx {
hipStream_t ctx_stream(const cqx_ctx *c);
int ctx_device(const cqx_ctx *c);
double ctx_thresh(const cqx_ctx *c);
const Basis &ctx_basis(const cqx_ctx *c);
int ctx_fail(cqx_ctx *c, int code, const std::string &msg);
// A batched context over the molecules of one-molecule contexts, from their
// pair data and Schwarz bounds (no ERI evaluated); the parts stay valid.
int ctx_merge(cqx_ctx *const *parts, int n, cqx_ctx **out);
}  // namespace cqx
