// Interface between the host driver and the HIP apply kernels.
#pragma once

#include <string>

#include "mw_internal.hpp"

namespace mwg {

struct DeviceCtx;

// Returns 0 on success; on failure a negative code and a message in err.
int dev_create(int device, DeviceCtx **out, std::string &err);
void dev_destroy(DeviceCtx *c);
// Drop the device copy of an operator's tables (called when it is destroyed).
void dev_forget_operator(DeviceCtx *c, const void *key);
// Run the per-node calculator for every output node of the plan and copy the
// raw non-standard blocks (n_out x 8(k+1)^3) to out_ns.  kernel_ms receives
// the device time of the kernel.
int dev_apply_plan(DeviceCtx *c, const void *op_key, const OperFlat &f, const ApplyPlan &plan, double *out_ns,
                   double *kernel_ms, std::string &err);

}  // namespace mwg
