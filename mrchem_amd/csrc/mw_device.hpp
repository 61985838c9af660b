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
// A plan resident in device memory: band tables, input slots, output blocks.
struct DevicePlan;
int dev_plan_create(DeviceCtx *c, const void *op_key, const OperFlat &f, const ApplyPlan &plan, DevicePlan **out,
                    std::string &err);
// Run the per-node calculator over the resident plan.  kernel_ms receives the
// device time of the launch, term_apps the number of term-applications that
// passed screening.
int dev_plan_run(DevicePlan *p, double *kernel_ms, unsigned long long *term_apps, std::string &err);
// Copy the raw non-standard blocks (n_out x 8(k+1)^3) to the host.
int dev_plan_download(DevicePlan *p, double *out_ns, std::string &err);
void dev_plan_destroy(DevicePlan *p);
// create + run + download + destroy
int dev_apply_plan(DeviceCtx *c, const void *op_key, const OperFlat &f, const ApplyPlan &plan, double *out_ns,
                   double *kernel_ms, unsigned long long *term_apps, std::string &err);

}  // namespace mwg
