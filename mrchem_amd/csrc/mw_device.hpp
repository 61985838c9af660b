// Interface between the host driver and the HIP apply kernels.
#pragma once

#include <string>
#include <vector>

#include "mw_internal.hpp"

namespace mwg {

struct DeviceCtx;

// Returns 0 on success; on failure a negative code and a message in err
// (-1: invalid argument, -2: device error).
int dev_create(int device, DeviceCtx **out, std::string &err);
void dev_destroy(DeviceCtx *c);
// Drop the device copy of an operator's tables (called when it is destroyed).
void dev_forget_operator(DeviceCtx *c, const void *key);

// One apply of a (possibly batched) device plan: the operator (its host
// tables, keyed for the context's resident copy) and the host plan.
struct PlanPart {
    const void *op_key;
    const OperFlat *flat;
    const ApplyPlan *plan;
};

// A plan resident in device memory: band tables, input slots and output
// blocks of one or more applies, run together in each launch.
struct DevicePlan;
int dev_plan_create(DeviceCtx *c, const std::vector<PlanPart> &parts, DevicePlan **out, std::string &err);
// Run the per-node calculator over the resident plan.  kernel_ms receives the
// device time of the launches, term_apps the number of term-applications that
// passed screening.
// mode 0: every term in FP64.  mode 1: terms with bound > tau * thr in FP64,
// the rest on FP32 MFMA, their sum added to the FP64 output.  lp_apps
// receives the number of FP32 term-applications.
// split_ms (optional, 2 values): device time of the FP64-class launch and of
// the low-precision launch (they run concurrently, on two streams).
int dev_plan_run(DevicePlan *p, int mode, double tau, double *kernel_ms, unsigned long long *term_apps,
                 unsigned long long *lp_apps, std::string &err, double *split_ms = nullptr);
// Copy the raw non-standard blocks of every part, in part order
// (sum of n_out x 8(k+1)^3), to the host.
int dev_plan_download(DevicePlan *p, double *out_ns, std::string &err);
void dev_plan_destroy(DevicePlan *p);

}  // namespace mwg
