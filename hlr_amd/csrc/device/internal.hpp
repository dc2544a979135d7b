// internal.hpp -- accessors between the translation units of libhmvm.so.
#pragma once
#include <cstdint>

#include "hmvm.h"
#include "kernels.hpp"
#include "plan_host.hpp"

namespace hmvm {

const DevPlan& plan_devplan(const hmvm_plan* p);
const HostLayout& plan_layout(const hmvm_plan* p);
int plan_device(const hmvm_plan* p);

// workspace doubles of a forward call with R right-hand sides
int64_t plan_fwd_ws_doubles(const hmvm_plan* p, int R);

} // namespace hmvm
