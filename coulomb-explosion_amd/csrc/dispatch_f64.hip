// f64 integrator kernels (the on-GPU accuracy oracle path).
#include "dispatch.hpp"

namespace ce {
hipError_t launch_dp5_f64(const LaunchReq& req, LaunchInfo* info, bool* bad_n) {
  return req.per_sim_charges ? launch_all<double, true, false>(req, info, bad_n)
                             : launch_all<double, false, false>(req, info, bad_n);
}
}  // namespace ce
