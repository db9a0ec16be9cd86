// f32 integrator kernels (the production path), instantiated here so they
// compile in parallel with the f64 ones.
#include "dispatch.hpp"

namespace ce {
hipError_t launch_dp5_f32(const LaunchReq& req, LaunchInfo* info, bool* bad_n) {
  return req.per_sim_charges ? launch_all<float, true, true>(req, info, bad_n)
                             : launch_all<float, false, true>(req, info, bad_n);
}
}  // namespace ce
