// Filter parameters: sigma_x, sigma_y (spatial), sigma_r (range) and N extra
// sigmas (sigmas[0] is the confidence variance scale).  Interface of the
// reference's src/domain_transform_filter_params.h:8-31.
#ifndef DOMAIN_TRANSFORM_FILTER_PARAMS_H_
#define DOMAIN_TRANSFORM_FILTER_PARAMS_H_

#include <string>

namespace domain_transform {

template <int N>
struct DomainFilterParamsVec {
  float sigma_x;
  float sigma_y;
  float sigma_r;
  float sigmas[N + 1];

  std::string ToString() const {
    std::string s = "DomainFilterParamsVec<" + std::to_string(N) + ">:\n";
    s += "sigma_x = " + std::to_string(sigma_x) + "\n";
    s += "sigma_y = " + std::to_string(sigma_y) + "\n";
    s += "sigma_r = " + std::to_string(sigma_r) + "\n";
    for (int i = 0; i < N; i++)
      s += "sigmas[" + std::to_string(i) + "] = " + std::to_string(sigmas[i]) + "\n";
    return s;
  }
};

using DomainFilterParams = DomainFilterParamsVec<0>;

}  // namespace domain_transform

#endif  // DOMAIN_TRANSFORM_FILTER_PARAMS_H_
