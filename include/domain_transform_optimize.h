// Solver parameters.  Interface and defaults of the reference's
// src/domain_transform_optimize.h:8-50; the free function is inline here so
// the header can be included from several translation units.
#ifndef DOMAIN_TRANSFORM_OPTIMIZE_H_
#define DOMAIN_TRANSFORM_OPTIMIZE_H_

#include <string>

namespace domain_transform {

enum class RobustLoss { L1, L2, CHARBONNIER };

inline std::string RobustLossToString(const RobustLoss& loss) {
  switch (loss) {
    case RobustLoss::L1: return "RobustLoss::L1";
    case RobustLoss::L2: return "RobustLoss::L2";
    case RobustLoss::CHARBONNIER: return "RobustLoss::CHARBONNIER";
  }
  return "RobustLoss::?";
}

struct DomainOptimizeParams {
  int num_iterations = 3;
  RobustLoss loss = RobustLoss::L2;
  float lambda = 1.0f;
  float sigma_z = 100;
  float step_size = 0.99;

  // Weights of the additional terms (photometric term of the stereo solver).
  float photo_consis_lambda = 1.0f;
  float DT_lambda = 0.2f;

  std::string ToString() const {
    std::string s = "Num iter: " + std::to_string(num_iterations) + "\n";
    s += "loss: " + RobustLossToString(loss) + "\n";
    s += "lambda: " + std::to_string(lambda) + "\n";
    s += "sigma_z: " + std::to_string(sigma_z) + "\n";
    s += "step_size: " + std::to_string(step_size) + "\n";
    s += "photo_consis_lambda: " + std::to_string(photo_consis_lambda) + "\n";
    s += "DT_lambda: " + std::to_string(DT_lambda) + "\n";
    return s;
  }
};

}  // namespace domain_transform

#endif  // DOMAIN_TRANSFORM_OPTIMIZE_H_
