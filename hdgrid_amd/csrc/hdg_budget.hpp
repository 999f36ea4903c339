// Learned per-query probe budget: fit parameters and report (hdg_budget.hip).
#pragma once

#include <cstdint>

namespace hdg {

constexpr int kBudgetFeatures = 48;

struct BudgetFitParams {
  int levels = 10;        // ladder length
  double t0 = 4.0;        // clusters probed at level 0
  double s0 = 8.0;        // sub-clusters kept at level 0
  double ratio = 1.41421356;  // growth per level
  float target = 0.95f;   // per-query recall@k that defines a query's label
  int epochs = 30;
  uint32_t batch = 1024;
  float lr = 2e-3f;
  double holdout = 0.1;   // share of the queries used to calibrate the margin
  float max_margin = 4.0f;
  uint64_t seed = 0x4844;
};

struct BudgetReport {
  float final_loss;       // training MSE of the last epoch (levels^2)
  float margin;           // added to the prediction before rounding up
  float holdout_recall;   // mean per-query recall on the holdout at the chosen margin
  float holdout_level;    // mean level used on the holdout
  float mean_label;       // mean minimal level over all training queries
  float fixed_recall;     // mean recall at the top level (the fixed-budget fallback)
  uint64_t n_train, n_holdout;
};

// Gradient of the training kernel on one batch (for checking against a CPU implementation).
void budget_debug_grads(const float* params, const float* X, const float* Y, uint32_t n, float* grads, float* loss);
int budget_param_count();

}  // namespace hdg
