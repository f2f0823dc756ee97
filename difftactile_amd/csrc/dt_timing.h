// Stage timer of the fused substep path (see timing.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace dt {

// Stage ids of dt_timing_read (include/difftactile.h DT_STAGE_*).
enum {
  STAGE_BIN = 0,
  STAGE_P2G = 1,
  STAGE_GRID_OP = 2,
  STAGE_G2P = 3,
  STAGE_G2P_ADJ = 4,
  STAGE_GRID_OP_ADJ = 5,
  STAGE_P2G_ADJ = 6,
};

// Records a HIP event pair around the launches issued during its lifetime on
// stream st when timing is enabled.
class StageTimer {
 public:
  StageTimer(int stage, hipStream_t st);
  ~StageTimer();

 private:
  int stage_;
  hipStream_t st_;
  hipEvent_t e1_;
};

}  // namespace dt
