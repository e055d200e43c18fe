// Lane-mode streamed kernels of libcnhip for element type double, blocks of 16 steps, up to 2
// border rows.
#define CN_LANE_T double
#define CN_LANE_SPX 16
#define CN_LANE_RB 2
#define CN_LANE_SOLVE lane_solve_d16_r2
#define CN_LANE_BORDER lane_border_d16_r2
#include "banded_lane.h"
