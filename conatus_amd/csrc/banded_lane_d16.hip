// Lane-mode streamed kernels of libcnhip for element type double, blocks of 16 steps.
#define CN_LANE_T double
#define CN_LANE_SPX 16
#define CN_LANE_SOLVE lane_solve_d16
#define CN_LANE_BORDER lane_border_d16
#include "banded_lane.h"
