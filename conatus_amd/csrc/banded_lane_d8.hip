// Lane-mode streamed kernels of libcnhip for element type double, blocks of 8 steps.
#define CN_LANE_T double
#define CN_LANE_SPX 8
#define CN_LANE_SOLVE lane_solve_d8
#define CN_LANE_BORDER lane_border_d8
#include "banded_lane.h"
