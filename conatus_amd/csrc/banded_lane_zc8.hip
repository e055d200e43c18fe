// Lane-mode streamed kernels of libcnhip for element type cn::zc, blocks of 8 steps.
#define CN_LANE_T cn::zc
#define CN_LANE_SPX 8
#define CN_LANE_SOLVE lane_solve_zc8
#define CN_LANE_BORDER lane_border_zc8
#include "banded_lane.h"
