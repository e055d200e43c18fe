// Lane-mode streamed kernels of libcnhip for element type cn::zc, blocks of 16 steps.
#define CN_LANE_T cn::zc
#define CN_LANE_SPX 16
#define CN_LANE_SOLVE lane_solve_zc16
#define CN_LANE_BORDER lane_border_zc16
#include "banded_lane.h"
