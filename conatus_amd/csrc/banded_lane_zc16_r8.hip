// Lane-mode streamed kernels of libcnhip for element type cn::zc, blocks of 16 steps, up to 8
// border rows.
#define CN_LANE_T cn::zc
#define CN_LANE_SPX 16
#define CN_LANE_RB 8
#define CN_LANE_SOLVE lane_solve_zc16_r8
#define CN_LANE_BORDER lane_border_zc16_r8
#include "banded_lane.h"
