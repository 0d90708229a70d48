// apsep_segmented_512.hip -- the segmented solver on tiles of 512 lanes x 16
// keys (apsep_segmented_impl.hpp).
#define APSEP_TILE_LANES 512
#define APSEP_SEG_NS seg512
#include "apsep_segmented_impl.hpp"
