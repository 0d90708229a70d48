// apsep_segmented_256.hip -- the segmented solver on tiles of 256 lanes x 16
// keys (apsep_segmented_impl.hpp).
#define APSEP_TILE_LANES 256
#define APSEP_SEG_NS seg256
#include "apsep_segmented_impl.hpp"
