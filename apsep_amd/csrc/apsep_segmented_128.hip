// apsep_segmented_128.hip -- the segmented solver on tiles of 128 lanes x 16
// keys (apsep_segmented_impl.hpp): twice the workgroups of the 256-lane
// shape for the same keys, for batches too small to fill the GPU otherwise.
#define APSEP_TILE_LANES 128
#define APSEP_SEG_NS seg128
#include "apsep_segmented_impl.hpp"
