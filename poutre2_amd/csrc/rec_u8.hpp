// rec_u8.hpp -- register-tiled uint8 geodesic reconstruction (rec_u8.hip).
#pragma once
#include "pm_common.hpp"

namespace pm {

// True for 2D images (rank 1 or 2) with a 3x3-class SE the kernel handles.
bool rec_u8_applicable(int se, Dims d);
// Reconstruction of marker under mask into out (out distinct from both), by dilation (DIL) or erosion.
template <bool DIL> int rec_u8(const uint8_t *marker, const uint8_t *mask, uint8_t *out, Dims d, int se, hipStream_t s);
// The same, run speculatively: its round 0 also checks whether some marker pixel lies beyond
// the mask (above it for a dilation, below it for an erosion), which the caller assumed not to
// happen; if one does, every round exits at once, *aborted is set and out is left unspecified.
template <bool DIL>
int rec_u8_guarded(const uint8_t *marker, const uint8_t *mask, uint8_t *out, Dims d, int se, bool *aborted,
                   hipStream_t s);

}  // namespace pm
