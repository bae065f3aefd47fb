// beadfinder.hpp -- bead detection behind QTrkFindBeads
// (cputrack/qtrk_c_api.cpp:33-58, cputrack/BeadFinder.h:9-28).
#pragma once

#include "qtrk_c_api.h"

#include <vector>

namespace qtrk {

// Largest power-of-two side the frame is padded to: one transform row in LDS.
constexpr int kMaxBeadFinderSide = 8192;
// Largest sample side: a candidate's roi x roi block is held in LDS (160 KB).
constexpr int kMaxBeadFinderROI = 200;

// BeadFinder::Find on the w x h float frame `img` with the roi x roi sample whose
// top-left corner is (sx, sy): the positions (top-left corners) of the beads
// found, in the reference's order.  img_distance is the minimum bead distance
// relative to roi, similarity the correlation threshold relative to the
// maximum.  When sample_out is not null the sample is copied into it.
// Throws std::runtime_error on bad arguments.
std::vector<ROIPosition> find_beads(const float* img, int w, int h, int sx, int sy, int roi, float img_distance,
                                    float similarity, float* sample_out);

}  // namespace qtrk
