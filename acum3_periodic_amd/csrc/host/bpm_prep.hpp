// Observer-independent per (surface, section, step) state of the BPM model, computed on the host
// in FP64 before the broadband kernels run (findBBnoise, bb_core.cu:60-440 and the section frame
// of bb_core.cu:473-513).
#pragma once
#include <string>
#include <vector>
#include "../../../include/acum.h"

namespace acum {

struct BpmSource {
  double xyzt[3];     // trailing-edge point, m
  double tr[9];       // rows: chordwise, spanwise, normal unit vectors (bb_core.cu:473-513)
  double u;           // chordwise flow speed, m/s (sign kept; core() takes |u|)
  double alpha;       // |alpha_eff - alpha(cl=0)|, degrees (bb_core.cu:340-341)
  double bl[4];       // delta_s, delta_p, delta*_s, delta*_p after the < 0 clamp (bb_core.cu:599-602)
  double c, lsect;    // chord and section length, m
  double R;           // span of the surface, m (tip aspect ratio, bb_core.cu:978)
  int trip, tip;      // LBL-VS disabled when tripped; tip noise on the last section
};

// Fills src[(s_offset + is) * nTime + it] for every surface; returns ACUM_OK or an error code
// with a message in *err.
int bpm_prepare(const acum_bpm_case* c, std::vector<BpmSource>* src, std::string* err);

}  // namespace acum
