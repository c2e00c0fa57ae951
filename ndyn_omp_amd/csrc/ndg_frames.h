// ndg_frames.h -- host-side formatter of trajectory frames (the reference's write_model).
#pragma once

#include <string>

#include "../../include/ndyn_gpu.h"

namespace ndg {

// One frame in caller order: pos(3,np) scaled, vel(3,np), per-atom energy(np), 1-based types.
struct FrameData {
  int format = NDG_FRAME_DUMP;
  int np = 0;
  const double *pos = nullptr, *vel = nullptr, *pene = nullptr;
  const int *typ = nullptr;
  const char *names = nullptr;   // ntypes names of name_len characters, blank padded
  int name_len = 0;
  double box[3] = {0, 0, 0};
  int step = 0;
};

void format_frame(const FrameData &f, std::string &out, int nthreads);
int format_threads();

}  // namespace ndg
