// controller.hpp -- the run driver: velocity preparation, shot loop, image
// stacking and finishing, output files.  Mirrors RTMAcousticProcess::rtm
// (rtm-3d/src/process/RTMAcousticProcess.cpp:8-60) and RTMController /
// RTMAcousticController (rtm-3d/src/controller/RTMController.cpp:27-166,
// RTMAcousticController.cpp:9-269) for a single process on one device.
#pragma once
#include <string>

#include "param.hpp"

namespace rtm3d {

// Runs modeling or migration as the parameters ask; throws on error.
void run(RTMParam &p, int device);

// Reads the JSON file and runs; prints errors as the reference does and
// returns the process exit status (0 or 1).
int run_json(const std::string &path, int device);

}  // namespace rtm3d
