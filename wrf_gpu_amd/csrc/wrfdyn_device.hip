// The dycore's device code in one translation unit: the RK-stage setup
// kernels, the acoustic kernels, and the persistent ensemble step, which
// runs the bodies of the other two from one launch.
#include "wrfdyn_stage.hip"
#include "wrfdyn_acoustic.hip"
#include "wrfdyn_persistent.h"
