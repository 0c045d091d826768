// mixed.h -- launch sequence of the FP32/FP64 mixed method.
#pragma once
#include "engine_host.h"

namespace shp {
int enqueue_mixed(Engine &E);
}
