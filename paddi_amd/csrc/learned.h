// Learned r proposal: features, MLP inference and training (see learned.cpp).
#pragma once
#include <stdint.h>

namespace paddi {
bool learned_available();
}  // namespace paddi
