// Learned r proposal (placeholder until the model is built).
#include "learned.h"
#include "engine.h"

namespace paddi {

bool learned_available() { return false; }

void Engine::propose_r2(Batch &B, int m, int k, std::vector<double> &out)
{
    (void)m;
    (void)k;
    out.assign((size_t)B.nser, 0.0);
}

}  // namespace paddi
