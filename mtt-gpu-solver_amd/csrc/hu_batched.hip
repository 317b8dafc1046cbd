// Batched heads-up path (filled in below the plain path).
#include "solver.h"

namespace mtt {

struct BatchedState {};

int batched_create(SolverImpl& s) {
    (void)s;
    set_error("batched mode not built yet");
    return MTT_ERR_STATE;
}
int batched_run(SolverImpl& s, int iters) { (void)s; (void)iters; return MTT_ERR_STATE; }
int batched_strategy_sums(SolverImpl& s, float* d) { (void)s; (void)d; return MTT_ERR_STATE; }
void batched_destroy(BatchedState* b) { delete b; }

}  // namespace mtt
