// Discrete adjoint (vector-Jacobian product) of the Euler residual.
#include "solver.h"

namespace fvk {

int adjoint_vjp(fvk_solver* s, const double* lambda, double* U_bar, double* X_bar) {
    (void)s; (void)lambda; (void)U_bar; (void)X_bar;
    return set_error(FVK_ERR_UNSUPPORTED, "residual VJP not built");
}

}  // namespace fvk
