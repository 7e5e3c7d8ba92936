// adjoint.hip -- gradient of a Mueller-matrix loss with respect to
// (Re m, Im m, ka) through one adjoint solve and a matrix-free contraction of
// dZ/dp.  (Filled in by the adjoint milestone.)
#include "../bem_internal.h"

int adjoint_launch(bem_ctx* c, const double* dL_dM, double* grad) {
    (void)dL_dM;
    (void)grad;
    return bem_fail(c, BEM_ERR_STATE, "mueller_grad: adjoint not built yet");
}
