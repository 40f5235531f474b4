// axs_adjoint.hip -- adjoint of the RHS and RK4 step.  Not built yet: every
// entry point reports AXS_ERR_UNSUPPORTED.
#include "axs_internal.hpp"

extern "C" {
int axs_adjoint_grad_energy(axs_level *, const axs_background *, double,
                            double, double, int, int, double *, double *,
                            double *)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "adjoint not built");
}
int axs_forward_grad_energy(axs_level *, const axs_background *, double,
                            double, double, int, double *)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "adjoint not built");
}
}
