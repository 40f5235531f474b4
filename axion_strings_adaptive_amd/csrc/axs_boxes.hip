// axs_boxes.hip -- AMR box levels (work-list launches).  Not built yet:
// every entry point reports AXS_ERR_UNSUPPORTED.
#include "axs_internal.hpp"

extern "C" {
int axs_boxlevel_create(axs_ctx *, int, const int *, const int *, const int *,
                        int, double, int, axs_boxlevel **out)
{
    if (out)
        *out = nullptr;
    return axs::set_error(AXS_ERR_UNSUPPORTED, "box levels not built");
}
int axs_boxlevel_destroy(axs_boxlevel *) { return AXS_OK; }
int64_t axs_boxlevel_num_cells(axs_boxlevel *) { return 0; }
int axs_boxlevel_upload_dense(axs_boxlevel *, const double *)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "box levels not built");
}
int axs_boxlevel_download_dense(axs_boxlevel *, double *)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "box levels not built");
}
int axs_boxlevel_fill_ghosts(axs_boxlevel *, const axs_level *, double)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "box levels not built");
}
int axs_boxlevel_rk4_step(axs_boxlevel *, const axs_rhs_coeffs *, double)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "box levels not built");
}
int axs_boxlevel_tag(axs_boxlevel *, const axs_tagger_params *, uint8_t *,
                     int64_t *)
{
    return axs::set_error(AXS_ERR_UNSUPPORTED, "box levels not built");
}
}
