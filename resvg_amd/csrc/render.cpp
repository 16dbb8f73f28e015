// resvg::render / resvg_render on the device (crates/resvg/src/lib.rs:34-43,
// crates/c-api/lib.rs:874-897).
#include "rvk_internal.h"
#include "host_geom.h"

extern "C" {

rvk_status rvk_render(rvk_context *ctx, const rvk_scene *scene, rvk_transform transform, uint32_t width,
                      uint32_t height, uint8_t *dev_pixmap) {
    if (!ctx || !scene || !dev_pixmap || !rvk::valid_dims(width, height)) return RVK_ERR_INVALID_ARGUMENT;
    (void)transform;
    return RVK_ERR_UNSUPPORTED;
}

rvk_status rvk_render_host(rvk_context *ctx, const rvk_scene *scene, rvk_transform transform, uint32_t width,
                           uint32_t height, uint8_t *host_pixmap) {
    if (!ctx || !scene || !host_pixmap || !rvk::valid_dims(width, height)) return RVK_ERR_INVALID_ARGUMENT;
    (void)transform;
    return RVK_ERR_UNSUPPORTED;
}

}  // extern "C"
