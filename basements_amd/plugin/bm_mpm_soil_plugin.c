/* bm_mpm_soil_plugin.c -- the MI355X soil substep as a plugin of the
 * reference engine's C plugin interface
 * (include/basements/engine/plugins/plugin_capi.h:34-102).
 *
 *   on_load       creates a fast-mode MPM handle over the soil domain;
 *   on_pre_step   advances the soil by one substep of dt -- coupled to the
 *                 tool body (the order of mpm_rigid_coupler.h:209-211) once
 *                 the engine has given one, plain otherwise; nothing waits;
 *   on_post_step  waits for the device, so the engine reads settled state;
 *   on_unload     destroys the handle.
 *
 * The engine fills the scene through bm_soil_plugin_handle()
 * (bm_mpm_add_particles / bm_mpm_upload_particles), hands the tool over with
 * bm_soil_plugin_set_body, and reads it back with bm_mpm_get_body.  The
 * callbacks return nothing, so the first failing status is kept for
 * bm_soil_plugin_status().
 *
 * Domain defaults below (the wheel bed: 256^3 nodes at dx = 1/120 m) are
 * compile-time settings: -DBM_SOIL_DX=... -DBM_SOIL_NODES=...
 * -DBM_SOIL_MAX_PARTICLES=....
 *
 * Build (C99):
 *   cc -std=c99 -shared -fPIC -I<reference>/include -I<this repository>/include \
 *      bm_mpm_soil_plugin.c -L<directory of libbm_mpm.so> -lbm_mpm \
 *      -o libbm_mpm_soil_plugin.so
 */
#include <stddef.h>
#include <string.h>

#define BASEMENTS_PLUGIN_EXPORT
#include "basements/engine/plugins/plugin_capi.h"
#include "bm_mpm.h"

#ifndef BM_SOIL_DX
#define BM_SOIL_DX (1.0f / 120.0f)
#endif
#ifndef BM_SOIL_NODES
#define BM_SOIL_NODES 256
#endif
#ifndef BM_SOIL_MAX_PARTICLES
#define BM_SOIL_MAX_PARTICLES (1 << 23)
#endif
#ifndef BM_SOIL_COUPLER
#define BM_SOIL_COUPLER BM_COUPLER_M1_PIC_RIGID
#endif

/* Built as C++ as well: the exports (the register function included, which the
 * header's macro defines outside its own extern "C" block) keep C names. */
#ifdef __cplusplus
extern "C" {
#endif

static bm_mpm_handle g_soil = NULL;
static int g_has_body = 0;
static int g_status = BM_OK;

static void keep_first(int rc) {
    if (rc != BM_OK && g_status == BM_OK) g_status = rc;
}

BASEMENTS_PLUGIN_API bm_mpm_handle bm_soil_plugin_handle(void) { return g_soil; }

BASEMENTS_PLUGIN_API int bm_soil_plugin_status(void) { return g_status; }

BASEMENTS_PLUGIN_API int bm_soil_plugin_set_body(const bm_body* body) {
    if (!g_soil) return BM_ERR_STATE;
    const int rc = bm_mpm_set_body(g_soil, body);
    if (rc == BM_OK) g_has_body = 1;
    return rc;
}

static BasementsPluginInfo info(void) {
    BasementsPluginInfo i = {"bm_mpm_soil", "1", "", "MLS-MPM soil substep on MI355X (HIP, gfx950)",
                             BASEMENTS_PLUGIN_API_VERSION};
    return i;
}

static int load(BasementsEngineContext ctx) {
    (void)ctx;
    bm_mpm_desc d;
    memset(&d, 0, sizeof d);
    d.dx = BM_SOIL_DX;
    d.grid_dims[0] = d.grid_dims[1] = d.grid_dims[2] = BM_SOIL_NODES;
    d.max_particles = BM_SOIL_MAX_PARTICLES;
    d.mode = BM_MODE_FAST;
    d.stream = NULL;
    g_status = BM_OK;
    g_has_body = 0;
    const int rc = bm_mpm_create(&d, &g_soil);
    if (rc != BM_OK) {
        g_soil = NULL;
        keep_first(rc);
        return -1;
    }
    return 0;
}

static void unload(void) {
    if (g_soil) bm_mpm_destroy(g_soil);
    g_soil = NULL;
    g_has_body = 0;
}

static void pre_step(float dt) {
    if (!g_soil || g_status != BM_OK) return;
    keep_first(g_has_body ? bm_mpm_step_coupled(g_soil, BM_SOIL_COUPLER, dt, 1) : bm_mpm_step(g_soil, dt));
}

static void post_step(float dt) {
    (void)dt;
    if (g_soil) keep_first(bm_mpm_synchronize(g_soil));
}

BASEMENTS_PLUGIN_DEFINE_FULL(info, load, unload, pre_step, post_step, NULL)

#ifdef __cplusplus
}
#endif
