/* fj_host_shim.c — plays the host side of the render library's optional
 * imports for the tests (test infrastructure): fj_diffusion_components_prepare
 * and fj_dir_boundary_{describe,acquire,view,release}, as the reference's Rust
 * plug-in defines them. Loaded with RTLD_GLOBAL before the library resolves
 * them; fjh_shim_mode selects the behaviour:
 *   components: 0 delegate to the FP64 oracle's expansion, 1 fail
 *     (preparation, field radius), 2 delegate with every lambda doubled;
 *   boundary: 0 certify (describe + acquire + view + release), 1 refuse at
 *     acquire (terminal tail escape), 2 break the describe envelope. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fj/juicer_diffusion_api.h"
#include "fj/juicer_dir_boundary_api.h"
#include "../../oracle/fj_oracle.h"

static int g_components_mode = 0, g_boundary_mode = 0;
static int g_components_calls = 0, g_describe_calls = 0, g_acquire_calls = 0, g_release_calls = 0;

void fjh_shim_mode(int components, int boundary) {
    g_components_mode = components;
    g_boundary_mode = boundary;
}

void fjh_shim_counts(int out[4]) {
    out[0] = g_components_calls;
    out[1] = g_describe_calls;
    out[2] = g_acquire_calls;
    out[3] = g_release_calls;
}

static void put(FjErrorBuffer* e, const char* text) {
    if (!e) return;
    if (!e->capacity || !e->data) {
        e->length = 0;
        return;
    }
    size_t n = strlen(text);
    if (n > e->capacity - 1) n = e->capacity - 1;
    memcpy(e->data, text, n);
    e->data[n] = '\0';
    e->length = n;
}

static FjStatus status(uint32_t category) {
    FjStatus s = {category, FJ_API_NONE, 0};
    return s;
}

FjStatus fj_diffusion_components_prepare(const FjDiffusionSample* in, FjDiffusionComponents* out,
                                         FjDiffusionFailure* failure, FjErrorBuffer* error) {
    ++g_components_calls;
    memset(out, 0, sizeof(*out));
    memset(failure, 0, sizeof(*failure));
    put(error, "");
    if (g_components_mode == 1) {
        failure->reason = FJ_DIFFUSION_REASON_PREPARATION;
        failure->field = FJ_DIFFUSION_FIELD_RADIUS_PIXELS;
        put(error, "host refused the diffusion sample field=radius_pixels");
        return status(FJ_STATUS_PREPARATION_FAILURE);
    }
    FjDiffusion d;
    memset(&d, 0, sizeof(d));
    d.family = in->family;
    memcpy(d.group_weights, in->group_weights, sizeof(d.group_weights));
    memcpy(d.group_lambda_um, in->group_lambda_um, sizeof(d.group_lambda_um));
    d.warmth = in->warmth;
    d.spatial_scale = in->spatial_scale;
    d.pixel_size_um = in->pixel_size_um;
    d.radius_px = in->radius_px;
    d.sample_hash = in->hash;
    double lam[9], w[3][9];
    if (fjo_diffusion_components(&d, lam, w) != 0) {
        failure->reason = FJ_DIFFUSION_REASON_PREPARATION;
        failure->field = FJ_DIFFUSION_FIELD_COMPONENT_EXPANSION;
        put(error, "host could not expand the sample");
        return status(FJ_STATUS_PREPARATION_FAILURE);
    }
    for (int k = 0; k < 9; ++k) {
        out->lambda_px[k] = g_components_mode == 2 ? 2.0 * lam[k] : lam[k];
        out->red_weights[k] = w[0][k];
        out->green_weights[k] = w[1][k];
        out->blue_weights[k] = w[2][k];
    }
    return status(FJ_STATUS_SUCCESS);
}

struct FjDirBoundaryOwner {
    FjDirBoundaryInput input;
    double* weights;
    size_t count;
};

static uint64_t identity_of(const FjDirBoundaryInput* in) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* p = (const unsigned char*)in;
    for (size_t i = 0; i < sizeof(*in); ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h ? h : 1;
}

static void no_failure(FjDirBoundaryFailure* f) {
    memset(f, 0, sizeof(*f));
    f->row = FJ_DIR_BOUNDARY_NO_ROW;
}

FjStatus fj_dir_boundary_describe(const FjDirBoundaryInput* in, uint64_t* key, FjDirBoundaryFailure* failure,
                                  FjErrorBuffer* error) {
    ++g_describe_calls;
    *key = 0;
    no_failure(failure);
    put(error, "");
    if (g_boundary_mode == 2) {
        *key = identity_of(in);
        put(error, "not the reference's text");
        return status(FJ_STATUS_SUCCESS);
    }
    *key = identity_of(in) ^ (uint64_t)g_boundary_mode;
    return status(FJ_STATUS_SUCCESS);
}

FjStatus fj_dir_boundary_acquire(const FjDirBoundaryInput* in, FjDirBoundaryOwner** owner,
                                 FjDirBoundaryFailure* failure, FjErrorBuffer* error) {
    ++g_acquire_calls;
    *owner = NULL;
    no_failure(failure);
    put(error, "");
    if (g_boundary_mode == 1) {
        failure->reason = FJ_DIR_BOUNDARY_TAIL_ESCAPE;
        failure->axis = FJ_DIR_BOUNDARY_VERTICAL;
        failure->length = in->height;
        put(error, "spatial DIR reflected terminal system escaped its tail");
        return status(FJ_STATUS_PREPARATION_FAILURE);
    }
    FjDirBoundaryOwner* o = (FjDirBoundaryOwner*)calloc(1, sizeof(*o));
    o->input = *in;
    const int h = in->width < 3 ? in->width : 3, v = in->height < 3 ? in->height : 3;
    o->count = 3 * (size_t)(h + v);
    o->weights = (double*)calloc(o->count, sizeof(double));
    *owner = o;
    return status(FJ_STATUS_SUCCESS);
}

FjStatus fj_dir_boundary_view(const FjDirBoundaryOwner* o, FjDirBoundaryView* view, FjDirBoundaryFailure* failure,
                              FjErrorBuffer* error) {
    memset(view, 0, sizeof(*view));
    no_failure(failure);
    put(error, "");
    const int h = o->input.width < 3 ? o->input.width : 3, v = o->input.height < 3 ? o->input.height : 3;
    view->weights = o->weights;
    view->count = o->count;
    view->identity = identity_of(&o->input) ^ (uint64_t)g_boundary_mode;
    view->width = o->input.width;
    view->height = o->input.height;
    view->horizontal_length = h;
    view->vertical_length = v;
    view->horizontal_terminal_size = h;
    view->vertical_terminal_size = v;
    return status(FJ_STATUS_SUCCESS);
}

FjStatus fj_dir_boundary_release(FjDirBoundaryOwner* o, FjErrorBuffer* error) {
    ++g_release_calls;
    put(error, "");
    if (o) {
        free(o->weights);
        free(o);
    }
    return status(FJ_STATUS_SUCCESS);
}
