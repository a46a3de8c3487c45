/*
 * juicer_dir_boundary_api.h — the spatial-DIR boundary import of the render
 * library (replaces native/juicer_dir_boundary_api.h; records, constants and
 * signatures are the reference's, for a host that links it).
 *
 * The four functions are defined by the host (the reference's Rust plug-in:
 * rust/film-juicer-plugin/src/dir_boundary_bridge.rs:423-470). They solve
 * and certify the reflected boundary of one recursive DIR component on a
 * frame: describe returns its identity (or the reason it cannot be
 * certified), acquire/view/release hand out its weights and terminal
 * matrices. The reference's native library needs those weights for its
 * recursive line filters (src/JuicerDirBoundary.cpp:140-181). This library
 * applies the DIR filter in the DCT domain of the reflected extension, which
 * needs no boundary weights; when the process defines
 * fj_dir_boundary_describe, the library certifies each recursive component
 * of a frame through the host: describe, and for an identity not certified
 * before in the process, acquire, view (its envelope is checked) and
 * release. A frame whose boundary the host cannot certify is refused with
 * the host's status and diagnostic, as the reference refuses it; a host
 * answer that breaks the call's contract is refused as an invalid view.
 *
 * Calls are synchronous on the calling thread, keep no input pointer and
 * clear each output before writing it; diagnostics follow FjErrorBuffer's
 * rules.
 */
#ifndef FJ_JUICER_DIR_BOUNDARY_API_H
#define FJ_JUICER_DIR_BOUNDARY_API_H

#include "juicer_cuda_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FJ_DIR_BOUNDARY_NONE UINT32_C(0)
#define FJ_DIR_BOUNDARY_RAW UINT32_C(1)
#define FJ_DIR_BOUNDARY_REQUEST UINT32_C(2)
#define FJ_DIR_BOUNDARY_AXIS_SIZE UINT32_C(3)
#define FJ_DIR_BOUNDARY_POLICY UINT32_C(4)
#define FJ_DIR_BOUNDARY_CAUSAL_SINGULAR UINT32_C(5)
#define FJ_DIR_BOUNDARY_NONFINITE_WEIGHTS UINT32_C(6)
#define FJ_DIR_BOUNDARY_CAUSAL_CERTIFICATION UINT32_C(7)
#define FJ_DIR_BOUNDARY_TAIL_ESCAPE UINT32_C(8)
#define FJ_DIR_BOUNDARY_TERMINAL_SINGULAR UINT32_C(9)
#define FJ_DIR_BOUNDARY_TERMINAL_CERTIFICATION UINT32_C(10)
#define FJ_DIR_BOUNDARY_STORAGE_SIZE UINT32_C(11)
#define FJ_DIR_BOUNDARY_ALLOCATION UINT32_C(12)
#define FJ_DIR_BOUNDARY_PANIC UINT32_C(13)
#define FJ_DIR_BOUNDARY_NO_AXIS UINT32_C(0)
#define FJ_DIR_BOUNDARY_HORIZONTAL UINT32_C(1)
#define FJ_DIR_BOUNDARY_VERTICAL UINT32_C(2)
#define FJ_DIR_BOUNDARY_NO_ROW UINT32_C(3)

/* One recursive component on a width x height frame. */
typedef struct FjDirBoundaryInput {
    FjDirGaussian component;
    int32_t width, height;
} FjDirBoundaryInput;

/* Row-major f64 weights: 3 x horizontal_length, then 3 x vertical_length;
 * terminal matrices use their terminal size as stride. */
typedef struct FjDirBoundaryView {
    const double* weights;
    size_t count;
    uint64_t identity;
    int32_t width, height, horizontal_length, vertical_length;
    int32_t horizontal_terminal_size, vertical_terminal_size;
    double horizontal_terminal[9], vertical_terminal[9];
} FjDirBoundaryView;

typedef struct FjDirBoundaryFailure {
    uint32_t reason, axis, row;
    int32_t length;
    float sigma_px;
    double sum, tolerance;
} FjDirBoundaryFailure;

typedef struct FjDirBoundaryOwner FjDirBoundaryOwner;

/* Imported (host-defined; resolved at run time, optional). */
FjStatus fj_dir_boundary_describe(const FjDirBoundaryInput*, uint64_t*, FjDirBoundaryFailure*, FjErrorBuffer*);
FjStatus fj_dir_boundary_acquire(const FjDirBoundaryInput*, FjDirBoundaryOwner**, FjDirBoundaryFailure*,
                                 FjErrorBuffer*);
FjStatus fj_dir_boundary_view(const FjDirBoundaryOwner*, FjDirBoundaryView*, FjDirBoundaryFailure*, FjErrorBuffer*);
FjStatus fj_dir_boundary_release(FjDirBoundaryOwner*, FjErrorBuffer*);

#ifdef __cplusplus
}
#endif

#endif
