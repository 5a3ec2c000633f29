/*
 * ufo_tasks.h -- the reference's task interface for the FBP hot path, in C,
 * over the ufomi C ABI.
 *
 * ufo-filters' plugins are GObjects implementing UfoTaskIface
 * (src/ufo-backproject-task.c:67-69, 327-336; ufo-general-backproject-task.c:2059-2068)
 * and configured through GObject properties.  ufo-core (the scheduler, buffers,
 * plugin manager) is not part of ufo-filters and is absent here, so this
 * header restates the interface in plain C with the same task names, property
 * names, types, ranges, defaults, requisition rules and error conditions, and
 * prefixes every symbol with ufomi_ so it can live in the same process as
 * ufo-core.  A GObject plugin that wants to be a literal drop-in keeps its
 * GType and property installation and forwards its UfoTaskIface vfuncs to
 * these functions.
 *
 * Buffers hold device memory of the task's ufomi context.
 */
#ifndef UFO_TASKS_H
#define UFO_TASKS_H

#include <stddef.h>

#include "ufomi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* UfoRequisition: n_dims and dims[3], dims[0] fastest (ufo-core). */
typedef struct {
    unsigned n_dims;
    size_t dims[3];
} UfomiRequisition;

/* UfoBufferLayout */
typedef enum { UFOMI_LAYOUT_REAL = 0, UFOMI_LAYOUT_COMPLEX_INTERLEAVED = 1 } UfomiBufferLayout;

/* A device buffer: float32 unless dtype says otherwise (filtered-sinogram storage). */
typedef struct {
    void *data;
    UfomiRequisition req;
    UfomiBufferLayout layout;
    int dtype;          /* ufomi_dtype for float data; general-backproject output uses its store type */
    int owned;          /* allocated by ufomi_buffer_new */
} UfomiBuffer;

/* UFO_TASK_ERROR codes of ufo-core, plus the ufomi status that caused them. */
typedef enum {
    UFOMI_TASK_ERROR_SETUP = 0,
    UFOMI_TASK_ERROR_GET_REQUISITION = 1,
    UFOMI_TASK_ERROR_PROCESS = 2
} UfomiTaskErrorCode;

typedef struct {
    int code;        /* UfomiTaskErrorCode */
    int status;      /* ufomi_status */
    char message[512];
} UfomiError;

/* UfoTaskMode bits */
enum { UFOMI_TASK_MODE_PROCESSOR = 1 << 0, UFOMI_TASK_MODE_REDUCTOR = 1 << 1, UFOMI_TASK_MODE_GENERATOR = 1 << 2,
       UFOMI_TASK_MODE_SINK = 1 << 3, UFOMI_TASK_MODE_GPU = 1 << 4, UFOMI_TASK_MODE_CPU = 1 << 5 };

typedef struct UfomiTask UfomiTask;

/* Task names: "fft", "filter", "ifft", "backproject", "general-backproject",
 * "cone-beam-projection-weight" (libufofilter<name>.so in the reference), and
 * "transpose-projections", "stacked-backproject", "lamino-backproject"
 * (src/ufo-lamino-backproject-task.c: a reductor emitting one 3-D volume),
 * "flat-field-correct" (three inputs), "average", "center-of-rotation", and
 * the file tasks "read" (src/ufo-read-task.c: a generator, TIFF / BigTIFF,
 * EDF and raw files) and "write" (src/ufo-write-task.c: a sink, TIFF / BigTIFF
 * and raw).  Returns NULL for an unknown name.
 *
 * Generators take `inputs` NULL in get_requisition; a generator with nothing
 * left leaves the requisition as it was and generate returns 0.  Sinks take
 * `output` NULL in process. */
UfomiTask *ufomi_task_new(const char *name);
void ufomi_task_free(UfomiTask *task);
const char *ufomi_task_name(const UfomiTask *task);

/* Properties, by their GObject names.  set_property parses the ufo-launch
 * text form ("1024", "ramp-fromreal", "true", "-2500", arrays "1.5,2,3");
 * out-of-range values are rejected and leave the property unchanged, as
 * g_object_set does with an invalid value. */
int ufomi_task_set_property(UfomiTask *task, const char *name, const char *value);
int ufomi_task_get_property(const UfomiTask *task, const char *name, char *buf, size_t len);
/* Names of all properties, NULL terminated. */
const char *const *ufomi_task_list_properties(const UfomiTask *task);

/* UfoTaskIface.  Tasks of UFOMI_TASK_MODE_CPU (read, write) may be set up
 * with a NULL context when they only see host memory (the *_host calls). */
int ufomi_task_setup(UfomiTask *task, ufomi_ctx *ctx, UfomiError *error);
int ufomi_task_get_requisition(UfomiTask *task, UfomiBuffer **inputs, UfomiRequisition *requisition,
                               UfomiError *error);
unsigned ufomi_task_get_num_inputs(const UfomiTask *task);
unsigned ufomi_task_get_num_dimensions(const UfomiTask *task, unsigned input);
unsigned ufomi_task_get_mode(const UfomiTask *task);
/* Returns 1 (TRUE) like the reference's process/generate, 0 on failure or end. */
int ufomi_task_process(UfomiTask *task, UfomiBuffer **inputs, UfomiBuffer *output, UfomiRequisition *requisition);
int ufomi_task_generate(UfomiTask *task, UfomiBuffer *output, UfomiRequisition *requisition);
/* CPU-mode tasks on host float32 memory, as ufo-core hands a CPU task its
 * buffers' host arrays (ufo_buffer_get_host_array): read's generate
 * (src/ufo-read-task.c:341-363) into `output` of `requisition`; the process
 * of write (src/ufo-write-task.c:301-392) and of center-of-rotation
 * (src/ufo-center-of-rotation-task.c:111-166) on `input`.  Return 1 / 0 as
 * above. */
int ufomi_task_generate_host(UfomiTask *task, float *output, const UfomiRequisition *requisition);
int ufomi_task_process_host(UfomiTask *task, const float *input, const UfomiRequisition *input_requisition);
/* The output layout and storage a task writes (ufo_buffer_set_layout). */
UfomiBufferLayout ufomi_task_output_layout(const UfomiTask *task);
int ufomi_task_output_dtype(const UfomiTask *task);
/* Message of the last failed process/generate call of this task. */
const char *ufomi_task_last_error(const UfomiTask *task);

/* Buffers on a context's device. */
int ufomi_buffer_new(ufomi_ctx *ctx, const UfomiRequisition *req, int dtype, UfomiBuffer *out);
void ufomi_buffer_release(ufomi_ctx *ctx, UfomiBuffer *buf);
size_t ufomi_requisition_elements(const UfomiRequisition *req);

/*
 * A linear pipeline in ufo-launch syntax, e.g.
 *   "fft dimensions=1 size-x=1024 ! filter ! ifft crop-width=512 ! backproject axis-pos=256"
 * run over n_items input items of `in_req` floats each (device memory,
 * contiguous).  Each stage is a task above; a processor chain
 * fft ! filter ! ifft runs as the fused ramp-filter kernel and consecutive
 * items of a backproject stage run as one multi-slice launch.  General
 * backprojection consumes all items (projections) and emits its slices.
 * The output items are written contiguously to d_out (at most out_capacity
 * bytes); *n_out and *out_req describe them.
 *
 * A first stage `read` produces the input items from files (d_in, in_req and
 * n_items are then ignored and may be NULL / 0); a last stage `write` writes
 * the output items to files (d_out may then be NULL; *n_out and *out_req
 * still describe what was written).
 */
typedef struct UfomiPipeline UfomiPipeline;
UfomiPipeline *ufomi_pipeline_new(const char *description, UfomiError *error);
void ufomi_pipeline_free(UfomiPipeline *pipeline);
int ufomi_pipeline_num_tasks(const UfomiPipeline *pipeline);
UfomiTask *ufomi_pipeline_task(UfomiPipeline *pipeline, int index);
int ufomi_pipeline_run(UfomiPipeline *pipeline, ufomi_ctx *ctx, const float *d_in, const UfomiRequisition *in_req,
                       int n_items, void *d_out, size_t out_capacity, int *n_out, UfomiRequisition *out_req,
                       UfomiError *error);
/* 1 when the last run used the fused fft ! filter ! ifft kernel. */
int ufomi_pipeline_used_fused_filter(const UfomiPipeline *pipeline);

#ifdef __cplusplus
}
#endif

#endif
