/* Private declarations of the C task layer (host/tasks.c, host/pipeline.c). */
#ifndef UFOMI_TASKS_INTERNAL_H
#define UFOMI_TASKS_INTERNAL_H

#include "ufo_tasks.h"

#define UFOMI_MAX_PROPS 40

typedef enum { PT_UINT, PT_DOUBLE, PT_FLOAT, PT_BOOL, PT_ENUM, PT_DARRAY, PT_STRING } PropType;

typedef struct {
    const char *name;
    PropType type;
    double min, max, dflt;      /* numeric range and default (enums: default index) */
    const char *const *nicks;   /* enum value nicks */
    int n_default;              /* arrays: number of default elements (all = dflt) */
    const char *sdflt;          /* strings: default value */
} PropSpec;

typedef struct {
    double num;
    double *arr;
    int n;
    char *str;
} PropValue;

typedef struct TaskClass {
    const char *name;
    const PropSpec *props;
    int n_props;
    unsigned mode;
    size_t priv_size;
    int (*setup)(UfomiTask *, UfomiError *);
    int (*requisition)(UfomiTask *, UfomiBuffer **, UfomiRequisition *, UfomiError *);
    int (*process)(UfomiTask *, UfomiBuffer **, UfomiBuffer *, UfomiRequisition *);
    int (*generate)(UfomiTask *, UfomiBuffer *, UfomiRequisition *);
    void (*finalize)(UfomiTask *);
} TaskClass;

struct UfomiTask {
    const TaskClass *cls;
    PropValue v[UFOMI_MAX_PROPS];
    const char *names[UFOMI_MAX_PROPS + 1];
    ufomi_ctx *ctx;
    void *priv;
    unsigned num_processed;
    char err[512];
    int err_status;
    int ready;                   /* setup succeeded */
};

int task_error(UfomiError *e, int code, int status, const char *fmt, ...);
/* Records a ufomi status + message for ufomi_last_error() and returns status. */
int ufomi_status_fail(int status, const char *fmt, ...);

/* Helpers the pipeline uses to fuse fft ! filter ! ifft and batch backproject. */
int fft_chain_fusable(const UfomiTask *fft, const UfomiTask *ifft, const UfomiBuffer *in);
int filter_task_coefficients(UfomiTask *t, size_t width, UfomiError *e);
const float *filter_task_device_coefficients(const UfomiTask *t);
void backproject_task_params(const UfomiTask *t, ufomi_bp_params *bp);
int gbp_task_store_type(const UfomiTask *t);
/* Output items (volume slices, or 1), element size and dtype of a generic reductor stage. */
int reductor_output_items(const UfomiTask *t, const UfomiRequisition *in, int *items, size_t *elem_size, int *dtype);
/* read: the next image into host memory (req from ufomi_task_get_requisition);
 * 1 = produced, 0 = no more images or an error (then t->err_status is set). */
int read_task_generate_host(UfomiTask *t, float *dst, const UfomiRequisition *req);
/* write: one item from host memory; UFOMI_OK or a status with t->err set. */
int write_task_write_host(UfomiTask *t, const float *src, const UfomiRequisition *req);
/* flat-field-correct over a batch: input sizes checked, then one launch. */
int ffc_check_inputs(const UfomiTask *t, const UfomiRequisition *data, const UfomiRequisition *dark,
                     const UfomiRequisition *flat, UfomiError *e);
int ffc_task_run(UfomiTask *t, ufomi_ctx *ctx, const float *data, const UfomiRequisition *r, int n,
                 const float *dark, int n_dark, const float *flat, int n_flat, float *out);
/* write: closes the open file (the end of a stream). */
void write_task_flush(UfomiTask *t);

#endif
