/* Private declarations of the C task layer (host/tasks.c, host/pipeline.c). */
#ifndef UFOMI_TASKS_INTERNAL_H
#define UFOMI_TASKS_INTERNAL_H

#include "ufo_tasks.h"

#define UFOMI_MAX_PROPS 40

typedef enum { PT_UINT, PT_DOUBLE, PT_FLOAT, PT_BOOL, PT_ENUM, PT_DARRAY } PropType;

typedef struct {
    const char *name;
    PropType type;
    double min, max, dflt;      /* numeric range and default (enums: default index) */
    const char *const *nicks;   /* enum value nicks */
    int n_default;              /* arrays: number of default elements (all = dflt) */
} PropSpec;

typedef struct {
    double num;
    double *arr;
    int n;
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
};

int task_error(UfomiError *e, int code, int status, const char *fmt, ...);
/* Records a ufomi status + message for ufomi_last_error() and returns status. */
int ufomi_status_fail(int status, const char *fmt, ...);

/* Helpers the pipeline uses to fuse fft ! filter ! ifft and batch backproject. */
int filter_task_coefficients(UfomiTask *t, size_t width, UfomiError *e);
const float *filter_task_device_coefficients(const UfomiTask *t);
void backproject_task_params(const UfomiTask *t, ufomi_bp_params *bp);
int gbp_task_store_type(const UfomiTask *t);

#endif
