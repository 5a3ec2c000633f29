/*
 * A linear pipeline in ufo-launch syntax over the C task layer.
 *
 * ufo-core's scheduler pushes one item at a time through every node
 * (docs/examples.rst:90-105).  This runner keeps the per-task semantics
 * (get_requisition before process, reductor generate after the last item) but
 * moves whole batches between stages, which is what a GPU wants:
 *   - `fft ! filter ! ifft` (1-D) becomes one launch of the fused LDS
 *     ramp-filter kernel over every row of every item;
 *   - `backproject` runs all items as one multi-slice launch;
 *   - `transpose-projections` turns all projections into sinograms in one
 *     launch;
 *   - other processors run item by item;
 *   - `general-backproject` takes every item as a projection and emits its
 *     slices through generate;
 *   - `average` and `flat-field-correct` run over the whole batch in one
 *     launch; a leading input group `[read ..., read ... ! average, ...] !
 *     flat-field-correct' runs each branch first (a finished input keeps
 *     feeding its last item);
 *   - a first stage `read` loads every image it generates into one host
 *     buffer and uploads it in one copy; a last stage `write` downloads the
 *     final batch in one copy and writes it item by item.
 */
#include "tasks_internal.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAX_STAGES 16
#define MAX_BRANCHES 4

struct UfomiPipeline {
    UfomiTask *tasks[MAX_STAGES];
    int n;
    int fused;
    struct UfomiPipeline *branches[MAX_BRANCHES];  /* [b0, b1, ...] ! tasks[0]: its inputs */
    int n_branches;
};

static int is_name(const UfomiTask *t, const char *name) { return !strcmp(ufomi_task_name(t), name); }

static char *trim(char *s)
{
    while (isspace((unsigned char) *s))
        s++;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char) e[-1]))
        *--e = '\0';
    return s;
}

/* Stages `name key=value ... ! name ...` appended to pl. */
static int parse_stages(UfomiPipeline *pl, char *text, UfomiError *e)
{
    char *save = NULL;
    for (char *stage = strtok_r(text, "!", &save); stage; stage = strtok_r(NULL, "!", &save)) {
        char *s = trim(stage), *save2 = NULL;
        char *name = strtok_r(s, " \t\n", &save2);
        if (!name || pl->n == MAX_STAGES)
            return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "invalid pipeline stage `%s'", s);
        UfomiTask *t = ufomi_task_new(name);
        if (!t)
            return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "no task named `%s'", name);
        pl->tasks[pl->n++] = t;
        for (char *kv = strtok_r(NULL, " \t\n", &save2); kv; kv = strtok_r(NULL, " \t\n", &save2)) {
            char *eq = strchr(kv, '=');
            if (!eq)
                return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "expected key=value, got `%s'", kv);
            *eq = '\0';
            if (ufomi_task_set_property(t, kv, eq + 1))
                return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "%s", ufomi_last_error());
        }
    }
    return UFOMI_OK;
}

/* ufo-launch syntax: stages joined by `!', optionally led by one group of
 * input branches `[a ! ..., b ! ..., c ! ...] ! task ! ...' feeding the
 * inputs of a multi-input task (each branch starts with `read'). */
UfomiPipeline *ufomi_pipeline_new(const char *description, UfomiError *e)
{
    if (!description) {
        task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "null pipeline");
        return NULL;
    }
    UfomiPipeline *pl = calloc(1, sizeof(UfomiPipeline));
    char *copy = strdup(description);
    char *text = trim(copy);
    if (*text == '[') {
        char *close = strchr(text, ']');
        char *nested = strchr(text + 1, '[');
        if (!close || (nested && nested < close)) {
            task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG,
                       close ? "nested input groups are not supported" : "unterminated `['");
            goto fail;
        }
        *close = '\0';
        char *rest = trim(close + 1);
        if (*rest != '!') {
            task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "an input group must be followed by `! task'");
            goto fail;
        }
        char *save = NULL;
        for (char *br = strtok_r(text + 1, ",", &save); br; br = strtok_r(NULL, ",", &save)) {
            if (pl->n_branches == MAX_BRANCHES) {
                task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "more than %d inputs", MAX_BRANCHES);
                goto fail;
            }
            UfomiPipeline *b = ufomi_pipeline_new(br, e);
            if (!b)
                goto fail;
            pl->branches[pl->n_branches++] = b;
            if (!is_name(b->tasks[0], "read") || is_name(b->tasks[b->n - 1], "write")) {
                task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG,
                           "input %d must start with `read' and must not end with `write'", pl->n_branches - 1);
                goto fail;
            }
        }
        text = rest + 1;
    }
    if (parse_stages(pl, text, e))
        goto fail;
    if (!pl->n) {
        task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "empty pipeline");
        goto fail;
    }
    if (pl->n_branches && (int) ufomi_task_get_num_inputs(pl->tasks[0]) != pl->n_branches) {
        task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "`%s' takes %u inputs, the group gives %d",
                   ufomi_task_name(pl->tasks[0]), ufomi_task_get_num_inputs(pl->tasks[0]), pl->n_branches);
        goto fail;
    }
    if (!pl->n_branches && ufomi_task_get_num_inputs(pl->tasks[0]) > 1) {
        task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "`%s' needs a group of %u inputs `[a, b, ...] ! %s'",
                   ufomi_task_name(pl->tasks[0]), ufomi_task_get_num_inputs(pl->tasks[0]),
                   ufomi_task_name(pl->tasks[0]));
        goto fail;
    }
    free(copy);
    return pl;
fail:
    free(copy);
    ufomi_pipeline_free(pl);
    return NULL;
}

void ufomi_pipeline_free(UfomiPipeline *pl)
{
    if (!pl)
        return;
    for (int i = 0; i < pl->n; i++)
        ufomi_task_free(pl->tasks[i]);
    for (int b = 0; b < pl->n_branches; b++)
        ufomi_pipeline_free(pl->branches[b]);
    free(pl);
}

int ufomi_pipeline_num_tasks(const UfomiPipeline *pl) { return pl ? pl->n : 0; }
UfomiTask *ufomi_pipeline_task(UfomiPipeline *pl, int i) { return pl && i >= 0 && i < pl->n ? pl->tasks[i] : NULL; }
int ufomi_pipeline_used_fused_filter(const UfomiPipeline *pl) { return pl ? pl->fused : 0; }

static int is(const UfomiTask *t, const char *name) { return !strcmp(ufomi_task_name(t), name); }

/* A batch of items between stages: n items of `req`, contiguous. */
typedef struct {
    UfomiBuffer buf;   /* req = one item; data holds n items */
    int n;
    size_t item_bytes;
} Batch;

static size_t elem_size(int dtype) { return dtype == UFOMI_F32 ? 4 : 2; }

static int batch_alloc(ufomi_ctx *ctx, Batch *b, const UfomiRequisition *req, int n, size_t esize, int dtype,
                       UfomiBufferLayout layout)
{
    memset(b, 0, sizeof(*b));
    b->n = n;
    b->item_bytes = ufomi_requisition_elements(req) * esize;
    b->buf.req = *req;
    b->buf.dtype = dtype;
    b->buf.layout = layout;
    int rc = ufomi_device_alloc(ctx, b->item_bytes * (size_t) (n ? n : 1), &b->buf.data);
    b->buf.owned = rc == 0;
    return rc;
}

static UfomiBuffer item_view(const Batch *b, int i)
{
    UfomiBuffer v = b->buf;
    v.data = (char *) b->buf.data + (size_t) i * b->item_bytes;
    v.owned = 0;
    return v;
}

static size_t rows_of(const UfomiRequisition *r)
{
    return (r->n_dims >= 2 ? r->dims[1] : 1) * (r->n_dims == 3 ? r->dims[2] : 1);
}

/* Every image of a `read' stage as one device batch (items of equal size). */
static int read_batch(ufomi_ctx *ctx, UfomiTask *t, Batch *out, UfomiError *e)
{
    float *host = NULL;
    size_t cap = 0, elems = 0;
    UfomiRequisition item;
    memset(&item, 0, sizeof(item));
    memset(out, 0, sizeof(*out));
    int n = 0, rc = UFOMI_OK;
    for (;;) {
        UfomiRequisition req;
        memset(&req, 0, sizeof(req));
        if ((rc = ufomi_task_get_requisition(t, NULL, &req, e)))
            goto done;
        if (!req.n_dims)
            break;
        if (!n) {
            item = req;
            elems = ufomi_requisition_elements(&req);
        } else if (req.dims[0] != item.dims[0] || req.dims[1] != item.dims[1]) {
            rc = task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                            "read: image %d is %zux%zu but the first was %zux%zu; a batch needs one size", n,
                            req.dims[0], req.dims[1], item.dims[0], item.dims[1]);
            goto done;
        }
        if ((size_t) (n + 1) * elems > cap) {
            const size_t want = cap * 2 > (size_t) (n + 1) * elems ? cap * 2 : (size_t) (n + 1) * elems;
            float *grown = realloc(host, want * sizeof(float));
            if (!grown) {
                rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_MEMORY, "read: out of host memory");
                goto done;
            }
            host = grown;
            cap = want;
        }
        t->err_status = 0;
        if (!read_task_generate_host(t, host + (size_t) n * elems, &req)) {
            if (t->err_status) {
                rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, t->err_status, "%s", ufomi_task_last_error(t));
                goto done;
            }
            break;
        }
        t->num_processed++;
        n++;
    }
    if (n) {
        if ((rc = batch_alloc(ctx, out, &item, n, sizeof(float), UFOMI_F32, UFOMI_LAYOUT_REAL))) {
            rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_MEMORY, "out of device memory: %s",
                            ufomi_last_error());
            goto done;
        }
        if ((rc = ufomi_memcpy(ctx, out->buf.data, host, out->item_bytes * (size_t) n, 0))) {
            ufomi_buffer_release(ctx, &out->buf);
            task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "read: %s", ufomi_last_error());
        }
    }
done:
    free(host);
    return rc;
}

/* Every item of the final batch through a `write' stage. */
static int write_batch(ufomi_ctx *ctx, UfomiTask *t, const Batch *cur, UfomiError *e)
{
    if (cur->buf.dtype != UFOMI_F32)
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_UNSUPPORTED, "write: input items must be float32");
    const size_t total = cur->item_bytes * (size_t) cur->n;
    float *host = malloc(total ? total : 1);
    if (!host)
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_MEMORY, "write: out of host memory");
    int rc = total ? ufomi_memcpy(ctx, host, cur->buf.data, total, 1) : UFOMI_OK;
    if (rc) {
        task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "write: %s", ufomi_last_error());
    } else {
        for (int i = 0; i < cur->n; i++) {
            if ((rc = write_task_write_host(t, (const float *) ((const char *) host + (size_t) i * cur->item_bytes),
                                            &cur->buf.req))) {
                task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "%s", ufomi_task_last_error(t));
                break;
            }
            t->num_processed++;
        }
    }
    write_task_flush(t);
    free(host);
    return rc;
}

/* The multi-input first stage over its input batches: flat-field-correct
 * pairs projection i with dark / flat item min(i, n - 1). */
static int run_multi_input(ufomi_ctx *ctx, UfomiTask *t, const Batch *in, Batch *out, UfomiError *e)
{
    for (int k = 0; k < 3; k++)
        if (in[k].buf.dtype != UFOMI_F32)
            return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_UNSUPPORTED,
                              "flat-field-correct: inputs must be float32");
    int rc = ffc_check_inputs(t, &in[0].buf.req, &in[1].buf.req, &in[2].buf.req, e);
    if (rc)
        return rc;
    if ((rc = batch_alloc(ctx, out, &in[0].buf.req, in[0].n, sizeof(float), UFOMI_F32, UFOMI_LAYOUT_REAL)))
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_MEMORY, "out of device memory: %s",
                          ufomi_last_error());
    rc = ffc_task_run(t, ctx, in[0].buf.data, &in[0].buf.req, in[0].n, in[1].buf.data, in[1].n, in[2].buf.data,
                      in[2].n, out->buf.data);
    if (rc) {
        ufomi_buffer_release(ctx, &out->buf);
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "flat-field-correct: %s", ufomi_last_error());
    }
    t->num_processed += (unsigned) in[0].n;
    return UFOMI_OK;
}

/* Runs every stage but a final `write' and leaves the last batch in *res
 * (*res_owned: res holds device memory the caller releases). */
static int run_batch(UfomiPipeline *pl, ufomi_ctx *ctx, const float *d_in, const UfomiRequisition *in_req,
                     int n_items, Batch *res, int *res_owned, UfomiError *e)
{
    const int has_read = is(pl->tasks[0], "read"), has_write = is(pl->tasks[pl->n - 1], "write");
    const int first = has_read || pl->n_branches, last = pl->n - has_write;
    memset(res, 0, sizeof(*res));
    *res_owned = 0;
    if (!has_read && !pl->n_branches && (!d_in || !in_req || n_items <= 0))
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_ARG, "invalid pipeline arguments");
    for (int i = first; i < last; i++)
        if (is(pl->tasks[i], "read") || is(pl->tasks[i], "write"))
            return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG,
                              "`%s' must be the %s stage of a pipeline", ufomi_task_name(pl->tasks[i]),
                              is(pl->tasks[i], "read") ? "first" : "last");
    int rc = UFOMI_OK;
    for (int i = 0; i < pl->n; i++)
        if ((rc = ufomi_task_setup(pl->tasks[i], ctx, e)))
            return rc;
    pl->fused = 0;

    Batch cur;
    int cur_owned = 0;
    if (pl->n_branches) {
        /* input branches (ufo-launch's [a, b, c] ! task): each one a batch, then the multi-input stage */
        Batch br[MAX_BRANCHES];
        int owned[MAX_BRANCHES] = {0};
        for (int b = 0; b < pl->n_branches && !rc; b++)
            rc = run_batch(pl->branches[b], ctx, NULL, NULL, 0, &br[b], &owned[b], e);
        for (int b = 0; b < pl->n_branches && !rc; b++)
            if (!br[b].n)
                rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_ARG, "input %d of `%s' produced no items", b,
                                ufomi_task_name(pl->tasks[0]));
        if (!rc)
            rc = run_multi_input(ctx, pl->tasks[0], br, &cur, e);
        for (int b = 0; b < pl->n_branches; b++)
            if (owned[b])
                ufomi_buffer_release(ctx, &br[b].buf);
        if (rc)
            return rc;
        cur_owned = 1;
    } else if (has_read) {
        if ((rc = read_batch(ctx, pl->tasks[0], &cur, e)))
            return rc;
        cur_owned = cur.n > 0;
        if (!cur.n) {  /* nothing matched the image range: an empty stream */
            *res = cur;
            return UFOMI_OK;
        }
    } else {
        memset(&cur, 0, sizeof(cur));
        cur.buf.data = (void *) d_in;
        cur.buf.req = *in_req;
        cur.buf.layout = UFOMI_LAYOUT_REAL;
        cur.buf.dtype = UFOMI_F32;
        cur.n = n_items;
        cur.item_bytes = ufomi_requisition_elements(in_req) * sizeof(float);
    }

    for (int s = first; s < last; s++) {
        UfomiTask *t = pl->tasks[s];
        Batch next;
        UfomiBuffer in0 = item_view(&cur, 0), *ins[1] = {&in0};
        UfomiRequisition req;

        if (is(t, "fft") && s + 2 < last && is(pl->tasks[s + 1], "filter") && is(pl->tasks[s + 2], "ifft") &&
            fft_chain_fusable(t, pl->tasks[s + 2], &in0)) {
            /* fused fft ! filter ! ifft: requisitions of the three tasks, one launch */
            UfomiRequisition r_fft, r_filt, r_ifft;
            UfomiBuffer spec = in0, filt;
            if ((rc = ufomi_task_get_requisition(t, ins, &r_fft, e)))
                goto out;
            spec.req = r_fft;
            spec.layout = UFOMI_LAYOUT_COMPLEX_INTERLEAVED;
            UfomiBuffer *sp[1] = {&spec};
            if ((rc = ufomi_task_get_requisition(pl->tasks[s + 1], sp, &r_filt, e)))
                goto out;
            filt = spec;
            filt.req = r_filt;
            UfomiBuffer *fp[1] = {&filt};
            if ((rc = ufomi_task_get_requisition(pl->tasks[s + 2], fp, &r_ifft, e)))
                goto out;
            if ((rc = batch_alloc(ctx, &next, &r_ifft, cur.n, sizeof(float), UFOMI_F32, UFOMI_LAYOUT_REAL)))
                goto fail_alloc;
            const int W = (int) cur.buf.req.dims[0];
            const int N = (int) (r_fft.dims[0] / 2);
            const long rows = (long) rows_of(&cur.buf.req) * cur.n;
            rc = ufomi_ramp_filter(ctx, cur.buf.data, W, (int) rows, N, filter_task_device_coefficients(pl->tasks[s + 1]),
                                   (int) r_ifft.dims[0], next.buf.data, UFOMI_F32);
            if (rc) {
                ufomi_buffer_release(ctx, &next.buf);
                task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "fused ramp filter: %s", ufomi_last_error());
                goto out;
            }
            for (int k = 0; k < 3; k++)
                pl->tasks[s + k]->num_processed += (unsigned) cur.n;
            pl->fused = 1;
            s += 2;
        } else if (is(t, "backproject")) {
            if ((rc = ufomi_task_get_requisition(t, ins, &req, e)))
                goto out;
            if ((rc = batch_alloc(ctx, &next, &req, cur.n, sizeof(float), UFOMI_F32, UFOMI_LAYOUT_REAL)))
                goto fail_alloc;
            ufomi_bp_params bp;
            backproject_task_params(t, &bp);
            rc = ufomi_bp_parallel(ctx, cur.buf.data, cur.buf.dtype, (int) cur.buf.req.dims[0],
                                   (int) cur.buf.req.dims[1], cur.n, cur.item_bytes / elem_size(cur.buf.dtype), &bp,
                                   next.buf.data, 0);
            if (rc) {
                ufomi_buffer_release(ctx, &next.buf);
                task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "backproject: %s", ufomi_last_error());
                goto out;
            }
            t->num_processed += (unsigned) cur.n;
        } else if (is(t, "average")) {
            /* the float32 sum of all items in order / count, `number' copies, one launch */
            const int copies = (int) t->v[0].num;
            req = cur.buf.req;
            if (cur.buf.dtype != UFOMI_F32) {
                rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_UNSUPPORTED, "average: input must be float32");
                goto out;
            }
            if ((rc = batch_alloc(ctx, &next, &req, copies, sizeof(float), UFOMI_F32, UFOMI_LAYOUT_REAL)))
                goto fail_alloc;
            if ((rc = ufomi_average(ctx, cur.buf.data, ufomi_requisition_elements(&req), cur.n, copies,
                                    next.buf.data))) {
                ufomi_buffer_release(ctx, &next.buf);
                task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "average: %s", ufomi_last_error());
                goto out;
            }
            t->num_processed += (unsigned) cur.n;
        } else if (is(t, "transpose-projections")) {
            /* all projections of the batch in one launch, zero rows for missing ones; the
             * requisition of ufo-transpose-projections-task.c:128-150 without the task's own store */
            req.n_dims = 2;
            req.dims[0] = cur.buf.req.dims[0];
            req.dims[1] = (size_t) t->v[0].num;
            req.dims[2] = 1;
            if (cur.buf.dtype != UFOMI_F32) {
                rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_UNSUPPORTED,
                                "transpose-projections: input must be float32");
                goto out;
            }
            const int W = (int) req.dims[0], number = (int) req.dims[1], H = (int) rows_of(&cur.buf.req);
            const int used = cur.n < number ? cur.n : number;
            if ((rc = batch_alloc(ctx, &next, &req, H, sizeof(float), UFOMI_F32, UFOMI_LAYOUT_REAL)))
                goto fail_alloc;
            rc = ufomi_transpose_projections(ctx, cur.buf.data, W, H, used, 0, number, next.buf.data);
            if (!rc && used < number)
                rc = ufomi_transpose_projections(ctx, NULL, W, H, number - used, used, number, next.buf.data);
            if (rc) {
                ufomi_buffer_release(ctx, &next.buf);
                task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "transpose-projections: %s", ufomi_last_error());
                goto out;
            }
            t->num_processed += (unsigned) used;
        } else if (ufomi_task_get_mode(t) & UFOMI_TASK_MODE_REDUCTOR) {
            if ((rc = ufomi_task_get_requisition(t, ins, &req, e)))
                goto out;
            for (int i = 0; i < cur.n; i++) {
                UfomiBuffer v = item_view(&cur, i), *vi[1] = {&v};
                UfomiBuffer dummy = v;
                if (!ufomi_task_process(t, vi, &dummy, &req)) {
                    if (!t->err_status)
                        break;  /* the reductor takes no more input */
                    rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, t->err_status, "%s", ufomi_task_last_error(t));
                    goto out;
                }
            }
            /* general-backproject emits its slices in the store type, lamino-backproject one volume */
            int nz = 0, dtype = UFOMI_F32;
            size_t esize = sizeof(float);
            if ((rc = reductor_output_items(t, &cur.buf.req, &nz, &esize, &dtype))) {
                task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, rc, "%s", ufomi_last_error());
                goto out;
            }
            if ((rc = batch_alloc(ctx, &next, &req, nz, esize, UFOMI_F32, UFOMI_LAYOUT_REAL)))
                goto fail_alloc;
            int produced = 0;
            for (;;) {
                UfomiBuffer v = item_view(&next, produced);
                if (produced >= nz || !ufomi_task_generate(t, &v, &req))
                    break;
                produced++;
            }
            if (t->err_status) {
                ufomi_buffer_release(ctx, &next.buf);
                rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, t->err_status, "%s", ufomi_task_last_error(t));
                goto out;
            }
            next.n = produced;
            next.buf.dtype = dtype;
        } else {
            if ((rc = ufomi_task_get_requisition(t, ins, &req, e)))
                goto out;
            if ((rc = batch_alloc(ctx, &next, &req, cur.n, sizeof(float), UFOMI_F32, ufomi_task_output_layout(t))))
                goto fail_alloc;
            /* the layout the task gives its outputs (filter keeps its input's) */
            UfomiBufferLayout layout = ufomi_task_output_layout(t);
            for (int i = 0; i < cur.n; i++) {
                UfomiBuffer v = item_view(&cur, i), o = item_view(&next, i), *vi[1] = {&v};
                if (!ufomi_task_process(t, vi, &o, &req)) {
                    ufomi_buffer_release(ctx, &next.buf);
                    rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, t->err_status ? t->err_status : UFOMI_ERR_ARG, "%s",
                                    ufomi_task_last_error(t));
                    goto out;
                }
                layout = o.layout;
            }
            next.buf.layout = layout;
        }
        if (cur_owned)
            ufomi_buffer_release(ctx, &cur.buf);
        cur = next;
        cur_owned = 1;
        continue;
    fail_alloc:
        task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_MEMORY, "out of device memory: %s", ufomi_last_error());
        goto out;
    }

    *res = cur;
    *res_owned = cur_owned;
    return UFOMI_OK;
out:
    if (cur_owned)
        ufomi_buffer_release(ctx, &cur.buf);
    return rc;
}

int ufomi_pipeline_run(UfomiPipeline *pl, ufomi_ctx *ctx, const float *d_in, const UfomiRequisition *in_req,
                       int n_items, void *d_out, size_t out_capacity, int *n_out, UfomiRequisition *out_req,
                       UfomiError *e)
{
    if (!pl || !ctx || !n_out || !out_req)
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_ARG, "invalid pipeline arguments");
    const int has_read = is(pl->tasks[0], "read"), has_write = is(pl->tasks[pl->n - 1], "write");
    if (!has_write && !d_out)
        return task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_ARG, "invalid pipeline arguments");
    if (has_read && pl->n == 1)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "a pipeline of `read' alone has no sink");
    Batch cur;
    int cur_owned = 0;
    int rc = run_batch(pl, ctx, d_in, in_req, n_items, &cur, &cur_owned, e);
    if (rc)
        return rc;
    if (has_write) {
        rc = cur.n ? write_batch(ctx, pl->tasks[pl->n - 1], &cur, e) : UFOMI_OK;
        if (!cur.n)
            write_task_flush(pl->tasks[pl->n - 1]);
        if (!rc) {
            *n_out = cur.n;
            *out_req = cur.buf.req;
        }
    } else {
        const size_t total = cur.item_bytes * (size_t) cur.n;
        if (total > out_capacity) {
            rc = task_error(e, UFOMI_TASK_ERROR_PROCESS, UFOMI_ERR_SIZE, "output needs %zu bytes, capacity %zu", total,
                            out_capacity);
        } else if (total && (rc = ufomi_memcpy(ctx, d_out, cur.buf.data, total, 2))) {
            task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "%s", ufomi_last_error());
        } else if ((rc = ufomi_synchronize(ctx))) {
            task_error(e, UFOMI_TASK_ERROR_PROCESS, rc, "%s", ufomi_last_error());
        } else {
            *n_out = cur.n;
            *out_req = cur.buf.req;
        }
    }
    if (cur_owned)
        ufomi_buffer_release(ctx, &cur.buf);
    return rc;
}
