/*
 * The reference's hot-path tasks in plain C over the ufomi C ABI
 * (include/ufo_tasks.h): property tables with the GObject names, ranges and
 * effective defaults, and the setup / get_requisition / process / generate
 * rules of
 *   fft                          src/ufo-fft-task.c
 *   filter                       src/ufo-filter-task.c
 *   ifft                         src/ufo-ifft-task.c
 *   backproject                  src/ufo-backproject-task.c
 *   cone-beam-projection-weight  src/ufo-cone-beam-projection-weight-task.c
 *   general-backproject          src/ufo-general-backproject-task.c
 *   lamino-backproject           src/ufo-lamino-backproject-task.c
 *   center-of-rotation           src/ufo-center-of-rotation-task.c
 *   flat-field-correct           src/ufo-flat-field-correct-task.c
 *   average                      src/ufo-average-task.c
 *   stacked-backproject          src/ufo-stacked-backproject-task.c
 *   transpose-projections        src/ufo-transpose-projections-task.c
 *   read                         src/ufo-read-task.c
 *   write                        src/ufo-write-task.c
 */
#include "tasks_internal.h"
#include "io.h"

#include <glob.h>
#include <libgen.h>
#include <sys/stat.h>
#include <unistd.h>

#include <ctype.h>
#include <errno.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ errors */

int task_error(UfomiError *e, int code, int status, const char *fmt, ...)
{
    if (e) {
        va_list ap;
        e->code = code;
        e->status = status;
        va_start(ap, fmt);
        vsnprintf(e->message, sizeof(e->message), fmt, ap);
        va_end(ap);
    }
    return status;
}

static int process_fail(UfomiTask *t, int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t->err, sizeof(t->err), fmt, ap);
    va_end(ap);
    t->err_status = status;
    return 0;
}

const char *ufomi_task_last_error(const UfomiTask *t) { return t ? t->err : ""; }

/* --------------------------------------------------------------- properties */

static const char *const filter_nicks[] = {"ramp", "ramp-fromreal", "butterworth", "faris-byer", "hamming", "bh3",
                                           "shepp-logan", NULL};
static const char *const bp_mode_nicks[] = {"nearest", "texture", NULL};
static const char *const gbp_parameter_nicks[] = {
    "axis-angle-x", "axis-angle-y", "axis-angle-z", "volume-angle-x", "volume-angle-y", "volume-angle-z",
    "detector-angle-x", "detector-angle-y", "detector-angle-z", "detector-position-x", "detector-position-y",
    "detector-position-z", "source-position-x", "source-position-y", "source-position-z", "center-position-x",
    "center-position-z", "z", NULL};
static const char *const compute_nicks[] = {"float", "double", NULL};
static const char *const ft_nicks[] = {"half", "float", "double", NULL};
static const char *const st_nicks[] = {"half", "float", "double", "uchar", "ushort", "uint", NULL};
static const char *const addr_nicks[] = {"none", "clamp_to_edge", "clamp", "repeat", "mirrored_repeat", NULL};

#define U32MAX 4294967295.0

/* fft (ufo-fft-task.c:498-531, init :541-568: auto-zeropadding is FALSE after init) */
static const PropSpec fft_props[] = {
    {"auto-zeropadding", PT_BOOL, 0, 1, 0, NULL, 0},
    {"dimensions", PT_UINT, 1, 3, 1, NULL, 0},
    {"size-x", PT_UINT, 0, 32768, 0, NULL, 0},
    {"size-y", PT_UINT, 0, 32768, 0, NULL, 0},
    {"size-z", PT_UINT, 0, 32768, 0, NULL, 0},
};
enum { FFT_ZEROPAD, FFT_DIMS, FFT_SX, FFT_SY, FFT_SZ };

/* filter (ufo-filter-task.c:481-521, init :531-545) */
static const PropSpec filter_props[] = {
    {"filter", PT_ENUM, 0, 6, 1, filter_nicks, 0},
    {"cutoff", PT_FLOAT, 0.0, 1.0, 0.5, NULL, 0},
    {"order", PT_FLOAT, 2.0, 32.0, 4.0, NULL, 0},
    {"tau", PT_FLOAT, -FLT_MAX, FLT_MAX, 0.1f, NULL, 0},
    {"theta", PT_FLOAT, -FLT_MAX, FLT_MAX, 1.0, NULL, 0},
    {"scale", PT_FLOAT, -FLT_MAX, FLT_MAX, 1.0, NULL, 0},
};
enum { FLT_TYPE, FLT_CUTOFF, FLT_ORDER, FLT_TAU, FLT_THETA, FLT_SCALE };

/* ifft (ufo-ifft-task.c:461-480) */
static const PropSpec ifft_props[] = {
    {"dimensions", PT_UINT, 1, 3, 1, NULL, 0},
    {"crop-width", PT_UINT, 0, 32768, 0, NULL, 0},
    {"crop-height", PT_UINT, 0, 32768, 0, NULL, 0},
};
enum { IFFT_DIMS, IFFT_CW, IFFT_CH };

/* backproject (ufo-backproject-task.c:442-510, init :520-541) */
static const PropSpec bp_props[] = {
    {"num-projections", PT_UINT, 0, 32768, 0, NULL, 0},
    {"offset", PT_UINT, 0, 32768, 0, NULL, 0},
    {"axis-pos", PT_DOUBLE, -1.0, 32768.0, -1.0, NULL, 0},
    {"angle-step", PT_DOUBLE, -DBL_MAX, DBL_MAX, -1.0, NULL, 0},
    {"angle-offset", PT_DOUBLE, 0.0, DBL_MAX, 0.0, NULL, 0},
    {"mode", PT_ENUM, 0, 1, 1, bp_mode_nicks, 0},
    {"roi-x", PT_UINT, 0, U32MAX, 0, NULL, 0},
    {"roi-y", PT_UINT, 0, U32MAX, 0, NULL, 0},
    {"roi-width", PT_UINT, 0, U32MAX, 0, NULL, 0},
    {"roi-height", PT_UINT, 0, U32MAX, 0, NULL, 0},
};
enum { BP_NPROJ, BP_OFFSET, BP_AXIS, BP_STEP, BP_AOFF, BP_MODE, BP_RX, BP_RY, BP_RW, BP_RH };

/* cone-beam-projection-weight (ufo-cone-beam-projection-weight-task.c:268-301, init :309-320) */
static const PropSpec cw_props[] = {
    {"center-position-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 0},
    {"center-position-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 0},
    {"source-distance", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 0},
    {"detector-distance", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 0},
    {"axis-angle-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
};
enum { CW_CX, CW_CZ, CW_SD, CW_DD, CW_AX };

/* general-backproject (ufo-general-backproject-task.c:2087-2311, init :2319-2359) */
static const PropSpec gbp_props[] = {
    {"burst", PT_UINT, 0, 128, 0, NULL, 0},
    {"parameter", PT_ENUM, 0, 17, 17, gbp_parameter_nicks, 0},
    {"z", PT_DOUBLE, -DBL_MAX, DBL_MAX, 0.0, NULL, 0},
    {"region", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 3},
    {"x-region", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 3},
    {"y-region", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 3},
    {"center-position-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"center-position-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"source-position-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"source-position-y", PT_DARRAY, -INFINITY, INFINITY, -INFINITY, NULL, 1},
    {"source-position-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"detector-position-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"detector-position-y", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"detector-position-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"detector-angle-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"detector-angle-y", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"detector-angle-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"axis-angle-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"axis-angle-y", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"axis-angle-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"volume-angle-x", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"volume-angle-y", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"volume-angle-z", PT_DARRAY, -INFINITY, INFINITY, 0, NULL, 1},
    {"num-projections", PT_UINT, 0, 32768, 0, NULL, 0},
    {"compute-type", PT_ENUM, 0, 1, 0, compute_nicks, 0},
    {"result-type", PT_ENUM, 0, 2, 1, ft_nicks, 0},
    {"store-type", PT_ENUM, 0, 5, 1, st_nicks, 0},
    {"overall-angle", PT_DOUBLE, -DBL_MAX, DBL_MAX, 2 * M_PI, NULL, 0},
    {"addressing-mode", PT_ENUM, 0, 4, 2, addr_nicks, 0},
    {"gray-map-min", PT_DOUBLE, -DBL_MAX, DBL_MAX, 0.0, NULL, 0},
    {"gray-map-max", PT_DOUBLE, -DBL_MAX, DBL_MAX, 0.0, NULL, 0},
};
enum { G_BURST, G_PARAM, G_Z, G_REGION, G_XREGION, G_YREGION, G_CX, G_CZ, G_SX, G_SY, G_SZ, G_DX, G_DY, G_DZ,
       G_DAX, G_DAY, G_DAZ, G_AX, G_AY, G_AZ, G_VX, G_VY, G_VZ, G_NPROJ, G_COMPUTE, G_RESULT, G_STORE, G_OVERALL,
       G_ADDR, G_GMIN, G_GMAX };

static void prop_reset(UfomiTask *t)
{
    for (int i = 0; i < t->cls->n_props; i++) {
        const PropSpec *s = &t->cls->props[i];
        PropValue *v = &t->v[i];
        free(v->arr);
        free(v->str);
        v->arr = NULL;
        v->str = s->type == PT_STRING ? strdup(s->sdflt ? s->sdflt : "") : NULL;
        v->n = 0;
        v->num = s->dflt;
        if (s->type == PT_DARRAY && s->n_default) {
            v->n = s->n_default;
            v->arr = calloc((size_t) v->n, sizeof(double));
            for (int k = 0; k < v->n; k++)
                v->arr[k] = s->dflt;
        }
    }
}

static int prop_index(const UfomiTask *t, const char *name)
{
    for (int i = 0; i < t->cls->n_props; i++) {
        if (!strcmp(t->cls->props[i].name, name))
            return i;
        /* accept '_' for '-' like GObject's canonical names */
        const char *a = t->cls->props[i].name, *b = name;
        while (*a && *b && (*a == *b || (*a == '-' && *b == '_'))) {
            a++;
            b++;
        }
        if (!*a && !*b)
            return i;
    }
    return -1;
}

static int parse_double(const char *s, double *out)
{
    char *end;
    while (isspace((unsigned char) *s))
        s++;
    if (!strncasecmp(s, "-inf", 4)) {
        *out = -INFINITY;
        return 1;
    }
    if (!strncasecmp(s, "inf", 3) || !strncasecmp(s, "+inf", 4)) {
        *out = INFINITY;
        return 1;
    }
    errno = 0;
    *out = strtod(s, &end);
    if (end == s || errno == ERANGE)
        return 0;
    while (isspace((unsigned char) *end))
        end++;
    return *end == '\0';
}

static int write_bits_property(UfomiTask *t, double bits);
static int lamino_region_property(const char *name, const double *vals, int n);

int ufomi_task_set_property(UfomiTask *t, const char *name, const char *value)
{
    if (!t || !name || !value)
        return ufomi_status_fail(UFOMI_ERR_ARG, "null argument");
    const int i = prop_index(t, name);
    if (i < 0)
        return ufomi_status_fail(UFOMI_ERR_ARG, "%s has no property named `%s'", t->cls->name, name);
    const PropSpec *s = &t->cls->props[i];
    PropValue *v = &t->v[i];
    double d;
    if (!strcmp(t->cls->name, "center-of-rotation") && i == 1)
        return ufomi_status_fail(UFOMI_ERR_ARG, "property `center' of center-of-rotation is not writable");
    switch (s->type) {
    case PT_BOOL:
        if (!strcasecmp(value, "true") || !strcmp(value, "1"))
            v->num = 1;
        else if (!strcasecmp(value, "false") || !strcmp(value, "0"))
            v->num = 0;
        else
            return ufomi_status_fail(UFOMI_ERR_ARG, "invalid boolean `%s' for %s", value, name);
        return UFOMI_OK;
    case PT_ENUM:
        for (int k = 0; s->nicks[k]; k++) {
            if (!strcmp(s->nicks[k], value)) {
                v->num = k;
                return UFOMI_OK;
            }
        }
        return ufomi_status_fail(UFOMI_ERR_ARG, "`%s' is not a valid value of %s", value, name);
    case PT_UINT:
    case PT_DOUBLE:
    case PT_FLOAT:
        if (!parse_double(value, &d))
            return ufomi_status_fail(UFOMI_ERR_ARG, "cannot parse `%s' for %s", value, name);
        if (s->type == PT_UINT && (d != floor(d)))
            return ufomi_status_fail(UFOMI_ERR_ARG, "%s needs an unsigned integer, got `%s'", name, value);
        if (s->type == PT_FLOAT)
            d = (float) d;
        if (!(d >= s->min && d <= s->max))
            return ufomi_status_fail(UFOMI_ERR_ARG, "value %s of %s is outside [%g, %g]", value, name, s->min, s->max);
        if (!strcmp(t->cls->name, "write") && !strcmp(s->name, "bits") && !write_bits_property(t, d))
            return UFOMI_OK;
        v->num = d;
        return UFOMI_OK;
    case PT_STRING:
        free(v->str);
        v->str = strdup(value);
        return UFOMI_OK;
    case PT_DARRAY: {
        double vals[65536];
        int n = 0;
        const char *p = value;
        while (*p) {
            char tok[128];
            size_t len = strcspn(p, ",");
            if (len >= sizeof(tok))
                return ufomi_status_fail(UFOMI_ERR_ARG, "array element too long in %s", name);
            memcpy(tok, p, len);
            tok[len] = '\0';
            if (n >= (int) (sizeof(vals) / sizeof(vals[0])) || !parse_double(tok, &vals[n]))
                return ufomi_status_fail(UFOMI_ERR_ARG, "cannot parse array `%s' for %s", value, name);
            n++;
            p += len;
            if (*p == ',')
                p++;
        }
        if (!strcmp(t->cls->name, "lamino-backproject") && (!strcmp(name, "x-region") || !strcmp(name, "y-region")) &&
            !lamino_region_property(name, vals, n))
            return UFOMI_OK;
        free(v->arr);
        v->arr = n ? malloc((size_t) n * sizeof(double)) : NULL;
        if (n)
            memcpy(v->arr, vals, (size_t) n * sizeof(double));
        v->n = n;
        return UFOMI_OK;
    }
    }
    return ufomi_status_fail(UFOMI_ERR_ARG, "unknown property type");
}

int ufomi_task_get_property(const UfomiTask *t, const char *name, char *buf, size_t len)
{
    if (!t || !name || !buf || !len)
        return ufomi_status_fail(UFOMI_ERR_ARG, "null argument");
    const int i = prop_index(t, name);
    if (i < 0)
        return ufomi_status_fail(UFOMI_ERR_ARG, "%s has no property named `%s'", t->cls->name, name);
    const PropSpec *s = &t->cls->props[i];
    const PropValue *v = &t->v[i];
    switch (s->type) {
    case PT_BOOL:
        snprintf(buf, len, "%s", v->num ? "true" : "false");
        break;
    case PT_ENUM:
        snprintf(buf, len, "%s", s->nicks[(int) v->num]);
        break;
    case PT_UINT:
        snprintf(buf, len, "%.0f", v->num);
        break;
    case PT_DOUBLE:
    case PT_FLOAT:
        snprintf(buf, len, "%.17g", v->num);
        break;
    case PT_STRING:
        snprintf(buf, len, "%s", v->str ? v->str : "");
        break;
    case PT_DARRAY: {
        size_t off = 0;
        buf[0] = '\0';
        for (int k = 0; k < v->n && off < len; k++)
            off += (size_t) snprintf(buf + off, len - off, k ? ",%.17g" : "%.17g", v->arr[k]);
        break;
    }
    }
    return UFOMI_OK;
}

const char *const *ufomi_task_list_properties(const UfomiTask *t)
{
    return t ? t->names : NULL;
}

static double arr_at(const PropValue *v, unsigned i)
{
    /* ufo_scarray_get_double: one element broadcasts (ufo-scarray.c:118-127) */
    if (v->n == 1)
        return v->arr[0];
    return (int) i < v->n ? v->arr[i] : NAN;
}

/* ------------------------------------------------------------ fft / ifft */

/* ufo_math_compute_closest_smaller_power_of_2 (ufo-math.c:145-151) */
static size_t closest_smaller_pow2(size_t x)
{
    double ip;
    modf(log2((double) x), &ip);
    return (size_t) pow(2, ip);
}

static int is_pow2(size_t x) { return x && !(x & (x - 1)); }

typedef struct {
    int n[3];      /* input extents (complex elements along x for complex input) */
    int m[3];      /* output extents */
    int in_complex;
} FftPriv;

/* get_requisition (ufo-fft-task.c:122-199): per axis the user size (not
 * smaller than the input), else the input size, rounded up to a power of two
 * on transformed axes when auto-zeropadding; the output is that size,
 * interleaved complex along x. */
static int fft_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    FftPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    const int dims = (int) t->v[FFT_DIMS].num;
    const int cplx = in[0]->layout == UFOMI_LAYOUT_COMPLEX_INTERLEAVED;
    const double user[3] = {t->v[FFT_SX].num, t->v[FFT_SY].num, t->v[FFT_SZ].num};
    for (int i = 0; i < 3; i++) {
        const size_t in_dim = i < (int) ir->n_dims ? ir->dims[i] : 1;
        size_t size = in_dim;
        if (i < (int) ir->n_dims) {
            if (user[i] != 0 && user[i] < (double) in_dim)
                return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_SIZE,
                                  "Specified size must be greater than or equal to input size");
            if (user[i] == 0) {
                if (i == 0 && cplx)
                    size >>= 1;
                if (t->v[FFT_ZEROPAD].num != 0 && i <= dims - 1)
                    size = 2 * closest_smaller_pow2(size - 1);
            } else {
                size = (size_t) user[i];
            }
        }
        p->m[i] = (int) size;
        p->n[i] = (int) (i == 0 && cplx ? in_dim / 2 : in_dim);
    }
    p->in_complex = cplx;
    req->n_dims = ir->n_dims;
    req->dims[0] = 2 * (size_t) p->m[0];
    req->dims[1] = p->m[1];
    req->dims[2] = p->m[2];
    return UFOMI_OK;
}

static int fft_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    FftPriv *p = t->priv;
    int rc = ufomi_fft_nd(t->ctx, in[0]->data, p->in_complex, p->n, (int) t->v[FFT_DIMS].num, p->m, out->data);
    if (rc)
        return process_fail(t, rc, "fft: %s", ufomi_last_error());
    out->layout = UFOMI_LAYOUT_COMPLEX_INTERLEAVED;
    return 1;
}

typedef struct {
    int m[3];      /* complex input extents */
    int cw, ch;    /* crops (0: none) */
} IfftPriv;

/* get_requisition (ufo-ifft-task.c:111-185): the input must be complex; the
 * output is the real part, crop-width x crop-height when set. */
static int ifft_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    IfftPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    /* :121-125 */
    if (in[0]->layout != UFOMI_LAYOUT_COMPLEX_INTERLEAVED)
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_LAYOUT, "ifft input must be complex");
    const size_t cw = (size_t) t->v[IFFT_CW].num, ch = (size_t) t->v[IFFT_CH].num;
    /* :133-138 compares the user size with the input requisition (the complex width along x) */
    if ((cw && cw > ir->dims[0]) || (ch && ir->n_dims >= 2 && ch > ir->dims[1]))
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_SIZE,
                          "Cropped size must be less than or equal to input size");
    p->m[0] = (int) (ir->dims[0] / 2);
    p->m[1] = ir->n_dims >= 2 ? (int) ir->dims[1] : 1;
    p->m[2] = ir->n_dims == 3 ? (int) ir->dims[2] : 1;
    if (cw > (size_t) p->m[0] || ch > (size_t) p->m[1])
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_SIZE,
                          "Cropped size must be less than or equal to input size");
    p->cw = (int) cw;
    p->ch = (int) ch;
    req->n_dims = ir->n_dims;
    req->dims[0] = cw ? cw : (size_t) p->m[0];
    req->dims[1] = ch ? ch : (size_t) p->m[1];
    req->dims[2] = p->m[2];
    return UFOMI_OK;
}

static int ifft_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    IfftPriv *p = t->priv;
    int rc = ufomi_ifft_nd(t->ctx, in[0]->data, p->m, (int) t->v[IFFT_DIMS].num, p->cw, p->ch, out->data);
    if (rc)
        return process_fail(t, rc, "ifft: %s", ufomi_last_error());
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

/* The chain fft ! filter ! ifft runs as the fused ramp kernel when it is a
 * 1-D transform of real rows with no padding or crop across rows. */
int fft_chain_fusable(const UfomiTask *fft, const UfomiTask *ifft, const UfomiBuffer *in)
{
    const UfomiRequisition *r = &in->req;
    const double sy = fft->v[FFT_SY].num, sz = fft->v[FFT_SZ].num, sx = fft->v[FFT_SX].num;
    /* the fused kernel covers power-of-two transforms of 4..8192 points (fft_requisition's size) */
    const size_t w = r->dims[0];
    const size_t n = sx != 0 ? (size_t) sx : (fft->v[FFT_ZEROPAD].num != 0 && w > 1 ? 2 * closest_smaller_pow2(w - 1) : w);
    return n >= 4 && n <= 8192 && is_pow2(n) && in->layout == UFOMI_LAYOUT_REAL && in->dtype == UFOMI_F32 &&
           (int) fft->v[FFT_DIMS].num == 1 &&
           (int) ifft->v[IFFT_DIMS].num == 1 && (sy == 0 || r->n_dims < 2 || sy == (double) r->dims[1]) &&
           (sz == 0 || r->n_dims < 3 || sz == (double) r->dims[2]) &&
           (ifft->v[IFFT_CH].num == 0 || r->n_dims < 2 || ifft->v[IFFT_CH].num == (double) r->dims[1]);
}

/* ---------------------------------------------------------------- filter */

typedef struct {
    float *d_coef;
    size_t width;
} FilterPriv;

int filter_task_coefficients(UfomiTask *t, size_t width, UfomiError *e)
{
    FilterPriv *p = t->priv;
    if (p->d_coef && p->width == width)
        return UFOMI_OK;
    /* ufo-filter-task.c:295-300 */
    if (!is_pow2(width))
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_NOT_POW2,
                          "Invalid filter input width %zu, must be a power of two", width);
    float *h = malloc(width * sizeof(float));
    int rc = ufomi_filter_coefficients((int) t->v[FLT_TYPE].num, (unsigned) width, (float) t->v[FLT_SCALE].num,
                                       (float) t->v[FLT_CUTOFF].num, (float) t->v[FLT_ORDER].num,
                                       (float) t->v[FLT_TAU].num, (float) t->v[FLT_THETA].num, h);
    if (!rc && p->d_coef) {
        ufomi_device_free(t->ctx, p->d_coef);
        p->d_coef = NULL;
    }
    if (!rc)
        rc = ufomi_device_alloc(t->ctx, width * sizeof(float), (void **) &p->d_coef);
    if (!rc)
        rc = ufomi_memcpy(t->ctx, p->d_coef, h, width * sizeof(float), 0);
    free(h);
    if (rc)
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, rc, "filter: %s", ufomi_last_error());
    p->width = width;
    return UFOMI_OK;
}

const float *filter_task_device_coefficients(const UfomiTask *t) { return ((FilterPriv *) t->priv)->d_coef; }

static int filter_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    *req = in[0]->req;
    return filter_task_coefficients(t, in[0]->req.dims[0], e);
}

static int filter_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    FilterPriv *p = t->priv;
    const size_t rows = (req->n_dims >= 2 ? req->dims[1] : 1) * (req->n_dims == 3 ? req->dims[2] : 1);
    int rc = ufomi_filter_apply(t->ctx, in[0]->data, out->data, p->d_coef, (int) p->width, (int) rows);
    if (rc)
        return process_fail(t, rc, "filter: %s", ufomi_last_error());
    out->layout = in[0]->layout;
    return 1;
}

static void filter_finalize(UfomiTask *t)
{
    FilterPriv *p = t->priv;
    if (p && p->d_coef && t->ctx)
        ufomi_device_free(t->ctx, p->d_coef);
}

/* ----------------------------------------------------------- backproject */

typedef struct {
    unsigned n_projections, burst;
} BpPriv;

void backproject_task_params(const UfomiTask *t, ufomi_bp_params *bp)
{
    const BpPriv *p = t->priv;
    ufomi_bp_params_init(bp);
    bp->num_projections = p->n_projections;
    bp->offset = (unsigned) t->v[BP_OFFSET].num;
    bp->axis_pos = t->v[BP_AXIS].num;
    bp->angle_step = t->v[BP_STEP].num;
    bp->angle_offset = t->v[BP_AOFF].num;
    bp->mode = (int) t->v[BP_MODE].num == 0 ? UFOMI_BP_NEAREST : UFOMI_BP_TEXTURE;
    bp->roi_x = (unsigned) t->v[BP_RX].num;
    bp->roi_y = (unsigned) t->v[BP_RY].num;
    bp->roi_width = (unsigned) t->v[BP_RW].num;
    bp->roi_height = (unsigned) t->v[BP_RH].num;
}

static int bp_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    BpPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    /* ufo-backproject-task.c:221-244: num-projections defaults to the first
     * sinogram's height and stays; the burst is every sinogram's height. */
    if (!p->n_projections)
        p->n_projections = (unsigned) t->v[BP_NPROJ].num ? (unsigned) t->v[BP_NPROJ].num : (unsigned) ir->dims[1];
    p->burst = (unsigned) ir->dims[1];
    if (p->burst > p->n_projections)
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_PROJECTIONS,
                          "Total number of projections (%u) must be greater than or equal to sinogram height (%u)",
                          p->n_projections, p->burst);
    req->n_dims = 2;
    req->dims[0] = t->v[BP_RW].num ? (size_t) t->v[BP_RW].num : ir->dims[0];
    req->dims[1] = t->v[BP_RH].num ? (size_t) t->v[BP_RH].num : ir->dims[0];
    req->dims[2] = 1;
    return UFOMI_OK;
}

static int bp_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    ufomi_bp_params bp;
    backproject_task_params(t, &bp);
    int rc = ufomi_bp_parallel(t->ctx, in[0]->data, in[0]->dtype, (int) in[0]->req.dims[0], (int) in[0]->req.dims[1],
                               1, 0, &bp, out->data, 0);
    if (rc)
        return process_fail(t, rc, "backproject: %s", ufomi_last_error());
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

/* ----------------------------------------------------- flat-field-correct */

/* flat-field-correct (ufo-flat-field-correct-task.c:262-297, init :300-309) */
static const PropSpec ffc_props[] = {
    {"fix-nan-and-inf", PT_BOOL, 0, 1, 0, NULL, 0, NULL},
    {"absorption-correct", PT_BOOL, 0, 1, 0, NULL, 0, NULL},
    {"sinogram-input", PT_BOOL, 0, 1, 0, NULL, 0, NULL},
    {"dark-scale", PT_FLOAT, -FLT_MAX, FLT_MAX, 1.0, NULL, 0, NULL},
    {"flat-scale", PT_FLOAT, -FLT_MAX, FLT_MAX, 1.0, NULL, 0, NULL},
};
enum { FFC_FIX, FFC_ABSORB, FFC_SINO, FFC_DSCALE, FFC_FSCALE };

typedef struct {
    int unused;
} FfcPriv;

static size_t req_rows(const UfomiRequisition *r)
{
    return (r->n_dims >= 2 ? r->dims[1] : 1) * (r->n_dims == 3 ? r->dims[2] : 1);
}

/* get_requisition (:76-88): the projection's size; darks and flats must match
 * it, or, with sinogram input, be one row of its width */
int ffc_check_inputs(const UfomiTask *t, const UfomiRequisition *data, const UfomiRequisition *dark,
                     const UfomiRequisition *flat, UfomiError *e)
{
    const int sino = (int) t->v[FFC_SINO].num;
    const UfomiRequisition *o[2] = {dark, flat};
    for (int k = 0; k < 2; k++) {
        const int ok = sino ? o[k]->dims[0] == data->dims[0] && req_rows(o[k]) == 1
                            : o[k]->dims[0] == data->dims[0] && req_rows(o[k]) == req_rows(data);
        if (!ok)
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_SIZE,
                              "flat-field-correct inputs must have the same size");
    }
    return UFOMI_OK;
}

static int ffc_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    if (!in[1] || !in[2])
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                          "flat-field-correct needs projections, darks and flats on inputs 0, 1 and 2");
    int rc = ffc_check_inputs(t, &in[0]->req, &in[1]->req, &in[2]->req, e);
    if (rc)
        return rc;
    *req = in[0]->req;
    return UFOMI_OK;
}

int ffc_task_run(UfomiTask *t, ufomi_ctx *ctx, const float *data, const UfomiRequisition *r, int n,
                 const float *dark, int n_dark, const float *flat, int n_flat, float *out)
{
    return ufomi_flat_field_correct(ctx, data, (int) r->dims[0], (int) req_rows(r), n, dark, n_dark, flat, n_flat,
                                    (int) t->v[FFC_SINO].num, (int) t->v[FFC_ABSORB].num, (int) t->v[FFC_FIX].num,
                                    (float) t->v[FFC_DSCALE].num, (float) t->v[FFC_FSCALE].num, out);
}

/* process (:107-160) */
static int ffc_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    for (int k = 0; k < 3; k++)
        if (in[k]->dtype != UFOMI_F32)
            return process_fail(t, UFOMI_ERR_UNSUPPORTED, "flat-field-correct: inputs must be float32");
    int rc = ffc_task_run(t, t->ctx, in[0]->data, &in[0]->req, 1, in[1]->data, 1, in[2]->data, 1, out->data);
    if (rc)
        return process_fail(t, rc, "flat-field-correct: %s", ufomi_last_error());
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

/* ---------------------------------------------------------------- average */

/* average (ufo-average-task.c:212-216, init :222-229) */
static const PropSpec avg_props[] = {
    {"number", PT_UINT, 1, U32MAX, 1, NULL, 0, NULL},
};

typedef struct {
    float *d_items;        /* the items in arrival order */
    float *d_avg;
    size_t elems, cap;     /* floats per item, items the store holds */
    unsigned count, generated;
    int averaged;
} AvgPriv;

static void avg_free(UfomiTask *t)
{
    AvgPriv *p = t->priv;
    if (t->ctx) {
        if (p->d_items)
            ufomi_device_free(t->ctx, p->d_items);
        if (p->d_avg)
            ufomi_device_free(t->ctx, p->d_avg);
    }
    memset(p, 0, sizeof(*p));
}

static int avg_setup(UfomiTask *t, UfomiError *e)
{
    (void) e;
    avg_free(t);
    return UFOMI_OK;
}

static int avg_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    AvgPriv *p = t->priv;
    const size_t elems = ufomi_requisition_elements(&in[0]->req);
    if (p->elems && elems != p->elems)
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_SIZE, "average: input size changed");
    p->elems = elems;
    *req = in[0]->req;
    return UFOMI_OK;
}

/* process (:102-122): the sum runs in arrival order when generate averages */
static int avg_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) out;
    (void) req;
    AvgPriv *p = t->priv;
    if (in[0]->dtype != UFOMI_F32)
        return process_fail(t, UFOMI_ERR_UNSUPPORTED, "average: input must be float32");
    int rc;
    if (p->count == p->cap) {
        const size_t cap = p->cap ? 2 * p->cap : 4;
        float *grown = NULL;
        if ((rc = ufomi_device_alloc(t->ctx, cap * p->elems * sizeof(float), (void **) &grown)))
            return process_fail(t, rc, "average: %s", ufomi_last_error());
        if (p->count && (rc = ufomi_memcpy(t->ctx, grown, p->d_items, p->count * p->elems * sizeof(float), 2))) {
            ufomi_device_free(t->ctx, grown);
            return process_fail(t, rc, "average: %s", ufomi_last_error());
        }
        if (p->d_items)
            ufomi_device_free(t->ctx, p->d_items);
        p->d_items = grown;
        p->cap = cap;
    }
    if ((rc = ufomi_memcpy(t->ctx, p->d_items + p->count * p->elems, in[0]->data, p->elems * sizeof(float), 2)))
        return process_fail(t, rc, "average: %s", ufomi_last_error());
    p->count++;
    return 1;
}

/* generate (:125-151): `number` copies of the average */
static int avg_generate(UfomiTask *t, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    AvgPriv *p = t->priv;
    if (!p->count || p->generated >= (unsigned) t->v[0].num)
        return 0;
    int rc;
    if (!p->averaged) {
        if ((rc = ufomi_device_alloc(t->ctx, p->elems * sizeof(float), (void **) &p->d_avg)) ||
            (rc = ufomi_average(t->ctx, p->d_items, p->elems, (int) p->count, 1, p->d_avg)))
            return process_fail(t, rc, "average: %s", ufomi_last_error());
        p->averaged = 1;
    }
    if ((rc = ufomi_memcpy(t->ctx, out->data, p->d_avg, p->elems * sizeof(float), 2)))
        return process_fail(t, rc, "average: %s", ufomi_last_error());
    p->generated++;
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

static void avg_finalize(UfomiTask *t)
{
    if (t->priv)
        avg_free(t);
}

/* ------------------------------------------------------ center-of-rotation */

/* center-of-rotation (ufo-center-of-rotation-task.c:232-246, init :250-257:
 * angle-step pi / 180; center is read-only) */
static const PropSpec cor_props[] = {
    {"angle-step", PT_DOUBLE, 0.00001f, 180.0f, M_PI / 180.0, NULL, 0, NULL},
    {"center", PT_DOUBLE, -DBL_MAX, DBL_MAX, 0.0, NULL, 0, NULL},
};
enum { COR_STEP, COR_CENTER };

typedef struct {
    float *stage;
    size_t cap;
} CorPriv;

static int cor_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    (void) t;
    (void) in;
    (void) e;
    req->n_dims = 0;
    return UFOMI_OK;
}

/* process (:111-166): squared differences of the first row against the last
 * row mirrored, for displacements -(W/2 - 1) .. W/2 - 1, the smallest score
 * gives center = (W + index - W/2 + 1) / 2.  The reference's mirror index
 * (width - x + 1) reaches one and two elements past the last row; those reads
 * are 0 here. */
static int cor_compute(UfomiTask *t, const float *sino, const UfomiRequisition *r)
{
    const unsigned width = (unsigned) r->dims[0], height = r->n_dims >= 2 ? (unsigned) r->dims[1] : 1;
    if (width < 2)
        return process_fail(t, UFOMI_ERR_ARG, "center-of-rotation: the sinogram needs at least 2 columns");
    const float *p0 = sino, *p180 = sino + (size_t) (height - 1) * width;
#define AT180(i) ((i) < width ? p180[(i)] : 0.0f)
    const unsigned maxd = width / 2, n = maxd * 2 - 1;
    float *scores = calloc(n, sizeof(float));
    if (!scores)
        return process_fail(t, UFOMI_ERR_MEMORY, "center-of-rotation: out of host memory");
    for (int d = -((int) maxd) + 1; d < 0; d++) {
        const unsigned index = (unsigned) (d + (int) maxd - 1), max_x = width - (unsigned) (-d);
        for (unsigned x = 0; x < max_x; x++) {
            const float diff = p0[x] - AT180(max_x - x + 1);
            scores[index] += diff * diff;
        }
    }
    for (unsigned d = 0; d < maxd; d++) {
        const unsigned index = d + maxd - 1;
        for (unsigned x = 0; x < width - d; x++) {
            const float diff = p0[x + d] - AT180(width - x + 1);
            scores[index] += diff * diff;
        }
    }
#undef AT180
    unsigned best = 0;
    for (unsigned i = 1; i < n; i++)
        if (scores[i] < scores[best])
            best = i;
    free(scores);
    t->v[COR_CENTER].num = (width + best - maxd + 1) / 2.0;
    return 1;
}

static int cor_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) out;
    (void) req;
    CorPriv *p = t->priv;
    if (in[0]->dtype != UFOMI_F32)
        return process_fail(t, UFOMI_ERR_UNSUPPORTED, "center-of-rotation: input must be float32");
    const size_t n = ufomi_requisition_elements(&in[0]->req);
    if (n > p->cap) {
        free(p->stage);
        p->stage = malloc(n * sizeof(float));
        p->cap = p->stage ? n : 0;
        if (!p->stage)
            return process_fail(t, UFOMI_ERR_MEMORY, "center-of-rotation: out of host memory");
    }
    int rc = ufomi_memcpy(t->ctx, p->stage, in[0]->data, n * sizeof(float), 1);
    if (rc)
        return process_fail(t, rc, "center-of-rotation: %s", ufomi_last_error());
    return cor_compute(t, p->stage, &in[0]->req);
}

static void cor_finalize(UfomiTask *t)
{
    CorPriv *p = t->priv;
    if (p)
        free(p->stage);
}

/* ---------------------------------------------------- stacked-backproject */

/* stacked-backproject (ufo-stacked-backproject-task.c:612-690, init :693-718:
 * axis-pos and angle-step are -1 after init; precision_values :37-42) */
static const char *const precision_nicks[] = {"int8", "half", "single", NULL};
static const PropSpec sbp_props[] = {
    {"num-projections", PT_UINT, 0, 32768, 0, NULL, 0, NULL},
    {"offset", PT_UINT, 0, 32768, 0, NULL, 0, NULL},
    {"axis-pos", PT_DOUBLE, -1.0, 32768.0, -1.0, NULL, 0, NULL},
    {"angle-step", PT_DOUBLE, -DBL_MAX, DBL_MAX, -1.0, NULL, 0, NULL},
    {"angle-offset", PT_DOUBLE, 0.0, DBL_MAX, 0.0, NULL, 0, NULL},
    {"precision-mode", PT_ENUM, 0, 2, 2, precision_nicks, 0, NULL},
    {"roi-x", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"roi-y", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"roi-width", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"roi-height", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
};
enum { SBP_NPROJ, SBP_OFFSET, SBP_AXIS, SBP_STEP, SBP_AOFF, SBP_PREC, SBP_RX, SBP_RY, SBP_RW, SBP_RH };

typedef struct {
    unsigned n_projections, burst;
} SbpPriv;

/* get_requisition (:186-240): num-projections defaults to the first stack's
 * sinogram height and stays; the output is [depth][roi-height][roi-width] */
static int sbp_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    SbpPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    if (!p->n_projections)
        p->n_projections = (unsigned) t->v[SBP_NPROJ].num ? (unsigned) t->v[SBP_NPROJ].num : (unsigned) ir->dims[1];
    p->burst = (unsigned) ir->dims[1];
    if (p->burst > p->n_projections)
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_PROJECTIONS,
                          "Total number of projections (%u) must be greater than or equal to sinogram height (%u)",
                          p->n_projections, p->burst);
    req->n_dims = 3;
    req->dims[0] = t->v[SBP_RW].num ? (size_t) t->v[SBP_RW].num : ir->dims[0];
    req->dims[1] = t->v[SBP_RH].num ? (size_t) t->v[SBP_RH].num : ir->dims[0];
    req->dims[2] = ir->n_dims == 3 ? ir->dims[2] : 1;
    return UFOMI_OK;
}

/* process (:300-430): every slice of the stack in the chosen precision */
static int sbp_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    SbpPriv *p = t->priv;
    if (in[0]->dtype != UFOMI_F32)
        return process_fail(t, UFOMI_ERR_UNSUPPORTED, "stacked-backproject: input must be float32");
    ufomi_bp_params bp;
    ufomi_bp_params_init(&bp);
    bp.num_projections = p->n_projections;
    bp.offset = (unsigned) t->v[SBP_OFFSET].num;
    bp.axis_pos = t->v[SBP_AXIS].num;
    bp.angle_step = t->v[SBP_STEP].num;
    bp.angle_offset = t->v[SBP_AOFF].num;
    bp.roi_x = (unsigned) t->v[SBP_RX].num;
    bp.roi_y = (unsigned) t->v[SBP_RY].num;
    bp.roi_width = (unsigned) req->dims[0];
    bp.roi_height = (unsigned) req->dims[1];
    const UfomiRequisition *ir = &in[0]->req;
    int rc = ufomi_bp_stacked(t->ctx, in[0]->data, (int) ir->dims[0], (int) ir->dims[1], (int) req->dims[2], &bp,
                              (int) t->v[SBP_PREC].num, out->data);
    if (rc)
        return process_fail(t, rc, "stacked-backproject: %s", ufomi_last_error());
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

/* -------------------------------------------------------- cone weighting */

typedef struct {
    unsigned count;
} CwPriv;

static int cw_setup(UfomiTask *t, UfomiError *e)
{
    static const int need[] = {CW_CX, CW_CZ, CW_SD, CW_DD};
    for (int k = 0; k < 4; k++)
        if (t->v[need[k]].n == 0)
            return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "cone-beam-projection-weight: %s must be set",
                              t->cls->props[need[k]].name);
    return UFOMI_OK;
}

static int cw_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    (void) t;
    (void) e;
    *req = in[0]->req;
    return UFOMI_OK;
}

static int cw_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    CwPriv *p = t->priv;
    const unsigned i = p->count;
    double vals[5];
    for (int k = 0; k < 5; k++) {
        const PropValue *v = &t->v[k];
        if (v->n != 1 && (int) i >= v->n)
            return process_fail(t, UFOMI_ERR_ARG, "cone-beam-projection-weight: %s has no value for projection %u",
                                t->cls->props[k].name, i);
        vals[k] = arr_at(v, i);
    }
    int rc = ufomi_cone_weight(t->ctx, in[0]->data, out->data, (int) req->dims[0], (int) req->dims[1], 1, &vals[0],
                               &vals[1], &vals[2], &vals[3], &vals[4], 1);
    if (rc)
        return process_fail(t, rc, "cone-beam-projection-weight: %s", ufomi_last_error());
    p->count++;
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

/* ---------------------------------------------------- general-backproject */

typedef struct {
    int width, height;
    unsigned count, generated;
    float *d_stack;       /* num_projections projections [H][W] */
    void *d_volume;       /* [nz][ny][nx] in the store type */
    int nx, ny, nz, reconstructed;
    ufomi_gbp_params params;
} GbpPriv;

static size_t store_size(int st)
{
    static const size_t sz[] = {2, 4, 8, 1, 2, 4};
    return sz[st];
}

int gbp_task_store_type(const UfomiTask *t) { return (int) t->v[G_STORE].num; }

static void gbp_fill_params(const UfomiTask *t, ufomi_gbp_params *p)
{
    ufomi_gbp_params_init(p);
    p->num_projections = (unsigned) t->v[G_NPROJ].num;
    p->overall_angle = t->v[G_OVERALL].num;
    p->z = t->v[G_Z].num;
    for (int k = 0; k < 3; k++) {
        p->region[k] = t->v[G_REGION].n > k ? t->v[G_REGION].arr[k] : 0.0;
        p->x_region[k] = t->v[G_XREGION].n > k ? t->v[G_XREGION].arr[k] : 0.0;
        p->y_region[k] = t->v[G_YREGION].n > k ? t->v[G_YREGION].arr[k] : 0.0;
    }
    p->store_type = (int) t->v[G_STORE].num;
    static const int addr_map[] = {UFOMI_ADDR_NONE, UFOMI_ADDR_CLAMP_TO_EDGE, UFOMI_ADDR_CLAMP, UFOMI_ADDR_REPEAT,
                                   UFOMI_ADDR_MIRRORED_REPEAT};
    p->addressing_mode = addr_map[(int) t->v[G_ADDR].num];
    p->gray_map_min = t->v[G_GMIN].num;
    p->gray_map_max = t->v[G_GMAX].num;
    p->compute_type = (int) t->v[G_COMPUTE].num;
    p->result_type = (int) t->v[G_RESULT].num;
    p->burst = (unsigned) t->v[G_BURST].num;
}

static void gbp_fill_geometry(const UfomiTask *t, ufomi_gbp_geometry *g)
{
    ufomi_gbp_geometry_init(g);
#define GEO(field, idx)                        \
    g->field = t->v[idx].arr;                  \
    g->n_##field = t->v[idx].n;
    GEO(axis_angle_x, G_AX) GEO(axis_angle_y, G_AY) GEO(axis_angle_z, G_AZ)
    GEO(volume_angle_x, G_VX) GEO(volume_angle_y, G_VY) GEO(volume_angle_z, G_VZ)
    GEO(detector_angle_x, G_DAX) GEO(detector_angle_y, G_DAY) GEO(detector_angle_z, G_DAZ)
    GEO(center_position_x, G_CX) GEO(center_position_z, G_CZ)
    GEO(source_position_x, G_SX) GEO(source_position_y, G_SY) GEO(source_position_z, G_SZ)
    GEO(detector_position_x, G_DX) GEO(detector_position_y, G_DY) GEO(detector_position_z, G_DZ)
#undef GEO
}

static int gbp_setup(UfomiTask *t, UfomiError *e)
{
    /* ufo-general-backproject-task.c:1429-1453 */
    if (!(unsigned) t->v[G_NPROJ].num)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_NUM_PROJECTIONS, "Number of projections not set");
    const int st = (int) t->v[G_STORE].num;
    if (t->v[G_GMIN].num >= t->v[G_GMAX].num && (st == UFOMI_ST_UCHAR || st == UFOMI_ST_USHORT || st == UFOMI_ST_UINT))
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_GRAY_MAP,
                          "Gray mapping minimum must be less then the maximum");
    /* compute-type double, result-type half / double and the repeat addressing
     * modes run on the precise kernel of ufomi_gbp (FP64 maps and weights,
     * per-projection rounding of the running sum to the result type, OpenCL's
     * normalised repeat rules applied to u / W).  A swept axis-angle-z follows
     * ufomi_gbp_sweep: an error without per-projection geometry arrays, every
     * slice the z reconstruction with them. */
    return UFOMI_OK;
}

/* Whether the reference generates its vectorized kernel: some geometry array
 * other than axis-angle-z holds one value per projection (:1321-1336). */
static int gbp_vectorized(const UfomiTask *t)
{
    static const int arrays[] = {G_AX, G_AY, G_VX, G_VY, G_VZ, G_DAX, G_DAY, G_DAZ, G_DX, G_DY, G_DZ,
                                 G_SX, G_SY, G_SZ, G_CX, G_CZ};
    const int n = (int) (unsigned) t->v[G_NPROJ].num;
    for (size_t k = 0; k < sizeof(arrays) / sizeof(arrays[0]); k++)
        if (n > 0 && t->v[arrays[k]].n == n)
            return 1;
    return 0;
}

static int gbp_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    GbpPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    /* parameter axis-angle-z with the scalar kernel: `axis_angle_z = region[idz]`
     * names an undeclared variable, the kernel does not build (:1509-1516) */
    if ((int) t->v[G_PARAM].num == 2 && !gbp_vectorized(t))
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_UNSUPPORTED,
                          "Error creating backprojection kernels");
    if (!p->d_stack) {
        gbp_fill_params(t, &p->params);
        int rc = ufomi_gbp_volume_shape(&p->params, (int) ir->dims[0], &p->nx, &p->ny, &p->nz);
        if (rc)
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, rc, "%s", ufomi_last_error());
        p->width = (int) ir->dims[0];
        p->height = ir->n_dims >= 2 ? (int) ir->dims[1] : 1;
        const size_t proj_bytes = (size_t) p->params.num_projections * p->width * p->height * sizeof(float);
        const size_t vol_bytes = (size_t) p->nx * p->ny * p->nz * store_size(p->params.store_type);
        size_t free_b = 0, total_b = 0;
        ufomi_device_memory(t->ctx, &free_b, &total_b);
        /* :1542-1546 */
        if (proj_bytes + vol_bytes > total_b)
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_MEMORY, "Volume size doesn't fit to memory");
        if ((rc = ufomi_device_alloc(t->ctx, proj_bytes, (void **) &p->d_stack)) ||
            (rc = ufomi_device_alloc(t->ctx, vol_bytes, &p->d_volume)))
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_MEMORY, "Volume size doesn't fit to memory");
    } else if ((int) ir->dims[0] != p->width || (ir->n_dims >= 2 ? (int) ir->dims[1] : 1) != p->height) {
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                          "general-backproject: projection size changed");
    }
    req->n_dims = 2;
    req->dims[0] = (size_t) p->nx;
    req->dims[1] = (size_t) p->ny;
    req->dims[2] = 1;
    return UFOMI_OK;
}

static int gbp_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) out;
    (void) req;
    GbpPriv *p = t->priv;
    if (p->count >= p->params.num_projections)
        return process_fail(t, UFOMI_ERR_ARG, "general-backproject: more projections than num-projections");
    const size_t plane = (size_t) p->width * p->height;
    int rc = ufomi_memcpy(t->ctx, p->d_stack + p->count * plane, in[0]->data, plane * sizeof(float), 2);
    if (rc)
        return process_fail(t, rc, "general-backproject: %s", ufomi_last_error());
    p->count++;
    return 1;
}

static int gbp_generate(UfomiTask *t, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    GbpPriv *p = t->priv;
    /* :1733-1742 */
    if (p->count != p->params.num_projections) {
        process_fail(t, UFOMI_ERR_INCOMPLETE,
                     "general-backproject received only %u projections out of %u specified, no outuput will be generated",
                     p->count, p->params.num_projections);
        fprintf(stderr, "warning: %s\n", t->err);
        return 0;
    }
    if (p->generated >= (unsigned) p->nz)
        return 0;
    if (!p->reconstructed) {
        ufomi_gbp_geometry g;
        gbp_fill_geometry(t, &g);
        int rc = ufomi_gbp_sweep(t->ctx, p->d_stack, p->width, p->height, (int) p->count, &g, &p->params,
                                 (int) t->v[G_PARAM].num, p->d_volume);
        if (rc)
            return process_fail(t, rc, "general-backproject: %s", ufomi_last_error());
        p->reconstructed = 1;
    }
    const size_t slice = (size_t) p->nx * p->ny * store_size(p->params.store_type);
    int rc = ufomi_memcpy(t->ctx, out->data, (char *) p->d_volume + p->generated * slice, slice, 2);
    if (rc)
        return process_fail(t, rc, "general-backproject: %s", ufomi_last_error());
    p->generated++;
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

static void gbp_finalize(UfomiTask *t)
{
    GbpPriv *p = t->priv;
    if (p && t->ctx) {
        if (p->d_stack)
            ufomi_device_free(t->ctx, p->d_stack);
        if (p->d_volume)
            ufomi_device_free(t->ctx, p->d_volume);
    }
}

/* ----------------------------------------------------- lamino-backproject */

static const char *const lamino_parameter_nicks[] = {"z", "x-center", "lamino-angle", "roll-angle", NULL};

/* lamino-backproject (ufo-lamino-backproject-task.c:660-790, init :796-831) */
static const PropSpec lbp_props[] = {
    {"x-region", PT_DARRAY, -2147483648.0, 2147483647.0, 0, NULL, 3, NULL},
    {"y-region", PT_DARRAY, -2147483648.0, 2147483647.0, 0, NULL, 3, NULL},
    {"z", PT_FLOAT, -FLT_MAX, FLT_MAX, 0.0, NULL, 0, NULL},
    {"region", PT_DARRAY, -FLT_MAX, FLT_MAX, 0, NULL, 3, NULL},
    {"projection-offset", PT_DARRAY, -2147483648.0, 2147483647.0, 0, NULL, 2, NULL},
    {"center", PT_DARRAY, -FLT_MAX, FLT_MAX, 0, NULL, 2, NULL},
    {"num-projections", PT_UINT, 0, 32768, 0, NULL, 0, NULL},
    {"overall-angle", PT_FLOAT, -FLT_MAX, FLT_MAX, M_PI, NULL, 0, NULL},
    {"tomo-angle", PT_FLOAT, -FLT_MAX, FLT_MAX, -FLT_MAX, NULL, 0, NULL},
    {"lamino-angle", PT_FLOAT, -FLT_MAX, FLT_MAX, 0.0, NULL, 0, NULL},
    {"roll-angle", PT_FLOAT, -FLT_MAX, FLT_MAX, 0.0, NULL, 0, NULL},
    {"parameter", PT_ENUM, 0, 3, 0, lamino_parameter_nicks, 0, NULL},
    {"addressing-mode", PT_ENUM, 0, 4, 2, addr_nicks, 0, NULL},
};
enum { L_XREGION, L_YREGION, L_Z, L_REGION, L_OFFSET, L_CENTER, L_NPROJ, L_OVERALL, L_TOMO, L_LAMINO, L_ROLL,
       L_PARAM, L_ADDR };

typedef struct {
    int width, height;
    unsigned count, generated;
    float *d_stack;       /* num_projections projections [H][W] */
    ufomi_lamino_params params;
    int nx, ny, nz;
} LbpPriv;

/* Element k of a value array, 0 past its end (no broadcasting). */
static double elem_at(const PropValue *v, int k) { return v->n > k ? v->arr[k] : 0.0; }

static void lbp_fill_params(const UfomiTask *t, ufomi_lamino_params *p)
{
    ufomi_lamino_params_init(p);
    p->num_projections = (unsigned) t->v[L_NPROJ].num;
    p->overall_angle = (float) t->v[L_OVERALL].num;
    p->tomo_angle = (float) t->v[L_TOMO].num;
    p->lamino_angle = (float) t->v[L_LAMINO].num;
    p->roll_angle = (float) t->v[L_ROLL].num;
    p->z = (float) t->v[L_Z].num;
    for (int k = 0; k < 3; k++) {
        p->x_region[k] = (int) elem_at(&t->v[L_XREGION], k);
        p->y_region[k] = (int) elem_at(&t->v[L_YREGION], k);
        p->region[k] = (float) elem_at(&t->v[L_REGION], k);
    }
    for (int k = 0; k < 2; k++) {
        p->center[k] = (float) elem_at(&t->v[L_CENTER], k);
        p->projection_offset[k] = (int) elem_at(&t->v[L_OFFSET], k);
    }
    p->parameter = (int) t->v[L_PARAM].num;
    static const int addr_map[] = {UFOMI_ADDR_NONE, UFOMI_ADDR_CLAMP_TO_EDGE, UFOMI_ADDR_CLAMP, UFOMI_ADDR_REPEAT,
                                   UFOMI_ADDR_MIRRORED_REPEAT};
    p->addressing_mode = addr_map[(int) t->v[L_ADDR].num];
}

/* set_region (:124-136): a region whose "from" exceeds "to" is ignored with a
 * warning; returns 0 then, 1 to store the value. */
static int lamino_region_property(const char *name, const double *vals, int n)
{
    if (n >= 2 && (int) vals[0] > (int) vals[1]) {
        fprintf(stderr, "warning: Invalid region [\"from\", \"to\", \"step\"]: [%d, %d, %d], \"from\" has to be "
                        "less than or equal to \"to\" (%s)\n", (int) vals[0], (int) vals[1], n > 2 ? (int) vals[2] : 0,
                name);
        return 0;
    }
    return 1;
}

static int lbp_setup(UfomiTask *t, UfomiError *e)
{
    /* :174-182 */
    if (!(unsigned) t->v[L_NPROJ].num)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_NUM_PROJECTIONS, "Number of projections has not been set");
    if ((float) elem_at(&t->v[L_REGION], 2) == 0.0f)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "Step in region is 0");
    const int addr = (int) t->v[L_ADDR].num;
    if (addr == 3 || addr == 4)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_UNSUPPORTED,
                          "lamino-backproject: repeat addressing needs normalized coordinates");
    return UFOMI_OK;
}

static int lbp_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    LbpPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    if (!p->d_stack) {
        lbp_fill_params(t, &p->params);
        int rc = ufomi_lamino_volume_shape(&p->params, &p->nx, &p->ny, &p->nz);
        if (rc)
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, rc, "%s", ufomi_last_error());
        p->width = (int) ir->dims[0];
        p->height = ir->n_dims >= 2 ? (int) ir->dims[1] : 1;
        const size_t proj_bytes = (size_t) p->params.num_projections * p->width * p->height * sizeof(float);
        if ((rc = ufomi_device_alloc(t->ctx, proj_bytes, (void **) &p->d_stack)))
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_MEMORY, "%s", ufomi_last_error());
    } else if ((int) ir->dims[0] != p->width || (ir->n_dims >= 2 ? (int) ir->dims[1] : 1) != p->height) {
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                          "lamino-backproject: projection size changed");
    }
    /* :247-262: one volume [nz][ny][nx] */
    req->n_dims = 3;
    req->dims[0] = (size_t) p->nx;
    req->dims[1] = (size_t) p->ny;
    req->dims[2] = (size_t) p->nz;
    return UFOMI_OK;
}

static int lbp_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) out;
    (void) req;
    LbpPriv *p = t->priv;
    if (p->count >= p->params.num_projections)
        return process_fail(t, UFOMI_ERR_ARG, "lamino-backproject: more projections than num-projections");
    const size_t plane = (size_t) p->width * p->height;
    int rc = ufomi_memcpy(t->ctx, p->d_stack + p->count * plane, in[0]->data, plane * sizeof(float), 2);
    if (rc)
        return process_fail(t, rc, "lamino-backproject: %s", ufomi_last_error());
    p->count++;
    return 1;
}

static int lbp_generate(UfomiTask *t, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    LbpPriv *p = t->priv;
    /* :480-496: the volume, once */
    if (p->generated)
        return 0;
    /* process (:353, :430-475) runs its 16-projection kernel only on a full
     * burst below num_projections / 16 * 16, the 1-projection kernel from
     * there on: short of that, a trailing partial burst never reaches the
     * volume. */
    const unsigned full = p->params.num_projections / 16 * 16;
    const unsigned used = p->count < full ? p->count / 16 * 16 : p->count;
    if (!used) {
        fprintf(stderr, "warning: lamino-backproject received %u projections, fewer than one burst of 16 out of "
                        "%u: no volume\n", p->count, p->params.num_projections);
        return 0;
    }
    int rc = ufomi_lamino_bp(t->ctx, p->d_stack, p->width, p->height, (int) used, &p->params,
                             (float *) out->data);
    if (rc)
        return process_fail(t, rc, "lamino-backproject: %s", ufomi_last_error());
    p->generated = 1;
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

static void lbp_finalize(UfomiTask *t)
{
    LbpPriv *p = t->priv;
    if (p && t->ctx && p->d_stack)
        ufomi_device_free(t->ctx, p->d_stack);
}

/* Output items and element size of a generic reductor stage. */
int reductor_output_items(const UfomiTask *t, const UfomiRequisition *in, int *items, size_t *elem_size,
                          int *dtype)
{
    static const size_t st_size[] = {2, 4, 8, 1, 2, 4};
    *items = 1;
    *elem_size = sizeof(float);
    *dtype = UFOMI_F32;
    if (!strcmp(t->cls->name, "general-backproject")) {
        ufomi_gbp_params gp;
        ufomi_gbp_params_init(&gp);
        for (int k = 0; k < 3; k++)
            gp.region[k] = elem_at(&t->v[G_REGION], k);
        int nx, ny, nz = 0;
        int rc = ufomi_gbp_volume_shape(&gp, (int) in->dims[0], &nx, &ny, &nz);
        if (rc)
            return rc;
        const int st = (int) t->v[G_STORE].num;
        *items = nz;
        *elem_size = st_size[st];
        *dtype = st == UFOMI_ST_FLOAT ? UFOMI_F32 : (st == UFOMI_ST_HALF ? UFOMI_F16 : -1 - st);
    }
    return UFOMI_OK;
}

/* --------------------------------------------------- transpose-projections */

/* transpose-projections (ufo-transpose-projections-task.c:256-261) */
static const PropSpec tp_props[] = {
    {"number", PT_UINT, 1, U32MAX, 1, NULL, 0, NULL},
};

typedef struct {
    float *d_sino;         /* [height][number][width] */
    unsigned width, height, number, count, generated;
    int zeroed;
} TpPriv;

static void tp_free(UfomiTask *t)
{
    TpPriv *p = t->priv;
    if (p->d_sino && t->ctx)
        ufomi_device_free(t->ctx, p->d_sino);
    p->d_sino = NULL;
}

static int tp_setup(UfomiTask *t, UfomiError *e)
{
    (void) e;
    tp_free(t);
    memset(t->priv, 0, sizeof(TpPriv));
    return UFOMI_OK;
}

/* get_requisition (:128-150): the first projection fixes width and height;
 * the output is one sinogram of `number` rows */
static int tp_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    TpPriv *p = t->priv;
    const UfomiRequisition *ir = &in[0]->req;
    const unsigned W = (unsigned) ir->dims[0], H = ir->n_dims >= 2 ? (unsigned) ir->dims[1] : 1;
    if (!p->d_sino) {
        p->number = (unsigned) t->v[0].num;
        const size_t bytes = (size_t) H * p->number * W * sizeof(float);
        int rc = ufomi_device_alloc(t->ctx, bytes, (void **) &p->d_sino);
        if (rc)
            return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_MEMORY,
                              "transpose-projections: %zu bytes of sinograms do not fit: %s", bytes,
                              ufomi_last_error());
        p->width = W;
        p->height = H;
        p->count = p->generated = 0;
        p->zeroed = 0;
    } else if (W != p->width || H != p->height) {
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                          "transpose-projections: projection size changed");
    }
    req->n_dims = 2;
    req->dims[0] = W;
    req->dims[1] = p->number;
    req->dims[2] = 1;
    return UFOMI_OK;
}

/* process (:64-99): the rows of one projection into the sinograms; a
 * projection beyond `number` ends the input (process returns FALSE) */
static int tp_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) out;
    (void) req;
    TpPriv *p = t->priv;
    if (p->count >= p->number)
        return 0;
    if (in[0]->dtype != UFOMI_F32)
        return process_fail(t, UFOMI_ERR_UNSUPPORTED, "transpose-projections: input must be float32");
    int rc = ufomi_transpose_projections(t->ctx, in[0]->data, (int) p->width, (int) p->height, 1, (int) p->count,
                                         (int) p->number, p->d_sino);
    if (rc)
        return process_fail(t, rc, "transpose-projections: %s", ufomi_last_error());
    p->count++;
    return 1;
}

/* generate (:101-126): sinogram after sinogram; rows of projections that
 * never arrived are zero */
static int tp_generate(UfomiTask *t, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) req;
    TpPriv *p = t->priv;
    if (!p->d_sino || p->generated >= p->height)
        return 0;
    int rc;
    if (!p->zeroed && p->count < p->number &&
        (rc = ufomi_transpose_projections(t->ctx, NULL, (int) p->width, (int) p->height,
                                          (int) (p->number - p->count), (int) p->count, (int) p->number,
                                          p->d_sino)))
        return process_fail(t, rc, "transpose-projections: %s", ufomi_last_error());
    p->zeroed = 1;
    const size_t sino = (size_t) p->number * p->width;
    if ((rc = ufomi_memcpy(t->ctx, out->data, p->d_sino + p->generated * sino, sino * sizeof(float), 2)))
        return process_fail(t, rc, "transpose-projections: %s", ufomi_last_error());
    p->generated++;
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

static void tp_finalize(UfomiTask *t)
{
    if (t->priv)
        tp_free(t);
}

/* ------------------------------------------------------------------- read */

/* read (ufo-read-task.c:551-663; init :671-700 sets path "."; type_values :53-63
 * without hdf5, which needs libhdf5) */
#define ULMAX 18446744073709551615.0
static const char *const read_type_nicks[] = {"edf", "raw", "tiff", "unspecified", NULL};
static const PropSpec read_props[] = {
    {"path", PT_STRING, 0, 0, 0, NULL, 0, "."},
    {"start", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"image-start", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"number", PT_UINT, 0, U32MAX, U32MAX, NULL, 0, NULL},
    {"step", PT_UINT, 1, U32MAX, 1, NULL, 0, NULL},
    {"y", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"height", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"y-step", PT_UINT, 1, U32MAX, 1, NULL, 0, NULL},
    {"image-step", PT_UINT, 1, U32MAX, 1, NULL, 0, NULL},
    {"convert", PT_BOOL, 0, 1, 1, NULL, 0, NULL},
    {"raw-width", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"raw-height", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"raw-bitdepth", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"raw-pre-offset", PT_UINT, 0, ULMAX, 0, NULL, 0, NULL},
    {"raw-post-offset", PT_UINT, 0, ULMAX, 0, NULL, 0, NULL},
    {"type", PT_ENUM, 0, 3, 3, read_type_nicks, 0, NULL},
};
enum { RD_PATH, RD_START, RD_IMAGE_START, RD_NUMBER, RD_STEP, RD_Y, RD_HEIGHT, RD_YSTEP, RD_IMAGE_STEP, RD_CONVERT,
       RD_RAW_W, RD_RAW_H, RD_RAW_BITS, RD_RAW_PRE, RD_RAW_POST, RD_TYPE };
enum { RT_EDF, RT_RAW, RT_TIFF, RT_UNSPECIFIED };

typedef struct {
    char **files;
    size_t n_files, cur;
    int single, open, done;
    ImgFile f;
    unsigned image_start;   /* images still to skip when the next file opens */
    unsigned roi_height;    /* `height', resolved against the first file */
    unsigned index;         /* next image of the open file */
    unsigned current;       /* images generated */
    float *stage;
    size_t stage_cap;
} ReadPriv;

static unsigned uprop(const UfomiTask *t, int i) { return (unsigned) t->v[i].num; }

/* get_reader (ufo-read-task.c:225-245): the extension or `type', TIFF first, then EDF, then raw */
static int read_format(const UfomiTask *t, const char *name)
{
    const int ext = img_format_of(name), type = (int) t->v[RD_TYPE].num;
    if (ext == IMG_TIFF || type == RT_TIFF)
        return IMG_TIFF;
    if (ext == IMG_EDF || type == RT_EDF)
        return IMG_EDF;
    if (ext == IMG_RAW || type == RT_RAW)
        return IMG_RAW;
    return IMG_NONE;
}

static void read_reset(ReadPriv *p)
{
    if (p->open)
        img_close(&p->f);
    p->open = 0;
    for (size_t i = 0; i < p->n_files; i++)
        free(p->files[i]);
    free(p->files);
    p->files = NULL;
    p->n_files = 0;
}

static int cmp_names(const void *a, const void *b) { return strcmp(*(char *const *) a, *(char *const *) b); }

/* read_filenames + setup (ufo-read-task.c:139-223): a regular file, or a glob
 * (a directory means every file in it), filtered by reader, sorted */
static int read_setup(UfomiTask *t, UfomiError *e)
{
    ReadPriv *p = t->priv;
    read_reset(p);
    const char *path = t->v[RD_PATH].str;
    struct stat st;
    p->single = stat(path, &st) == 0 && S_ISREG(st.st_mode);
    const size_t len = strlen(path);
    char *pattern = malloc(len + 3);
    if (p->single || strchr(path, '*'))
        memcpy(pattern, path, len + 1);
    else
        snprintf(pattern, len + 3, "%s%s*", path, len && path[len - 1] == '/' ? "" : "/");
    glob_t g;
    memset(&g, 0, sizeof(g));
    if (glob(pattern, GLOB_MARK | GLOB_TILDE, NULL, &g) == 0) {
        p->files = calloc(g.gl_pathc + 1, sizeof(char *));
        for (size_t i = 0; i < g.gl_pathc; i++)
            if (read_format(t, g.gl_pathv[i]) != IMG_NONE)
                p->files[p->n_files++] = strdup(g.gl_pathv[i]);
    }
    globfree(&g);
    free(pattern);
    if (!p->n_files)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "`%s' does not match any files", path);
    qsort(p->files, p->n_files, sizeof(char *), cmp_names);
    p->cur = p->single ? 0 : uprop(t, RD_START);
    if (p->cur >= p->n_files)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "start=%u skips too many files",
                          uprop(t, RD_START));
    p->image_start = uprop(t, RD_IMAGE_START);
    p->roi_height = uprop(t, RD_HEIGHT);
    p->index = p->current = 0;
    p->done = 0;
    return UFOMI_OK;
}

/* get_requisition (ufo-read-task.c:247-325): the next file with an image at
 * the running image-start, the vertical ROI checked against it */
static int read_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    (void) in;
    ReadPriv *p = t->priv;
    const unsigned y = uprop(t, RD_Y);
    if (!p->open || p->index >= p->f.n_images) {
        for (;;) {
            if (p->open) {
                img_close(&p->f);
                p->open = 0;
                p->cur += uprop(t, RD_STEP);
            }
            if (p->cur >= p->n_files) {
                p->done = 1;
                return UFOMI_OK;
            }
            const char *name = p->files[p->cur];
            const ImgRawSpec raw = {uprop(t, RD_RAW_W), uprop(t, RD_RAW_H), uprop(t, RD_RAW_BITS),
                                    (unsigned long) t->v[RD_RAW_PRE].num, (unsigned long) t->v[RD_RAW_POST].num};
            char msg[320];
            if (img_open(name, read_format(t, name), &raw, &p->f, msg, sizeof msg))
                return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_IO, "%s", msg);
            p->open = 1;
            if (y >= p->f.height)
                return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                                  "read: vertical ROI start %u >= height %u", y, p->f.height);
            if (!p->roi_height)
                p->roi_height = p->f.height - y;
            else if ((unsigned long long) y + p->roi_height > p->f.height)
                return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG,
                                  "read: vertical ROI start + height %llu >= height %u",
                                  (unsigned long long) y + p->roi_height, p->f.height);
            if (p->image_start >= p->f.n_images) {
                p->image_start -= p->f.n_images;
                continue;
            }
            p->index = p->image_start;
            p->image_start = 0;
            break;
        }
    }
    req->n_dims = 2;
    req->dims[0] = p->f.width;
    req->dims[1] = (p->roi_height - 1) / uprop(t, RD_YSTEP) + 1;
    req->dims[2] = 1;
    return UFOMI_OK;
}

/* generate (ufo-read-task.c:341-363) with the readers' image stepping
 * (ufo-tiff-reader.c:188-216, ufo-edf-reader.c:160-166, ufo-raw-reader.c:150-156):
 * one image read, then image-step images consumed or the rest of the file */
int read_task_generate_host(UfomiTask *t, float *dst, const UfomiRequisition *req)
{
    ReadPriv *p = t->priv;
    if (p->current == uprop(t, RD_NUMBER) || p->done || !p->open || p->index >= p->f.n_images)
        return 0;
    const unsigned ystep = uprop(t, RD_YSTEP), rows = (p->roi_height - 1) / ystep + 1;
    if (req->dims[0] != p->f.width || req->dims[1] != rows)
        return process_fail(t, UFOMI_ERR_ARG, "read: requisition does not match the open file");
    if (!t->v[RD_CONVERT].num && !(p->f.kind == 'f'))
        return process_fail(t, UFOMI_ERR_ARG,
                            "read: convert=false keeps %d-bit integer samples, but buffers here are float32",
                            8 * p->f.bytes);
    char msg[320];
    if (img_read(&p->f, p->index, uprop(t, RD_Y), ystep, rows, dst, msg, sizeof msg))
        return process_fail(t, UFOMI_ERR_IO, "read: %s", msg);
    const unsigned image_step = uprop(t, RD_IMAGE_STEP), left = p->f.n_images - p->index;
    const unsigned consumed = image_step < left ? image_step : left;
    p->index += consumed;
    p->image_start = image_step - consumed;
    p->current++;
    return 1;
}

static int read_generate(UfomiTask *t, UfomiBuffer *out, UfomiRequisition *req)
{
    ReadPriv *p = t->priv;
    const size_t n = req->dims[0] * req->dims[1];
    if (n > p->stage_cap) {
        free(p->stage);
        p->stage = malloc(n * sizeof(float));
        p->stage_cap = p->stage ? n : 0;
        if (!p->stage)
            return process_fail(t, UFOMI_ERR_MEMORY, "read: out of host memory");
    }
    if (!read_task_generate_host(t, p->stage, req))
        return 0;
    int rc = ufomi_memcpy(t->ctx, out->data, p->stage, n * sizeof(float), 0);
    if (rc)
        return process_fail(t, rc, "read: %s", ufomi_last_error());
    out->layout = UFOMI_LAYOUT_REAL;
    return 1;
}

static void read_finalize(UfomiTask *t)
{
    ReadPriv *p = t->priv;
    if (p) {
        read_reset(p);
        free(p->stage);
    }
}

/* ------------------------------------------------------------------ write */

/* write (ufo-write-task.c:606-690, init :694-720); jpeg-quality exists only
 * with libjpeg and HDF5 output needs libhdf5, neither of which is built here */
static const PropSpec write_props[] = {
    {"filename", PT_STRING, 0, 0, 0, NULL, 0, ""},
    {"counter-start", PT_UINT, 0, U32MAX, 0, NULL, 0, NULL},
    {"counter-step", PT_UINT, 1, U32MAX, 1, NULL, 0, NULL},
    {"bytes-per-file", PT_UINT, 0, ULMAX, 137438953472.0, NULL, 0, NULL},
    {"append", PT_BOOL, 0, 1, 0, NULL, 0, NULL},
    {"bits", PT_UINT, 8, 32, 32, NULL, 0, NULL},
    {"minimum", PT_FLOAT, -FLT_MAX, FLT_MAX, FLT_MAX, NULL, 0, NULL},
    {"maximum", PT_FLOAT, -FLT_MAX, FLT_MAX, -FLT_MAX, NULL, 0, NULL},
    {"rescale", PT_BOOL, 0, 1, 1, NULL, 0, NULL},
    {"tiff-bigtiff", PT_BOOL, 0, 1, 1, NULL, 0, NULL},
};
enum { WR_FILENAME, WR_CSTART, WR_CSTEP, WR_BPF, WR_APPEND, WR_BITS, WR_MIN, WR_MAX, WR_RESCALE, WR_BIGTIFF };

typedef struct {
    ImgWriter w;
    int opened, n_fmt, format;
    unsigned counter;
    double written;
    /* bits per sample of the bytes-per-file accounting: an explicit bits=32
     * counts 16 there (ufo-write-task.c:438-441), the default counts 32 */
    unsigned acct_bits;
    float *stage;
    size_t stage_cap;
} WritePriv;

/* set_property of bits (ufo-write-task.c:419-443): values other than 8, 16
 * and 32 are ignored with a warning; returns 0 then, 1 to store the value. */
static int write_bits_property(UfomiTask *t, double bits)
{
    WritePriv *p = t->priv;
    if (bits != 8 && bits != 16 && bits != 32) {
        fprintf(stderr, "warning: Write::bits can only 8, 16 or 32\n");
        return 0;
    }
    p->acct_bits = bits == 32 ? 16 : (unsigned) bits;
    return 1;
}

/* The one conversion allowed in `filename': %[flags][width][.precision] of d, i, o, u, x or X. */
static int int_format_ok(const char *s)
{
    const char *q = strchr(s, '%');
    if (!q)
        return 1;
    q++;
    q += strspn(q, "-+ #0");
    q += strspn(q, "0123456789");
    if (*q == '.') {
        q++;
        q += strspn(q, "0123456789");
    }
    return *q && strchr("diouxX", *q);
}

static void write_name(const UfomiTask *t, const WritePriv *p, char *buf, size_t len)
{
    const char *fmt = t->v[WR_FILENAME].str;
    if (p->n_fmt)
        snprintf(buf, len, fmt, p->counter);
    else
        snprintf(buf, len, "%s", fmt);
}

static int exists(const char *path)
{
    struct stat st;
    return stat(path, &st) == 0;
}

static int mkdir_parents(const char *dir)
{
    char tmp[4096];
    snprintf(tmp, sizeof tmp, "%s", dir);
    for (char *c = tmp + 1; *c; c++) {
        if (*c != '/')
            continue;
        *c = '\0';
        if (!exists(tmp) && mkdir(tmp, 0755) && errno != EEXIST)
            return -1;
        *c = '/';
    }
    return !exists(tmp) && mkdir(tmp, 0755) && errno != EEXIST ? -1 : 0;
}

/* setup (ufo-write-task.c:167-278) */
static int write_setup(UfomiTask *t, UfomiError *e)
{
    WritePriv *p = t->priv;
    if (p->opened)
        img_writer_close(&p->w);
    p->opened = 0;
    p->written = 0;
    p->counter = 0;
    if (!p->acct_bits)
        p->acct_bits = 32;
    const char *fn = t->v[WR_FILENAME].str;
    if (!fn || !*fn) {  /* no filename: raw to stdout */
        p->format = IMG_RAW;
        p->n_fmt = 0;
        return UFOMI_OK;
    }
    p->n_fmt = 0;
    for (const char *c = fn; *c; c++)
        p->n_fmt += *c == '%';
    if (p->n_fmt > 1)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "`%s' has too many format specifiers", fn);
    if (!int_format_ok(fn))
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG,
                          "`%s' needs an integer format specifier such as %%i or %%05d", fn);
    if (!p->n_fmt && exists(fn) && access(fn, W_OK) < 0)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_IO, "Cannot access `%s': %s.", fn, strerror(errno));
    p->format = img_format_of(fn);
    if (p->format != IMG_TIFF && p->format != IMG_RAW)
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG,
                          "`%s' does not have a valid file extension or requires format specifiers", fn);
    char *copy = strdup(fn), *dir = dirname(copy);
    if (!exists(dir) && mkdir_parents(dir)) {
        task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_IO, "Could not create `%s'.", dir);
        free(copy);
        return UFOMI_ERR_IO;
    }
    free(copy);
    p->counter = uprop(t, WR_CSTART);
    if (t->v[WR_APPEND].num && p->n_fmt) {
        char name[4096];
        for (;;) {
            write_name(t, p, name, sizeof name);
            if (!exists(name))
                break;
            p->counter++;
        }
    }
    return UFOMI_OK;
}

static int write_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    (void) t;
    (void) in;
    (void) e;
    req->n_dims = 0;
    return UFOMI_OK;
}

/* process (ufo-write-task.c:301-392): each plane of the item to the open file,
 * a new file (counter += counter-step) once bytes-per-file would be exceeded */
int write_task_write_host(UfomiTask *t, const float *src, const UfomiRequisition *req)
{
    WritePriv *p = t->priv;
    const unsigned w = (unsigned) req->dims[0], h = req->n_dims >= 2 ? (unsigned) req->dims[1] : 1;
    const unsigned frames = req->n_dims == 3 ? (unsigned) req->dims[2] : 1;
    const double out_size = (double) w * h * p->acct_bits / 8;
    const char *fn = t->v[WR_FILENAME].str;
    char msg[320];
    for (unsigned i = 0; i < frames; i++) {
        if (!p->opened) {
            char name[4096];
            const char *path = NULL;
            if (fn && *fn) {
                for (;;) {
                    write_name(t, p, name, sizeof name);
                    if (!(exists(name) && access(name, W_OK) < 0))
                        break;
                    fprintf(stderr, "warning: Cannot access `%s': %s.\n", name, strerror(errno));
                    p->counter += uprop(t, WR_CSTEP);
                }
                path = name;
            }
            if (img_writer_open(&p->w, path, p->format, (int) t->v[WR_BIGTIFF].num, msg, sizeof msg)) {
                process_fail(t, UFOMI_ERR_IO, "write: %s", msg);
                return UFOMI_ERR_IO;
            }
            p->opened = 1;
        }
        if (img_writer_write(&p->w, src + (size_t) i * w * h, w, h, uprop(t, WR_BITS), (float) t->v[WR_MIN].num,
                             (float) t->v[WR_MAX].num, (int) t->v[WR_RESCALE].num, msg, sizeof msg)) {
            process_fail(t, UFOMI_ERR_IO, "write: %s", msg);
            return UFOMI_ERR_IO;
        }
        fflush(p->w.fp);
        p->written += out_size;
        if (p->n_fmt && p->written + out_size > t->v[WR_BPF].num) {
            img_writer_close(&p->w);
            p->opened = 0;
            p->written = 0;
            p->counter += uprop(t, WR_CSTEP);
        }
    }
    return UFOMI_OK;
}

static int write_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    (void) out;
    (void) req;
    WritePriv *p = t->priv;
    const UfomiBuffer *b = in[0];
    if (b->dtype != UFOMI_F32)
        return process_fail(t, UFOMI_ERR_UNSUPPORTED, "write: input items must be float32");
    const size_t n = ufomi_requisition_elements(&b->req);
    if (n > p->stage_cap) {
        free(p->stage);
        p->stage = malloc(n * sizeof(float));
        p->stage_cap = p->stage ? n : 0;
        if (!p->stage)
            return process_fail(t, UFOMI_ERR_MEMORY, "write: out of host memory");
    }
    int rc = ufomi_memcpy(t->ctx, p->stage, b->data, n * sizeof(float), 1);
    if (rc)
        return process_fail(t, rc, "write: %s", ufomi_last_error());
    return write_task_write_host(t, p->stage, &b->req) == UFOMI_OK;
}

void write_task_flush(UfomiTask *t)
{
    WritePriv *p = t->priv;
    if (p->opened)
        img_writer_close(&p->w);
    p->opened = 0;
}

static void write_finalize(UfomiTask *t)
{
    WritePriv *p = t->priv;
    if (p) {
        write_task_flush(t);
        free(p->stage);
    }
}

/* ---------------------------------------------------------------- classes */

#define NPROPS(a) ((int) (sizeof(a) / sizeof((a)[0])))

static const TaskClass classes[] = {
    {"fft", fft_props, NPROPS(fft_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU, sizeof(FftPriv), NULL,
     fft_requisition, fft_process, NULL, NULL},
    {"filter", filter_props, NPROPS(filter_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU,
     sizeof(FilterPriv), NULL, filter_requisition, filter_process, NULL, filter_finalize},
    {"ifft", ifft_props, NPROPS(ifft_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU, sizeof(IfftPriv), NULL,
     ifft_requisition, ifft_process, NULL, NULL},
    {"backproject", bp_props, NPROPS(bp_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU, sizeof(BpPriv), NULL,
     bp_requisition, bp_process, NULL, NULL},
    {"cone-beam-projection-weight", cw_props, NPROPS(cw_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU,
     sizeof(CwPriv), cw_setup, cw_requisition, cw_process, NULL, NULL},
    {"general-backproject", gbp_props, NPROPS(gbp_props), UFOMI_TASK_MODE_REDUCTOR | UFOMI_TASK_MODE_GPU,
     sizeof(GbpPriv), gbp_setup, gbp_requisition, gbp_process, gbp_generate, gbp_finalize},
    {"center-of-rotation", cor_props, NPROPS(cor_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_CPU,
     sizeof(CorPriv), NULL, cor_requisition, cor_process, NULL, cor_finalize},
    {"flat-field-correct", ffc_props, NPROPS(ffc_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU,
     sizeof(FfcPriv), NULL, ffc_requisition, ffc_process, NULL, NULL},
    {"average", avg_props, NPROPS(avg_props), UFOMI_TASK_MODE_REDUCTOR | UFOMI_TASK_MODE_GPU, sizeof(AvgPriv),
     avg_setup, avg_requisition, avg_process, avg_generate, avg_finalize},
    {"stacked-backproject", sbp_props, NPROPS(sbp_props), UFOMI_TASK_MODE_PROCESSOR | UFOMI_TASK_MODE_GPU,
     sizeof(SbpPriv), NULL, sbp_requisition, sbp_process, NULL, NULL},
    {"lamino-backproject", lbp_props, NPROPS(lbp_props), UFOMI_TASK_MODE_REDUCTOR | UFOMI_TASK_MODE_GPU,
     sizeof(LbpPriv), lbp_setup, lbp_requisition, lbp_process, lbp_generate, lbp_finalize},
    {"transpose-projections", tp_props, NPROPS(tp_props), UFOMI_TASK_MODE_REDUCTOR | UFOMI_TASK_MODE_GPU,
     sizeof(TpPriv), tp_setup, tp_requisition, tp_process, tp_generate, tp_finalize},
    {"read", read_props, NPROPS(read_props), UFOMI_TASK_MODE_GENERATOR | UFOMI_TASK_MODE_CPU, sizeof(ReadPriv),
     read_setup, read_requisition, NULL, read_generate, read_finalize},
    {"write", write_props, NPROPS(write_props), UFOMI_TASK_MODE_SINK | UFOMI_TASK_MODE_CPU, sizeof(WritePriv),
     write_setup, write_requisition, write_process, NULL, write_finalize},
};

UfomiTask *ufomi_task_new(const char *name)
{
    if (!name)
        return NULL;
    for (size_t c = 0; c < sizeof(classes) / sizeof(classes[0]); c++) {
        if (strcmp(classes[c].name, name))
            continue;
        UfomiTask *t = calloc(1, sizeof(UfomiTask));
        t->cls = &classes[c];
        t->priv = calloc(1, classes[c].priv_size);
        for (int i = 0; i < t->cls->n_props; i++)
            t->names[i] = t->cls->props[i].name;
        t->names[t->cls->n_props] = NULL;
        prop_reset(t);
        return t;
    }
    ufomi_status_fail(UFOMI_ERR_ARG, "no task named `%s'", name);
    return NULL;
}

void ufomi_task_free(UfomiTask *t)
{
    if (!t)
        return;
    if (t->cls->finalize)
        t->cls->finalize(t);
    for (int i = 0; i < t->cls->n_props; i++) {
        free(t->v[i].arr);
        free(t->v[i].str);
    }
    free(t->priv);
    free(t);
}

const char *ufomi_task_name(const UfomiTask *t) { return t ? t->cls->name : NULL; }

int ufomi_task_setup(UfomiTask *t, ufomi_ctx *ctx, UfomiError *e)
{
    if (!t || (!ctx && !(t->cls->mode & UFOMI_TASK_MODE_CPU)))
        return task_error(e, UFOMI_TASK_ERROR_SETUP, UFOMI_ERR_ARG, "null task or context");
    t->ctx = ctx;
    t->ready = 0;
    const int rc = t->cls->setup ? t->cls->setup(t, e) : UFOMI_OK;
    t->ready = rc == UFOMI_OK;
    return rc;
}

int ufomi_task_get_requisition(UfomiTask *t, UfomiBuffer **in, UfomiRequisition *req, UfomiError *e)
{
    if (!t || !req || (!(t->cls->mode & UFOMI_TASK_MODE_GENERATOR) && (!in || !in[0])))
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG, "null argument");
    if (!t->ready)
        return task_error(e, UFOMI_TASK_ERROR_GET_REQUISITION, UFOMI_ERR_ARG, "task not set up");
    return t->cls->requisition(t, in, req, e);
}

unsigned ufomi_task_get_num_inputs(const UfomiTask *t)
{
    if (!t || (t->cls->mode & UFOMI_TASK_MODE_GENERATOR))
        return 0;
    return !strcmp(t->cls->name, "flat-field-correct") ? 3 : 1;
}

unsigned ufomi_task_get_num_dimensions(const UfomiTask *t, unsigned input)
{
    if (!t || (t->cls->mode & UFOMI_TASK_MODE_GENERATOR))
        return 0;
    if (!strcmp(t->cls->name, "fft"))
        return (int) t->v[FFT_DIMS].num > 2 ? 3 : 2;
    if (!strcmp(t->cls->name, "ifft"))
        return (int) t->v[IFFT_DIMS].num > 2 ? 3 : 2;
    if (!strcmp(t->cls->name, "stacked-backproject"))
        return 3;
    if (!strcmp(t->cls->name, "flat-field-correct") && input > 0 && t->v[FFC_SINO].num)
        return 1;
    return 2;
}

unsigned ufomi_task_get_mode(const UfomiTask *t) { return t ? t->cls->mode : 0; }

int ufomi_task_process(UfomiTask *t, UfomiBuffer **in, UfomiBuffer *out, UfomiRequisition *req)
{
    if (!t || !t->cls->process || !in || !in[0] || (!out && !(t->cls->mode & UFOMI_TASK_MODE_SINK)) || !req ||
        !t->ctx)
        return 0;
    t->err[0] = '\0';
    t->err_status = 0;
    const int ok = t->cls->process(t, in, out, req);
    if (ok)
        t->num_processed++;
    return ok;
}

int ufomi_task_generate_host(UfomiTask *t, float *out, const UfomiRequisition *req)
{
    if (!t || !out || !req || !t->ready || strcmp(t->cls->name, "read"))
        return 0;
    t->err[0] = '\0';
    t->err_status = 0;
    return read_task_generate_host(t, out, req);
}

int ufomi_task_process_host(UfomiTask *t, const float *in, const UfomiRequisition *req)
{
    if (!t || !in || !req || !t->ready)
        return 0;
    t->err[0] = '\0';
    t->err_status = 0;
    int ok;
    if (!strcmp(t->cls->name, "write"))
        ok = write_task_write_host(t, in, req) == UFOMI_OK;
    else if (!strcmp(t->cls->name, "center-of-rotation"))
        ok = cor_compute(t, in, req);
    else
        return 0;
    if (ok)
        t->num_processed++;
    return ok;
}

int ufomi_task_generate(UfomiTask *t, UfomiBuffer *out, UfomiRequisition *req)
{
    if (!t || !out || !req || !t->cls->generate || !t->ctx)
        return 0;
    t->err[0] = '\0';
    t->err_status = 0;
    return t->cls->generate(t, out, req);
}

UfomiBufferLayout ufomi_task_output_layout(const UfomiTask *t)
{
    return t && !strcmp(t->cls->name, "fft") ? UFOMI_LAYOUT_COMPLEX_INTERLEAVED : UFOMI_LAYOUT_REAL;
}

int ufomi_task_output_dtype(const UfomiTask *t)
{
    (void) t;
    return UFOMI_F32;
}

/* ---------------------------------------------------------------- buffers */

size_t ufomi_requisition_elements(const UfomiRequisition *r)
{
    size_t n = 1;
    for (unsigned i = 0; i < r->n_dims && i < 3; i++)
        n *= r->dims[i];
    return n;
}

int ufomi_buffer_new(ufomi_ctx *ctx, const UfomiRequisition *req, int dtype, UfomiBuffer *out)
{
    memset(out, 0, sizeof(*out));
    out->req = *req;
    out->dtype = dtype;
    const size_t bytes = ufomi_requisition_elements(req) * (dtype == UFOMI_F32 ? 4 : 2);
    int rc = ufomi_device_alloc(ctx, bytes, &out->data);
    if (!rc)
        out->owned = 1;
    return rc;
}

void ufomi_buffer_release(ufomi_ctx *ctx, UfomiBuffer *buf)
{
    if (buf && buf->owned && buf->data)
        ufomi_device_free(ctx, buf->data);
    if (buf) {
        buf->data = NULL;
        buf->owned = 0;
    }
}
