/*
 * Image files of the `read` and `write` tasks: see io.h.  Readers follow the
 * behaviour of src/readers/ufo-tiff-reader.c, ufo-edf-reader.c and
 * ufo-raw-reader.c (formats, data types, frame layout); the writers that of
 * src/writers/ufo-tiff-writer.c, ufo-raw-writer.c and the conversion of
 * src/writers/ufo-writer.c:55-230.  The TIFF code is written against the
 * TIFF 6.0 baseline and the BigTIFF extension, written without libtiff.
 */
#include "io.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

static int fail(char *err, size_t n, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, n, fmt, ap);
    va_end(ap);
    return -1;
}

static int ends_with(const char *s, const char *suffix)
{
    size_t a = strlen(s), b = strlen(suffix);
    return a >= b && !strcmp(s + a - b, suffix);
}

/* The readers' can_open tests (case sensitive, as g_str_has_suffix). */
int img_format_of(const char *path)
{
    if (ends_with(path, ".tif") || ends_with(path, ".tiff"))
        return IMG_TIFF;
    if (ends_with(path, ".edf"))
        return IMG_EDF;
    if (ends_with(path, ".raw"))
        return IMG_RAW;
    return IMG_NONE;
}

static uint64_t rdn(const unsigned char *p, int n, int be)
{
    uint64_t v = 0;
    for (int i = 0; i < n; i++)
        v |= (uint64_t) p[be ? n - 1 - i : i] << (8 * i);
    return v;
}

static long file_size(FILE *fp)
{
    fseek(fp, 0, SEEK_END);
    long s = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    return s;
}

/* ------------------------------------------------------------------ TIFF */

typedef struct {
    unsigned width, height, bits, compression, spp, format, planar;
    uint64_t rps;
    long strip_entry;      /* file offset of the StripOffsets entry */
    int tiled;
} TiffPage;

/* Entry layout: classic 12 bytes (tag, type, count u32, value u32), BigTIFF 20 (count u64, value u64). */
static int tiff_value(FILE *fp, int be, int big, const unsigned char *e, uint64_t index, uint64_t *out)
{
    const unsigned type = (unsigned) rdn(e + 2, 2, be);
    const uint64_t count = rdn(e + 4, big ? 8 : 4, be);
    const unsigned size = type == 1 ? 1 : type == 3 ? 2 : type == 4 ? 4 : type == 16 ? 8 : 0;
    const unsigned inline_cap = big ? 8 : 4;
    const unsigned char *val = e + (big ? 12 : 8);
    if (!size || index >= count)
        return -1;
    unsigned char buf[8];
    /* count comes from the file: compare by division so that no product can wrap */
    if (count <= inline_cap / size) {
        memcpy(buf, val + index * size, size);
    } else {
        const uint64_t base = rdn(val, big ? 8 : 4, be);
        if (index > (UINT64_MAX - base) / size)
            return -1;
        const uint64_t at = base + index * size;
        if (at > (uint64_t) LONG_MAX - size)
            return -1;
        if (fseek(fp, (long) at, SEEK_SET) || fread(buf, 1, size, fp) != size)
            return -1;
    }
    *out = rdn(buf, (int) size, be);
    return 0;
}

static int tiff_page(ImgFile *f, long ifd, TiffPage *pg, char *err, size_t errlen)
{
    const int be = f->swap, big = f->big, es = big ? 20 : 12;
    unsigned char bn[8];
    if (fseek(f->fp, ifd, SEEK_SET) || fread(bn, 1, big ? 8 : 2, f->fp) != (size_t) (big ? 8 : 2))
        return fail(err, errlen, "truncated TIFF directory");
    const uint64_t n = rdn(bn, big ? 8 : 2, be);
    if (n > 4096)
        return fail(err, errlen, "corrupt TIFF directory");
    unsigned char *ents = malloc((size_t) n * es + 1);
    if (!ents || fread(ents, (size_t) es, (size_t) n, f->fp) != n) {
        free(ents);
        return fail(err, errlen, "truncated TIFF directory");
    }
    memset(pg, 0, sizeof(*pg));
    pg->bits = 1;
    pg->compression = 1;
    pg->spp = 1;
    pg->format = 1;
    pg->planar = 1;
    pg->rps = UINT64_MAX;
    int rc = 0;
    for (uint64_t i = 0; i < n && !rc; i++) {
        const unsigned char *e = ents + es * i;
        const unsigned tag = (unsigned) rdn(e, 2, be);
        uint64_t v = 0;
        switch (tag) {
        case 256: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->width = (unsigned) v; break;
        case 257: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->height = (unsigned) v; break;
        case 258: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->bits = (unsigned) v; break;
        case 259: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->compression = (unsigned) v; break;
        case 277: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->spp = (unsigned) v; break;
        case 278: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->rps = v; break;
        case 284: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->planar = (unsigned) v; break;
        case 339: rc = tiff_value(f->fp, be, big, e, 0, &v); pg->format = (unsigned) v; break;
        case 322: case 323: case 324: case 325: pg->tiled = 1; break;
        case 273: pg->strip_entry = ifd + (big ? 8 : 2) + (long) (es * i); break;
        default: break;
        }
    }
    free(ents);
    if (rc)
        return fail(err, errlen, "unreadable TIFF tag");
    if (pg->tiled)
        return fail(err, errlen, "tiled TIFF is not supported");
    if (pg->compression != 1)
        return fail(err, errlen, "compressed TIFF (compression %u) is not supported", pg->compression);
    if (pg->spp != 1)
        return fail(err, errlen, "TIFF with %u samples per pixel is not supported", pg->spp);
    if (!pg->width || !pg->height || !pg->strip_entry)
        return fail(err, errlen, "TIFF page without size or strips");
    if (pg->rps > pg->height)
        pg->rps = pg->height;
    if (!pg->rps)
        return fail(err, errlen, "TIFF page with zero rows per strip");
    return 0;
}

static int tiff_open(ImgFile *f, char *err, size_t errlen)
{
    unsigned char h[16];
    if (fread(h, 1, 8, f->fp) != 8)
        return fail(err, errlen, "not a TIFF file");
    int be;
    if (h[0] == 'I' && h[1] == 'I')
        be = 0;
    else if (h[0] == 'M' && h[1] == 'M')
        be = 1;
    else
        return fail(err, errlen, "not a TIFF file");
    const unsigned magic = (unsigned) rdn(h + 2, 2, be);
    f->swap = be;  /* the host is little endian */
    uint64_t ofs;
    if (magic == 43) {
        if (rdn(h + 4, 2, be) != 8 || fread(h + 8, 1, 8, f->fp) != 8)
            return fail(err, errlen, "unsupported BigTIFF header");
        f->big = 1;
        ofs = rdn(h + 8, 8, be);
    } else if (magic == 42) {
        ofs = rdn(h + 4, 4, be);
    } else {
        return fail(err, errlen, "not a TIFF file");
    }
    size_t cap = 64;
    f->ifd = malloc(cap * sizeof(long));
    if (!f->ifd)
        return fail(err, errlen, "out of host memory");
    const uint64_t fsz = (uint64_t) file_size(f->fp);
    const int es = f->big ? 20 : 12, cn = f->big ? 8 : 2, on = f->big ? 8 : 4;
    while (ofs > 0 && ofs < fsz) {
        if (f->n_images && (long) ofs <= f->ifd[f->n_images - 1]) {
            /* pages out of file order are valid; a page seen before is a cycle */
            int seen = 0;
            for (unsigned k = 0; k < f->n_images && !seen; k++)
                seen = f->ifd[k] == (long) ofs;
            if (seen)
                break;
        }
        if (f->n_images == cap) {
            long *grown = realloc(f->ifd, 2 * cap * sizeof(long));
            if (!grown)
                return fail(err, errlen, "out of host memory");
            f->ifd = grown;
            cap *= 2;
        }
        f->ifd[f->n_images++] = (long) ofs;
        unsigned char bc[8], bo[8];
        if (fseek(f->fp, (long) ofs, SEEK_SET) || fread(bc, 1, (size_t) cn, f->fp) != (size_t) cn)
            break;
        if (fseek(f->fp, (long) (ofs + cn + (uint64_t) es * rdn(bc, cn, be)), SEEK_SET) ||
            fread(bo, 1, (size_t) on, f->fp) != (size_t) on)
            break;
        ofs = rdn(bo, on, be);
    }
    if (!f->n_images)
        return fail(err, errlen, "TIFF without pages");
    TiffPage pg;
    if (tiff_page(f, f->ifd[0], &pg, err, errlen))
        return -1;
    f->width = pg.width;
    f->height = pg.height;
    f->bytes = (int) (pg.bits / 8);
    f->kind = pg.format == 3 ? 'f' : (pg.format == 2 ? 's' : 'u');
    if (pg.bits % 8 || !(f->bytes == 1 || f->bytes == 2 || f->bytes == 4 || f->bytes == 8) ||
        (f->kind == 'f' && f->bytes < 4) || (f->kind != 'f' && f->bytes > 4))
        return fail(err, errlen, "TIFF with %u-bit %s samples is not supported", pg.bits,
                    f->kind == 'f' ? "float" : "integer");
    f->cached_page = -1;
    return 0;
}

static int tiff_load_page(ImgFile *f, unsigned page, char *err, size_t errlen)
{
    if (f->cached_page == (long) page)
        return 0;
    TiffPage pg;
    if (tiff_page(f, f->ifd[page], &pg, err, errlen))
        return -1;
    if (pg.width != f->width || pg.height != f->height || pg.bits != (unsigned) f->bytes * 8)
        return fail(err, errlen, "TIFF page %u differs from page 0 in size or depth", page);
    unsigned char e[20];
    const int es = f->big ? 20 : 12;
    if (fseek(f->fp, pg.strip_entry, SEEK_SET) || fread(e, 1, (size_t) es, f->fp) != (size_t) es)
        return fail(err, errlen, "truncated TIFF directory");
    const uint64_t count = rdn(e + 4, f->big ? 8 : 4, f->swap);
    const uint64_t need = (pg.height + pg.rps - 1) / pg.rps;
    if (count < need)
        return fail(err, errlen, "TIFF page %u has %llu strips, needs %llu", page, (unsigned long long) count,
                    (unsigned long long) need);
    if (need > (uint64_t) file_size(f->fp))
        return fail(err, errlen, "TIFF page %u has more strips than the file has bytes", page);
    free(f->strips);
    f->cached_page = -1;  /* the strips of the page cached before are gone */
    f->n_strips = (unsigned) need;
    f->strips = malloc(f->n_strips * sizeof(long));
    if (!f->strips) {
        f->n_strips = 0;
        return fail(err, errlen, "out of host memory");
    }
    for (unsigned i = 0; i < f->n_strips; i++) {
        uint64_t v;
        if (tiff_value(f->fp, f->swap, f->big, e, i, &v))
            return fail(err, errlen, "unreadable TIFF strip offsets");
        f->strips[i] = (long) v;
    }
    f->rows_per_strip = (unsigned) pg.rps;
    f->cached_page = page;
    return 0;
}

/* ------------------------------------------------------------------- EDF */

static int edf_open(ImgFile *f, char *err, size_t errlen)
{
    const long fsz = file_size(f->fp);
    const long cap = fsz < (1 << 20) ? fsz : (1 << 20);
    char *h = malloc((size_t) cap + 1);
    if (!h || fread(h, 1, (size_t) cap, f->fp) != (size_t) cap) {
        free(h);
        return fail(err, errlen, "Could not read EDF header.");
    }
    h[cap] = '\0';
    char *close = memchr(h, '}', (size_t) cap);
    if (!close || (close - h + 2) % 512) {
        free(h);
        return fail(err, errlen, "Corrupt EDF header or not an EDF file.");
    }
    *close = '\0';
    long dim1 = 0, dim2 = 0, size = -1;
    char type[64] = "", order[64] = "";
    for (char *save = NULL, *tok = strtok_r(h, ";", &save); tok; tok = strtok_r(NULL, ";", &save)) {
        char *eq = strchr(tok, '=');
        if (!eq)
            continue;
        *eq = '\0';
        char key[64], val[64];
        if (sscanf(tok, " %63[^ \t\r\n{]", key) != 1 || sscanf(eq + 1, " %63s", val) != 1)
            continue;
        if (!strcmp(key, "Dim_1")) dim1 = atol(val);
        else if (!strcmp(key, "Dim_2")) dim2 = atol(val);
        else if (!strcmp(key, "DataType")) snprintf(type, sizeof type, "%s", val);
        else if (!strcmp(key, "ByteOrder")) snprintf(order, sizeof order, "%s", val);
        else if (!strcmp(key, "Size")) size = atol(val);
    }
    const long header = close - h + 2;
    free(h);
    /* the reference's DataType map (ufo-edf-reader.c:178-199) plus the other ESRF names */
    static const struct { const char *name; int kind, bytes; } map[] = {
        {"UnsignedShort", 'u', 2}, {"SignedInteger", 's', 4}, {"UnsignedLong", 'u', 4}, {"Float", 'f', 4},
        {"FloatValue", 'f', 4}, {"UnsignedByte", 'u', 1}, {"SignedByte", 's', 1}, {"SignedShort", 's', 2},
        {"UnsignedInteger", 'u', 4}, {"SignedLong", 's', 4}, {"DoubleValue", 'f', 8}, {"Double", 'f', 8},
        {NULL, 0, 0}};
    f->kind = 0;
    for (int i = 0; map[i].name; i++)
        if (!strcmp(type, map[i].name)) {
            f->kind = map[i].kind;
            f->bytes = map[i].bytes;
        }
    if (!f->kind)
        return fail(err, errlen, "unsupported EDF DataType `%s'", type);
    if (dim1 <= 0 || dim2 <= 0)
        return fail(err, errlen, "EDF header without Dim_1 / Dim_2");
    f->width = (unsigned) dim1;
    f->height = (unsigned) dim2;
    f->swap = !strcmp(order, "HighByteFirst");
    f->data0 = header;
    f->stride = (long) f->width * f->height * f->bytes;
    long data = fsz - header;
    if (size >= 0 && size < data)
        data = size;
    f->n_images = (unsigned) (data / f->stride);
    if (!f->n_images)
        return fail(err, errlen, "EDF file without a complete frame");
    return 0;
}

/* ------------------------------------------------------------------- raw */

static int raw_open(ImgFile *f, const ImgRawSpec *raw, char *err, size_t errlen)
{
    if (!raw || !raw->width || !raw->height)
        return fail(err, errlen, "raw: `raw-width' and `raw-height' must be set");
    if (raw->bitdepth != 8 && raw->bitdepth != 16 && raw->bitdepth != 32)
        return fail(err, errlen, "Cannot set bitdepth other than 8, 16 or 32.");
    f->width = raw->width;
    f->height = raw->height;
    f->bytes = (int) raw->bitdepth / 8;
    f->kind = raw->bitdepth == 32 ? 'f' : 'u';
    const long frame = (long) f->width * f->height * f->bytes;
    f->data0 = (long) raw->pre_offset;
    f->stride = frame + (long) raw->pre_offset + (long) raw->post_offset;
    const long fsz = file_size(f->fp);
    f->n_images = fsz >= (long) raw->pre_offset + frame
                      ? (unsigned) ((fsz - (long) raw->pre_offset - frame) / f->stride + 1) : 0;
    if (!f->n_images)
        return fail(err, errlen, "raw file smaller than one frame");
    return 0;
}

/* ------------------------------------------------------------- common API */

int img_open(const char *path, int format, const ImgRawSpec *raw, ImgFile *f, char *err, size_t errlen)
{
    memset(f, 0, sizeof(*f));
    f->format = format;
    f->fp = fopen(path, "rb");
    if (!f->fp)
        return fail(err, errlen, "Cannot open %s", path);
    int rc;
    switch (format) {
    case IMG_TIFF: rc = tiff_open(f, err, errlen); break;
    case IMG_EDF: rc = edf_open(f, err, errlen); break;
    case IMG_RAW: rc = raw_open(f, raw, err, errlen); break;
    default: rc = fail(err, errlen, "`%s' has no known image format", path); break;
    }
    if (rc)
        img_close(f);
    return rc;
}

void img_close(ImgFile *f)
{
    if (f->fp)
        fclose(f->fp);
    free(f->ifd);
    free(f->strips);
    memset(f, 0, sizeof(*f));
}

static void swap_bytes(unsigned char *p, size_t n, int bytes)
{
    for (size_t i = 0; i < n; i++, p += bytes)
        for (int a = 0, b = bytes - 1; a < b; a++, b--) {
            const unsigned char t = p[a];
            p[a] = p[b];
            p[b] = t;
        }
}

static void to_float(const unsigned char *src, size_t n, int kind, int bytes, float *dst)
{
    for (size_t i = 0; i < n; i++) {
        const unsigned char *p = src + i * (size_t) bytes;
        float v;
        if (kind == 'f') {
            if (bytes == 4) {
                memcpy(&v, p, 4);
            } else {
                double d;
                memcpy(&d, p, 8);
                v = (float) d;
            }
        } else if (bytes == 1) {
            v = kind == 's' ? (float) (int8_t) p[0] : (float) p[0];
        } else if (bytes == 2) {
            uint16_t u;
            memcpy(&u, p, 2);
            v = kind == 's' ? (float) (int16_t) u : (float) u;
        } else {
            uint32_t u;
            memcpy(&u, p, 4);
            v = kind == 's' ? (float) (int32_t) u : (float) u;
        }
        dst[i] = v;
    }
}

int img_read(ImgFile *f, unsigned index, unsigned y0, unsigned ystep, unsigned nrows, float *out, char *err,
             size_t errlen)
{
    if (index >= f->n_images)
        return fail(err, errlen, "image %u beyond the %u of the file", index, f->n_images);
    if (!ystep || !nrows || (unsigned long long) y0 + (unsigned long long) (nrows - 1) * ystep >= f->height)
        return fail(err, errlen, "row range beyond the image height %u", f->height);
    if (f->format == IMG_TIFF && tiff_load_page(f, index, err, errlen))
        return -1;
    const size_t row_bytes = (size_t) f->width * f->bytes;
    unsigned char *raw = malloc(row_bytes * nrows);
    if (!raw)
        return fail(err, errlen, "out of memory");
    /* rows that follow each other in the file are read in one run */
    long run_ofs = -1;
    size_t run_rows = 0, run_first = 0;
    for (unsigned k = 0; k <= nrows; k++) {
        long ofs = -1;
        if (k < nrows) {
            const unsigned y = y0 + k * ystep;
            if (f->format == IMG_TIFF)
                ofs = f->strips[y / f->rows_per_strip] + (long) (y % f->rows_per_strip) * (long) row_bytes;
            else
                ofs = f->data0 + (long) index * f->stride + (long) y * (long) row_bytes;
            if (run_rows && ofs == run_ofs + (long) (run_rows * row_bytes)) {
                run_rows++;
                continue;
            }
        }
        if (run_rows) {
            const size_t want = run_rows * row_bytes;
            if (fseek(f->fp, run_ofs, SEEK_SET) || fread(raw + run_first * row_bytes, 1, want, f->fp) != want) {
                free(raw);
                return fail(err, errlen, "truncated image data");
            }
        }
        run_ofs = ofs;
        run_first = k;
        run_rows = 1;
    }
    if (f->swap && f->bytes > 1)
        swap_bytes(raw, (size_t) f->width * nrows, f->bytes);
    to_float(raw, (size_t) f->width * nrows, f->kind, f->bytes, out);
    free(raw);
    return 0;
}

/* ------------------------------------------------------------ conversion */

/* A float to a 32-bit integer the way x86-64's cvttss2si does it (the
 * reference's (guint8) / (guint16) casts compile to that and keep the low
 * bits): truncation toward zero, INT32_MIN for NaN and out-of-range values. */
static uint32_t trunc_i32(float v)
{
    if (!(v > -2147483904.0f && v < 2147483648.0f))
        return 0x80000000u;
    return (uint32_t) (int32_t) v;
}

void img_convert(const float *src, size_t n, unsigned bits, float minimum, float maximum, int rescale, void *dst)
{
    unsigned char *d8 = dst;
    uint16_t *d16 = dst;
    if (!rescale) {
        for (size_t i = 0; i < n; i++) {
            const uint32_t u = trunc_i32(src[i]);
            if (bits == 8)
                d8[i] = (unsigned char) u;
            else
                d16[i] = (uint16_t) u;
        }
        return;
    }
    /* get_min_max: a bound the user left at its default comes from the image */
    const int user_min = minimum < FLT_MAX, user_max = maximum > -FLT_MAX;
    float mn = minimum, mx = maximum;
    if (!(user_min && user_max)) {
        float cmin = FLT_MAX, cmax = -FLT_MAX;
        for (size_t i = 0; i < n; i++) {
            if (!user_min && src[i] < cmin)
                cmin = src[i];
            if (!user_max && src[i] > cmax)
                cmax = src[i];
        }
        mn = user_min ? minimum : cmin;
        mx = user_max ? maximum : cmax;
    }
    const float scale = (bits == 8 ? 255.0f : 65535.0f) / (mx - mn);
    for (size_t i = 0; i < n; i++) {
        float v = src[i];
        if (mn < mx)
            v = v < mn ? mn : (v > mx ? mx : v);
        else
            v = v < mx ? mx : (v > mn ? mn : v);
        const uint32_t u = trunc_i32((v - mn) * scale);
        if (bits == 8)
            d8[i] = (unsigned char) u;
        else
            d16[i] = (uint16_t) u;
    }
}

/* --------------------------------------------------------------- writers */

int img_writer_open(ImgWriter *w, const char *path, int format, int bigtiff, char *err, size_t errlen)
{
    memset(w, 0, sizeof(*w));
    if (!path) {
        w->format = IMG_RAW;
        w->fp = stdout;
        w->to_stdout = 1;
        return 0;
    }
    w->format = format;
    if (format != IMG_TIFF && format != IMG_RAW)
        return fail(err, errlen, "`%s' does not have a valid file extension or requires format specifiers", path);
    w->fp = fopen(path, format == IMG_TIFF ? "w+b" : "wb");
    if (!w->fp)
        return fail(err, errlen, "Cannot open %s for writing", path);
    if (format == IMG_TIFF) {
        w->big = bigtiff;
        if (bigtiff) {
            const unsigned char h[16] = {'I', 'I', 43, 0, 8, 0, 0, 0};
            fwrite(h, 1, 16, w->fp);
            w->last_next = 8;
        } else {
            const unsigned char h[8] = {'I', 'I', 42, 0};
            fwrite(h, 1, 8, w->fp);
            w->last_next = 4;
        }
    }
    return 0;
}

static void putn(unsigned char *p, uint64_t v, int n)
{
    for (int i = 0; i < n; i++)
        p[i] = (unsigned char) (v >> (8 * i));
}

int img_writer_write(ImgWriter *w, const float *data, unsigned width, unsigned height, unsigned bits, float minimum,
                     float maximum, int rescale, char *err, size_t errlen)
{
    if (bits != 8 && bits != 16 && bits != 32)
        return fail(err, errlen, "bits must be 8, 16 or 32");
    const size_t n = (size_t) width * height;
    const size_t bytes = n * (bits / 8);
    unsigned char *buf = NULL;
    const unsigned char *payload = (const unsigned char *) data;
    if (bits != 32) {
        buf = malloc(bytes);
        if (!buf)
            return fail(err, errlen, "out of memory");
        img_convert(data, n, bits, minimum, maximum, rescale, buf);
        payload = buf;
    }
    int rc = 0;
    if (w->format == IMG_RAW) {
        if (fwrite(payload, 1, bytes, w->fp) != bytes)
            rc = fail(err, errlen, "write failed");
    } else {
        const int big = w->big, es = big ? 20 : 12, cn = big ? 8 : 2, on = big ? 8 : 4, vn = big ? 8 : 4;
        fseek(w->fp, 0, SEEK_END);
        long data_ofs = ftell(w->fp);
        if (data_ofs & 1) {  /* word alignment */
            fputc(0, w->fp);
            data_ofs++;
        }
        if (fwrite(payload, 1, bytes, w->fp) != bytes)
            rc = fail(err, errlen, "write failed");
        long ifd = ftell(w->fp);
        if (ifd & 1) {
            fputc(0, w->fp);
            ifd++;
        }
        if (!rc && !big && (uint64_t) ifd + 256 > 0xffffffffull)
            rc = fail(err, errlen, "classic TIFF is limited to 4 GiB (set tiff-bigtiff=true)");
        /* the tags of src/writers/ufo-tiff-writer.c:94-150 in ascending order, one strip per page */
        enum { NE = 13 };
        const uint64_t tags[NE][4] = {
            {254, 4, 1, 2},                          /* SubfileType: page */
            {256, 4, 1, width},
            {257, 4, 1, height},
            {258, 3, 1, bits},
            {259, 3, 1, 1},                          /* no compression */
            {262, 3, 1, 1},                          /* min-is-black */
            {273, big ? 16u : 4u, 1, (uint64_t) data_ofs},
            {277, 3, 1, 1},
            {278, 4, 1, height},
            {279, big ? 16u : 4u, 1, bytes},
            {284, 3, 1, 1},                          /* contiguous */
            {297, 3, 2, 0},                          /* PageNumber (page, page) */
            {339, 3, 1, bits == 32 ? 3u : 1u},       /* IEEE float or unsigned */
        };
        unsigned char d[8 + 20 * NE + 8];
        memset(d, 0, sizeof d);
        putn(d, NE, cn);
        for (int i = 0; i < NE; i++) {
            unsigned char *e = d + cn + es * i;
            putn(e, tags[i][0], 2);
            putn(e + 2, tags[i][1], 2);
            putn(e + 4, tags[i][2], vn);
            unsigned char *v = e + 4 + vn;
            if (tags[i][0] == 297) {
                putn(v, w->pages, 2);
                putn(v + 2, w->pages, 2);
            } else if (tags[i][1] == 3) {
                putn(v, tags[i][3], 2);
            } else {
                putn(v, tags[i][3], tags[i][1] == 16 ? 8 : 4);
            }
        }
        const size_t dsize = (size_t) (cn + es * NE + on);
        if (!rc && fwrite(d, 1, dsize, w->fp) != dsize)
            rc = fail(err, errlen, "write failed");
        if (!rc) {
            /* link: the previous next-IFD field (or the header) points here */
            unsigned char o[8];
            putn(o, (uint64_t) ifd, on);
            fseek(w->fp, w->last_next, SEEK_SET);
            if (fwrite(o, 1, (size_t) on, w->fp) != (size_t) on)
                rc = fail(err, errlen, "write failed");
            w->last_next = ifd + cn + es * NE;
        }
        fseek(w->fp, 0, SEEK_END);
    }
    free(buf);
    if (!rc)
        w->pages++;
    return rc;
}

void img_writer_close(ImgWriter *w)
{
    if (w->fp) {
        if (w->to_stdout)
            fflush(w->fp);
        else
            fclose(w->fp);
    }
    memset(w, 0, sizeof(*w));
}
