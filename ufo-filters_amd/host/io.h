/*
 * Image files of the `read` and `write` tasks (host side, no GPU):
 *   - TIFF: classic (32-bit offsets) and BigTIFF (64-bit offsets) baseline
 *     files, uncompressed strips, one sample per pixel, 8/16/32-bit unsigned or
 *     signed integers, 32/64-bit floats, either byte order, any number of pages;
 *   - EDF: one ESRF header ('{' ... '}' padded to a multiple of 512 bytes;
 *     Dim_1, Dim_2, DataType, ByteOrder, Size) followed by equally sized frames;
 *   - raw: frames of width x height samples of 8, 16 (unsigned) or 32 (float)
 *     bits, each after `pre_offset` and before `post_offset` bytes.
 * Every read converts to float32.  The writers produce TIFF (classic or
 * BigTIFF; float32, or 8/16-bit unsigned after the reference's conversion) and
 * raw.
 */
#ifndef UFOMI_IO_H
#define UFOMI_IO_H

#include <stdint.h>
#include <stdio.h>

enum { IMG_NONE = 0, IMG_TIFF, IMG_EDF, IMG_RAW };

typedef struct {
    unsigned width, height, bitdepth;
    unsigned long pre_offset, post_offset;
} ImgRawSpec;

typedef struct {
    FILE *fp;
    int format;
    unsigned width, height, n_images;
    int kind;          /* 'u' unsigned, 's' signed, 'f' float */
    int bytes;         /* bytes per sample */
    int swap;          /* file byte order differs from the host's */
    int big;           /* TIFF: BigTIFF */
    long data0;        /* EDF / raw: offset of frame 0 */
    long stride;       /* EDF / raw: bytes from one frame to the next */
    long *ifd;         /* TIFF: IFD offset of each page */
    long *strips;      /* TIFF: strip offsets of the cached page */
    unsigned n_strips, rows_per_strip;
    long cached_page;
} ImgFile;

/* Format from the file name (.tif/.tiff, .edf, .raw), IMG_NONE if none. */
int img_format_of(const char *path);
/* 0 on success; on failure a message in err. */
int img_open(const char *path, int format, const ImgRawSpec *raw, ImgFile *f, char *err, size_t errlen);
/* Rows y0, y0 + ystep, ... (nrows of them) of image `index` as float32. */
int img_read(ImgFile *f, unsigned index, unsigned y0, unsigned ystep, unsigned nrows, float *out, char *err,
             size_t errlen);
void img_close(ImgFile *f);

/* Writers: a new file (truncated), pages appended one by one. */
typedef struct {
    FILE *fp;
    int format;
    int big;           /* TIFF: BigTIFF */
    int to_stdout;
    long last_next;    /* TIFF: offset of the previous page's next-IFD field */
    unsigned pages;
} ImgWriter;

/* path NULL writes raw to stdout. */
int img_writer_open(ImgWriter *w, const char *path, int format, int bigtiff, char *err, size_t errlen);
/* bits: 32 (float32) or 8 / 16 (unsigned, converted as src/writers/ufo-writer.c does). */
int img_writer_write(ImgWriter *w, const float *data, unsigned width, unsigned height, unsigned bits, float minimum,
                     float maximum, int rescale, char *err, size_t errlen);
void img_writer_close(ImgWriter *w);

/* The reference's float -> 8 / 16-bit conversion of one image (n samples). */
void img_convert(const float *src, size_t n, unsigned bits, float minimum, float maximum, int rescale, void *dst);

#endif
