/*
 * mtr_io.h -- MRC volume / tilt-series files and TXBR geometry (C ABI).
 *
 * Replaces the reference's file boundary (Autumn1998/Multi-tilt-Reconstuction):
 *   mtr_mrc_header_t        struct MrcHeader            MrcFileIOBaseFunction.h:99-174 (1024 bytes, packed)
 *   mtr_mrc_read_header     mrcReadhead / readMrcHeader MrcFileIOBaseFunction.cpp:11-16, MrcFileIO.cpp:107-125
 *   mtr_mrc_read_data       mrcReadAllData              MrcFileIOBaseFunction.cpp:18-70
 *   mtr_mrc_write_volume    updataAndWriteModel         MrcFileIO.cpp:63-88, 180-201; MrcFileIOBaseFunction.cpp:136-160
 *   mtr_txbr_read           readCoef / readTXBRData     MrcFileIOBaseFunction.cpp:72-134, MrcFileIO.cpp:144-160
 * Differences from the reference, on purpose:
 *   - data start at 1024 + next (extended header), for every mode; modes 0
 *     (unsigned byte, as the reference reads it), 1 (int16), 2 (float32) and
 *     6 (uint16) are read in full (the reference's offset and its float /
 *     ushort loops read the wrong bytes);
 *   - the detector row count is ny (the reference sets prj.Y = nx);
 *   - header statistics are seeded from the written data;
 *   - TXBR coefficient lines are split on any whitespace (a trailing blank
 *     does not add a zero coefficient).
 */
#ifndef MTR_IO_H_
#define MTR_IO_H_

#include <stdint.h>
#include "mtr.h"

#ifdef __cplusplus
extern "C" {
#endif

#pragma pack(push, 1)
typedef struct {
    int32_t nx, ny, nz, mode;
    int32_t nxstart, nystart, nzstart;
    int32_t mx, my, mz;
    float xlen, ylen, zlen;
    float alpha, beta, gamma;
    int32_t mapc, mapr, maps;
    float amin, amax, amean;
    int16_t ispg, nsymbt;
    int32_t next;
    int16_t creatid;
    char blank[30];
    int16_t nint, nreal;
    int16_t sub, zfac;
    float min2, max2, min3, max3, min4, max4;
    int16_t idtype, lens, nd1, nd2, vd1, vd2;
    float tiltangles[6];
    float xorg, yorg, zorg;
    char cmap[4];
    char stamp[4];
    float rms;
    int32_t nlabl;
    char labels[10][80];
} mtr_mrc_header_t;
#pragma pack(pop)

enum { MTR_MRC_BYTE = 0, MTR_MRC_SHORT = 1, MTR_MRC_FLOAT = 2, MTR_MRC_USHORT = 6 };

int mtr_mrc_read_header(const char *path, mtr_mrc_header_t *hdr);
/* Reads all nx*ny*nz values as float into out (capacity count). */
int mtr_mrc_read_data(const char *path, float *out, int64_t count);
/* Writes a float (mode 2) MRC volume [Z][Y][X] with the box as origin. */
int mtr_mrc_write_volume(const char *path, const mtr_volume_size *vol, const float *data);
/* Writes a float MRC stack nx x ny x nz (tilt series / generic). */
int mtr_mrc_write_stack(const char *path, int nx, int ny, int nz, const float *data);
/* Parses a TXBR file: coefficients into xc/yc (capacity 4*max_views),
 * the number of views into *views, and widens *vol (union of the x->, y->,
 * z-> ranges; X/Y/Z recomputed).  Initialise vol with an empty box
 * (start = +huge, end = -huge) to take the file's box alone. */
int mtr_txbr_read(const char *path, mtr_volume_size *vol, double *xc, double *yc, int max_views,
                  int *views);
/* Writes a TXBR file in the grammar mtr_txbr_read accepts. */
int mtr_txbr_write(const char *path, const mtr_volume_size *vol, const double *xc, const double *yc,
                   int views);

#ifdef __cplusplus
}
#endif

#endif /* MTR_IO_H_ */
