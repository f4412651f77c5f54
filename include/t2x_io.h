/* t2x_io.h — on-disk formats around the E-step: MRC images/stacks/volumes and STAR metadata.
 *
 * These let the library run against real relion_refine inputs: particles.star (data_optics +
 * data_particles, RELION 3.1 layout) with images in MRC stacks, and write the per-particle
 * results back as the run_itNNN_data.star columns. Host code only (no GPU calls); status codes
 * as in t2x.h, text in t2x_last_error().
 */
#ifndef T2X_IO_H
#define T2X_IO_H

#include "t2x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------
 * MRC (rwMRC.h; MRC2014 header of 1024 bytes plus nsymbt bytes of extended header)
 * ------------------------------------------------------------------------------------- */
typedef struct t2x_mrc_info {
    int    nx, ny, nz;      /* columns, rows, sections (a stack of n images has nz = n)        */
    int    mode;            /* 0 int8, 1 int16, 2 float32, 6 uint16, 12 float16                */
    double angpix;          /* cella.x / mx (0 when the header carries none)                   */
    long   data_offset;     /* 1024 + nsymbt                                                  */
    int    swapped;         /* 1: file written on the other endianness                         */
    double dmin, dmax, dmean, rms;
} t2x_mrc_info;

/* Replaces the header part of readMRC (rwMRC.h:146-265): dimensions, mode, pixel size. */
int t2x_mrc_info_read(const char* path, t2x_mrc_info* info);
/* Sections [first, first + count) as float32 (readData for the modes above); out holds
 * count * nx * ny floats, x fastest. */
int t2x_mrc_read_sections(const char* path, long first, long count, float* out);
/* writeMRC (rwMRC.h:271-487) in mode 2: nx*ny*nz floats (a volume, or a stack of nz images),
 * pixel size angpix; header as writeMRC fills it (mx/my/mz = nx/ny/nz, cella = angpix * n, 90
 * degree angles, axes 1 2 3, min/max/mean/rms of the data, "MAP ", little-endian stamp). */
int t2x_mrc_write(const char* path, int nx, int ny, int nz, double angpix, const float* data);

/* ---------------------------------------------------------------------------------------
 * STAR (MetaDataTable::read/write, metadata_table.cpp:959-1420): all data blocks of a file,
 * loop or list form, values kept as text.
 * ------------------------------------------------------------------------------------- */
typedef struct t2x_star t2x_star;

int t2x_star_read(const char* path, t2x_star** out);
int t2x_star_free(t2x_star* s);
/* Writes every block in MetaDataTable::write's layout ("# version 30001", data_NAME, loop_
 * with "_label #i" lines, rows, blank line). */
int t2x_star_write(const t2x_star* s, const char* path);
int t2x_star_num_rows(const t2x_star* s, const char* block, long* n);
/* label without the leading underscore, e.g. "rlnDefocusU"; T2X_EINVAL if absent */
int t2x_star_get_double(const t2x_star* s, const char* block, const char* label, long row, double* v);
int t2x_star_get_string(const t2x_star* s, const char* block, const char* label, long row, char* buf,
                        long buflen);
/* Sets a value, adding the column (filled with the value's zero) when it is missing. */
int t2x_star_set_double(t2x_star* s, const char* block, const char* label, long row, double v);
int t2x_star_set_int(t2x_star* s, const char* block, const char* label, long row, long v);

/* E-step inputs of rows [first, first + n) of data_particles, as MlOptimiser::
 * getMetaAndImageDataSubset reads them (ml_optimiser.cpp:10295-10400): CTF from the row and its
 * optics group (data_optics), offsets and priors converted from Angstrom to pixels with the
 * group's pixel size, angles and angle priors (rlnAngleRot/Tilt/PsiPrior), norm correction (1
 * when absent), priors 999 when absent, group_id / optics_group from rlnGroupNumber /
 * rlnOpticsGroup (1-based in the file, 0-based here; 0 when absent), part_id = the row.
 * angpix[i] receives the pixel size of each row (may be NULL). */
int t2x_star_particles(const t2x_star* s, long first, long n, t2x_particle_in* pin, double* angpix);
/* rlnImageName of a row, "index@path": the 1-based index and the stack path. */
int t2x_star_image_ref(const t2x_star* s, long row, long* index, char* path, long pathlen);
/* E-step results into rows [first, first + n), as MlOptimiser::setMetaDataSubset writes them
 * (ml_optimiser.cpp:10034-10090): rlnAngleRot/Tilt (3D references only) and rlnAnglePsi,
 * rlnOriginXAngst/YAngst = pixel size * offset, rlnClassNumber, rlnLogLikeliContribution,
 * rlnMaxValueProbDistribution, rlnNrOfSignificantSamples, rlnNormCorrection. */
int t2x_star_store_results(t2x_star* s, long first, long n, const t2x_particle_out* out, int ref_dim);

#ifdef __cplusplus
}
#endif
#endif /* T2X_IO_H */
