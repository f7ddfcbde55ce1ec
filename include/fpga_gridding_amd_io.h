/*
 * Plain host I/O of the MI355X gridder library: OSKAR binary visibility
 * files in, FITS kernel cubes in, FITS grid cubes out.  These are the file
 * formats the reference harness uses (main.cpp:100-250, :560-601), read and
 * written by code of this library (no oskar_binary or cfitsio needed).
 * Status codes are OSKAR's (utils/oskar_global.h:43-64, and the binary
 * reader's utils/binary/oskar_binary.h:96-117).
 */
#ifndef FPGA_GRIDDING_AMD_IO_H_
#define FPGA_GRIDDING_AMD_IO_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#  define GW_IO_EXPORT __attribute__((visibility("default")))
#else
#  define GW_IO_EXPORT
#endif

/* Data type codes of the OSKAR binary format
 * (utils/binary/oskar_binary_data_types.h:44-74). */
enum gw_oskar_type
{
    GW_OSKAR_CHAR = 0x01, GW_OSKAR_INT = 0x02, GW_OSKAR_SINGLE = 0x04,
    GW_OSKAR_DOUBLE = 0x08, GW_OSKAR_COMPLEX = 0x20, GW_OSKAR_MATRIX = 0x40
};

/* Visibility file tag groups and tags (utils/binary/oskar_binary.h:69-70,
 * utils/read_vis.h:39-59). */
enum gw_vis_tags
{
    GW_VIS_GROUP_HEADER = 11,
    GW_VIS_GROUP_BLOCK = 12,
    GW_VIS_HEADER_TAG_NUM_TAGS_PER_BLOCK = 2,
    GW_VIS_HEADER_TAG_AMP_TYPE = 5,
    GW_VIS_HEADER_TAG_COORD_PRECISION = 6,
    GW_VIS_HEADER_TAG_MAX_TIMES_PER_BLOCK = 7,
    GW_VIS_HEADER_TAG_NUM_TIMES_TOTAL = 8,
    GW_VIS_HEADER_TAG_MAX_CHANNELS_PER_BLOCK = 9,
    GW_VIS_HEADER_TAG_NUM_CHANNELS_TOTAL = 10,
    GW_VIS_HEADER_TAG_NUM_STATIONS = 11,
    GW_VIS_HEADER_TAG_FREQ_START_HZ = 23,
    GW_VIS_BLOCK_TAG_DIM_START_AND_SIZE = 1,
    GW_VIS_BLOCK_TAG_CROSS_CORRELATIONS = 3,
    GW_VIS_BLOCK_TAG_BASELINE_UU = 4,
    GW_VIS_BLOCK_TAG_BASELINE_VV = 5,
    GW_VIS_BLOCK_TAG_BASELINE_WW = 6
};

typedef struct gw_binary gw_binary;

/* Opens and indexes an OSKAR binary file for reading (the read side of
 * oskar_binary_create, utils/binary/oskar_binary_create.h:55-56). */
GW_IO_EXPORT gw_binary* gw_binary_open(const char* filename, int* status);
GW_IO_EXPORT void gw_binary_free(gw_binary* handle);
GW_IO_EXPORT int gw_binary_num_tags(const gw_binary* handle);
/* Queries look at chunks from index start on (oskar_binary_set_query_
 * search_start, used by main.cpp:232 and :422 to start each block's reads at its
 * own tags). */
GW_IO_EXPORT void gw_binary_set_query_search_start(gw_binary* handle,
        int start, int* status);
/* Index of the first standard tag with these IDs (data_type 0: any type);
 * sets OSKAR_ERR_BINARY_TAG_NOT_FOUND if there is none
 * (utils/binary/oskar_binary_query.h). */
GW_IO_EXPORT int gw_binary_query(const gw_binary* handle,
        unsigned char data_type, unsigned char id_group, unsigned char id_tag,
        int user_index, size_t* payload_size, int* status);
/* Copies a tag's payload, as stored, into data (data_size >= payload), and
 * checks its CRC-32C when the file holds one (oskar_binary_read,
 * utils/binary/oskar_binary_read.h:81-84). */
GW_IO_EXPORT void gw_binary_read(gw_binary* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, void* data, int* status);
GW_IO_EXPORT void gw_binary_read_ext(gw_binary* handle,
        unsigned char data_type, const char* name_group, const char* name_tag,
        int user_index, size_t data_size, void* data, int* status);
GW_IO_EXPORT void gw_binary_read_int(gw_binary* handle, unsigned char id_group,
        unsigned char id_tag, int user_index, int* value, int* status);
GW_IO_EXPORT void gw_binary_read_double(gw_binary* handle,
        unsigned char id_group, unsigned char id_tag, int user_index,
        double* value, int* status);

/* Replaces read_coords (utils/read_vis.h:62-63, read_vis.c:30-46): the
 * baseline coordinates of visibility block i_block. */
GW_IO_EXPORT void gw_vis_read_coords(gw_binary* handle, int i_block,
        int precision, size_t num_coords, void* uu, void* vv, void* ww,
        int* status);
/* Replaces scale_coords (utils/read_vis.h:66-67, read_vis.c:49-73): metres
 * to wavelengths, scale = freq_hz / c computed in FP64 and applied in the
 * precision of the coordinates (GW_OSKAR_SINGLE: rounded to float first). */
GW_IO_EXPORT void gw_vis_scale_coords(double freq_hz, int precision,
        size_t num_coords, void* uu, void* vv, void* ww);

/* Replaces read_kernel (utils/read_kernel.h, read_kernel.c:35-118): reads a
 * 3-D FITS kernel cube (NAXIS1 = conv_size_half, NAXIS3 = planes) as float
 * or double (precision GW_OSKAR_SINGLE / GW_OSKAR_DOUBLE; BSCALE / BZERO
 * applied), its OVERSAMP, GRIDSIZE, CELLSIZE and W_SCALE keywords and, if
 * support is not NULL, column 1 of the W_KERNELS binary table as ints.
 * Returns the cube (free with gw_free); *support is allocated too. */
GW_IO_EXPORT void* gw_read_kernel(const char* file_name, int precision,
        int* conv_size_half, int* num_w_planes, int** support,
        int* oversample, int* grid_size, double* cellsize_rad,
        double* w_scale, int* status);
GW_IO_EXPORT void gw_free(void* ptr);

/* Replaces write_fits_cube (utils/write_fits_cube.h, write_fits_cube.c:
 * 113-168): data_type GW_OSKAR_SINGLE or GW_OSKAR_DOUBLE, optionally
 * | GW_OSKAR_COMPLEX (then ROOT_REAL.fits and ROOT_IMAG.fits are written);
 * a width x height x num_planes primary image, plane i_plane written
 * (i_plane < 0: the whole cube) into an existing file of the same
 * dimensions or a new zeroed one. */
GW_IO_EXPORT void gw_write_fits_cube(int data_type, const void* data,
        const char* root_name, int width, int height, int num_planes,
        int i_plane, int* status);

#ifdef __cplusplus
}
#endif

#endif /* FPGA_GRIDDING_AMD_IO_H_ */
