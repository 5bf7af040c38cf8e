/*
 * osb_io.h -- OPS-layout HDF5 files for the OpenSBLI hot path's state.
 *
 * libosb_io.so (host code, linked against the HDF5 C library) writes and
 * reads the files the generated OPS program writes and restarts from:
 *
 *   /<block>                group, attributes dims (int32[1]), index (int32[1]),
 *                           ops_type "ops_block" (fixed 9-byte string)
 *   /<block>/<name>_B<i>    dataset: the full padded array, slowest axis first
 *                           (x fastest in memory), float64, attributes
 *                           d_m, d_p, size, base (int32[ndim]), dim (int32[1]),
 *                           block_index (int32[1]), ops_type "ops_dat" (10 B),
 *                           type "double" (15 B), block (25 B), strings
 *                           null-padded ASCII
 *   /<constant>             dataset of one element at the file root
 *                           (float64, or int32 for integer constants)
 *
 * It replaces the OPS calls the generated code makes for output and restart
 * (opensbli/core/io_hdf5.py:165-250, opensbli/code_generation/opsc.py:962-977,
 * :1198-1252): ops_fetch_block_hdf5_file + ops_fetch_dat_hdf5_file +
 * ops_write_const_hdf5 (write_constants), and ops_decl_dat_hdf5 +
 * ops_get_const_hdf5 (restart).  The attribute set and string sizes are those
 * of opensbli/utilities/helperfunctions.py:99-147 (set_hdf5_metadata,
 * output_hdf5).
 *
 * All functions return 0 on success or a negative OSB_E_* code; the message
 * of the last failure on the calling thread is osb_io_last_error().  Calls
 * are serialised internally (the HDF5 library is not thread-safe).
 */
#ifndef OSB_IO_H
#define OSB_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One block's datasets.  size[d] is the interior points per direction, halo
 * the padding on both sides (d_m = -halo, d_p = +halo); every array holds
 * prod(size[d] + 2 halo) doubles, x fastest. */
typedef struct osb_io_block {
    const char* name;          /* group name, "opensbliblock00"          */
    int32_t index;             /* block number (block_index, index)      */
    int32_t ndim;              /* 1..3                                   */
    int32_t size[3];
    int32_t halo;
} osb_io_block;

/* Create (or truncate) `path` and write the block group, the arrays
 * names[i] <- data[i], and the root constants (dvals: float64, ivals: int32).
 * Replaces ops_fetch_block_hdf5_file + ops_fetch_dat_hdf5_file + write_constants. */
int osb_io_write(const char* path, const osb_io_block* block,
                 int narrays, const char* const* names, const double* const* data,
                 int ndconst, const char* const* dnames, const double* dvals,
                 int niconst, const char* const* inames, const int32_t* ivals);

/* Shape of dataset <block>/<name>: ndim, size (interior), d_m, d_p as stored.
 * Replaces the metadata read of ops_decl_dat_hdf5. */
int osb_io_dat_info(const char* path, const char* block, const char* name,
                    int32_t* ndim, int32_t size[3], int32_t d_m[3], int32_t d_p[3]);

/* Read dataset <block>/<name> (the full padded array) into out, which holds
 * `count` doubles; count must equal the stored element count. */
int osb_io_read_dat(const char* path, const char* block, const char* name, double* out, int64_t count);

/* Read a root constant (float64 or integer) as a double (ops_get_const_hdf5). */
int osb_io_read_const(const char* path, const char* name, double* out);

const char* osb_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OSB_IO_H */
